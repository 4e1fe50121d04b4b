# Builds the DualSPHysics v5.4 CPU solver from its own sources, in place under
# /root/reference/src/source, into oracle/_ref/ (objects and executables only).
# Used only to pin the oracle and as the timed CPU baseline: never shipped,
# never part of the product path. Closed libraries (Chrono, WaveGen,
# MoorDynPlus) compiled out.
#  - DualSPHysics5.4CPU_ieee: IEEE arithmetic (no -ffast-math) so FP64 cell
#    hashing and FP32 pair sums are reproducible (fixtures);
#  - DualSPHysics5.4CPU_fast: the reference's own release flags (Makefile_cpu:
#    -O3 -fopenmp -ffast-math) for timing, with a portable AVX2/FMA target
#    instead of -march=native (the timing host is not this machine).
#  - ref_forces: oracle/ref_forces.cpp (a driver of this repository) linked
#    with the IEEE objects minus the reference's main: the solver's own
#    interaction results on a PART state, for the fixtures.
#   make -f oracle/ref.mk -j8 [fast] [forces]
REFSRC ?= /root/reference/src/source
OUT    ?= oracle/_ref
OBJDIR := $(OUT)/obj
EXE    := $(OUT)/DualSPHysics5.4CPU_ieee
CXX    ?= g++
CXXFLAGS := -c -O2 -fopenmp -std=c++0x -D_WITHMR -DDISABLE_CHRONO -DDISABLE_WAVEGEN \
            -DDISABLE_MOORDYNPLUS -w
OBJS := JXml tinystr tinyxml tinyxmlerror tinyxmlparser \
  JMotion JMotionData JMotionList JMotionMov JMotionObj JMotionPos JDsMotion \
  Functions FunGeo3d FunSphKernelsCfg JAppInfo JBinaryData JCfgRunBase JDataArrays \
  JException JLinearValue JLog2 JObject JOutputCsv JRadixSort \
  JRangeFilter JReadDatafile JSaveCsv2 JSpVtkData JSpVtkShape \
  JTimeControl JTrianglesMesh randomc \
  JDsphConfig JDsPips JPartDataBi4 JPartDataHead JDsPartFloatSave JDsPartMotionSave \
  JPartFloatInfoBi4 JPartMotRefBi4Save JPartOutBi4Save JCaseCtes \
  JCaseEParms JCaseParts JCaseProperties \
  JArraysCpu JCellDivCpu JDebugSphCpu JSphCfgRun JComputeMotionRef JDsDcell JDsDamping \
  JDsExtraData JDsGaugeItem JDsGaugeSystem JDsPartsOut JDsSaveDt JSphShifting \
  JSphShiftingAdv JSph JDsAccInput JSphCpu JSphCpu_mdbc JSphCpu_preloop \
  JDsInitialize JSphMk JDsOutputParts JDsPartsInit JDsFixedDt JDsViscoInput \
  JDsOutputTime JDsTimers JWaveSpectrumGpu main \
  JCellDivCpuSingle JPartsLoad4 JSphCpuSingle \
  JMLPistonsGpu JRelaxZonesGpu JChronoGeoObj JChronoObjects JDsMooredFloatings JDsFtForcePoints \
  JSphCpu_InOut JSphCpuSingle_InOut JSphInOut JSphInOutZone JSphInOutGridData \
  JSphInOutPoints JSphInOutVel JSphInOutVelAwas JSphInOutZsurf JSimpleNeigs \
  JMeshData JMeshTDatas JMeshTDatasLoad JMeshTDatasSave JMeshTDatasDsVel JMeshTDatasXml \
  JSphVRes JSphVResZone JSphCpuSingle_VRes JSphCpu_VRes JCaseVRes JVResDataLimits JDsVresData JBoxDef \
  JSphFlexStruc
OBJFILES := $(addprefix $(OBJDIR)/,$(addsuffix .o,$(OBJS)))

OBJDIR_FAST := $(OUT)/obj_fast
EXE_FAST := $(OUT)/DualSPHysics5.4CPU_fast
CXXFLAGS_FAST := -c -O3 -fopenmp -ffast-math -march=x86-64-v3 -std=c++0x -D_WITHMR \
            -DDISABLE_CHRONO -DDISABLE_WAVEGEN -DDISABLE_MOORDYNPLUS -w
OBJFILES_FAST := $(addprefix $(OBJDIR_FAST)/,$(addsuffix .o,$(OBJS)))

EXE_FORCES := $(OUT)/ref_forces
OBJFILES_NOMAIN := $(filter-out $(OBJDIR)/main.o,$(OBJFILES))

all: $(EXE)
fast: $(EXE_FAST)
forces: $(EXE_FORCES)
$(EXE_FORCES): $(OBJFILES_NOMAIN) oracle/ref_forces.cpp | $(OBJDIR)
	$(CXX) -O2 -fopenmp -std=c++0x -D_WITHMR -DDISABLE_CHRONO -DDISABLE_WAVEGEN \
	  -DDISABLE_MOORDYNPLUS -w -I$(REFSRC) oracle/ref_forces.cpp $(OBJFILES_NOMAIN) -fopenmp -lgomp -o $@
$(EXE): $(OBJFILES)
	$(CXX) $(OBJFILES) -fopenmp -lgomp -o $@
$(EXE_FAST): $(OBJFILES_FAST)
	$(CXX) $(OBJFILES_FAST) -fopenmp -lgomp -o $@
$(OBJDIR)/%.o: $(REFSRC)/%.cpp | $(OBJDIR)
	$(CXX) $(CXXFLAGS) $< -o $@
$(OBJDIR_FAST)/%.o: $(REFSRC)/%.cpp | $(OBJDIR_FAST)
	$(CXX) $(CXXFLAGS_FAST) $< -o $@
$(OBJDIR):
	mkdir -p $(OBJDIR)
$(OBJDIR_FAST):
	mkdir -p $(OBJDIR_FAST)
.PHONY: all fast forces
