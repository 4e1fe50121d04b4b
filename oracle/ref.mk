# Builds the DualSPHysics v5.4 CPU solver from its own sources, in place under
# /root/reference/src/source, into oracle/_ref/ (objects and executable only).
# Used only to pin the oracle: never shipped, never part of the product path.
# IEEE arithmetic (no -ffast-math) so FP64 cell hashing and FP32 pair sums are
# reproducible; closed libraries (Chrono, WaveGen, MoorDynPlus) compiled out.
#   make -f oracle/ref.mk -j8
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

all: $(EXE)
$(EXE): $(OBJFILES)
	$(CXX) $(OBJFILES) -fopenmp -lgomp -o $@
$(OBJDIR)/%.o: $(REFSRC)/%.cpp | $(OBJDIR)
	$(CXX) $(CXXFLAGS) $< -o $@
$(OBJDIR):
	mkdir -p $(OBJDIR)
.PHONY: all
