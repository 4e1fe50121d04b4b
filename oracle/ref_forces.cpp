// Forces of the reference CPU solver on a saved state (TEST INFRASTRUCTURE
// ONLY: builds with oracle/ref.mk against the reference's own objects into
// oracle/_ref/ and writes fixtures; nothing of the product links it).
//
// The reference's PART files hold positions, velocities and densities, not
// the interaction's results. This driver starts the reference solver
// (DualSPHysics v5.4 CPU build, JSphCpuSingle) on a case resumed from a
// PART (-partbegin), runs its initialisation exactly as
// JSphCpuSingle::Run does (JSphCpuSingle.cpp:1144-1168: LoadConfig,
// LoadCaseParticles, ConfigDomain, ConfigRunMode, InitRunCpu with its first
// cell divide), then the first interaction of a Symplectic step
// (ComputeStep_Sym, JSphCpuSingle.cpp:902-907: mDBC correction,
// PreInteraction_Forces, Interaction_Forces with INTERSTEP_SymPredictor), and
// writes the per-particle results in the solver's particle order:
//   uint32 np, then np x {uint32 idp, float ace[3], float ar}
// Usage: ref_forces <case> <dirout> -cpu -partbegin:<k> <dirpart> ... <outfile>
// (every argument but the last is the reference's own command line).
#include <cstdio>
#include <exception>
#include <string>
#include <vector>

#include "JAppInfo.h"
#include "JLog2.h"
#include "JSphCfgRun.h"
#include "JSphCpuSingle.h"

// The application descriptor the reference's main.cpp defines (main.cpp:72).
JAppInfo AppInfo("DualSPHysics5", "v5.4.355", "08-04-2025");

namespace {

class RefForces : public JSphCpuSingle {
 public:
  int Forces(const std::string& appname, const JSphCfgRun* cfg, JLog2* log,
             const std::string& outfile) {
    AppName = appname;
    Log = log;
    CfgRun = cfg;
    Arrays_Cpu = new JArraysCpu(Log);
    Timersc->Config(cfg->SvTimers);
    LoadConfig(cfg);
    LoadCaseParticles();
    ConfigDomain();
    ConfigRunMode();
    InitRunCpu();
    InterStep = INTERSTEP_SymPredictor;
    MdbcBoundCorrection(InterStep);
    PreInteraction_Forces(InterStep);
    Interaction_Forces(InterStep);
    const unsigned np = Np;
    const unsigned* idp = Idp_c->cptr();
    const tfloat3* ace = Ace_c->cptr();
    const float* ar = Ar_c->cptr();
    std::FILE* f = std::fopen(outfile.c_str(), "wb");
    if (!f) return 2;
    std::fwrite(&np, sizeof(np), 1, f);
    for (unsigned p = 0; p < np; p++) {
      const float rec[4] = {ace[p].x, ace[p].y, ace[p].z, ar[p]};
      std::fwrite(idp + p, sizeof(unsigned), 1, f);
      std::fwrite(rec, sizeof(float), 4, f);
    }
    std::fclose(f);
    PosInteraction_Forces();
    return 0;
  }
};

}  // namespace

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s <reference arguments...> <outfile>\n", argv[0]);
    return 2;
  }
  AppInfo.ConfigRunPaths(argv[0]);
  JSphCfgRun cfg;
  cfg.SetFeatureList(JSph::GetFeatureList());
  try {
    cfg.LoadArgv(argc - 1, argv);
    AppInfo.ConfigOutput(cfg.CreateDirs, cfg.CsvSepComa, cfg.DirOut, cfg.DirDataOut);
    AppInfo.LogInit(AppInfo.GetDirOut() + "/Run.out");
    RefForces sph;
    return sph.Forces(AppInfo.GetFullName(), &cfg, AppInfo.LogPtr(), argv[argc - 1]);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "ref_forces: %s\n", e.what());
  } catch (...) {
    std::fprintf(stderr, "ref_forces: reference exception\n");
  }
  return 1;
}
