// kepler_disk -- the shearing-box MRI driver on the C++ facade, mirroring the
// reference application (application/kepler_disk_demo/main.curb:19-118):
// --new starts from the initial condition, --continue resumes from the "c"
// checkpoint; the run stops at the goal time or at the wallclock limit, in
// which case it saves the "c" checkpoint and prints "ToBeContinued".
//
//   kepler_disk [--new|--continue] [-o DIR] [-n N] [-g GOAL] [-w SECONDS]
//               [--beta B] [--steps K] [--fixes BITS]
//
// --fixes takes enum nmhd_fix bits (corrections of reference quirks; the
// default 0 runs the reference's behaviour).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "nmhd/mhd_solver.h"

using namespace nmhd;

int main(int argc, char **argv) {
  bool fresh = true;
  std::string outdir = "stat/kepler";
  int n = 64, max_steps = -1, fixes = 0;
  double goal = 1.0, wall = -1, beta = 400.0;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() { return std::string(i + 1 < argc ? argv[++i] : ""); };
    if (a == "--new") fresh = true;
    else if (a == "--continue") fresh = false;
    else if (a == "-o") outdir = next();
    else if (a == "-n") n = std::atoi(next().c_str());
    else if (a == "-g") goal = std::atof(next().c_str());
    else if (a == "-w") wall = std::atof(next().c_str());
    else if (a == "--beta") beta = std::atof(next().c_str());
    else if (a == "--steps") max_steps = std::atoi(next().c_str());
    else if (a == "--fixes") fixes = std::atoi(next().c_str());
    else {
      std::fprintf(stderr, "unknown option %s\n", a.c_str());
      return 2;
    }
  }
  SimulationConfig cfg;  // application/kepler_disk_demo/simulation_config.hrb
  cfg.global_size[0] = cfg.global_size[1] = cfg.global_size[2] = n;
  cfg.margin[1] = 4;
  cfg.cfl_factor = 0.4f;
  cfg.source = NMHD_SOURCE_SHEARING_BOX;
  cfg.shearing_boundary = true;
  cfg.fixes = fixes;
  LocalSetup local(0, outdir);
  MHDSolver solver(cfg, local);
  const Real b0 = Real(std::sqrt(2.0 / beta));
  // uniform state with the shear flow and a seeded velocity perturbation
  solver.set_initial_condition = [&](const PositioningSetup &, const RCo<7> &o, BodyScalar &rho,
                                     BodyVector &v, BodyScalar &p, FaceVector &B) {
    const Triplet<Real> r = o.position(Triplet<Real>(0, -solver.local_setup().computational_extent_[1] / 2, 0));
    unsigned long h = (unsigned long)o.addr() * 2654435761ul + 12345ul;
    auto u = [&]() {
      h ^= h << 13;
      h ^= h >> 7;
      h ^= h << 17;
      return Real((h >> 11) % 2000001) * Real(1e-8) - Real(0.01);
    };
    rho[o] = 1;
    p[o] = 1;
    v[X][o] = -cfg.qshear * cfg.omega * r[Y] + u();
    v[Y][o] = u();
    v[Z][o] = u();
    B[X][o - HX] = 0;
    B[Y][o - HY] = 0;
    B[Z][o - HZ] = b0;
  };
  int generation_limit = max_steps;
  // the reference's monitor (mhd_solver.curb:107-130) with the averages
  // reduced on the GPU instead of copying B_x, B_y to the host
  solver.monitor_hook = [&](MHDSolver &s) {
    const nmhd_diagnostics d = s.diagnostics();
    s.local_setup().ofs_log() << "t: " << s.current_time() << " B_init: " << b0
                              << " Bx_squared_average: " << d.b2_mean[0]
                              << " By_squared_average: " << d.b2_mean[1] << std::endl;
    return (generation_limit >= 0 && s.generation() >= generation_limit) ? 1 : 0;
  };
  if (fresh) {
    solver.initialize_state();
  } else {
    solver.load_state("c");
    if (generation_limit >= 0) generation_limit += solver.generation();
  }
  solver.initial_event();
  const bool reached = solver.proceed(Real(goal), wall);
  std::printf("t = %.9g generation = %d\n", solver.current_time(), solver.generation());
  if (!reached && (wall > 0 || max_steps >= 0)) {
    solver.save_state("c");
    std::printf("ToBeContinued\n");
    return 0;
  }
  solver.final_event();
  solver.save_state("e");
  return 0;
}
