// GPU test of the C++ facade (include/nmhd/mhd_solver.h) through libnmhd.so:
//  1. proceed() with a monitor that stops at a generation, save_state("c"),
//     load_state("c") into a fresh solver, continue: the state equals an
//     uninterrupted run bit for bit (Orszag-Tang: dt does not depend on t);
//  2. read() rejects a state file of another size;
//  3. a wallclock limit in the past makes proceed() return false at once;
//  4. the setup text holds the reference's 20 + 21 members.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "nmhd/mhd_solver.h"

using namespace nmhd;

static int failures = 0;
#define EXPECT(c)                                                              \
  do {                                                                         \
    if (!(c)) {                                                                \
      ++failures;                                                              \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);                  \
    }                                                                          \
  } while (0)

static SimulationConfig ot_config(int n) {
  SimulationConfig c;
  c.global_size[0] = c.global_size[1] = c.global_size[2] = n;
  c.cfl_factor = 0.4f;
  return c;
}

static void ot_ic(MHDSolver &s) {
  s.set_initial_condition = [](const PositioningSetup &pos, const RCo<7> &o, BodyScalar &rho,
                               BodyVector &v, BodyScalar &p, FaceVector &B) {
    const Triplet<Real> r = o.position(pos.corner);
    const double pi = 3.14159265358979323846;
    rho[o] = Real(25.0 / (36.0 * pi));
    p[o] = Real(5.0 / (12.0 * pi));
    v[X][o] = Real(-std::sin(2 * pi * r[Y]));
    v[Y][o] = Real(std::sin(2 * pi * r[X]));
    v[Z][o] = Real(0.01 * std::sin(2 * pi * r[Z]));
    B[X][o - HX] = Real(-std::sin(2 * pi * r[Y]) / std::sqrt(4 * pi));
    B[Y][o - HY] = Real(std::sin(4 * pi * r[X]) / std::sqrt(4 * pi));
    B[Z][o - HZ] = 0;
  };
}

static std::vector<char> slurp(const std::string &f) {
  std::ifstream i(f.c_str(), std::ios::binary);
  return std::vector<char>((std::istreambuf_iterator<char>(i)), std::istreambuf_iterator<char>());
}

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : "facade_test_out";
  const SimulationConfig cfg = ot_config(16);
  // 1. uninterrupted 6 steps
  {
    LocalSetup la(0, dir + "/a");
    MHDSolver a(cfg, la);
    ot_ic(a);
    a.initialize_state();
    a.monitor_hook = [](MHDSolver &s) { return s.generation() >= 6 ? 1 : 0; };
    const bool reached = a.proceed(Real(10.0));
    EXPECT(!reached);
    EXPECT(a.generation() == 6);
    a.save_state("e");
  }
  {
    LocalSetup lb(0, dir + "/b");
    MHDSolver b(cfg, lb);
    ot_ic(b);
    b.initialize_state();
    b.monitor_hook = [](MHDSolver &s) { return s.generation() >= 3 ? 1 : 0; };
    b.proceed(Real(10.0));
    EXPECT(b.generation() == 3);
    b.save_state("c");
  }
  {
    LocalSetup lc(0, dir + "/b");
    MHDSolver c(cfg, lc);
    c.load_state("c");
    EXPECT(c.generation() == 3);
    c.monitor_hook = [](MHDSolver &s) { return s.generation() >= 6 ? 1 : 0; };
    c.proceed(Real(10.0));
    EXPECT(c.generation() == 6);
    c.save_state("e");
  }
  const std::vector<char> ea = slurp(dir + "/a/e/0000.bin"), eb = slurp(dir + "/b/e/0000.bin");
  EXPECT(!ea.empty() && ea.size() == 12 + 8 * 4 * 16 * 16 * 16);
  EXPECT(ea == eb);
  // 2. size mismatch
  {
    SimulationConfig other = ot_config(8);
    LocalSetup ld(0, dir + "/d");
    MHDSolver d(other, ld);
    EXPECT(!d.read(dir + "/a/e/0000.bin"));
    EXPECT(!d.read(dir + "/does/not/exist.bin"));
  }
  // 3. wallclock limit already passed
  {
    LocalSetup le(0, dir + "/e");
    MHDSolver e(cfg, le);
    ot_ic(e);
    e.initialize_state();
    const bool reached = e.proceed(Real(10.0), 1e-9);
    EXPECT(!reached);
    EXPECT(e.generation() <= 1);
  }
  // 4. setup text layout
  {
    std::ifstream f((dir + "/a/e/0000_setup.txt").c_str());
    int n1 = 0, n2 = 0;
    std::string line, spacer;
    f >> n1;
    for (int i = 0; i < n1; ++i) std::getline(f >> std::ws, line);
    f >> spacer >> n2;
    EXPECT(n1 == 20 && spacer == "################" && n2 == 21);
  }
  // 5. goal clipping through proceed(): the run ends exactly at the goal
  {
    LocalSetup lg(0, dir + "/g");
    MHDSolver g(cfg, lg);
    ot_ic(g);
    g.initialize_state();
    const bool reached = g.proceed(Real(0.05));
    EXPECT(reached);
    EXPECT(std::fabs(g.current_time() - 0.05) < 1e-7);
  }
  std::printf("facade_test: %d failures\n", failures);
  return failures ? 1 : 0;
}
