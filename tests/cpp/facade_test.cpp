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
  // 6. the internal step calls: add_flux_kernel_caller adds the right-hand
  //    side to the paces (twice = doubled; a uniform state has none),
  //    communicate() fills the shearing box's radial ghost rows from the
  //    interior (at t = 0 the lower rows are the top interior rows shifted
  //    by 0 cells, with v_az - relvel), boundary_condition() without a hook
  //    leaves the state alone
  {
    LocalSetup lh(0, dir + "/h");
    MHDSolver h(cfg, lh);
    ot_ic(h);
    h.initialize_state();
    const size_t n = 16 * 16 * 16;
    std::vector<std::vector<Real> > pace(8, std::vector<Real>(n, Real(0)));
    BodyScalar dp(&pace[0][0]), ep(&pace[4][0]);
    BodyScalar m0(&pace[1][0]), m1(&pace[2][0]), m2(&pace[3][0]);
    BodyVector mp(m0, m1, m2);
    FaceVector bp(&pace[5][0], &pace[6][0], &pace[7][0]);
    h.add_flux_kernel_caller(Real(0.001), h.density(), h.velocity(), h.pressure(), h.magnet(),
                             dp, mp, ep, bp);
    std::vector<std::vector<Real> > once = pace;
    h.add_flux_kernel_caller(Real(0.001), h.density(), h.velocity(), h.pressure(), h.magnet(),
                             dp, mp, ep, bp);
    double mx = 0;
    bool doubled = true;
    for (int f = 0; f < 8; ++f)
      for (size_t a = 0; a < n; ++a) {
        mx = std::fmax(mx, std::fabs(once[f][a]));
        doubled = doubled && pace[f][a] == once[f][a] + once[f][a];
      }
    EXPECT(mx > 0.1 && doubled);
    std::vector<std::vector<Real> > uni(8, std::vector<Real>(n, Real(0.5)));
    std::vector<std::vector<Real> > zp(8, std::vector<Real>(n, Real(0)));
    h.add_flux_kernel_caller(
        Real(0.001), BodyScalar(&uni[0][0]),
        BodyVector(BodyScalar(&uni[1][0]), BodyScalar(&uni[2][0]), BodyScalar(&uni[3][0])),
        BodyScalar(&uni[4][0]), FaceVector(&uni[5][0], &uni[6][0], &uni[7][0]),
        BodyScalar(&zp[0][0]),
        BodyVector(BodyScalar(&zp[1][0]), BodyScalar(&zp[2][0]), BodyScalar(&zp[3][0])),
        BodyScalar(&zp[4][0]), FaceVector(&zp[5][0], &zp[6][0], &zp[7][0]));
    bool zero = true;
    for (int f = 0; f < 8; ++f)
      for (size_t a = 0; a < n; ++a) zero = zero && zp[f][a] == 0;
    EXPECT(zero);
    h.sync_to_host();
    EXPECT(h.host_field(0)[0] == Real(0.5));  // the uniform state became the solver's
  }
  {
    SimulationConfig sb;
    sb.global_size[0] = 16;
    sb.global_size[1] = 16;
    sb.global_size[2] = 8;
    sb.margin[1] = 4;
    sb.global_corner[0] = sb.global_corner[1] = Real(-0.5);
    sb.cfl_factor = 0.4f;
    sb.source = NMHD_SOURCE_SHEARING_BOX;
    sb.shearing_boundary = true;
    LocalSetup ls(0, dir + "/s");
    MHDSolver s(sb, ls);
    s.set_initial_condition = [](const PositioningSetup &, const RCo<7> &o, BodyScalar &rho,
                                 BodyVector &v, BodyScalar &p, FaceVector &B) {
      rho[o] = Real(1 + 0.01 * o[X] + 0.001 * o[Y] + 0.0001 * o[Z]);
      p[o] = 1;
      v[X][o] = Real(0.01 * o[Y]);
      v[Y][o] = 0;
      v[Z][o] = 0;
      B[X][o - HX] = 0;
      B[Y][o - HY] = 0;
      B[Z][o - HZ] = Real(0.07);
    };
    s.initialize_state();
    s.boundary_condition();  // no hook: nothing happens
    s.communicate();
    s.sync_to_host();
    const int nx = 16, ny = 24, M = 4;
    const Real relvel = Real(-1.5f * 1.0f * 1.0f);
    bool ok = true;
    for (int z = 0; z < 8; ++z)
      for (int j = 0; j < M; ++j)
        for (int x = 0; x < nx; ++x) {
          const size_t dst = ((size_t)z * ny + j) * nx + x;
          const size_t src = ((size_t)z * ny + (ny - 2 * M + j)) * nx + x;
          ok = ok && s.host_field(0)[dst] == s.host_field(0)[src];
          ok = ok && s.host_field(1)[dst] == s.host_field(1)[src] - relvel;
        }
    EXPECT(ok);
  }
  std::printf("facade_test: %d failures\n", failures);
  return failures ? 1 : 0;
}
