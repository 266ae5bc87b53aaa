// nmhd/mhd_solver.h -- the reference's C++ solver interface over the C ABI.
//
// Drop-in counterpart of class MHDSolver with GlobalSetup / LocalSetup
// (library/mhd_solver.h:37-640, template library/mhd_solver.hrb:47-276):
// the same class, member and method names and semantics, with the
// compile-time configuration of the reference (simulation_config.h) replaced
// by a run-time SimulationConfig and the link-time application hooks
// replaced by std::function members.  Every step runs on the GPU through
// libnmhd.so (include/nmhd.h); this header adds only host logic: the proceed
// loop with its goal / wallclock / monitor semantics, the .bin state files
// and the PersistentPod text of the setups.
//
// Header-only; link with -lnmhd.
#ifndef NMHD_MHD_SOLVER_H
#define NMHD_MHD_SOLVER_H

#include <sys/stat.h>
#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../nmhd.h"
#include "crystal.h"

namespace nmhd {

// The constants of the reference's simulation_config.h as run-time values
// (simulation_config_default.h:25-122; KEP/simulation_config.hrb:15-74).
struct SimulationConfig {
  int global_size[3] = {10, 10, 10};      // kGlobalSizeX..
  int margin[3] = {0, 0, 0};              // kMarginSizeX..
  Real global_extent[3] = {1, 1, 1};      // kGlobalExtentX..
  Real global_corner[3] = {0, 0, 0};      // kGlobalCornerX..
  Real gamma = Real(5.0f / 3.0f);         // kGamma
  Real cfl_factor = 0.5f;                 // kCflFactor
  int source = NMHD_SOURCE_NONE;          // external_acceleration hook
  Real omega = 1.0f;                      // kOmega
  Real qshear = 1.5f;                     // kDifferentialRotationIndex
  bool shearing_boundary = false;         // ConfigOptions[:shearing_boundary]
  int recon = NMHD_RECON_MINBEE2;         // SpaceInterpolation
  int integrator = NMHD_MIDPOINT2;        // TimeIntegral
  int path = NMHD_PATH_AUTO;              // kernel path
  int fixes = 0;                          // enum nmhd_fix bits (0: as the reference)
};

inline PreciseReal wallclock_now() {  // get_time.h
  timeval tv;
  gettimeofday(&tv, 0);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

// PersistentPod text encoding (nushio_script/persistent_pod.hrb:80-120):
// numbers through default ostream formatting, strings with %xx escapes.
namespace pod {
template <class T>
inline std::string enc(const T &x) {
  std::ostringstream o;
  o << x;
  return o.str();
}
inline std::string enc(const std::string &x) {
  std::ostringstream r;
  char buf[4];
  for (size_t i = 0; i < x.size(); ++i) {
    const unsigned char c = x[i];
    if (33 <= c && c <= 126 && c != '%') {
      r << c;
    } else {
      std::snprintf(buf, sizeof buf, "%02x", c);
      r << '%' << buf;
    }
  }
  return r.str();
}
template <class T>
inline void dec(const std::string &s, T &x) {
  std::istringstream i(s);
  i >> x;
}
inline void dec(const std::string &x, std::string &out) {
  std::ostringstream r;
  for (size_t i = 0; i < x.size(); ++i) {
    if (x[i] != '%') {
      r << x[i];
    } else if (i + 2 < x.size()) {
      const std::string hx = x.substr(i + 1, 2);
      r << char(std::strtol(hx.c_str(), 0, 16));
      i += 2;
    }
  }
  out = r.str();
}
inline std::map<std::string, std::string> read_block(std::istream &in) {
  int n = 0;
  in >> n;
  std::map<std::string, std::string> kv;
  for (int i = 0; i < n; ++i) {
    std::string k, v;
    in >> k >> v;
    kv[k] = v;
  }
  return kv;
}
}  // namespace pod

// GlobalSetup (mhd_solver.hrb:47-61): generation, time and the global
// topology of the run.
class GlobalSetup {
 public:
  int generation_ = 0;
  PreciseReal current_time_ = 0;
  int real_size_ = sizeof(Real);
  std::string real_precision_ = "float";
  int parallel_size_ = 1;
  int global_size_[3] = {0, 0, 0};
  int parallel_size_d_[3] = {1, 1, 1};
  Real global_extent_[3] = {0, 0, 0};
  Real global_corner_[3] = {0, 0, 0};
  Real d_[3] = {0, 0, 0};

  int &generation() { return generation_; }
  int generation() const { return generation_; }
  PreciseReal &current_time() { return current_time_; }
  PreciseReal current_time() const { return current_time_; }
  void proceed(int generation_increase, PreciseReal time_increase) {
    generation_ += generation_increase;
    current_time_ += time_increase;
  }
  friend std::ostream &operator<<(std::ostream &o, const GlobalSetup &g) {
    static const char *ax[3] = {"x", "y", "z"}, *AX[3] = {"X", "Y", "Z"};
    o << 20 << std::endl;
    o << "generation " << pod::enc(g.generation_) << std::endl;
    o << "real_size " << pod::enc(g.real_size_) << std::endl;
    o << "real_precision " << pod::enc(g.real_precision_) << std::endl;
    o << "current_time " << pod::enc(g.current_time_) << std::endl;
    o << "parallel_size " << pod::enc(g.parallel_size_) << std::endl;
    for (int d = 0; d < 3; ++d) {
      o << "global_size_" << ax[d] << " " << pod::enc(g.global_size_[d]) << std::endl;
      o << "parallel_size_" << ax[d] << " " << pod::enc(g.parallel_size_d_[d]) << std::endl;
      o << "global_extent_" << ax[d] << " " << pod::enc(g.global_extent_[d]) << std::endl;
      o << "global_corner_" << ax[d] << " " << pod::enc(g.global_corner_[d]) << std::endl;
      o << "d" << AX[d] << " " << pod::enc(g.d_[d]) << std::endl;
    }
    return o;
  }
  // only the members the reference restores (non-ephemeral, writable)
  friend std::istream &operator>>(std::istream &in, GlobalSetup &g) {
    std::map<std::string, std::string> kv = pod::read_block(in);
    if (kv.count("generation")) pod::dec(kv["generation"], g.generation_);
    if (kv.count("current_time")) pod::dec(kv["current_time"], g.current_time_);
    return in;
  }
};

// LocalSetup (mhd_solver.hrb:63-144): the rank's view, output paths and the
// shearing displacement.
class LocalSetup {
 public:
  int mpi_rank_ = 0;
  std::string output_directory_ = "stat";
  std::string mpi_prefix_ = "0000";
  PreciseReal relative_displacement_ = 0;
  PreciseReal relative_velocity_ = 0;
  size_t computational_size_ = 0;
  int computational_size_d_[3] = {0, 0, 0};
  int margin_[3] = {0, 0, 0};
  Real computational_corner_[3] = {0, 0, 0};
  Real computational_extent_[3] = {0, 0, 0};
  Real d_[3] = {0, 0, 0};
  std::ofstream ofs_log_, ofs_bench_;

  LocalSetup() {}
  // opens <dir>/log/<rank>.txt and <dir>/bench/<rank>.txt for appending, as
  // the reference's constructor does (mhd_solver.hrb:81-88, 133-138)
  LocalSetup(int mpi_rank0, std::string output_dir0) : mpi_rank_(mpi_rank0) {
    char work[32];
    std::snprintf(work, sizeof work, "%04d", mpi_rank0);
    output_directory_ = output_dir0;
    mpi_prefix_ = work;
    ofs_log_.open(log_filename().c_str(), std::ios_base::out | std::ios_base::app);
    ofs_bench_.open(bench_filename().c_str(), std::ios_base::out | std::ios_base::app);
  }
  LocalSetup(const LocalSetup &) = delete;
  LocalSetup &operator=(const LocalSetup &) = delete;
  std::ofstream &ofs_log() { return ofs_log_; }
  std::ofstream &ofs_bench() { return ofs_bench_; }
  std::string log_filename() const { return mkdir_prefix("log") + ".txt"; }
  std::string bench_filename() const { return mkdir_prefix("bench") + ".txt"; }
  int &mpi_rank() { return mpi_rank_; }
  std::string &output_directory() { return output_directory_; }
  const std::string &output_directory() const { return output_directory_; }
  std::string &mpi_prefix() { return mpi_prefix_; }
  const std::string &mpi_prefix() const { return mpi_prefix_; }
  PreciseReal &relative_displacement() { return relative_displacement_; }
  PreciseReal &relative_velocity() { return relative_velocity_; }
  Triplet<Real> computational_corner() const {
    return Triplet<Real>(computational_corner_[0], computational_corner_[1],
                         computational_corner_[2]);
  }
  // mhd_solver.hrb:127-131
  std::string mkdir_prefix(const std::string &tag) const {
    const std::string dir = output_directory_ + "/" + tag;
    std::string partial;
    for (size_t i = 0; i <= dir.size(); ++i) {
      if (i == dir.size() || dir[i] == '/') {
        partial = dir.substr(0, i);
        if (!partial.empty()) ::mkdir(partial.c_str(), 0755);
      }
    }
    return dir + "/" + mpi_prefix_;
  }
  friend std::ostream &operator<<(std::ostream &o, const LocalSetup &l) {
    static const char *ax[3] = {"x", "y", "z"}, *AX[3] = {"X", "Y", "Z"};
    o << 21 << std::endl;
    o << "mpi_rank " << pod::enc(l.mpi_rank_) << std::endl;
    o << "output_directory " << pod::enc(l.output_directory_) << std::endl;
    o << "mpi_prefix " << pod::enc(l.mpi_prefix_) << std::endl;
    o << "relative_displacement " << pod::enc(l.relative_displacement_) << std::endl;
    o << "relative_velocity " << pod::enc(l.relative_velocity_) << std::endl;
    o << "computational_size " << pod::enc(l.computational_size_) << std::endl;
    for (int d = 0; d < 3; ++d) {
      o << "computational_size_" << ax[d] << " " << pod::enc(l.computational_size_d_[d])
        << std::endl;
      o << "margin_" << ax[d] << " " << pod::enc(l.margin_[d]) << std::endl;
      o << "computational_corner_" << ax[d] << " " << pod::enc(l.computational_corner_[d])
        << std::endl;
      o << "computational_extent_" << ax[d] << " " << pod::enc(l.computational_extent_[d])
        << std::endl;
      o << "d" << AX[d] << " " << pod::enc(l.d_[d]) << std::endl;
    }
    return o;
  }
  friend std::istream &operator>>(std::istream &in, LocalSetup &l) {
    std::map<std::string, std::string> kv = pod::read_block(in);
    if (kv.count("mpi_rank")) pod::dec(kv["mpi_rank"], l.mpi_rank_);
    if (kv.count("output_directory")) pod::dec(kv["output_directory"], l.output_directory_);
    if (kv.count("mpi_prefix")) pod::dec(kv["mpi_prefix"], l.mpi_prefix_);
    if (kv.count("relative_displacement"))
      pod::dec(kv["relative_displacement"], l.relative_displacement_);
    if (kv.count("relative_velocity")) pod::dec(kv["relative_velocity"], l.relative_velocity_);
    return in;
  }
};

// The setup text of save_state / dump (INL:1250-1259): the GlobalSetup
// block, an empty line, a line of 16 '#', the LocalSetup block and an empty
// line.  read_setup is load_state's reader (INL:1261-1267): the two blocks
// around the spacer word, each a count and that many "key value" pairs in any
// order; only the members the reference restores are taken.
inline void write_setup(std::ostream &o, const GlobalSetup &g, const LocalSetup &l) {
  o << g << std::endl << "################" << std::endl << l << std::endl;
}
inline void read_setup(std::istream &in, GlobalSetup &g, LocalSetup &l) {
  in >> g;
  std::string spacer;
  in >> spacer;
  in >> l;
}

struct PositioningSetup {  // crystal.hrb:11-15
  Triplet<Real> corner;
  int mpi_rank_y;
};

class MHDSolver {
 public:
  // Application hooks (the reference links them in at build time).
  // Initial condition on the host, called for every cell including ghosts
  // (set_initial_condition, INL:43-54).
  std::function<void(const PositioningSetup &, const RCo<7> &, BodyScalar &, BodyVector &,
                     BodyScalar &, FaceVector &)>
      set_initial_condition;
  // Boundary condition on the host, called for every cell including ghosts
  // (set_boundary_condition through boundary_condition_kernel, INL:56-66);
  // every reference application leaves it empty.
  std::function<void(const PositioningSetup &, const RCo<7> &, BodyScalar &, BodyVector &,
                     BodyScalar &, FaceVector &)>
      set_boundary_condition;
  std::function<int(MHDSolver &)> monitor_hook;      // non-zero breaks proceed()
  std::function<int(MHDSolver &)> initial_event_hook;
  std::function<int(MHDSolver &)> final_event_hook;
  std::function<void(MHDSolver &, const std::string &)> dump_hook;

  explicit MHDSolver(const SimulationConfig &cfg, LocalSetup &local_setup = default_local())
      : cfg_(cfg), local_setup_(local_setup) {
    nmhd_desc d{};
    for (int i = 0; i < 3; ++i) {
      d.global_size[i] = cfg.global_size[i];
      d.margin[i] = cfg.margin[i];
      d.global_extent[i] = cfg.global_extent[i];
    }
    d.gamma = cfg.gamma;
    d.cfl_factor = cfg.cfl_factor;
    d.source = cfg.source;
    d.omega = cfg.omega;
    d.qshear = cfg.qshear;
    d.shear_y = cfg.shearing_boundary ? 1 : 0;
    d.recon = cfg.recon;
    d.integrator = cfg.integrator;
    d.path = cfg.path;
    d.fixes = cfg.fixes;
    if (nmhd_create(&d, &ctx_) != NMHD_OK)
      throw std::runtime_error(std::string("nmhd_create: ") + nmhd_last_error(0));
    nmhd_get_geometry(ctx_, &geo_);
    for (int i = 0; i < 3; ++i) {
      lat_.n[i] = geo_.size[i];
      lat_.d[i] = geo_.d[i];
      global_setup_.global_size_[i] = cfg.global_size[i];
      global_setup_.global_extent_[i] = cfg.global_extent[i];
      global_setup_.global_corner_[i] = cfg.global_corner[i];
      global_setup_.d_[i] = geo_.d[i];
      local_setup_.computational_size_d_[i] = geo_.size[i];
      local_setup_.margin_[i] = cfg.margin[i];
      local_setup_.computational_extent_[i] = geo_.extent[i];
      local_setup_.d_[i] = geo_.d[i];
      // LocalSetup constructor, mhd_solver.hrb:75-78 (rank 0)
      local_setup_.computational_corner_[i] =
          cfg.global_corner[i] + Real(-cfg.margin[i]) * geo_.d[i];
    }
    local_setup_.computational_size_ = geo_.ncell;
    host_.assign(8, std::vector<Real>(geo_.ncell, Real(0)));
  }
  ~MHDSolver() { nmhd_destroy(ctx_); }
  MHDSolver(const MHDSolver &) = delete;
  MHDSolver &operator=(const MHDSolver &) = delete;

  GlobalSetup &global_setup() { return global_setup_; }
  const GlobalSetup &global_setup() const { return global_setup_; }
  LocalSetup &local_setup() { return local_setup_; }
  const LocalSetup &local_setup() const { return local_setup_; }
  const Lattice &lattice() const { return lat_; }
  PositioningSetup positioning_setup() const {
    return PositioningSetup{local_setup_.computational_corner(), 0};
  }
  template <class T>
  Triplet<Real> positioning(const T &positioner) const {
    return positioner.position(local_setup_.computational_corner());
  }
  PreciseReal current_time() const { return global_setup_.current_time(); }
  int generation() const { return global_setup_.generation(); }
  nmhd_ctx *context() { return ctx_; }

  // Host copies of the state (downloaded on request, uploaded by sync_to_device).
  BodyScalar density() { return BodyScalar(&host_[0][0]); }
  BodyVector velocity() {
    return BodyVector(BodyScalar(&host_[1][0]), BodyScalar(&host_[2][0]),
                      BodyScalar(&host_[3][0]));
  }
  BodyScalar pressure() { return BodyScalar(&host_[4][0]); }
  FaceVector magnet() { return FaceVector(&host_[5][0], &host_[6][0], &host_[7][0]); }
  std::vector<Real> &host_field(int f) { return host_[f]; }
  void sync_to_host() { check(nmhd_get_state(ctx_, ptrs()), "nmhd_get_state"); }
  void sync_to_device() { check(nmhd_set_state(ctx_, cptrs()), "nmhd_set_state"); }

  // initialize_state / initial_condition (INL:43-54, 1246-1248)
  void initialize_state() {
    if (set_initial_condition) {
      BodyScalar rho = density(), p = pressure();
      BodyVector v = velocity();
      FaceVector b = magnet();
      const PositioningSetup pos = positioning_setup();
      for (size_t a = 0; a < geo_.ncell; ++a) set_initial_condition(pos, RCo<7>(lat_, (long)a),
                                                                   rho, v, p, b);
    }
    sync_to_device();
    sync_time();
  }

  // proceed (INL:1086-1204): integrate until goal_time; true when it was
  // reached, false when monitor() broke the loop or the wallclock limit
  // passed first (ignored when negative).
  bool proceed(const Real goal_time, const PreciseReal wallclock_time_limit = -1) {
    const PreciseReal begin = wallclock_now();
    const bool per_step = (bool)monitor_hook || wallclock_time_limit > 0;
    check(nmhd_reset_run(ctx_), "nmhd_reset_run");
    bool reached = false;
    while (!reached) {
      if (monitor()) break;
      if (wallclock_time_limit > 0 && wallclock_now() > begin + wallclock_time_limit) break;
      int steps = 0, r = 0;
      check(nmhd_enqueue_steps(ctx_, goal_time, per_step ? 1 : 32), "nmhd_enqueue_steps");
      check(nmhd_synchronize(ctx_, &steps, &r), "nmhd_synchronize");
      reached = r != 0;
      pull_time();
    }
    update_displacement();
    return reached;
  }

  // read / write (INL:1207-1244): three int sizes (ghosts included) then the
  // eight fields, rho, vx, vy, vz, p, Bx, By, Bz, as raw Real arrays.
  bool read(const std::string &filename) {
    FILE *fp = std::fopen(filename.c_str(), "rb");
    if (!fp) return false;
    int sz[3];
    for (int i = 0; i < 3; ++i)
      if (std::fread(&sz[i], sizeof(int), 1, fp) != 1 || sz[i] != geo_.size[i]) {
        std::fclose(fp);
        return false;
      }
    for (int f = 0; f < 8; ++f)
      if (std::fread(&host_[f][0], sizeof(Real), geo_.ncell, fp) != geo_.ncell) {
        std::fclose(fp);
        return false;
      }
    std::fclose(fp);
    sync_to_device();
    return true;
  }
  bool write(const std::string &filename) {
    sync_to_host();
    FILE *fp = std::fopen(filename.c_str(), "wb");
    if (!fp) return false;
    for (int i = 0; i < 3; ++i) std::fwrite(&geo_.size[i], sizeof(int), 1, fp);
    for (int f = 0; f < 8; ++f) std::fwrite(&host_[f][0], sizeof(Real), geo_.ncell, fp);
    std::fclose(fp);
    return true;
  }

  // save_state / load_state (INL:1250-1272): <dir>/<tag>/<rank>.bin and
  // <rank>_setup.txt (GlobalSetup, a line of '#', LocalSetup).
  void save_state(const std::string &tag) {
    const std::string prefix = local_setup_.mkdir_prefix(tag);
    write(prefix + ".bin");
    update_displacement();
    std::ofstream ofs((prefix + "_setup.txt").c_str());
    write_setup(ofs, global_setup_, local_setup_);
  }
  void load_state(const std::string &tag) {
    const std::string prefix = local_setup_.mkdir_prefix(tag);
    std::ifstream ifs((prefix + "_setup.txt").c_str());
    read_setup(ifs, global_setup_, local_setup_);
    if (!read(prefix + ".bin"))
      std::cerr << "error reading file: " << prefix << ".bin" << std::endl;
    sync_time();
  }
  void dump(const std::string &tag) {
    if (dump_hook) {
      dump_hook(*this, tag);
      return;
    }
    const std::string prefix = local_setup_.mkdir_prefix(tag);
    write(prefix + ".bin");
    std::ofstream ofs((prefix + "_setup.txt").c_str());
    write_setup(ofs, global_setup_, local_setup_);
  }

  int monitor() { return monitor_hook ? monitor_hook(*this) : 0; }

  // Run diagnostics reduced on the GPU: <B_d^2> over every face (what the
  // Kepler-disk monitor computes on the host, application/kepler_disk_demo/
  // mhd_solver.curb:107-130), conservation totals and max |div B|.
  nmhd_diagnostics diagnostics() {
    nmhd_diagnostics d;
    check(nmhd_get_diagnostics(ctx_, &d), "nmhd_get_diagnostics");
    return d;
  }
  int initial_event() { return initial_event_hook ? initial_event_hook(*this) : 0; }
  int final_event() { return final_event_hook ? final_event_hook(*this) : 0; }

  // communicate (INL:307-711): the ghost layers of the state from its
  // interior at the current time -- shearing-periodic along the radial
  // axis Y when the box is a shearing box, plain periodic otherwise, and
  // plain periodic along Z.  proceed() does this at the start of every step.
  void communicate() { check(nmhd_fill_ghosts(ctx_), "nmhd_fill_ghosts"); }

  // boundary_condition (INL:56-66): set_boundary_condition for every cell.
  // Without a hook (every reference application) this is a no-op and, like
  // the reference's empty kernel, touches nothing; with one, the state makes
  // a host round trip.
  void boundary_condition() {
    if (!set_boundary_condition) return;
    sync_to_host();
    BodyScalar rho = density(), p = pressure();
    BodyVector v = velocity();
    FaceVector b = magnet();
    const PositioningSetup pos = positioning_setup();
    for (size_t a = 0; a < geo_.ncell; ++a)
      set_boundary_condition(pos, RCo<7>(lat_, (long)a), rho, v, p, b);
    sync_to_device();
  }

  // add_flux_kernel_caller (INL:918-988): adds the right-hand side of the
  // state (density, velocity, pressure, magnet) -- the X, Y, Z sweeps of
  // face fluxes, their divergence and the constrained-transport EMF
  // differences, with the source term's face predictor taken over the step
  // dt -- to the eight pace arrays (the reference calls it after clear_pace,
  // so the paces then hold the right-hand side).  Views are host arrays of
  // the solver's size; a state other than the solver's own views
  // (density(), velocity(), pressure(), magnet()) becomes the solver's state.
  void add_flux_kernel_caller(Real dt, BodyScalar density, BodyVector velocity,
                              BodyScalar pressure, FaceVector magnet, BodyScalar density_pace,
                              BodyVector momentum_pace, BodyScalar energy_pace,
                              FaceVector magnet_pace) {
    const Real *in[8] = {density.ptr(),  velocity[X].ptr(), velocity[Y].ptr(),
                         velocity[Z].ptr(), pressure.ptr(),  magnet[X].ptr(),
                         magnet[Y].ptr(), magnet[Z].ptr()};
    for (int f = 0; f < 8; ++f)
      if (in[f] != &host_[f][0]) std::copy(in[f], in[f] + geo_.ncell, host_[f].begin());
    sync_to_device();
    std::vector<std::vector<Real> > pace(8, std::vector<Real>(geo_.ncell));
    float *pp[8];
    for (int f = 0; f < 8; ++f) pp[f] = &pace[f][0];
    check(nmhd_rhs(ctx_, dt, pp), "nmhd_rhs");
    Real *out[8] = {density_pace.ptr(),  momentum_pace[X].ptr(), momentum_pace[Y].ptr(),
                    momentum_pace[Z].ptr(), energy_pace.ptr(),    magnet_pace[X].ptr(),
                    magnet_pace[Y].ptr(), magnet_pace[Z].ptr()};
    for (int f = 0; f < 8; ++f)
      for (size_t a = 0; a < geo_.ncell; ++a) out[f][a] += pace[f][a];
  }

  // relative_displacement / update_displacement (INL:288-303)
  void relative_displacement(PreciseReal t, int index, Real &normalized, int &int_displacement,
                             Real &minor_displacement) {
    (void)index;
    check(nmhd_relative_displacement(ctx_, t, &normalized, &int_displacement,
                                     &minor_displacement),
          "nmhd_relative_displacement");
  }
  void update_displacement() {
    if (!cfg_.shearing_boundary) return;
    Real nd, md;
    int n;
    relative_displacement(current_time(), 0, nd, n, md);
    local_setup_.relative_displacement() = nd;
    local_setup_.relative_velocity() =
        -cfg_.qshear * cfg_.omega * Real(cfg_.global_extent[1] / Real(cfg_.global_size[1]) *
                                         Real(cfg_.global_size[1]));
  }

 private:
  static LocalSetup &default_local() {
    static LocalSetup blanc;
    return blanc;
  }
  void check(int rc, const char *what) {
    if (rc != NMHD_OK)
      throw std::runtime_error(std::string(what) + ": " + nmhd_last_error(ctx_));
  }
  float *const *ptrs() {
    for (int f = 0; f < 8; ++f) p_[f] = &host_[f][0];
    return p_;
  }
  const float *const *cptrs() {
    for (int f = 0; f < 8; ++f) p_[f] = &host_[f][0];
    return p_;
  }
  void sync_time() {
    check(nmhd_set_time(ctx_, global_setup_.current_time(), global_setup_.generation()),
          "nmhd_set_time");
  }
  void pull_time() {
    double t;
    int g;
    check(nmhd_get_time(ctx_, &t, &g), "nmhd_get_time");
    global_setup_.current_time() = t;
    global_setup_.generation() = g;
  }

  SimulationConfig cfg_;
  GlobalSetup global_setup_;
  LocalSetup &local_setup_;
  nmhd_ctx *ctx_ = nullptr;
  nmhd_geometry geo_;
  Lattice lat_;
  std::vector<std::vector<Real> > host_;
  float *p_[8];
};

}  // namespace nmhd

#endif  // NMHD_MHD_SOLVER_H
