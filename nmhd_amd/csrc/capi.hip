// capi.hip -- the C ABI of include/nmhd.h: context lifetime, state transfer,
// step enqueueing on the context's stream and per-kernel event timing.
//
// A context holds three rotating copies of the 8 state fields (W^n, W*,
// W^{n+1}); a step reads W^n, writes W* in stage 1 and W^{n+1} in stage 2,
// and the roles rotate, so no kernel updates a field that a neighbouring
// thread still reads.  dt never leaves the device: the CFL minimum is
// reduced into a StepState record that the stage kernels read.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nmhd.h"
#include "capi_internal.h"

using namespace nmhd;

namespace nmhd {

bool desc_valid(const nmhd_desc *d, std::string &why) {
  for (int i = 0; i < 3; ++i) {
    if (d->global_size[i] <= 0) return why = "global_size must be positive", false;
    if (d->margin[i] < 0) return why = "margin must be >= 0", false;
    if (!(d->global_extent[i] > 0.0f)) return why = "global_extent must be positive", false;
  }
  if (d->margin[0] != 0) return why = "margin[0] must be 0 (X ghosts are never exchanged)", false;
  if (d->margin[1] > 0 && d->global_size[1] < d->margin[1])
    return why = "Y interior must hold at least margin[1] rows", false;
  if (d->margin[2] > 0 && d->global_size[2] < d->margin[2])
    return why = "Z interior must hold at least margin[2] planes", false;
  if (!(d->gamma > 1.0f)) return why = "gamma must be > 1", false;
  if (!(d->cfl_factor > 0.0f)) return why = "cfl_factor must be positive", false;
  if (d->source != NMHD_SOURCE_NONE && d->source != NMHD_SOURCE_SHEARING_BOX)
    return why = "unknown source", false;
  if (d->shear_y && d->margin[1] <= 0)
    return why = "shearing-periodic Y fill needs margin[1] > 0", false;
  if (d->recon != NMHD_RECON_MINBEE2 && d->recon != NMHD_RECON_PIECEWISECONST1)
    return why = "unknown recon", false;
  if (d->integrator != NMHD_MIDPOINT2 && d->integrator != NMHD_EULER1)
    return why = "unknown integrator", false;
  if (d->path < NMHD_PATH_AUTO || d->path > NMHD_PATH_STRICT) return why = "unknown path", false;
  if (d->fixes & ~NMHD_FIX_ALL) return why = "unknown bit in fixes", false;
  long n = 1;
  for (int i = 0; i < 3; ++i) n *= (long)(d->global_size[i] + 2 * d->margin[i]);
  if (n >= (1L << 31)) return why = "box exceeds 2^31 cells", false;
  // the kernels address a plane with 32-bit byte offsets
  if ((long)(d->global_size[0] + 2 * d->margin[0]) * (d->global_size[1] + 2 * d->margin[1]) >=
      (1L << 30))
    return why = "an XY plane exceeds 2^30 cells", false;
  return true;
}

// Derived constants, in FP32 exactly as the reference's constant
// expressions (simulation_config_default.h:79-103).
void derive_box(const nmhd_desc *d, Box &b, nmhd_geometry &g) {
  std::memset(&b, 0, sizeof(b));
  std::memset(&g, 0, sizeof(g));
  for (int i = 0; i < 3; ++i) {
    const int n = d->global_size[i] + 2 * d->margin[i];
    const float cell = d->global_extent[i] / (float)d->global_size[i];
    const float ext = cell * (float)n;
    b.n[i] = g.size[i] = n;
    b.m[i] = d->margin[i];
    b.ext[i] = g.extent[i] = ext;
    b.d[i] = g.d[i] = ext / (float)n;
    b.inv_d[i] = (float)(1.0 / (double)b.d[i]);
    b.local_ext[i] = g.local_extent[i] = cell * (float)d->global_size[i];
  }
  b.ncell = (long)b.n[0] * b.n[1] * b.n[2];
  g.ncell = (size_t)b.ncell;
  b.gamma = d->gamma;
  b.gm1 = d->gamma - 1.0f;
  b.inv_gm1 = (float)(1.0 / (double)b.gm1);
  b.cfl = d->cfl_factor;
  b.source = d->source;
  b.omega = d->omega;
  b.qshear = d->qshear;
  b.shear_y = d->shear_y;
  b.recon = d->recon;
  b.fix_cfl = (d->fixes & NMHD_FIX_CFL_SPEED) != 0;
  b.fix_interp = (d->fixes & NMHD_FIX_SHEAR_INTERP) != 0;
  b.face_off = (d->fixes & NMHD_FIX_FACE_POSITION) ? 0.0f : 0.5f;
}

}  // namespace nmhd

namespace {

enum Slot {
  S_BEGIN = 0,
  S_GHOST_Y,
  S_GHOST_Z,
  S_CFL,
  S_DT,
  S_FLUX_X,
  S_FLUX_Y,
  S_FLUX_Z,
  S_UPDATE1,
  S_UPDATE2,
  S_COPY,
  S_COMMIT,
  S_FACES1,
  S_FACES2,
  S_CFL_GHOSTS,
  S_NSLOT
};
const char *kSlotName[S_NSLOT] = {"begin_step", "ghost_y",   "ghost_z", "cfl",      "dt",
                                  "flux_x",     "flux_y",    "flux_z",  "update_s1", "update_s2",
                                  "copy",       "commit",    "faces_s1", "faces_s2", "cfl_ghosts"};

struct Pending {
  int slot;
  hipEvent_t a, b;
};

// per thread: the message of the calling thread's last failed create
thread_local std::string g_create_error;

}  // namespace

struct nmhd_ctx {
  nmhd_desc desc;
  Box box;
  nmhd_geometry geo;
  int device = 0;
  hipStream_t stream = nullptr;
  float *state[3][8] = {};
  int cur = 0;
  float *flux[3][7] = {};
  float *scratch8[8] = {};
  StepState *st = nullptr;
  StepState *diag = nullptr;
  double *diag_part = nullptr;  // run diagnostics: per-block partials, result
  double *diag_out = nullptr;
  int path = NMHD_PATH_TILED;
  // tile shapes of the tiled path (fastest measured at 256^3-512^3 on MI355X)
  int ftile = TV_32x8, utile = TV_64x4;
  bool cfl_valid = false;  // st->cfl_next_bits holds the CFL minimum of the state
  std::string err;
  bool timing = false;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  double slot_ms[S_NSLOT] = {};
  long slot_n[S_NSLOT] = {};
};

namespace {

int fail(nmhd_ctx *c, int code, const std::string &msg) {
  if (c)
    c->err = msg;
  else
    g_create_error = msg;
  return code;
}

#define HIPCHECK(c, call)                                                          \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail((c), NMHD_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

FieldPtrs fields(nmhd_ctx *c, int k) {
  FieldPtrs p;
  for (int f = 0; f < 8; ++f) p.f[f] = c->state[k][f];
  return p;
}

FluxSet fluxes(nmhd_ctx *c) {
  FluxSet F;
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < 7; ++k) F.dir[d].f[k] = c->flux[d][k];
  return F;
}

hipEvent_t get_event(nmhd_ctx *c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

template <class F>
void timed(nmhd_ctx *c, int slot, F &&launch) {
  if (!c->timing) {
    launch();
    return;
  }
  Pending p{slot, get_event(c), get_event(c)};
  (void)hipEventRecord(p.a, c->stream);
  launch();
  (void)hipEventRecord(p.b, c->stream);
  c->pending.push_back(p);
}

void collect_timing(nmhd_ctx *c) {
  for (auto &p : c->pending) {
    (void)hipEventSynchronize(p.b);
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      c->slot_ms[p.slot] += ms;
      c->slot_n[p.slot] += 1;
    }
    c->pool.push_back(p.a);
    c->pool.push_back(p.b);
  }
  c->pending.clear();
}

int resolve_path(int p) { return p == NMHD_PATH_AUTO ? NMHD_PATH_TILED : p; }

void flux_dispatch(nmhd_ctx *c, int d, FieldPtrs W, FluxOut F, const StepState *st, int hsel) {
  if (c->path == NMHD_PATH_STRICT)
    launch_flux_strict(c->stream, d, c->box, W, F, st, hsel);
  else
    launch_flux(c->stream, d, c->box, W, F, st, hsel);
}

void update_dispatch(nmhd_ctx *c, FieldPtrs Win, const FluxSet &F, FieldPtrs Wb, FieldPtrs Wo,
                     const StepState *st, int hsel, bool pace_only) {
  if (c->path == NMHD_PATH_STRICT)
    launch_update_strict(c->stream, c->box, Win, F, Wb, Wo, st, hsel, pace_only);
  else
    launch_update(c->stream, c->box, Win, F, Wb, Wo, st, hsel, pace_only);
}

// One iteration of the proceed() loop (INL:1094-1201), enqueued.
// after_step: the previous step of the same enqueue call left its commit to
// this step's opening launch; defer_commit: leave this step's commit to the
// next one (steps enqueued back to back share one bookkeeping launch).
void enqueue_step(nmhd_ctx *c, float goal, bool after_step = false, bool defer_commit = false) {
  const Box &b = c->box;
  hipStream_t s = c->stream;
  const int A = c->cur;
  const bool tiled = c->path == NMHD_PATH_TILED;
  const bool full_cfl = !(tiled && c->cfl_valid);
  const bool ghosts = b.m[1] > 0 || b.m[2] > 0;
  // with the carried CFL minimum and no ghost cells to add to it, dt is
  // decided in the opening launch
  const bool dt_in_open = !full_cfl && !ghosts;
  timed(c, S_BEGIN,
        [&] { launch_open_step(s, c->st, after_step, full_cfl, dt_in_open, goal); });
  if (b.m[1] > 0) timed(c, S_GHOST_Y, [&] { launch_ghost_y(s, b, fields(c, A), c->st); });
  if (b.m[2] > 0) timed(c, S_GHOST_Z, [&] { launch_ghost_z(s, b, fields(c, A), c->st); });
  if (full_cfl)
    timed(c, S_CFL, [&] { launch_cfl(s, b, fields(c, A), goal, c->st); });
  else if (ghosts)
    timed(c, S_CFL_GHOSTS, [&] { launch_cfl_ghosts(s, b, fields(c, A), c->st); });
  if (!dt_in_open) timed(c, S_DT, [&] { launch_dt(s, c->st, goal); });
  const FluxSet F = fluxes(c);
  const int slots[3] = {S_FLUX_X, S_FLUX_Y, S_FLUX_Z};
  // one RK stage: fluxes of W[in], update of the conserved base W[base] into W[out]
  auto stage = [&](int in, int base, int out, int hsel, bool last) {
    if (tiled) {
      timed(c, last ? S_FACES2 : S_FACES1,
            [&] { launch_faces(s, b, fields(c, in), F, c->st, hsel, c->ftile); });
      timed(c, last ? S_UPDATE2 : S_UPDATE1, [&] {
        launch_update_tiled(s, b, fields(c, in), F, fields(c, base), fields(c, out), c->st, hsel,
                            false, last, c->utile);
      });
      return;
    }
    for (int d = 0; d < 3; ++d)
      timed(c, slots[d], [&] { flux_dispatch(c, d, fields(c, in), F.dir[d], c->st, hsel); });
    timed(c, last ? S_UPDATE2 : S_UPDATE1, [&] {
      update_dispatch(c, fields(c, in), F, fields(c, base), fields(c, out), c->st, hsel, false);
    });
  };
  int next;
  if (c->desc.integrator == NMHD_EULER1) {
    next = (A + 1) % 3;
    stage(A, A, next, H_FULL, true);
  } else {
    const int B = (A + 1) % 3;
    next = (A + 2) % 3;
    stage(A, A, B, H_HALF, false);
    stage(B, A, next, H_FULL, true);
  }
  // an inactive step carries the state into the next buffer: on the tiled
  // path the last stage's update kernel does it, the other paths copy
  if (!tiled)
    timed(c, S_COPY, [&] { launch_copy_state(s, b, fields(c, A), fields(c, next), c->st); });
  if (!defer_commit) timed(c, S_COMMIT, [&] { launch_commit(s, c->st); });
  c->cur = next;
  c->cfl_valid = tiled;
}

int read_step_state(nmhd_ctx *c, StepState *h) {
  HIPCHECK(c, hipMemcpyAsync(h, c->st, sizeof(StepState), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return NMHD_OK;
}

int write_step_state(nmhd_ctx *c, const StepState *h) {
  HIPCHECK(c, hipMemcpyAsync(c->st, h, sizeof(StepState), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return NMHD_OK;
}

int sync_and_check(nmhd_ctx *c) {
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (c->timing) collect_timing(c);
  return NMHD_OK;
}

void free_all(nmhd_ctx *c) {
  for (auto &row : c->state)
    for (auto &p : row)
      if (p) (void)hipFree(p), p = nullptr;
  for (auto &row : c->flux)
    for (auto &p : row)
      if (p) (void)hipFree(p), p = nullptr;
  for (auto &p : c->scratch8)
    if (p) (void)hipFree(p), p = nullptr;
  if (c->st) (void)hipFree(c->st), c->st = nullptr;
  if (c->diag) (void)hipFree(c->diag), c->diag = nullptr;
  if (c->diag_part) (void)hipFree(c->diag_part), c->diag_part = nullptr;
  if (c->diag_out) (void)hipFree(c->diag_out), c->diag_out = nullptr;
  for (auto &p : c->pending) c->pool.push_back(p.a), c->pool.push_back(p.b);
  c->pending.clear();
  for (auto e : c->pool) (void)hipEventDestroy(e);
  c->pool.clear();
  if (c->stream) (void)hipStreamDestroy(c->stream), c->stream = nullptr;
}

}  // namespace

extern "C" {

const char *nmhd_last_error(const nmhd_ctx *ctx) {
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int nmhd_create(const nmhd_desc *desc, nmhd_ctx **out) {
  if (!desc || !out) return fail(nullptr, NMHD_EINVAL, "null argument");
  *out = nullptr;
  std::string why;
  if (!desc_valid(desc, why)) return fail(nullptr, NMHD_EINVAL, why);
  nmhd_ctx *c = new nmhd_ctx();
  c->desc = *desc;
  derive_box(desc, c->box, c->geo);
  c->path = resolve_path(desc->path);
  auto bail = [&](int code, const std::string &msg) {
    free_all(c);
    delete c;
    return fail(nullptr, code, msg);
  };
  if (hipGetDevice(&c->device) != hipSuccess) return bail(NMHD_EHIP, "no HIP device");
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(NMHD_EHIP, "hipStreamCreate failed");
  const size_t bytes = sizeof(float) * (size_t)c->box.ncell;
  for (auto &row : c->state)
    for (auto &p : row)
      if (hipMalloc(&p, bytes) != hipSuccess) return bail(NMHD_ENOMEM, "state allocation failed");
  for (auto &row : c->flux)
    for (auto &p : row)
      if (hipMalloc(&p, bytes) != hipSuccess) return bail(NMHD_ENOMEM, "flux allocation failed");
  if (hipMalloc(&c->st, sizeof(StepState)) != hipSuccess ||
      hipMalloc(&c->diag, sizeof(StepState)) != hipSuccess)
    return bail(NMHD_ENOMEM, "step record allocation failed");
  for (int k = 0; k < 3; ++k)
    for (int f = 0; f < 8; ++f)
      if (hipMemsetAsync(c->state[k][f], 0, bytes, c->stream) != hipSuccess)
        return bail(NMHD_EHIP, "state clear failed");
  StepState h;
  std::memset(&h, 0, sizeof(h));
  if (hipMemcpyAsync(c->st, &h, sizeof(h), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(c->diag, &h, sizeof(h), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return bail(NMHD_EHIP, "step record init failed");
  *out = c;
  return NMHD_OK;
}

int nmhd_destroy(nmhd_ctx *ctx) {
  if (!ctx) return NMHD_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_all(ctx);
  delete ctx;
  return NMHD_OK;
}

int nmhd_get_geometry(const nmhd_ctx *ctx, nmhd_geometry *geo) {
  if (!ctx || !geo) return NMHD_EINVAL;
  *geo = ctx->geo;
  return NMHD_OK;
}

static int copy_state(nmhd_ctx *c, const float *const src[8], float *const dst[8], bool in,
                      hipMemcpyKind kind) {
  if (!c) return NMHD_EINVAL;
  const size_t bytes = sizeof(float) * (size_t)c->box.ncell;
  for (int f = 0; f < 8; ++f) {
    if (in && !src[f]) return fail(c, NMHD_EINVAL, "null field pointer");
    if (!in && !dst[f]) return fail(c, NMHD_EINVAL, "null field pointer");
  }
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (in) c->cfl_valid = false;
  for (int f = 0; f < 8; ++f) {
    if (in)
      HIPCHECK(c, hipMemcpyAsync(c->state[c->cur][f], src[f], bytes, kind, c->stream));
    else
      HIPCHECK(c, hipMemcpyAsync(dst[f], c->state[c->cur][f], bytes, kind, c->stream));
  }
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  return NMHD_OK;
}

int nmhd_set_state(nmhd_ctx *ctx, const float *const fields[8]) {
  if (!fields) return NMHD_EINVAL;
  return copy_state(ctx, fields, nullptr, true, hipMemcpyHostToDevice);
}
int nmhd_get_state(nmhd_ctx *ctx, float *const fields[8]) {
  if (!fields) return NMHD_EINVAL;
  return copy_state(ctx, nullptr, fields, false, hipMemcpyDeviceToHost);
}
int nmhd_set_state_device(nmhd_ctx *ctx, const float *const f[8]) {
  if (!f) return NMHD_EINVAL;
  return copy_state(ctx, f, nullptr, true, hipMemcpyDeviceToDevice);
}
int nmhd_get_state_device(nmhd_ctx *ctx, float *const f[8]) {
  if (!f) return NMHD_EINVAL;
  return copy_state(ctx, nullptr, f, false, hipMemcpyDeviceToDevice);
}

int nmhd_set_time(nmhd_ctx *ctx, double t, int generation) {
  if (!ctx) return NMHD_EINVAL;
  HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
  StepState h;
  std::memset(&h, 0, sizeof(h));
  h.t = t;
  h.t_next = t;
  h.generation = generation;
  h.cfl_bits = 0x7f800000u;
  h.cfl_next_bits = 0x7f800000u;
  ctx->cfl_valid = false;
  return write_step_state(ctx, &h);
}

int nmhd_get_time(nmhd_ctx *ctx, double *t, int *generation) {
  if (!ctx) return NMHD_EINVAL;
  StepState h;
  int r = read_step_state(ctx, &h);
  if (r) return r;
  if (t) *t = h.t;
  if (generation) *generation = h.generation;
  return NMHD_OK;
}

int nmhd_relative_displacement(const nmhd_ctx *ctx, double t, float *normalized,
                               int *int_displacement, float *minor_displacement) {
  if (!ctx || !normalized || !int_displacement || !minor_displacement) return NMHD_EINVAL;
  host_displacement(ctx->box, t, normalized, int_displacement, minor_displacement);
  return NMHD_OK;
}

int nmhd_fill_ghosts(nmhd_ctx *c) {
  if (!c) return NMHD_EINVAL;
  launch_diag_prepare(c->stream, c->diag, c->st, 0.0f);
  if (c->box.m[1] > 0) launch_ghost_y(c->stream, c->box, fields(c, c->cur), c->diag);
  if (c->box.m[2] > 0) launch_ghost_z(c->stream, c->box, fields(c, c->cur), c->diag);
  return sync_and_check(c);
}

int nmhd_cfl_dt(nmhd_ctx *c, float goal_time, float *dt) {
  if (!c || !dt) return NMHD_EINVAL;
  launch_diag_prepare(c->stream, c->diag, c->st, 0.0f);
  launch_cfl(c->stream, c->box, fields(c, c->cur), goal_time, c->diag);
  StepState h;
  HIPCHECK(c, hipMemcpyAsync(&h, c->diag, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  int r = sync_and_check(c);
  if (r) return r;
  float v;
  std::memcpy(&v, &h.cfl_bits, sizeof(v));
  *dt = v;
  return NMHD_OK;
}

int nmhd_rhs(nmhd_ctx *c, float h, float *const pace[8]) {
  if (!c || !pace) return NMHD_EINVAL;
  const size_t bytes = sizeof(float) * (size_t)c->box.ncell;
  for (auto &p : c->scratch8)
    if (!p && hipMalloc(&p, bytes) != hipSuccess)
      return fail(c, NMHD_ENOMEM, "scratch allocation failed");
  FieldPtrs out;
  for (int f = 0; f < 8; ++f) out.f[f] = c->scratch8[f];
  launch_diag_prepare(c->stream, c->diag, c->st, h);
  const FluxSet F = fluxes(c);
  if (c->path == NMHD_PATH_TILED) {
    launch_faces(c->stream, c->box, fields(c, c->cur), F, c->diag, H_DIAG, c->ftile);
    launch_update_tiled(c->stream, c->box, fields(c, c->cur), F, fields(c, c->cur), out, c->diag,
                        H_DIAG, true, false, c->utile);
  } else {
    for (int d = 0; d < 3; ++d)
      flux_dispatch(c, d, fields(c, c->cur), F.dir[d], c->diag, H_DIAG);
    update_dispatch(c, fields(c, c->cur), F, fields(c, c->cur), out, c->diag, H_DIAG, true);
  }
  for (int f = 0; f < 8; ++f)
    HIPCHECK(c, hipMemcpyAsync(pace[f], out.f[f], bytes, hipMemcpyDeviceToHost, c->stream));
  return sync_and_check(c);
}

int nmhd_reset_run(nmhd_ctx *c) {
  if (!c) return NMHD_EINVAL;
  StepState h;
  int r = read_step_state(c, &h);
  if (r) return r;
  h.finished = 0;
  h.active = 0;
  h.goal_reached = 0;
  h.steps_done = 0;
  return write_step_state(c, &h);
}

int nmhd_enqueue_steps(nmhd_ctx *c, float goal_time, int nsteps) {
  if (!c || nsteps < 0) return NMHD_EINVAL;
  for (int k = 0; k < nsteps; ++k) enqueue_step(c, goal_time, k > 0, k + 1 < nsteps);
  HIPCHECK(c, hipGetLastError());
  return NMHD_OK;
}

int nmhd_synchronize(nmhd_ctx *c, int *steps_taken, int *goal_reached) {
  if (!c) return NMHD_EINVAL;
  int r = sync_and_check(c);
  if (r) return r;
  StepState h;
  r = read_step_state(c, &h);
  if (r) return r;
  if (steps_taken) *steps_taken = h.steps_done;
  if (goal_reached) *goal_reached = h.goal_reached;
  return NMHD_OK;
}

int nmhd_step(nmhd_ctx *c, float goal_time, float *dt, int *goal_reached) {
  if (!c) return NMHD_EINVAL;
  int r = nmhd_reset_run(c);
  if (r) return r;
  enqueue_step(c, goal_time);
  r = sync_and_check(c);
  if (r) return r;
  StepState h;
  r = read_step_state(c, &h);
  if (r) return r;
  if (dt) *dt = h.dt;
  if (goal_reached) *goal_reached = h.goal_reached;
  return NMHD_OK;
}

int nmhd_run(nmhd_ctx *c, float goal_time, int max_steps, int *steps_taken, int *goal_reached) {
  if (!c || max_steps < 0) return NMHD_EINVAL;
  int r = nmhd_reset_run(c);
  if (r) return r;
  const int chunk = 32;
  int done = 0, reached = 0, taken = 0;
  while (done < max_steps && !reached) {
    const int k = max_steps - done < chunk ? max_steps - done : chunk;
    r = nmhd_enqueue_steps(c, goal_time, k);
    if (r) return r;
    done += k;
    r = nmhd_synchronize(c, &taken, &reached);
    if (r) return r;
  }
  if (steps_taken) *steps_taken = taken;
  if (goal_reached) *goal_reached = reached;
  return NMHD_OK;
}

int nmhd_hlld_batch(size_t n, float gamma, int path, const float *const in[15],
                    float *const out[7]) {
  if (!in || !out) return fail(nullptr, NMHD_EINVAL, "null argument");
  if (n == 0) return NMHD_OK;
  const size_t bytes = n * sizeof(float);
  float *buf = nullptr;
  if (hipMalloc(&buf, bytes * 22) != hipSuccess) return fail(nullptr, NMHD_ENOMEM, "hipMalloc");
  FluxIn fi;
  FluxOut fo;
  int rc = NMHD_OK;
  for (int i = 0; i < 15; ++i) {
    fi.f[i] = buf + i * n;
    if (hipMemcpy(buf + i * n, in[i], bytes, hipMemcpyHostToDevice) != hipSuccess) rc = NMHD_EHIP;
  }
  for (int i = 0; i < 7; ++i) fo.f[i] = buf + (15 + i) * n;
  if (rc == NMHD_OK) {
    launch_hlld_batch(nullptr, (long)n, gamma, path == NMHD_PATH_STRICT, fi, fo);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = NMHD_EHIP;
  }
  for (int i = 0; i < 7 && rc == NMHD_OK; ++i)
    if (hipMemcpy(out[i], fo.f[i], bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = NMHD_EHIP;
  (void)hipFree(buf);
  if (rc) return fail(nullptr, rc, "HIP failure in nmhd_hlld_batch");
  return NMHD_OK;
}

int nmhd_timing_enable(nmhd_ctx *c, int enable) {
  if (!c) return NMHD_EINVAL;
  int r = sync_and_check(c);
  c->timing = enable != 0;
  return r;
}

int nmhd_timing_reset(nmhd_ctx *c) {
  if (!c) return NMHD_EINVAL;
  int r = sync_and_check(c);
  for (int i = 0; i < S_NSLOT; ++i) c->slot_ms[i] = 0.0, c->slot_n[i] = 0;
  return r;
}

int nmhd_timing_get(nmhd_ctx *c, int i, const char **name, double *total_ms, long *launches) {
  if (!c || i < 0 || i >= S_NSLOT) return NMHD_EINVAL;
  int r = sync_and_check(c);
  if (r) return r;
  if (name) *name = kSlotName[i];
  if (total_ms) *total_ms = c->slot_ms[i];
  if (launches) *launches = c->slot_n[i];
  return NMHD_OK;
}

int nmhd_set_path(nmhd_ctx *c, int path) {
  if (!c || path < NMHD_PATH_AUTO || path > NMHD_PATH_STRICT) return NMHD_EINVAL;
  c->path = resolve_path(path);
  c->cfl_valid = false;
  return NMHD_OK;
}

int nmhd_get_diagnostics(nmhd_ctx *c, nmhd_diagnostics *out) {
  if (!c || !out) return NMHD_EINVAL;
  if (!c->diag_part) {
    if (hipMalloc(&c->diag_part, sizeof(double) * kDiagVals * diag_blocks(c->box)) != hipSuccess ||
        hipMalloc(&c->diag_out, sizeof(double) * kDiagVals) != hipSuccess)
      return fail(c, NMHD_ENOMEM, "diagnostics allocation failed");
  }
  launch_diag(c->stream, c->box, fields(c, c->cur), c->diag_part, c->diag_out);
  double v[kDiagVals];
  HIPCHECK(c, hipMemcpyAsync(v, c->diag_out, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  int r = sync_and_check(c);
  if (r) return r;
  for (int d = 0; d < 3; ++d) out->b2_mean[d] = v[d] / (double)c->box.ncell;
  for (int k = 0; k < 8; ++k) out->total[k] = v[3 + k];
  out->divb_max = v[11];
  return NMHD_OK;
}

int nmhd_set_tiles(nmhd_ctx *c, int faces_tile, int update_tile) {
  if (!c || faces_tile < 0 || faces_tile >= TV_COUNT || update_tile < 0 ||
      update_tile >= TV_COUNT)
    return NMHD_EINVAL;
  int r = sync_and_check(c);
  c->ftile = faces_tile;
  c->utile = update_tile;
  return r;
}

int nmhd_get_tiles(const nmhd_ctx *c, int *faces_tile, int *update_tile) {
  if (!c) return NMHD_EINVAL;
  if (faces_tile) *faces_tile = c->ftile;
  if (update_tile) *update_tile = c->utile;
  return NMHD_OK;
}

int nmhd_get_path(const nmhd_ctx *c, int *path) {
  if (!c || !path) return NMHD_EINVAL;
  *path = c->path;
  return NMHD_OK;
}

}  // extern "C"
