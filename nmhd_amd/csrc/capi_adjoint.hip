// capi_adjoint.hip -- the nmhd_adj_* entry points of include/nmhd.h: the
// forward step with a frozen dt and its hand-written reverse, on caller-owned
// device state arrays (e.g. torch tensors), launched on a caller-chosen
// stream.  The reverse step needs the step's stage states (W*, W^{n+1}):
// nmhd_adj_reverse recomputes them from W^n with the forward kernels,
// nmhd_adj_reverse_stages takes them from the caller (as
// nmhd_adj_forward_stages stored them); both then run the stage reverses
// (kernels_adjoint.hip) from the last stage to the first.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "capi_internal.h"

using namespace nmhd;

namespace nmhd {

void enqueue_frozen_step(hipStream_t s, const Box &b, int integrator, FieldPtrs Win,
                         FieldPtrs Wstar, FieldPtrs Wout, const FluxSet &F, StepState *st,
                         float dt) {
  launch_adj_set_dt(s, st, dt);
  if (integrator == NMHD_EULER1) {
    launch_faces(s, b, Win, F, st, H_FULL);
    launch_update_tiled(s, b, Win, F, Win, Wout, st, H_FULL, false, false);
    return;
  }
  launch_faces(s, b, Win, F, st, H_HALF);
  launch_update_tiled(s, b, Win, F, Win, Wstar, st, H_HALF, false, false);
  launch_faces(s, b, Wstar, F, st, H_FULL);
  launch_update_tiled(s, b, Wstar, F, Win, Wout, st, H_FULL, false, false);
}

}  // namespace nmhd

namespace {
// per thread: the message of the calling thread's last object-less failure
thread_local std::string g_adj_create_error;
constexpr int kWork = 5 + 3 + 3 + 3 + 45;  // gU, lE, lbb, abn, face
}  // namespace

struct nmhd_adj {
  nmhd_desc desc;
  Box box;
  nmhd_geometry geo;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  float *wstar[8] = {};   // recomputed stage-1 state
  float *wnext[8] = {};   // recomputed stage-2 state
  float *lstar[8] = {};   // dL/dW*
  float *flux[21] = {};   // face fluxes of the forward stages
  float *work[kWork] = {};
  float *scratch_out[8] = {};
  StepState *st = nullptr;
  std::string err;
};

namespace {

int afail(nmhd_adj *a, int code, const std::string &msg) {
  if (a)
    a->err = msg;
  else
    g_adj_create_error = msg;
  return code;
}

#define AHIP(a, call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return afail((a), NMHD_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

FieldPtrs to_fields(float *const p[8]) {
  FieldPtrs f;
  for (int i = 0; i < 8; ++i) f.f[i] = p[i];
  return f;
}

FieldPtrs to_fields_c(const float *const p[8]) {
  FieldPtrs f;
  for (int i = 0; i < 8; ++i) f.f[i] = const_cast<float *>(p[i]);
  return f;
}

FluxSet flux_set(nmhd_adj *a) {
  FluxSet F;
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < 7; ++k) F.dir[d].f[k] = a->flux[d * 7 + k];
  return F;
}

// W_out = S(W_in; dt) with the production stage kernels.
void enqueue_forward(nmhd_adj *a, FieldPtrs Win, FieldPtrs Wstar, FieldPtrs Wout, float dt) {
  enqueue_frozen_step(a->stream, a->box, a->desc.integrator, Win, Wstar, Wout, flux_set(a), a->st,
                      dt);
}

AdjStageArgs stage_args(nmhd_adj *a) {
  AdjStageArgs A;
  std::memset(&A, 0, sizeof(A));
  float **w = a->work;
  for (int i = 0; i < 5; ++i) A.gU[i] = *w++;
  for (int i = 0; i < 3; ++i) A.lE[i] = *w++;
  for (int i = 0; i < 3; ++i) A.lbb[i] = *w++;
  for (int i = 0; i < 3; ++i) A.abn[i] = *w++;
  for (int i = 0; i < 45; ++i) A.face[i] = *w++;
  return A;
}

void free_adj(nmhd_adj *a) {
  float **groups[] = {a->wstar, a->wnext, a->lstar, a->scratch_out};
  for (float **g : groups)
    for (int i = 0; i < 8; ++i)
      if (g[i]) (void)hipFree(g[i]), g[i] = nullptr;
  for (auto &p : a->flux)
    if (p) (void)hipFree(p), p = nullptr;
  for (auto &p : a->work)
    if (p) (void)hipFree(p), p = nullptr;
  if (a->st) (void)hipFree(a->st), a->st = nullptr;
  if (a->own) (void)hipStreamDestroy(a->own), a->own = nullptr;
}

bool ptrs_ok(const float *const p[8]) {
  if (!p) return false;
  for (int i = 0; i < 8; ++i)
    if (!p[i]) return false;
  return true;
}

}  // namespace

namespace nmhd {
void set_adj_global_error(const std::string &msg) { g_adj_create_error = msg; }
}  // namespace nmhd

extern "C" {

const char *nmhd_adj_last_error(const nmhd_adj *a) {
  return a ? a->err.c_str() : g_adj_create_error.c_str();
}

int nmhd_adj_create(const nmhd_desc *desc, nmhd_adj **out) {
  if (!desc || !out) return afail(nullptr, NMHD_EINVAL, "null argument");
  *out = nullptr;
  std::string why;
  if (!desc_valid(desc, why)) return afail(nullptr, NMHD_EINVAL, why);
  if (desc->margin[0] || desc->margin[1] || desc->margin[2] ||
      desc->source != NMHD_SOURCE_NONE || desc->shear_y)
    return afail(nullptr, NMHD_EINVAL,
                 "the adjoint step is for periodic boxes (margin 0, source NONE, shear_y 0)");
  nmhd_adj *a = new nmhd_adj();
  a->desc = *desc;
  derive_box(desc, a->box, a->geo);
  auto bail = [&](int code, const std::string &msg) {
    free_adj(a);
    delete a;
    return afail(nullptr, code, msg);
  };
  if (hipStreamCreateWithFlags(&a->own, hipStreamNonBlocking) != hipSuccess)
    return bail(NMHD_EHIP, "hipStreamCreate failed");
  a->stream = a->own;
  const size_t bytes = sizeof(float) * (size_t)a->box.ncell;
  float **groups[] = {a->wstar, a->wnext, a->lstar, a->scratch_out};
  for (float **g : groups)
    for (int i = 0; i < 8; ++i)
      if (hipMalloc(&g[i], bytes) != hipSuccess) return bail(NMHD_ENOMEM, "allocation failed");
  for (auto &p : a->flux)
    if (hipMalloc(&p, bytes) != hipSuccess) return bail(NMHD_ENOMEM, "allocation failed");
  for (auto &p : a->work)
    if (hipMalloc(&p, bytes) != hipSuccess) return bail(NMHD_ENOMEM, "allocation failed");
  if (hipMalloc(&a->st, sizeof(StepState)) != hipSuccess)
    return bail(NMHD_ENOMEM, "allocation failed");
  StepState h;
  std::memset(&h, 0, sizeof(h));
  if (hipMemcpy(a->st, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess)
    return bail(NMHD_EHIP, "step record init failed");
  *out = a;
  return NMHD_OK;
}

int nmhd_adj_destroy(nmhd_adj *a) {
  if (!a) return NMHD_OK;
  if (a->stream) (void)hipStreamSynchronize(a->stream);
  free_adj(a);
  delete a;
  return NMHD_OK;
}

int nmhd_adj_get_geometry(const nmhd_adj *a, nmhd_geometry *geo) {
  if (!a || !geo) return NMHD_EINVAL;
  *geo = a->geo;
  return NMHD_OK;
}

int nmhd_adj_set_stream(nmhd_adj *a, void *stream) {
  if (!a) return NMHD_EINVAL;
  if (a->stream != (hipStream_t)stream) AHIP(a, hipStreamSynchronize(a->stream));
  a->stream = (hipStream_t)stream;  // NULL: HIP's null stream
  // the object's own stream is used only until the caller names one: release
  // it, so that it does not share the process's few hardware queues with
  // the caller's streams (streams beyond them share a queue and serialise)
  if (a->own && a->own != a->stream) {
    AHIP(a, hipStreamDestroy(a->own));
    a->own = nullptr;
  }
  return NMHD_OK;
}

int nmhd_adj_synchronize(nmhd_adj *a) {
  if (!a) return NMHD_EINVAL;
  AHIP(a, hipGetLastError());
  AHIP(a, hipStreamSynchronize(a->stream));
  return NMHD_OK;
}

int nmhd_adj_cfl_dt(nmhd_adj *a, const float *const w[8], float goal_time, float *dt) {
  if (!a || !ptrs_ok(w) || !dt) return afail(a, NMHD_EINVAL, "bad argument");
  StepState h;
  std::memset(&h, 0, sizeof(h));
  h.cfl_bits = 0x7f800000u;
  AHIP(a, hipMemcpyAsync(a->st, &h, sizeof(h), hipMemcpyHostToDevice, a->stream));
  launch_cfl(a->stream, a->box, to_fields_c(w), goal_time, a->st);
  AHIP(a, hipMemcpyAsync(&h, a->st, sizeof(h), hipMemcpyDeviceToHost, a->stream));
  AHIP(a, hipStreamSynchronize(a->stream));
  std::memcpy(dt, &h.cfl_bits, sizeof(float));
  return NMHD_OK;
}

int nmhd_adj_forward(nmhd_adj *a, const float *const w_in[8], float dt, float *const w_out[8]) {
  if (!a || !ptrs_ok(w_in) || !w_out || !ptrs_ok(w_out)) return afail(a, NMHD_EINVAL, "bad argument");
  for (int i = 0; i < 8; ++i)
    if (w_in[i] == w_out[i]) return afail(a, NMHD_EINVAL, "w_out must not alias w_in");
  enqueue_forward(a, to_fields_c(w_in), to_fields(a->wstar), to_fields(w_out), dt);
  AHIP(a, hipGetLastError());
  return NMHD_OK;
}

int nmhd_adj_forward_stages(nmhd_adj *a, const float *const w_in[8], float dt,
                            float *const w_stage[8], float *const w_out[8]) {
  if (!a || !ptrs_ok(w_in) || !w_out || !ptrs_ok(w_out)) return afail(a, NMHD_EINVAL, "bad argument");
  const bool two = a->desc.integrator != NMHD_EULER1;
  if (two && (!w_stage || !ptrs_ok(w_stage))) return afail(a, NMHD_EINVAL, "bad argument");
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j)
      if (w_in[i] == w_out[j] || (two && (w_stage[i] == w_in[j] || w_stage[i] == w_out[j])))
        return afail(a, NMHD_EINVAL, "w_in, w_stage and w_out must not alias");
  enqueue_forward(a, to_fields_c(w_in), two ? to_fields(w_stage) : to_fields(a->wstar),
                  to_fields(w_out), dt);
  AHIP(a, hipGetLastError());
  return NMHD_OK;
}

namespace {

// lam_in = (dS/dW at Wn)^T lam_out given the step's stage states Ws, Wnext.
void enqueue_reverse(nmhd_adj *a, FieldPtrs Wn, FieldPtrs Ws, FieldPtrs Wnext, float dt,
                     FieldPtrs lam_out, FieldPtrs lam_in) {
  AdjStageArgs A = stage_args(a);
  if (a->desc.integrator == NMHD_EULER1) {
    A.Win = Wn, A.Wbase = Wn, A.Wout = Wnext;
    A.lam_out = lam_out;
    A.lam_base = lam_in;
    A.lam_in = lam_in;
    A.h = dt, A.acc_base = false, A.acc_in = true;
    launch_adj_stage(a->stream, a->box, A);
    return;
  }
  // stage 2: W^{n+1} = prim(cons(W^n) + dt P(W*))
  A.Win = Ws, A.Wbase = Wn, A.Wout = Wnext;
  A.lam_out = lam_out;
  A.lam_base = lam_in;
  A.lam_in = to_fields(a->lstar);
  A.h = dt, A.acc_base = false, A.acc_in = false;
  launch_adj_stage(a->stream, a->box, A);
  // stage 1: W* = prim(cons(W^n) + dt/2 P(W^n))
  A.Win = Wn, A.Wbase = Wn, A.Wout = Ws;
  A.lam_out = to_fields(a->lstar);
  A.lam_base = lam_in;
  A.lam_in = lam_in;
  A.h = dt / 2, A.acc_base = true, A.acc_in = true;
  launch_adj_stage(a->stream, a->box, A);
}

}  // namespace

int nmhd_adj_reverse(nmhd_adj *a, const float *const w_in[8], float dt,
                     const float *const lam_out[8], float *const lam_in[8]) {
  if (!a || !ptrs_ok(w_in) || !ptrs_ok(lam_out) || !lam_in || !ptrs_ok(lam_in))
    return afail(a, NMHD_EINVAL, "bad argument");
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j)
      if (lam_in[i] == lam_out[j] || lam_in[i] == w_in[j])
        return afail(a, NMHD_EINVAL, "lam_in must not alias w_in or lam_out");
  const FieldPtrs Wn = to_fields_c(w_in);
  const FieldPtrs Ws = to_fields(a->wstar);
  const FieldPtrs Wnext = to_fields(a->wnext);
  enqueue_forward(a, Wn, Ws, Wnext, dt);
  enqueue_reverse(a, Wn, Ws, Wnext, dt, to_fields_c(lam_out), to_fields(lam_in));
  AHIP(a, hipGetLastError());
  return NMHD_OK;
}

int nmhd_adj_reverse_stages(nmhd_adj *a, const float *const w_in[8],
                            const float *const w_stage[8], const float *const w_out[8],
                            float dt, const float *const lam_out[8], float *const lam_in[8]) {
  const bool two = a && a->desc.integrator != NMHD_EULER1;
  if (!a || !ptrs_ok(w_in) || !ptrs_ok(w_out) || (two && !ptrs_ok(w_stage)) ||
      !ptrs_ok(lam_out) || !lam_in || !ptrs_ok(lam_in))
    return afail(a, NMHD_EINVAL, "bad argument");
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j)
      if (lam_in[i] == lam_out[j] || lam_in[i] == w_in[j] || lam_in[i] == w_out[j] ||
          (two && lam_in[i] == w_stage[j]))
        return afail(a, NMHD_EINVAL, "lam_in must not alias a state or lam_out");
  enqueue_reverse(a, to_fields_c(w_in), two ? to_fields_c(w_stage) : to_fields(a->wstar),
                  to_fields_c(w_out), dt, to_fields_c(lam_out), to_fields(lam_in));
  AHIP(a, hipGetLastError());
  return NMHD_OK;
}

}  // extern "C"
