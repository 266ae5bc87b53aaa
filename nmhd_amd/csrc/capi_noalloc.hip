// capi_noalloc.hip -- nmhd_adj_forward_noalloc: the frozen-dt step of
// nmhd_adj_forward for callers that cannot allocate device memory (work
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  The step's scratch -- the step record, the midpoint stage
// state W* and the 11 arrays the face kernel hands to the update kernel --
// lives in one file-scope device array of fixed size instead of buffers
// allocated by nmhd_adj_create.  Every GPU the library runs on holds its
// own copy of the array until the process ends.
//
// Layout of the array (floats): [0, 64) the step record (StepState, 256 B
// slot); then 19 arrays of `stride` floats each (ncell rounded up to a
// multiple of 64, so that every array starts 256-byte aligned): W* (8),
// the hydro divergence (5), the six face B fluxes (6).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "capi_internal.h"

using namespace nmhd;

namespace {
constexpr size_t kAlignFloats = 64;   // 256 bytes
constexpr size_t kRecordFloats = 64;  // step record slot
constexpr size_t kArrays = 8 + 11;    // W* + the face -> update hand-off
constexpr size_t kMaxCells = NMHD_NOALLOC_MAX_CELLS;
constexpr size_t kScratchFloats = kRecordFloats + kArrays * kMaxCells;
static_assert(kMaxCells % kAlignFloats == 0, "array strides stay within the scratch");
static_assert(sizeof(StepState) <= kRecordFloats * sizeof(float), "record slot too small");
}  // namespace

// The scratch.  In host code `scratch` names a host-side stand-in, not the
// device address: host code uses only the per-device pointer that
// hipGetSymbolAddress returns for it (scratch_on_current_device).
static __device__ __align__(256) float scratch[kScratchFloats];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
float *g_dev_scratch[kMaxDevices] = {};

int nfail(int code, const std::string &msg) {
  set_adj_global_error(msg);
  return code;
}

bool ptrs_ok(const float *const p[8]) {
  if (!p) return false;
  for (int i = 0; i < 8; ++i)
    if (!p[i]) return false;
  return true;
}

// The current device's address of `scratch`, looked up once per device.
int scratch_on_current_device(float **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return nfail(NMHD_EHIP, "no HIP device");
  if (dev < 0 || dev >= kMaxDevices) return nfail(NMHD_EHIP, "device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_dev_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return nfail(NMHD_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_dev_scratch[dev] = static_cast<float *>(p);
  }
  *out = g_dev_scratch[dev];
  return NMHD_OK;
}

}  // namespace

extern "C" {

int nmhd_adj_forward_noalloc(const nmhd_desc *desc, const float *const w_in[8], float dt,
                             float *const w_out[8], void *stream) {
  if (!desc || !ptrs_ok(w_in) || !w_out || !ptrs_ok(w_out))
    return nfail(NMHD_EINVAL, "bad argument");
  std::string why;
  if (!desc_valid(desc, why)) return nfail(NMHD_EINVAL, why);
  if (desc->margin[0] || desc->margin[1] || desc->margin[2] ||
      desc->source != NMHD_SOURCE_NONE || desc->shear_y)
    return nfail(NMHD_EINVAL,
                 "the adjoint step is for periodic boxes (margin 0, source NONE, shear_y 0)");
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j)
      if (w_in[i] == w_out[j]) return nfail(NMHD_EINVAL, "w_out must not alias w_in");
  Box b;
  nmhd_geometry g;
  derive_box(desc, b, g);
  if ((size_t)b.ncell > kMaxCells)
    return nfail(NMHD_EINVAL, "box of " + std::to_string((long long)b.ncell) +
                                  " cells exceeds the no-allocation scratch (at most " +
                                  std::to_string((long long)kMaxCells) + " cells)");
  float *base = nullptr;
  const int rc = scratch_on_current_device(&base);
  if (rc != NMHD_OK) return rc;

  const size_t stride = ((size_t)b.ncell + kAlignFloats - 1) / kAlignFloats * kAlignFloats;
  float *arr = base + kRecordFloats;
  auto slot = [&](size_t k) { return arr + k * stride; };
  FieldPtrs Win, Wstar, Wout;
  for (int f = 0; f < 8; ++f) {
    Win.f[f] = const_cast<float *>(w_in[f]);
    Wstar.f[f] = slot(f);
    Wout.f[f] = w_out[f];
  }
  // the tiled hand-off: hydro divergence in the x-sweep's hydro slots, the
  // six face B fluxes in the B slots of the three sweeps (tiled::stage_buf)
  FluxSet F;
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < 7; ++k) F.dir[d].f[k] = nullptr;
  for (int k = 0; k < 7; ++k) F.dir[0].f[k] = slot(8 + k);
  F.dir[1].f[5] = slot(15);
  F.dir[1].f[6] = slot(16);
  F.dir[2].f[5] = slot(17);
  F.dir[2].f[6] = slot(18);

  // the record: dt, h, active are set by the step's first kernel; nothing
  // else of it is read, so no clearing is needed
  StepState *st = reinterpret_cast<StepState *>(base);
  enqueue_frozen_step((hipStream_t)stream, b, desc->integrator, Win, Wstar, Wout, F, st, dt);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return nfail(NMHD_EHIP, std::string("launch: ") + hipGetErrorString(e));
  return NMHD_OK;
}

}  // extern "C"
