// nmhd_kernels.h -- internal interface between the C ABI (capi.hip) and the
// kernel translation units.  Not installed; the public ABI is include/nmhd.h.
#pragma once

#include "nmhd_device.h"

namespace nmhd {

struct FieldPtrs {
  float *f[8];  // rho, vx, vy, vz, p, Bx, By, Bz
};
struct FluxIn {
  const float *f[15];
};
struct FluxOut {
  float *f[7];
};
// Face fluxes of the three sweeps: dir[d].f[k], k = rho, m_x, m_y, m_z, E,
// B_next(d), B_prev(d) (the reference's flux buffers, INL:26-27).
struct FluxSet {
  FluxOut dir[3];
};

// Which stage step a stage kernel uses: StepState::h, ::dt or ::h_diag.
enum HSel { H_HALF = 0, H_FULL = 1, H_DIAG = 2 };

constexpr int kBlock = 256;

inline unsigned grid_for(long n, int block = kBlock) {
  return (unsigned)((n + block - 1) / block);
}

void host_displacement(const Box &b, double t, float *normalized, int *n, float *minor);

// kernels_common.hip
void launch_begin_step(hipStream_t s, StepState *st, bool full);
void launch_cfl_ghosts(hipStream_t s, const Box &b, FieldPtrs W, StepState *st);
void launch_ghost_y(hipStream_t s, const Box &b, FieldPtrs W, const StepState *st);
void launch_ghost_z(hipStream_t s, const Box &b, FieldPtrs W, const StepState *st);
void launch_cfl(hipStream_t s, const Box &b, FieldPtrs W, float goal, StepState *st);
void launch_dt(hipStream_t s, StepState *st, float goal);
void launch_commit(hipStream_t s, StepState *st);
void launch_copy_state(hipStream_t s, const Box &b, FieldPtrs src, FieldPtrs dst,
                       const StepState *st);
void launch_hlld_batch(hipStream_t s, long n, float gamma, bool strict, FluxIn in,
                       FluxOut out);
void launch_diag_prepare(hipStream_t s, StepState *diag, const StepState *st, float h);

// kernels_unfused.hip (fast arithmetic) and kernels_strict.hip (reference
// rounding, bit-exact against the CPU oracle)
void launch_flux_strict(hipStream_t s, int dir, const Box &b, FieldPtrs W, FluxOut F,
                        const StepState *st, int hsel);
void launch_update_strict(hipStream_t s, const Box &b, FieldPtrs Win, const FluxSet &F,
                          FieldPtrs Wbase, FieldPtrs Wout, const StepState *st, int hsel,
                          bool pace_only);
void launch_hlld_batch_strict(hipStream_t s, long n, float gamma, FluxIn in, FluxOut out);
void launch_flux(hipStream_t s, int dir, const Box &b, FieldPtrs W, FluxOut F,
                 const StepState *st, int hsel);
void launch_update(hipStream_t s, const Box &b, FieldPtrs Win, const FluxSet &F,
                   FieldPtrs Wbase, FieldPtrs Wout, const StepState *st, int hsel,
                   bool pace_only);

// kernels_tiled.hip
int tiled_zchunk(const Box &b);
void launch_faces(hipStream_t s, const Box &b, FieldPtrs W, const FluxSet &F, const StepState *st,
                  int hsel);
void launch_update_tiled(hipStream_t s, const Box &b, FieldPtrs Win, const FluxSet &F,
                         FieldPtrs Wbase, FieldPtrs Wout, StepState *st, int hsel, bool pace_only,
                         bool cfl);

}  // namespace nmhd
