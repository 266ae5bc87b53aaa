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

// ------------------------------------------------ step record transitions
// (shared by the single-box and the multi-box bookkeeping kernels)

// Start a step.  full: the CFL minimum is rebuilt over every cell by a full
// CFL pass; otherwise it starts from the stage-2 epilogue's minimum of the
// current state (plus, with ghosts, the cells the ghost fill rewrote).
// (x, y, z) of the cell at linear index a (x fastest), in 32-bit unsigned
// arithmetic: a box holds fewer than 2^31 cells.
__device__ __forceinline__ void cell_coords(const Box &b, long a, int &x, int &y, int &z) {
  const unsigned u = (unsigned)a, n0 = (unsigned)b.n[0], n1 = (unsigned)b.n[1];
  const unsigned row = u / n0;
  x = (int)(u - row * n0);
  z = (int)(row / n1);
  y = (int)(row - (unsigned)z * n1);
}

// A step enqueued after the run reached its goal is inactive and leaves the
// record alone: cfl_next_bits then still holds the last active step's
// minimum, which the first step of the next run starts from.
__device__ __forceinline__ void step_begin(StepState &s, int full) {
  if (s.finished) return;
  s.cfl_bits = full ? 0x7f800000u : s.cfl_next_bits;
  s.cfl_next_bits = 0x7f800000u;  // +inf
}

// Decide the step: dt = CFL minimum, clipped so that t + dt <= goal
// (proceed(), INL:1126-1137).
__device__ __forceinline__ void step_dt(StepState &s, float goal) {
  if (s.finished) {
    s.active = 0;
    return;
  }
  float dt = fminf(goal, __uint_as_float(s.cfl_bits));
  const double next = s.t + (double)dt;
  if (next > (double)goal) {
    dt = (float)((double)goal - s.t);
    s.finished = 1;
    s.goal_reached = 1;
  }
  s.dt = dt;
  s.h = dt / 2;
  s.t_next = s.t + (double)dt;
  s.active = 1;
}

// GlobalSetup::proceed (mhd_solver.hrb:56-59)
__device__ __forceinline__ void step_commit(StepState &s) {
  if (!s.active) return;
  s.t = s.t_next;
  s.generation += 1;
  s.steps_done += 1;
  s.active = 0;
}

// CFL time of one cell (cfl_condition_inner, INL:714-720):
// min_d C*d_d/|v_d + sqrt((gamma p + b2)/rho)|, b2 = sum_d max(B_lo^2, B_hi^2)
// (|v_d| + sqrt(..) with fix_cfl), without FMA contraction so that every CFL
// pass rounds the same way.
__device__ __forceinline__ float cell_cfl(const Box &b, const FieldPtrs &W, long a, int x, int y,
                                          int z) {
#pragma clang fp contract(off)
  const float bxl = W.f[5][a], bxh = W.f[5][addr(b, wrap(x + 1, b.n[0]), y, z)];
  const float byl = W.f[6][a], byh = W.f[6][addr(b, x, wrap(y + 1, b.n[1]), z)];
  const float bzl = W.f[7][a], bzh = W.f[7][addr(b, x, y, wrap(z + 1, b.n[2]))];
  const float b2 = fmaxf(bxl * bxl, bxh * bxh) + fmaxf(byl * byl, byh * byh) +
                   fmaxf(bzl * bzl, bzh * bzh);
  const float c = sqrtf((b.gamma * W.f[4][a] + b2) / W.f[0][a]);
  float t = 3.0e38f;
  for (int d = 0; d < 3; ++d) t = fminf(t, b.cfl * b.d[d] / cfl_speed(b, W.f[1 + d][a], c));
  return t;
}

// ------------------------------------------------------- multi-box batches

// One box of a batch: its parameters and the offsets (in floats) of its
// 8 state fields and 21 face-flux arrays in the batch's pools.
struct BatchBox {
  Box b;
  long off8;
  long off21;
};

__host__ __device__ inline FieldPtrs box_fields(float *pool, const BatchBox &bb) {
  FieldPtrs p;
  for (int f = 0; f < 8; ++f) p.f[f] = pool + bb.off8 + (long)f * bb.b.ncell;
  return p;
}

__host__ __device__ inline FluxSet box_fluxes(float *pool, const BatchBox &bb) {
  FluxSet F;
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < 7; ++k) F.dir[d].f[k] = pool + bb.off21 + (long)(d * 7 + k) * bb.b.ncell;
  return F;
}

// One tile of a persistent launch: box, tile origin, z range, tile shape.
struct WorkItem {
  int box;
  int xy;       // x0 | y0 << 16
  int z;        // z0 | zend << 16
  int variant;  // TileVariant
};

// Tile shapes of the z-marching kernels (TX x TY, 256 threads each).
enum TileVariant { TV_64x4 = 0, TV_32x8 = 1, TV_16x16 = 2, TV_COUNT = 3 };
constexpr int kTileTX[TV_COUNT] = {64, 32, 16};
constexpr int kTileTY[TV_COUNT] = {4, 8, 16};

// kernels_common.hip
void launch_begin_step(hipStream_t s, StepState *st, bool full);
// commit the previous step (commit_prev), begin the next (full CFL pass or
// the carried minimum) and, with dt_too, decide its dt, in one launch
void launch_open_step(hipStream_t s, StepState *st, bool commit_prev, bool full, bool dt_too,
                      float goal);
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

// kernels_batch.hip (persistent multi-box launches over a work table)
void launch_batch_begin(hipStream_t s, StepState *st, int nbox, bool full);
void launch_batch_cfl(hipStream_t s, const BatchBox *boxes, int nbox, float *state, float goal,
                      StepState *st);
void launch_batch_dt(hipStream_t s, StepState *st, int nbox, const float *goal);
void launch_batch_commit(hipStream_t s, StepState *st, int nbox);
void launch_batch_copy(hipStream_t s, const BatchBox *boxes, int nbox, float *src, float *dst,
                       const StepState *st);
// ghost fills and their CFL minimum for the boxes with ghost layers (grid
// blocks per box, blockIdx.y the box; boxes without return at once)
void launch_batch_ghost_y(hipStream_t s, const BatchBox *boxes, int nbox, int grid, float *state,
                          const StepState *st);
void launch_batch_ghost_z(hipStream_t s, const BatchBox *boxes, int nbox, int grid, float *state,
                          const StepState *st);
void launch_batch_cfl_ghosts(hipStream_t s, const BatchBox *boxes, int nbox, int grid,
                             float *state, StepState *st);
void launch_batch_faces(hipStream_t s, int grid, const BatchBox *boxes, const WorkItem *items,
                        int nitems, float *state, float *flux, const StepState *st, int hsel);
void launch_batch_update(hipStream_t s, int grid, const BatchBox *boxes, const WorkItem *items,
                         int nitems, float *state_in, float *flux, float *state_base,
                         float *state_out, StepState *st, int hsel, bool cfl);
// One workgroup per item, all items of one tile shape (variant).
void launch_batch_faces_items(hipStream_t s, int variant, const BatchBox *boxes,
                              const WorkItem *items, int nitems, float *state, float *flux,
                              const StepState *st, int hsel);
void launch_batch_update_items(hipStream_t s, int variant, const BatchBox *boxes,
                               const WorkItem *items, int nitems, float *state_in, float *flux,
                               float *state_base, float *state_out, StepState *st, int hsel,
                               bool cfl);

// kernels_diag.hip (run diagnostics: 12 FP64 values, see k_diag_partial)
constexpr int kDiagVals = 12;
int diag_blocks(const Box &b);
void launch_diag(hipStream_t s, const Box &b, FieldPtrs W, double *part, double *out);

// kernels_adjoint.hip (reverse of one RK stage of the periodic step)
struct AdjStageArgs {
  FieldPtrs Win, Wbase, Wout;  // stage input, conserved base, stage output
  FieldPtrs lam_out;           // dL/dW_out
  FieldPtrs lam_base;          // dL/dW_base (written, or added to with acc_base)
  FieldPtrs lam_in;            // dL/dW_in through P (written, or added to with acc_in)
  float h;
  bool acc_base, acc_in;
  // work arrays (ncell floats each); face: 15 per direction, the adjoints
  // of a face's left and right states (7 + 7) and its limiter branches
  // (one 32-bit word per face)
  float *gU[5], *lE[3], *lbb[3], *abn[3], *face[45];
};
void launch_adj_set_dt(hipStream_t s, StepState *st, float dt);
void launch_adj_stage(hipStream_t s, const Box &b, const AdjStageArgs &A);

// kernels_tiled.hip
int tiled_zchunk(const Box &b, int tx, int ty, bool faces);
// tile: TileVariant of the z-marching tiles
void launch_faces(hipStream_t s, const Box &b, FieldPtrs W, const FluxSet &F, const StepState *st,
                  int hsel, int tile = TV_32x8);
void launch_update_tiled(hipStream_t s, const Box &b, FieldPtrs Win, const FluxSet &F,
                         FieldPtrs Wbase, FieldPtrs Wout, StepState *st, int hsel, bool pace_only,
                         bool cfl, int tile = TV_64x4);

}  // namespace nmhd
