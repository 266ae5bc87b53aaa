// kernels_common.hip -- kernels shared by every path: step bookkeeping on the
// device, the shearing-periodic ghost fill, the periodic Z ghost copy, the
// CFL reduction and the HLLD diagnostic.
//
// Reference correspondence (INL = library/mhd_solver.inl):
//   k_ghost_y   communicate_gather_kernel_y + communicate_scatter_kernel_y
//               (INL:90-286) with relative_displacement (INL:288-295)
//   k_ghost_z   the Z halo copies of communicate() (INL:446-708)
//   k_cfl       cfl_condition_kernel (INL:714-739) + thrust::min_element
//               (INL:1111-1116), as one pass with a wave/block/grid min
//   k_dt        dt clipping of proceed() (INL:1126-1137)
//   k_commit    GlobalSetup::proceed (mhd_solver.hrb:56-59)
#include "ghost_fill.h"

namespace nmhd {

__global__ void k_begin_step(StepState *st, int full) { step_begin(*st, full); }

// The bookkeeping between two steps enqueued back to back, in one launch:
// commit the previous step (when commit_prev), begin the next one and, when
// its CFL minimum needs no pass over cells (dt_too), decide its dt.
__global__ void k_open_step(StepState *st, int commit_prev, int full, int dt_too, float goal) {
  if (commit_prev) step_commit(*st);
  step_begin(*st, full);
  if (dt_too) step_dt(*st, goal);
}

void host_displacement(const Box &b, double t, float *normalized, int *n, float *minor) {
  displacement(b, t, normalized, n, minor);
}

// Y ghost rows (ghost_y_cell), one thread per (x, row, z) of the slab.
__global__ void k_ghost_y(Box b, FieldPtrs W, const StepState *st) {
  if (st->finished) return;
  const long unit = (long)b.n[0] * b.m[1] * b.n[2];
  const long a = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= unit) return;
  ghost_y_cell(b, W, st->t, a);
}

// Z ghost planes (ghost_z_cell).
__global__ void k_ghost_z(Box b, FieldPtrs W, const StepState *st) {
  if (st->finished) return;
  const long unit = (long)b.n[0] * b.n[1] * b.m[2];
  const long a = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= unit) return;
  ghost_z_cell(b, W, a);
}

__device__ __forceinline__ float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}

// Per-cell CFL time min(goal, C*d/|v_d + sqrt((gamma p + b2)/rho)|) with
// b2 = sum_d max(B_d,lo^2, B_d,hi^2), reduced to one float with a wave
// shuffle min, an LDS block min and one atomic min per block on the float
// bits (valid: all candidates are >= 0).
__device__ __forceinline__ void cfl_block_fold(float t, StepState *st) {
  __shared__ float red[4];
  t = wave_min(t);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = fminf(m, red[i]);
    atomicMin(&st->cfl_bits, __float_as_uint(m));
  }
}

// Grid-stride over the cells with a capped grid: the per-block atomics on the
// one result word serialise, so a grid of one block per 256 cells would
// spend most of its time in them.
__global__ void __launch_bounds__(256) k_cfl(Box b, FieldPtrs W, float goal, StepState *st) {
  if (st->finished) return;
  float t = goal;
  for (long a = (long)blockIdx.x * blockDim.x + threadIdx.x; a < b.ncell;
       a += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    cell_coords(b, a, x, y, z);
    t = fminf(t, cell_cfl(b, W, a, x, y, z));
  }
  cfl_block_fold(t, st);
}

// CFL over the cells a ghost fill rewrites (ghost_cfl_cell), slab 0 or 1.
__global__ void __launch_bounds__(256) k_cfl_ghosts(Box b, FieldPtrs W, int slab, StepState *st) {
  if (st->finished) return;
  const long n = ghost_cfl_count(b, slab);
  float t = 3.0e38f;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    t = fminf(t, ghost_cfl_cell(b, W, slab, i));
  cfl_block_fold(t, st);
}

__global__ void k_dt(StepState *st, float goal) { step_dt(*st, goal); }

__global__ void k_commit(StepState *st) { step_commit(*st); }

__global__ void k_hlld_batch(long n, float gamma, FluxIn in, FluxOut out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float inv_gm1 = (float)(1.0 / (double)(gamma - 1.0f));
  Flux7 F = hlld(gamma, inv_gm1, in.f[0][i], in.f[1][i], in.f[2][i], in.f[3][i], in.f[4][i],
                 in.f[5][i], in.f[6][i], in.f[7][i], in.f[8][i], in.f[9][i], in.f[10][i],
                 in.f[11][i], in.f[12][i], in.f[13][i], in.f[14][i]);
  out.f[0][i] = F.rho;
  out.f[1][i] = F.mn;
  out.f[2][i] = F.mt1;
  out.f[3][i] = F.mt2;
  out.f[4][i] = F.e;
  out.f[5][i] = F.bt1;
  out.f[6][i] = F.bt2;
}

// Copy the state (used when an enqueued step is inactive: the rotation of
// the state buffers still has to carry the state forward).  A small grid
// strides over the box, so that the usual case (the step was active and the
// kernel returns at once) costs one short launch.
__global__ void k_copy_state(Box b, FieldPtrs src, FieldPtrs dst, const StepState *st) {
  if (st->active) return;
  for (long a = (long)blockIdx.x * blockDim.x + threadIdx.x; a < b.ncell;
       a += (long)gridDim.x * blockDim.x)
    for (int f = 0; f < 8; ++f) dst.f[f][a] = src.f[f][a];
}

// Diagnostic entry points run on their own step record: current time of the
// run, active, not finished, explicit step h.
__global__ void k_diag_prepare(StepState *diag, const StepState *st, float h) {
  diag->t = st->t;
  diag->t_next = st->t;
  diag->dt = h;
  diag->h = h;
  diag->h_diag = h;
  diag->cfl_bits = 0x7f800000u;
  diag->finished = 0;
  diag->active = 1;
  diag->generation = st->generation;
  diag->steps_done = 0;
  diag->goal_reached = 0;
  diag->cfl_next_bits = 0x7f800000u;
}

// ------------------------------------------------------------- launchers

void launch_begin_step(hipStream_t s, StepState *st, bool full) {
  hipLaunchKernelGGL(k_begin_step, dim3(1), dim3(1), 0, s, st, (int)full);
}
void launch_open_step(hipStream_t s, StepState *st, bool commit_prev, bool full, bool dt_too,
                      float goal) {
  hipLaunchKernelGGL(k_open_step, dim3(1), dim3(1), 0, s, st, (int)commit_prev, (int)full,
                     (int)dt_too, goal);
}
// Blocks of a reduction kernel: one per 1024 cells, at most 512.
unsigned reduce_grid(long n) {
  const unsigned g = grid_for(n, 4 * kBlock);
  return g < 1 ? 1 : (g > 512 ? 512 : g);
}

void launch_cfl_ghosts(hipStream_t s, const Box &b, FieldPtrs W, StepState *st) {
  if (b.m[1] > 0) {
    const long n = (long)b.n[0] * 2 * b.m[1] * b.n[2];
    hipLaunchKernelGGL(k_cfl_ghosts, dim3(reduce_grid(n)), dim3(kBlock), 0, s, b, W, 0, st);
  }
  if (b.m[2] > 0) {
    const long n = (long)b.n[0] * b.n[1] * (2 * b.m[2] + 1);
    hipLaunchKernelGGL(k_cfl_ghosts, dim3(reduce_grid(n)), dim3(kBlock), 0, s, b, W, 1, st);
  }
}
void launch_ghost_y(hipStream_t s, const Box &b, FieldPtrs W, const StepState *st) {
  const long unit = (long)b.n[0] * b.m[1] * b.n[2];
  hipLaunchKernelGGL(k_ghost_y, dim3(grid_for(unit)), dim3(kBlock), 0, s, b, W, st);
}
void launch_ghost_z(hipStream_t s, const Box &b, FieldPtrs W, const StepState *st) {
  const long unit = (long)b.n[0] * b.n[1] * b.m[2];
  hipLaunchKernelGGL(k_ghost_z, dim3(grid_for(unit)), dim3(kBlock), 0, s, b, W, st);
}
void launch_cfl(hipStream_t s, const Box &b, FieldPtrs W, float goal, StepState *st) {
  hipLaunchKernelGGL(k_cfl, dim3(reduce_grid(b.ncell)), dim3(kBlock), 0, s, b, W, goal, st);
}
void launch_dt(hipStream_t s, StepState *st, float goal) {
  hipLaunchKernelGGL(k_dt, dim3(1), dim3(1), 0, s, st, goal);
}
void launch_commit(hipStream_t s, StepState *st) {
  hipLaunchKernelGGL(k_commit, dim3(1), dim3(1), 0, s, st);
}
void launch_copy_state(hipStream_t s, const Box &b, FieldPtrs src, FieldPtrs dst,
                       const StepState *st) {
  const unsigned g = grid_for(b.ncell);
  hipLaunchKernelGGL(k_copy_state, dim3(g < 2048 ? g : 2048), dim3(kBlock), 0, s, b, src, dst,
                     st);
}
void launch_hlld_batch(hipStream_t s, long n, float gamma, bool strict, FluxIn in,
                       FluxOut out) {
  if (strict)
    launch_hlld_batch_strict(s, n, gamma, in, out);
  else
    hipLaunchKernelGGL(k_hlld_batch, dim3(grid_for(n)), dim3(kBlock), 0, s, n, gamma, in, out);
}
void launch_diag_prepare(hipStream_t s, StepState *diag, const StepState *st, float h) {
  hipLaunchKernelGGL(k_diag_prepare, dim3(1), dim3(1), 0, s, diag, st, h);
}

}  // namespace nmhd
