// kernels_batch.hip -- many independent boxes of uneven sizes
// advanced together.  Each RK stage runs the face fluxes, then the update,
// over a work table of tiles from every box, either as one workgroup per
// work item with one launch per tile shape (the kernel then needs only that
// shape's registers and keeps the single-box occupancy), or as two
// persistent launches whose workgroups walk the whole table; a tile's box
// record gives its sizes, its per-box step (own dt) and the offsets of its
// fields in the batch's pools.  The tile functions are the
// single-box path's (tiled_tiles.h), so a box advanced in a batch is
// bit-identical to the same box advanced alone.  The bookkeeping kernels run
// one thread per box on the per-box step records; boxes with ghost layers
// (the shearing box) are filled, and their ghost cells' CFL reduced, by
// per-box grids over the single-box cell functions (ghost_fill.h).
//
// The reference advances one compile-time-sized box per process and has no
// counterpart of this path; its per-box arithmetic is that of proceed()
// (library/mhd_solver.inl:1086-1204).
#include "ghost_fill.h"
#include "tiled_tiles.h"

namespace nmhd {

namespace {

constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int kFacesLds = cmax(cmax(tiled::FacesShape<64, 4>::kLds, tiled::FacesShape<32, 8>::kLds),
                               tiled::FacesShape<16, 16>::kLds);
constexpr int kUpdateLds =
    cmax(cmax(tiled::UpdateShape<64, 4>::kLds, tiled::UpdateShape<32, 8>::kLds),
         tiled::UpdateShape<16, 16>::kLds);

__device__ __forceinline__ void item_tile(const WorkItem &w, int &x0, int &y0, int &z0,
                                          int &zend) {
  x0 = w.xy & 0xffff;
  y0 = (w.xy >> 16) & 0xffff;
  z0 = w.z & 0xffff;
  zend = (w.z >> 16) & 0xffff;
}

__global__ void __launch_bounds__(tiled::kTileThreads)
    k_batch_faces(const BatchBox *boxes, const WorkItem *items, int nitems, float *state,
                  float *flux, const StepState *st, int hsel) {
  __shared__ float lds[kFacesLds];
  const int G = gridDim.x;
  for (int it = xcd_logical_block(blockIdx.x, G); it < nitems; it += G) {
    const WorkItem w = items[it];
    const StepState *s = st + w.box;
    if (!s->active) continue;
    const BatchBox bb = boxes[w.box];
    const FieldPtrs W = box_fields(state, bb);
    const FluxSet F = box_fluxes(flux, bb);
    const float h = tiled::stage_h(s, hsel);
    int x0, y0, z0, zend;
    item_tile(w, x0, y0, z0, zend);
    switch (w.variant) {
      case TV_64x4:
        tiled::faces_tile_any<64, 4>(bb.b, W, F, h, x0, y0, z0, zend, lds);
        break;
      case TV_32x8:
        tiled::faces_tile_any<32, 8>(bb.b, W, F, h, x0, y0, z0, zend, lds);
        break;
      default:
        tiled::faces_tile_any<16, 16>(bb.b, W, F, h, x0, y0, z0, zend, lds);
        break;
    }
  }
}

template <bool Cfl>
__global__ void __launch_bounds__(tiled::kTileThreads)
    k_batch_update(const BatchBox *boxes, const WorkItem *items, int nitems, float *state_in,
                   float *flux, float *state_base, float *state_out, StepState *st, int hsel) {
  __shared__ float lds[kUpdateLds];
  __shared__ float red[4];
  const int G = gridDim.x;
  for (int it = xcd_logical_block(blockIdx.x, G); it < nitems; it += G) {
    const WorkItem w = items[it];
    StepState *s = st + w.box;
    if (!s->active) continue;
    const BatchBox bb = boxes[w.box];
    const FieldPtrs Win = box_fields(state_in, bb);
    const FieldPtrs Wb = box_fields(state_base, bb);
    const FieldPtrs Wo = box_fields(state_out, bb);
    const FluxSet F = box_fluxes(flux, bb);
    const float h = tiled::stage_h(s, hsel);
    int x0, y0, z0, zend;
    item_tile(w, x0, y0, z0, zend);
    float cfl_t = 3.0e38f;
    switch (w.variant) {
      case TV_64x4:
        tiled::update_tile<64, 4, false, Cfl>(bb.b, Win, F, Wb, Wo, h, x0, y0, z0, zend, lds,
                                              cfl_t);
        break;
      case TV_32x8:
        tiled::update_tile<32, 8, false, Cfl>(bb.b, Win, F, Wb, Wo, h, x0, y0, z0, zend, lds,
                                              cfl_t);
        break;
      default:
        tiled::update_tile<16, 16, false, Cfl>(bb.b, Win, F, Wb, Wo, h, x0, y0, z0, zend, lds,
                                               cfl_t);
        break;
    }
    if (Cfl) tiled::block_min_into(cfl_t, &s->cfl_next_bits, red);
  }
}

// One workgroup per work item of one tile shape.
template <int TX, int TY>
__global__ void __launch_bounds__(tiled::kTileThreads, (tiled::FacesShape<TX, TY>::kWavesPerSimd))
    k_batch_faces_items(const BatchBox *boxes, const WorkItem *items, int nitems, float *state,
                        float *flux, const StepState *st, int hsel) {
  __shared__ float lds[tiled::FacesShape<TX, TY>::kLds];
  const int it = xcd_logical_block(blockIdx.x, gridDim.x);
  if (it >= nitems) return;
  const WorkItem w = items[it];
  const StepState *s = st + w.box;
  if (!s->active) return;
  const BatchBox bb = boxes[w.box];
  int x0, y0, z0, zend;
  item_tile(w, x0, y0, z0, zend);
  tiled::faces_tile_any<TX, TY>(bb.b, box_fields(state, bb), box_fluxes(flux, bb),
                            tiled::stage_h(s, hsel), x0, y0, z0, zend, lds);
}

template <int TX, int TY, bool Cfl>
__global__ void __launch_bounds__(tiled::kTileThreads)
    k_batch_update_items(const BatchBox *boxes, const WorkItem *items, int nitems,
                         float *state_in, float *flux, float *state_base, float *state_out,
                         StepState *st, int hsel) {
  __shared__ float lds[tiled::UpdateShape<TX, TY>::kLds];
  __shared__ float red[4];
  const int it = xcd_logical_block(blockIdx.x, gridDim.x);
  if (it >= nitems) return;
  const WorkItem w = items[it];
  StepState *s = st + w.box;
  if (!s->active) return;
  const BatchBox bb = boxes[w.box];
  int x0, y0, z0, zend;
  item_tile(w, x0, y0, z0, zend);
  float cfl_t = 3.0e38f;
  tiled::update_tile<TX, TY, false, Cfl>(bb.b, box_fields(state_in, bb), box_fluxes(flux, bb),
                                         box_fields(state_base, bb), box_fields(state_out, bb),
                                         tiled::stage_h(s, hsel), x0, y0, z0, zend, lds, cfl_t);
  if (Cfl) tiled::block_min_into(cfl_t, &s->cfl_next_bits, red);
}

__global__ void k_batch_begin(StepState *st, int nbox, int full) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbox) step_begin(st[i], full);
}

// goal is read from device memory so that a captured step serves any goal.
__global__ void k_batch_dt(StepState *st, int nbox, const float *goal) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbox) step_dt(st[i], *goal);
}

__global__ void k_batch_commit(StepState *st, int nbox) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbox) step_commit(st[i]);
}

// Full CFL pass: blockIdx.y is the box, blocks stride over its cells.
__global__ void __launch_bounds__(256) k_batch_cfl(const BatchBox *boxes, float *state,
                                                   float goal, StepState *st) {
  const int box = blockIdx.y;
  if (st[box].finished) return;
  __shared__ float red[4];
  const BatchBox bb = boxes[box];
  const FieldPtrs W = box_fields(state, bb);
  const Box &b = bb.b;
  float t = goal;
  for (long a = (long)blockIdx.x * blockDim.x + threadIdx.x; a < b.ncell;
       a += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    cell_coords(b, a, x, y, z);
    t = fminf(t, cell_cfl(b, W, a, x, y, z));
  }
  const float m = tiled::wave_min64(t);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float r = fminf(fminf(red[0], red[1]), fminf(red[2], red[3]));
    atomicMin(&st[box].cfl_bits, __float_as_uint(r));
  }
}

// Ghost fills of the boxes that have ghost layers (shearing-periodic or
// periodic Y rows, periodic Z planes): blockIdx.y is the box, blocks stride
// over its slab; the cell operations are the single-box kernels'.
__global__ void k_batch_ghost_y(const BatchBox *boxes, float *state, const StepState *st) {
  const int box = blockIdx.y;
  const BatchBox bb = boxes[box];
  if (bb.b.m[1] == 0 || st[box].finished) return;
  const FieldPtrs W = box_fields(state, bb);
  const long unit = (long)bb.b.n[0] * bb.b.m[1] * bb.b.n[2];
  const double t = st[box].t;
  for (long a = (long)blockIdx.x * blockDim.x + threadIdx.x; a < unit;
       a += (long)gridDim.x * blockDim.x)
    ghost_y_cell(bb.b, W, t, a);
}

__global__ void k_batch_ghost_z(const BatchBox *boxes, float *state, const StepState *st) {
  const int box = blockIdx.y;
  const BatchBox bb = boxes[box];
  if (bb.b.m[2] == 0 || st[box].finished) return;
  const FieldPtrs W = box_fields(state, bb);
  const long unit = (long)bb.b.n[0] * bb.b.n[1] * bb.b.m[2];
  for (long a = (long)blockIdx.x * blockDim.x + threadIdx.x; a < unit;
       a += (long)gridDim.x * blockDim.x)
    ghost_z_cell(bb.b, W, a);
}

// CFL minimum over the cells the ghost fills rewrote, added to each box's
// running minimum (the stage-2 epilogue reduced the others).
__global__ void __launch_bounds__(256) k_batch_cfl_ghosts(const BatchBox *boxes, float *state,
                                                          StepState *st) {
  const int box = blockIdx.y;
  const BatchBox bb = boxes[box];
  const long n0 = ghost_cfl_count(bb.b, 0), n1 = ghost_cfl_count(bb.b, 1);
  if ((n0 == 0 && n1 == 0) || st[box].finished) return;
  __shared__ float red[4];
  const FieldPtrs W = box_fields(state, bb);
  float t = 3.0e38f;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n0;
       i += (long)gridDim.x * blockDim.x)
    t = fminf(t, ghost_cfl_cell(bb.b, W, 0, i));
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n1;
       i += (long)gridDim.x * blockDim.x)
    t = fminf(t, ghost_cfl_cell(bb.b, W, 1, i));
  const float m = tiled::wave_min64(t);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float r = fminf(fminf(red[0], red[1]), fminf(red[2], red[3]));
    atomicMin(&st[box].cfl_bits, __float_as_uint(r));
  }
}

// Carry the state of boxes whose step is inactive (goal reached) into the
// next rotation slot.
__global__ void k_batch_copy(const BatchBox *boxes, float *src, float *dst, const StepState *st) {
  const int box = blockIdx.y;
  if (st[box].active) return;
  const BatchBox bb = boxes[box];
  const long n = 8 * bb.b.ncell;
  for (long a = (long)blockIdx.x * blockDim.x + threadIdx.x; a < n;
       a += (long)gridDim.x * blockDim.x)
    dst[bb.off8 + a] = src[bb.off8 + a];
}

}  // namespace

void launch_batch_begin(hipStream_t s, StepState *st, int nbox, bool full) {
  hipLaunchKernelGGL(k_batch_begin, dim3(grid_for(nbox, 64)), dim3(64), 0, s, st, nbox,
                     (int)full);
}
void launch_batch_cfl(hipStream_t s, const BatchBox *boxes, int nbox, float *state, float goal,
                      StepState *st) {
  hipLaunchKernelGGL(k_batch_cfl, dim3(64, nbox), dim3(256), 0, s, boxes, state, goal, st);
}
void launch_batch_dt(hipStream_t s, StepState *st, int nbox, const float *goal) {
  hipLaunchKernelGGL(k_batch_dt, dim3(grid_for(nbox, 64)), dim3(64), 0, s, st, nbox, goal);
}
void launch_batch_commit(hipStream_t s, StepState *st, int nbox) {
  hipLaunchKernelGGL(k_batch_commit, dim3(grid_for(nbox, 64)), dim3(64), 0, s, st, nbox);
}
void launch_batch_ghost_y(hipStream_t s, const BatchBox *boxes, int nbox, int grid, float *state,
                          const StepState *st) {
  hipLaunchKernelGGL(k_batch_ghost_y, dim3(grid, nbox), dim3(256), 0, s, boxes, state, st);
}
void launch_batch_ghost_z(hipStream_t s, const BatchBox *boxes, int nbox, int grid, float *state,
                          const StepState *st) {
  hipLaunchKernelGGL(k_batch_ghost_z, dim3(grid, nbox), dim3(256), 0, s, boxes, state, st);
}
void launch_batch_cfl_ghosts(hipStream_t s, const BatchBox *boxes, int nbox, int grid,
                             float *state, StepState *st) {
  hipLaunchKernelGGL(k_batch_cfl_ghosts, dim3(grid, nbox), dim3(256), 0, s, boxes, state, st);
}
void launch_batch_copy(hipStream_t s, const BatchBox *boxes, int nbox, float *src, float *dst,
                       const StepState *st) {
  hipLaunchKernelGGL(k_batch_copy, dim3(32, nbox), dim3(256), 0, s, boxes, src, dst, st);
}
void launch_batch_faces(hipStream_t s, int grid, const BatchBox *boxes, const WorkItem *items,
                        int nitems, float *state, float *flux, const StepState *st, int hsel) {
  hipLaunchKernelGGL(k_batch_faces, dim3(grid), dim3(tiled::kTileThreads), 0, s, boxes, items,
                     nitems, state, flux, st, hsel);
}
void launch_batch_faces_items(hipStream_t s, int variant, const BatchBox *boxes,
                              const WorkItem *items, int nitems, float *state, float *flux,
                              const StepState *st, int hsel) {
  if (nitems <= 0) return;
  const dim3 grid(nitems), blk(tiled::kTileThreads);
  if (variant == TV_64x4)
    hipLaunchKernelGGL((k_batch_faces_items<64, 4>), grid, blk, 0, s, boxes, items, nitems, state,
                       flux, st, hsel);
  else if (variant == TV_32x8)
    hipLaunchKernelGGL((k_batch_faces_items<32, 8>), grid, blk, 0, s, boxes, items, nitems, state,
                       flux, st, hsel);
  else
    hipLaunchKernelGGL((k_batch_faces_items<16, 16>), grid, blk, 0, s, boxes, items, nitems,
                       state, flux, st, hsel);
}

template <bool Cfl>
void update_items_as(hipStream_t s, int variant, const BatchBox *boxes, const WorkItem *items,
                     int nitems, float *state_in, float *flux, float *state_base,
                     float *state_out, StepState *st, int hsel) {
  const dim3 grid(nitems), blk(tiled::kTileThreads);
  if (variant == TV_64x4)
    hipLaunchKernelGGL((k_batch_update_items<64, 4, Cfl>), grid, blk, 0, s, boxes, items, nitems,
                       state_in, flux, state_base, state_out, st, hsel);
  else if (variant == TV_32x8)
    hipLaunchKernelGGL((k_batch_update_items<32, 8, Cfl>), grid, blk, 0, s, boxes, items, nitems,
                       state_in, flux, state_base, state_out, st, hsel);
  else
    hipLaunchKernelGGL((k_batch_update_items<16, 16, Cfl>), grid, blk, 0, s, boxes, items,
                       nitems, state_in, flux, state_base, state_out, st, hsel);
}

void launch_batch_update_items(hipStream_t s, int variant, const BatchBox *boxes,
                               const WorkItem *items, int nitems, float *state_in, float *flux,
                               float *state_base, float *state_out, StepState *st, int hsel,
                               bool cfl) {
  if (nitems <= 0) return;
  if (cfl)
    update_items_as<true>(s, variant, boxes, items, nitems, state_in, flux, state_base,
                          state_out, st, hsel);
  else
    update_items_as<false>(s, variant, boxes, items, nitems, state_in, flux, state_base,
                           state_out, st, hsel);
}

void launch_batch_update(hipStream_t s, int grid, const BatchBox *boxes, const WorkItem *items,
                         int nitems, float *state_in, float *flux, float *state_base,
                         float *state_out, StepState *st, int hsel, bool cfl) {
  if (cfl)
    hipLaunchKernelGGL(k_batch_update<true>, dim3(grid), dim3(tiled::kTileThreads), 0, s, boxes,
                       items, nitems, state_in, flux, state_base, state_out, st, hsel);
  else
    hipLaunchKernelGGL(k_batch_update<false>, dim3(grid), dim3(tiled::kTileThreads), 0, s, boxes,
                       items, nitems, state_in, flux, state_base, state_out, st, hsel);
}

}  // namespace nmhd
