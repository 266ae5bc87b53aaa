// kernels_tiled.hip -- the production kernel path for one box: two kernels
// per RK stage, each a grid of z-marching column tiles (the tile functions
// are in tiled_tiles.h).
//
//  k_faces         all three face fluxes of every cell (21 floats per cell)
//  k_update_march  divergence, CT, conserved update, primitive recovery, the
//                  shearing-box kick and (stage 2) the CFL minimum of the
//                  new state, one atomic per workgroup
//
// Tiles are mapped to workgroups XCD-aware (xcd_logical_block) so that
// neighbouring tiles, which re-read each other's halos, share an L2.
#include "tiled_tiles.h"

namespace nmhd {
namespace tiled {

// Recon = b.recon (the launcher picks the instantiation)
template <int TX, int TY, int Recon>
__global__ void __launch_bounds__(kTileThreads, (FacesShape<TX, TY>::kWavesPerSimd))
    k_faces(Box b, FieldPtrs W, FluxSet F,
                                                        const StepState *st, int hsel, int zc,
                                                        int gx, int gy) {
  if (!st->active) return;
  __shared__ float lds[FacesShape<TX, TY>::kLds];
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int x0 = (lb % gx) * TX, y0 = ((lb / gx) % gy) * TY, z0 = (lb / gx / gy) * zc;
  faces_tile<TX, TY, Recon>(b, W, F, stage_h(st, hsel), x0, y0, z0, min(z0 + zc, b.n[2]), lds);
}

template <int TX, int TY, bool PaceOnly, bool Cfl>
__global__ void __launch_bounds__(kTileThreads) k_update_march(Box b, FieldPtrs Win, FluxSet F,
                                                               FieldPtrs Wb, FieldPtrs Wo,
                                                               StepState *st, int hsel, int zc,
                                                               int gx, int gy) {
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int x0 = (lb % gx) * TX, y0 = ((lb / gx) % gy) * TY, z0 = (lb / gx / gy) * zc;
  if (!st->active) {
    // a step enqueued after the run reached its goal: the last stage's
    // output carries the step's base state through the buffer rotation
    if (Cfl && !PaceOnly) {
      const int x = x0 + (int)threadIdx.x % TX, y = y0 + (int)threadIdx.x / TX;
      const int zend = min(z0 + zc, b.n[2]);
      if (x < b.n[0] && y < b.n[1])
        for (int r = z0; r < zend; ++r) {
          const long a = ((long)r * b.n[1] + y) * b.n[0] + x;
          for (int f = 0; f < 8; ++f) Wo.f[f][a] = Wb.f[f][a];
        }
    }
    return;
  }
  __shared__ float lds[UpdateShape<TX, TY>::kLds];
  __shared__ float red[4];
  float cfl_t = 3.0e38f;
  update_tile<TX, TY, PaceOnly, Cfl>(b, Win, F, Wb, Wo, stage_h(st, hsel), x0, y0, z0,
                                     min(z0 + zc, b.n[2]), lds, cfl_t);
  if (Cfl && !PaceOnly) block_min_into(cfl_t, &st->cfl_next_bits, red);
}

}  // namespace tiled

// ------------------------------------------------------------- launchers

int tiled_zchunk(const Box &b, int tx, int ty, bool faces) {
  // Planes per workgroup march.  Measured on MI355X (256^3 and 512^3): the
  // update is fastest with 16-plane chunks at both sizes; the face march
  // with 32 planes once that gives at least 4096 workgroups (4 per CU
  // resident: enough rounds that the last one is short), else 16.  Smaller
  // chunks balance the grid better but repeat more prologue planes.  Small
  // boxes halve the chunk until the grid fills the chip: one full round of
  // 4 workgroups per CU (1024) for marches of 8 or more planes, half of it
  // for shorter ones, whose prologue planes weigh more.  Measured step
  // times (periodic cubes, both kernels at the same chunk): 32^3 best at 1
  // plane (2.7x faster than 8), 48^3 at 1, 64^3 at 2 (1.8x faster than 8;
  // 1 plane is 6% slower), 96^3 at 4, 128^3 at 8 (16 is 12% slower).
  const long cols = (long)((b.n[0] + tx - 1) / tx) * ((b.n[1] + ty - 1) / ty);
  auto groups = [&](int zc) { return cols * ((b.n[2] + zc - 1) / zc); };
  int zc = faces ? (groups(32) >= 4096 ? 32 : 16) : 16;
  while (zc > 1 && groups(zc) < (zc >= 8 ? 1024 : 512)) zc /= 2;
  return zc;
}

namespace {

template <int TX, int TY>
void faces_as(hipStream_t s, const Box &b, FieldPtrs W, const FluxSet &F, const StepState *st,
              int hsel) {
  const int zc = tiled_zchunk(b, TX, TY, true);
  const int gx = (b.n[0] + TX - 1) / TX, gy = (b.n[1] + TY - 1) / TY, gz = (b.n[2] + zc - 1) / zc;
  if (b.recon == 1)
    hipLaunchKernelGGL((tiled::k_faces<TX, TY, 1>), dim3(gx * gy * gz), dim3(tiled::kTileThreads),
                       0, s, b, W, F, st, hsel, zc, gx, gy);
  else
    hipLaunchKernelGGL((tiled::k_faces<TX, TY, 0>), dim3(gx * gy * gz), dim3(tiled::kTileThreads),
                       0, s, b, W, F, st, hsel, zc, gx, gy);
}

template <int TX, int TY>
void update_as(hipStream_t s, const Box &b, FieldPtrs Win, const FluxSet &F, FieldPtrs Wbase,
               FieldPtrs Wout, StepState *st, int hsel, bool pace_only, bool cfl) {
  const int zc = tiled_zchunk(b, TX, TY, false);
  const int gx = (b.n[0] + TX - 1) / TX, gy = (b.n[1] + TY - 1) / TY, gz = (b.n[2] + zc - 1) / zc;
  const dim3 grid(gx * gy * gz), blk(tiled::kTileThreads);
  if (pace_only)
    hipLaunchKernelGGL((tiled::k_update_march<TX, TY, true, false>), grid, blk, 0, s, b, Win, F,
                       Wbase, Wout, st, hsel, zc, gx, gy);
  else if (cfl)
    hipLaunchKernelGGL((tiled::k_update_march<TX, TY, false, true>), grid, blk, 0, s, b, Win, F,
                       Wbase, Wout, st, hsel, zc, gx, gy);
  else
    hipLaunchKernelGGL((tiled::k_update_march<TX, TY, false, false>), grid, blk, 0, s, b, Win, F,
                       Wbase, Wout, st, hsel, zc, gx, gy);
}

}  // namespace

void launch_faces(hipStream_t s, const Box &b, FieldPtrs W, const FluxSet &F, const StepState *st,
                  int hsel, int tile) {
  if (tile == TV_64x4)
    faces_as<64, 4>(s, b, W, F, st, hsel);
  else if (tile == TV_16x16)
    faces_as<16, 16>(s, b, W, F, st, hsel);
  else
    faces_as<32, 8>(s, b, W, F, st, hsel);
}

void launch_update_tiled(hipStream_t s, const Box &b, FieldPtrs Win, const FluxSet &F,
                         FieldPtrs Wbase, FieldPtrs Wout, StepState *st, int hsel, bool pace_only,
                         bool cfl, int tile) {
  if (tile == TV_64x4)
    update_as<64, 4>(s, b, Win, F, Wbase, Wout, st, hsel, pace_only, cfl);
  else if (tile == TV_16x16)
    update_as<16, 16>(s, b, Win, F, Wbase, Wout, st, hsel, pace_only, cfl);
  else
    update_as<32, 8>(s, b, Win, F, Wbase, Wout, st, hsel, pace_only, cfl);
}

}  // namespace nmhd
