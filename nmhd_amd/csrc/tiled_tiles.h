// tiled_tiles.h -- the two stage kernels of the production path as tile
// functions, shared by the single-box launches (kernels_tiled.hip) and the
// persistent multi-box launches (kernels_batch.hip).
//
//  faces_tile   all three face fluxes of every cell of a TX x TY column tile
//               over planes [z0, zend), one HLLD per face and no halo
//               recomputation.  The x and y stencils (cells i-2..i+1) are
//               read from an LDS ring of two planes that the workgroup stages
//               cooperatively; the z stencil (planes k-2..k+1) lives in a
//               per-thread register queue filled as the march passes each
//               plane.  Writes the seven face fluxes of each direction.
//  update_tile  per cell of the tile: flux divergence, the 24 upwinded edge
//               EMFs of the cell's six faces, the conserved update, primitive
//               recovery and the shearing-box kick, as a z march with the
//               neighbour data staged in LDS rings one plane ahead; in stage
//               2 it also returns the CFL time of the new state of the
//               thread's cells (interior cells only).
//
// Both run on 256-thread workgroups (TX * TY == 256, 1-D thread index) and
// take their LDS as a pointer so that one persistent kernel can run tiles of
// several shapes out of one buffer.  The arithmetic of a cell does not depend
// on the tile shape or the z chunk, so every tiling gives bit-identical
// results.
//
// Reference correspondence (INL = library/mhd_solver.inl): calc_flux_kernel
// (INL:760-816) for the three sweeps -> faces_tile; add_flux_kernel x3
// (INL:851-897), corner_magnet_flux (INL:819-840), to_conserved
// (INL:990-1007), add_pace / update (INL:1029-1071), to_primitive
// (INL:1009-1027), external_acceleration_kernel (INL:1073-1084) and
// cfl_condition_kernel (INL:714-739) -> update_tile.
#pragma once

#include "nmhd_kernels.h"

namespace nmhd {
namespace tiled {

__device__ __forceinline__ float stage_h(const StepState *st, int hsel) {
  return hsel == H_HALF ? st->h : (hsel == H_FULL ? st->dt : st->h_diag);
}

// true modulo for staging coordinates, which may lie several box lengths
// outside [0, n) when a tile is wider than a small box
__device__ __forceinline__ int pmod(int i, int n) {
  const int r = i % n;
  return r < 0 ? r + n : r;
}

__device__ __forceinline__ void recon(int mode, float x0, float x1, float x2, float x3, float &l,
                                      float &r) {
  if (mode == 1) {
    l = x1;
    r = x2;
  } else {
    recon_minbee2(x0, x1, x2, x3, l, r);
  }
}

__device__ __forceinline__ float upwind(float s, float f0, float f1) {
  return s > 0.0f ? f0 : (s < 0.0f ? f1 : 0.5f * (f0 + f1));
}

__device__ __forceinline__ float wave_min64(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}

// Minimum of v over the 256-thread workgroup, folded into *target (float
// bits, all candidates >= 0) by one atomic.  red: 4 floats of LDS.
__device__ __forceinline__ void block_min_into(float v, unsigned *target, float *red) {
  const float m = wave_min64(v);
  const int tid = threadIdx.x;
  if ((tid & 63) == 0) red[tid >> 6] = m;
  __syncthreads();
  if (tid == 0) {
    const float r = fminf(fminf(red[0], red[1]), fminf(red[2], red[3]));
    atomicMin(target, __float_as_uint(r));
  }
  __syncthreads();
}

constexpr int kTileThreads = 256;

template <int TX, int TY>
struct FacesShape {
  static_assert(TX * TY == kTileThreads, "tiles have 256 threads");
  static constexpr int LX = TX + 4, LY = TY + 4, LN = LX * LY;
  static constexpr int NS = (LN + kTileThreads - 1) / kTileThreads;  // slots per thread
  static constexpr int kLds = 2 * 8 * LN;                              // floats
};

template <int TX, int TY>
__device__ __forceinline__ void faces_tile(const Box &b, const FieldPtrs &W, const FluxSet &F,
                                           float h, int x0, int y0, int z0, int zend,
                                           float *lds) {
  using S = FacesShape<TX, TY>;
  constexpr int NT = kTileThreads, LX = S::LX, LN = S::LN, NS = S::NS;
  const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
  const int nx = b.n[0], ny = b.n[1], nz = b.n[2];
  const int plane = nx * ny;
  const int x = x0 + tx, y = y0 + ty;
  const bool valid = (x < nx) && (y < ny);
  const int mode = b.recon;
  const bool src = b.source == 1;

  // global in-plane offset of every LDS slot this thread stages
  int gxy[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int li = tid + s * NT;
    const int lx = li % LX, ly = li / LX;
    gxy[s] = (li < LN) ? pmod(y0 - 2 + ly, ny) * nx + pmod(x0 - 2 + lx, nx) : 0;
  }
  float stage[NS][8];
  auto load_plane = [&](int p) {
    const int base = wrap(p, nz) * plane;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      if (tid + s * NT < LN) {
#pragma unroll
        for (int f = 0; f < 8; ++f) stage[s][f] = W.f[f][base + gxy[s]];
      }
    }
  };
  auto store_plane = [&](int buf) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int li = tid + s * NT;
      if (li < LN) {
#pragma unroll
        for (int f = 0; f < 8; ++f) lds[(buf * 8 + f) * LN + li] = stage[s][f];
      }
    }
  };

  // own cell in the LDS tile
  const int c = (ty + 2) * LX + tx + 2;
  // z-stencil queue: q[s] holds plane p-3+s: rho, vz, vx, vy, p, Bx_avg, By_avg
  float q[4][7];
  float bz_prev = 0.0f;  // B_z on the lower z-face of plane p-1
  float bz_cur = 0.0f;

  const int zs = z0 - 2;
  load_plane(zs);
  store_plane(zs & 1);
  load_plane(zs + 1);
  store_plane((zs + 1) & 1);
  __syncthreads();

  for (int p = zs; p <= zend; ++p) {
    const int cb = p & 1, nb = (p + 1) & 1;
    const bool prefetch = (p + 2 <= zend);
    if (prefetch) load_plane(p + 2);
    const float *R = lds + (cb * 8 + 0) * LN;
    const float *VX = lds + (cb * 8 + 1) * LN;
    const float *VY = lds + (cb * 8 + 2) * LN;
    const float *VZ = lds + (cb * 8 + 3) * LN;
    const float *PR = lds + (cb * 8 + 4) * LN;
    const float *BX = lds + (cb * 8 + 5) * LN;
    const float *BY = lds + (cb * 8 + 6) * LN;
    const float *BZ = lds + (cb * 8 + 7) * LN;
    const float *BZN = lds + (nb * 8 + 7) * LN;

    // push plane p into the z queue
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
      for (int v = 0; v < 7; ++v) q[s][v] = q[s + 1][v];
    q[3][0] = R[c];
    q[3][1] = VZ[c];
    q[3][2] = VX[c];
    q[3][3] = VY[c];
    q[3][4] = PR[c];
    q[3][5] = 0.5f * (BX[c] + BX[c + 1]);
    q[3][6] = 0.5f * (BY[c] + BY[c + LX]);
    bz_prev = bz_cur;
    bz_cur = BZ[c];

    if (valid && p >= z0 && p < zend) {
      const int a = wrap(p, nz) * plane + y * nx + x;
      // ---- x-face (lower x-face of cell (x, y, p))
      {
        float L[7], Rr[7];
        const int s0 = c - 2, s1 = c - 1, s2 = c, s3 = c + 1;
        recon(mode, R[s0], R[s1], R[s2], R[s3], L[0], Rr[0]);
        recon(mode, VX[s0], VX[s1], VX[s2], VX[s3], L[1], Rr[1]);
        recon(mode, VY[s0], VY[s1], VY[s2], VY[s3], L[2], Rr[2]);
        recon(mode, VZ[s0], VZ[s1], VZ[s2], VZ[s3], L[3], Rr[3]);
        recon(mode, PR[s0], PR[s1], PR[s2], PR[s3], L[4], Rr[4]);
        recon(mode, 0.5f * (BY[s0] + BY[s0 + LX]), 0.5f * (BY[s1] + BY[s1 + LX]),
              0.5f * (BY[s2] + BY[s2 + LX]), 0.5f * (BY[s3] + BY[s3 + LX]), L[5], Rr[5]);
        recon(mode, 0.5f * (BZ[s0] + BZN[s0]), 0.5f * (BZ[s1] + BZN[s1]),
              0.5f * (BZ[s2] + BZN[s2]), 0.5f * (BZ[s3] + BZN[s3]), L[6], Rr[6]);
        if (src) {
          const float mbv = 0.5f * h * accel_x(b, 0.5f * (VY[s1] + VY[s2]));
          L[1] += mbv;
          Rr[1] += mbv;
        }
        const Flux7 r = hlld(b.gamma, b.inv_gm1, BX[c], L[0], Rr[0], L[1], Rr[1], L[2], Rr[2],
                             L[3], Rr[3], L[4], Rr[4], L[5], Rr[5], L[6], Rr[6]);
        F.dir[0].f[0][a] = r.rho;
        F.dir[0].f[1][a] = r.mn;
        F.dir[0].f[2][a] = r.mt1;
        F.dir[0].f[3][a] = r.mt2;
        F.dir[0].f[4][a] = r.e;
        F.dir[0].f[5][a] = r.bt1;  // B_y
        F.dir[0].f[6][a] = r.bt2;  // B_z
      }
      // ---- y-face
      {
        float L[7], Rr[7];
        const int s0 = c - 2 * LX, s1 = c - LX, s2 = c, s3 = c + LX;
        recon(mode, R[s0], R[s1], R[s2], R[s3], L[0], Rr[0]);
        recon(mode, VY[s0], VY[s1], VY[s2], VY[s3], L[1], Rr[1]);
        recon(mode, VZ[s0], VZ[s1], VZ[s2], VZ[s3], L[2], Rr[2]);
        recon(mode, VX[s0], VX[s1], VX[s2], VX[s3], L[3], Rr[3]);
        recon(mode, PR[s0], PR[s1], PR[s2], PR[s3], L[4], Rr[4]);
        recon(mode, 0.5f * (BZ[s0] + BZN[s0]), 0.5f * (BZ[s1] + BZN[s1]),
              0.5f * (BZ[s2] + BZN[s2]), 0.5f * (BZ[s3] + BZN[s3]), L[5], Rr[5]);
        recon(mode, 0.5f * (BX[s0] + BX[s0 + 1]), 0.5f * (BX[s1] + BX[s1 + 1]),
              0.5f * (BX[s2] + BX[s2 + 1]), 0.5f * (BX[s3] + BX[s3 + 1]), L[6], Rr[6]);
        if (src) {
          const float mbv = 0.5f * h * accel_y(b, 0.5f * (VX[s1] + VX[s2]), y);
          L[1] += mbv;
          Rr[1] += mbv;
        }
        const Flux7 r = hlld(b.gamma, b.inv_gm1, BY[c], L[0], Rr[0], L[1], Rr[1], L[2], Rr[2],
                             L[3], Rr[3], L[4], Rr[4], L[5], Rr[5], L[6], Rr[6]);
        F.dir[1].f[0][a] = r.rho;
        F.dir[1].f[2][a] = r.mn;
        F.dir[1].f[3][a] = r.mt1;
        F.dir[1].f[1][a] = r.mt2;
        F.dir[1].f[4][a] = r.e;
        F.dir[1].f[5][a] = r.bt1;  // B_z
        F.dir[1].f[6][a] = r.bt2;  // B_x
      }
    }
    // ---- z-face of plane p-1 from the queue (planes p-3 .. p)
    if (valid && p - 1 >= z0 && p - 1 < zend) {
      const int a = wrap(p - 1, nz) * plane + y * nx + x;
      float L[7], Rr[7];
#pragma unroll
      for (int v = 0; v < 7; ++v) recon(mode, q[0][v], q[1][v], q[2][v], q[3][v], L[v], Rr[v]);
      // the z acceleration is zero: the face predictor adds nothing
      const Flux7 r = hlld(b.gamma, b.inv_gm1, bz_prev, L[0], Rr[0], L[1], Rr[1], L[2], Rr[2],
                           L[3], Rr[3], L[4], Rr[4], L[5], Rr[5], L[6], Rr[6]);
      F.dir[2].f[0][a] = r.rho;
      F.dir[2].f[3][a] = r.mn;
      F.dir[2].f[1][a] = r.mt1;
      F.dir[2].f[2][a] = r.mt2;
      F.dir[2].f[4][a] = r.e;
      F.dir[2].f[5][a] = r.bt1;  // B_x
      F.dir[2].f[6][a] = r.bt2;  // B_y
    }
    __syncthreads();
    if (prefetch) store_plane(cb);  // plane p is no longer needed
    __syncthreads();
  }
}

template <int TX, int TY>
struct UpdateShape {
  static_assert(TX * TY == kTileThreads, "tiles have 256 threads");
  static constexpr int RX = TX + 2, RN = RX * (TY + 2);
  static constexpr int NR = (RN + kTileThreads - 1) / kTileThreads;  // ring slots per thread
  // LDS layout (floats): S3[5][3][RN] | SZ[2][2][RN] | SP[2][14][RN]
  static constexpr int kS3 = 0, kSZ = 15 * RN, kSP = 19 * RN;
  static constexpr int kLds = 47 * RN;
};

// The cell update as a z march: a TX x TY column tile plus a one-cell ring
// is staged in LDS rings, one plane ahead (prefetched into registers while
// the current plane is computed), so every neighbour access of the EMFs and
// the divergence is an LDS read and each HBM value is fetched once per tile.
//   S3 (planes r-1, r, r+1): v_x, v_y, v_z, the x-flux of B_z, the y-flux of B_z;
//   SZ (faces r, r+1): the z-fluxes of B_x and B_y;
//   SP (plane r): the hydro x/y fluxes, the x-flux of B_y, the y-flux of B_x
//   and the base B_x, B_y (for the upper faces).
// The z-flux hydro components and the base cell values are the thread's own
// and stay in registers.  cfl_t is lowered to the CFL time of every interior
// cell the thread updates (stage 2 with Cfl).
template <int TX, int TY, bool PaceOnly, bool Cfl>
__device__ __forceinline__ void update_tile(const Box &b, const FieldPtrs &Win, const FluxSet &F,
                                            const FieldPtrs &Wb, const FieldPtrs &Wo, float h,
                                            int x0, int y0, int z0, int zend, float *lds,
                                            float &cfl_t) {
  using S = UpdateShape<TX, TY>;
  constexpr int NT = kTileThreads, RX = S::RX, RN = S::RN, NR = S::NR;
  float *const S3 = lds + S::kS3;
  float *const SZ = lds + S::kSZ;
  float *const SP = lds + S::kSP;

  const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
  const int nx = b.n[0], ny = b.n[1], nz = b.n[2];
  const int plane = nx * ny;
  const int x = x0 + tx, y = y0 + ty;
  const bool valid = x < nx && y < ny;

  int gxy[NR];
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    const int li = tid + s * NT;
    gxy[s] = (li < RN) ? pmod(y0 - 1 + li / RX, ny) * nx + pmod(x0 - 1 + li % RX, nx) : 0;
  }
  const float *a3[5] = {Win.f[1], Win.f[2], Win.f[3], F.dir[0].f[6], F.dir[1].f[5]};
  const float *az[2] = {F.dir[2].f[5], F.dir[2].f[6]};
  const float *ap[14] = {F.dir[0].f[0], F.dir[0].f[1], F.dir[0].f[2], F.dir[0].f[3],
                         F.dir[0].f[4], F.dir[0].f[5], F.dir[1].f[0], F.dir[1].f[1],
                         F.dir[1].f[2], F.dir[1].f[3], F.dir[1].f[4], F.dir[1].f[6],
                         Wb.f[5],       Wb.f[6]};
  // staging registers
  float r3[NR][5], rz[NR][2], rp[NR][14];
  auto load3 = [&](int p) {
    const int base = wrap(p, nz) * plane;
#pragma unroll
    for (int s = 0; s < NR; ++s)
      if (tid + s * NT < RN)
#pragma unroll
        for (int k = 0; k < 5; ++k) r3[s][k] = a3[k][base + gxy[s]];
  };
  auto store3 = [&](int p) {
    const int slot = (p + 3) % 3;
#pragma unroll
    for (int s = 0; s < NR; ++s)
      if (tid + s * NT < RN)
#pragma unroll
        for (int k = 0; k < 5; ++k) S3[(k * 3 + slot) * RN + tid + s * NT] = r3[s][k];
  };
  auto loadz = [&](int p) {
    const int base = wrap(p, nz) * plane;
#pragma unroll
    for (int s = 0; s < NR; ++s)
      if (tid + s * NT < RN)
#pragma unroll
        for (int k = 0; k < 2; ++k) rz[s][k] = az[k][base + gxy[s]];
  };
  auto storez = [&](int p) {
#pragma unroll
    for (int s = 0; s < NR; ++s)
      if (tid + s * NT < RN)
#pragma unroll
        for (int k = 0; k < 2; ++k) SZ[(k * 2 + (p & 1)) * RN + tid + s * NT] = rz[s][k];
  };
  auto loadp = [&](int p) {
    const int base = wrap(p, nz) * plane;
#pragma unroll
    for (int s = 0; s < NR; ++s)
      if (tid + s * NT < RN)
#pragma unroll
        for (int k = 0; k < 14; ++k) rp[s][k] = ap[k][base + gxy[s]];
  };
  auto storep = [&](int p) {
#pragma unroll
    for (int s = 0; s < NR; ++s)
      if (tid + s * NT < RN)
#pragma unroll
        for (int k = 0; k < 14; ++k) SP[((p & 1) * 14 + k) * RN + tid + s * NT] = rp[s][k];
  };
  // own-column registers: z-flux hydro at faces r, r+1 and base cell values
  const int own = valid ? y * nx + x : 0;
  float fz_lo[5], fz_hi[5], fz_nx[5], wb[8], wb_nx[8], bz_hi = 0.0f, bz_nx = 0.0f;
  auto load_own_fz = [&](int p, float *dst) {
    const int base = wrap(p, nz) * plane + own;
#pragma unroll
    for (int k = 0; k < 5; ++k) dst[k] = F.dir[2].f[k][base];
  };
  auto load_own_w = [&](int p, float *dst) {
    const int base = wrap(p, nz) * plane + own;
#pragma unroll
    for (int k = 0; k < 8; ++k) dst[k] = Wb.f[k][base];
  };

  // prologue: planes z0-1, z0, z0+1 of the 3-rings, faces z0, z0+1, plane z0
  load3(z0 - 1);
  store3(z0 - 1);
  load3(z0);
  store3(z0);
  load3(z0 + 1);
  store3(z0 + 1);
  loadz(z0);
  storez(z0);
  loadz(z0 + 1);
  storez(z0 + 1);
  loadp(z0);
  storep(z0);
  load_own_fz(z0, fz_lo);
  load_own_fz(z0 + 1, fz_hi);
  load_own_w(z0, wb);
  bz_hi = Wb.f[7][wrap(z0 + 1, nz) * plane + own];
  __syncthreads();

  const int c = (ty + 1) * RX + tx + 1;  // own slot in the ring tile
  const float idx = b.inv_d[0], idy = b.inv_d[1], idz = b.inv_d[2];

  for (int r = z0; r < zend; ++r) {
    const bool more = r + 1 < zend;
    if (more) {  // prefetch the data of the next iteration
      load3(r + 2);
      loadz(r + 2);
      loadp(r + 1);
      load_own_fz(r + 2, fz_nx);
      load_own_w(r + 1, wb_nx);
      bz_nx = Wb.f[7][wrap(r + 2, nz) * plane + own];
    }
    const int sm = (r + 2) % 3, s0 = (r + 3) % 3, sp = (r + 4) % 3;  // planes r-1, r, r+1
    const int f0 = r & 1, f1 = (r + 1) & 1;                         // faces r, r+1
    const int pr_ = r & 1;                                          // plane r
    const float *VX[3] = {S3 + (0 * 3 + sm) * RN, S3 + (0 * 3 + s0) * RN, S3 + (0 * 3 + sp) * RN};
    const float *VY[3] = {S3 + (1 * 3 + sm) * RN, S3 + (1 * 3 + s0) * RN, S3 + (1 * 3 + sp) * RN};
    const float *VZ[3] = {S3 + (2 * 3 + sm) * RN, S3 + (2 * 3 + s0) * RN, S3 + (2 * 3 + sp) * RN};
    const float *FX6[3] = {S3 + (3 * 3 + sm) * RN, S3 + (3 * 3 + s0) * RN,
                           S3 + (3 * 3 + sp) * RN};
    const float *FY5[3] = {S3 + (4 * 3 + sm) * RN, S3 + (4 * 3 + s0) * RN,
                           S3 + (4 * 3 + sp) * RN};
    const float *FZ5[2] = {SZ + (0 * 2 + f0) * RN, SZ + (0 * 2 + f1) * RN};
    const float *FZ6[2] = {SZ + (1 * 2 + f0) * RN, SZ + (1 * 2 + f1) * RN};
    const float *FXk[6], *FYk[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      FXk[k] = SP + (pr_ * 14 + k) * RN;
      FYk[k] = SP + (pr_ * 14 + 6 + k) * RN;
    }
    const float *BXB = SP + (pr_ * 14 + 12) * RN, *BYB = SP + (pr_ * 14 + 13) * RN;

    if (valid) {
      const int a = r * plane + own;
      // divergence (X, Y, Z order)
      float Pc[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        float sacc = (FXk[k][c] - FXk[k][c + 1]) * idx;
        sacc += (FYk[k][c] - FYk[k][c + RX]) * idy;
        sacc += (fz_lo[k] - fz_hi[k]) * idz;
        Pc[k] = sacc;
      }
      // Edge EMFs; u (0: lower, 1: upper) indexes the edge along each axis.
      // EXY[u][v]: x-flux of B_y at x-face i+u, upwinded in y across y-face j+v
      // EYX[u][v]: y-flux of B_x at y-face j+v, upwinded in x across x-face i+u
      // (and likewise for the XZ and YZ pairs)
      float EXY[2][2], EYX[2][2], EXZ[2][2], EZX[2][2], EYZ[2][2], EZY[2][2];
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          // Z-edge at x-face i+u, y-face j+v (plane r): cells x in {i+u-1, i+u},
          // y in {j+v-1, j+v}
          const int e = c + u + v * RX;  // cell (i+u, j+v)
          const float sy = VY[1][e] + VY[1][e - RX] + VY[1][e - 1] + VY[1][e - 1 - RX];
          const float sx = VX[1][e] + VX[1][e - 1] + VX[1][e - RX] + VX[1][e - 1 - RX];
          EXY[u][v] = upwind(sy, FXk[5][e - RX], FXk[5][e]);
          EYX[u][v] = upwind(sx, FYk[5][e - 1], FYk[5][e]);
          // Y-edge at x-face i+u, z-face r+v: cells x in {i+u-1, i+u}, planes {r+v-1, r+v}
          const int ex = c + u;
          const float sz = VZ[v + 1][ex] + VZ[v][ex] + VZ[v + 1][ex - 1] + VZ[v][ex - 1];
          const float sx2 = VX[v + 1][ex] + VX[v + 1][ex - 1] + VX[v][ex] + VX[v][ex - 1];
          EXZ[u][v] = upwind(sz, FX6[v][ex], FX6[v + 1][ex]);
          EZX[u][v] = upwind(sx2, FZ5[v][ex - 1], FZ5[v][ex]);
          // X-edge at y-face j+u, z-face r+v
          const int ey = c + u * RX;
          const float sz2 = VZ[v + 1][ey] + VZ[v][ey] + VZ[v + 1][ey - RX] + VZ[v][ey - RX];
          const float sy2 = VY[v + 1][ey] + VY[v + 1][ey - RX] + VY[v][ey] + VY[v][ey - RX];
          EYZ[u][v] = upwind(sz2, FY5[v][ey], FY5[v + 1][ey]);
          EZY[u][v] = upwind(sy2, FZ6[v][ey - RX], FZ6[v][ey]);
        }
      float pbx[2], pby[2], pbz[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        pbx[u] = 0.5f * ((EXY[u][1] - EXY[u][0]) * idy + (EXZ[u][1] - EXZ[u][0]) * idz) -
                 0.5f * (EYX[u][1] - EYX[u][0]) * idy - 0.5f * (EZX[u][1] - EZX[u][0]) * idz;
        pby[u] = -0.5f * (EXY[1][u] - EXY[0][u]) * idx +
                 0.5f * ((EYZ[u][1] - EYZ[u][0]) * idz + (EYX[1][u] - EYX[0][u]) * idx) -
                 0.5f * (EZY[u][1] - EZY[u][0]) * idz;
        pbz[u] = -0.5f * (EXZ[1][u] - EXZ[0][u]) * idx - 0.5f * (EYZ[1][u] - EYZ[0][u]) * idy +
                 0.5f * ((EZX[1][u] - EZX[0][u]) * idx + (EZY[1][u] - EZY[0][u]) * idy);
      }
      if (PaceOnly) {
#pragma unroll
        for (int k = 0; k < 5; ++k) Wo.f[k][a] = Pc[k];
        Wo.f[5][a] = pbx[0];
        Wo.f[6][a] = pby[0];
        Wo.f[7][a] = pbz[0];
      } else {
        const float rho0 = wb[0], vx0 = wb[1], vy0 = wb[2], vz0 = wb[3];
        const float bxl0 = wb[5], bxh0 = BXB[c + 1];
        const float byl0 = wb[6], byh0 = BYB[c + RX];
        const float bzl0 = wb[7], bzh0 = bz_hi;
        const float e0 = wb[4] * b.inv_gm1 + 0.5f * rho0 * (vx0 * vx0 + vy0 * vy0 + vz0 * vz0) +
                         0.25f * (bxl0 * bxl0 + bxh0 * bxh0 + byl0 * byl0 + byh0 * byh0 +
                                  bzl0 * bzl0 + bzh0 * bzh0);
        const float rho = rho0 + h * Pc[0];
        const float mx = rho0 * vx0 + h * Pc[1];
        const float my = rho0 * vy0 + h * Pc[2];
        const float mz = rho0 * vz0 + h * Pc[3];
        const float e = e0 + h * Pc[4];
        const float bxl = bxl0 + h * pbx[0], bxh = bxh0 + h * pbx[1];
        const float byl = byl0 + h * pby[0], byh = byh0 + h * pby[1];
        const float bzl = bzl0 + h * pbz[0], bzh = bzh0 + h * pbz[1];
        const float irho = frcp(rho);
        float ux = mx * irho, uy = my * irho;
        const float uz = mz * irho;
        const float eint = e - 0.5f * (mx * mx + my * my + mz * mz) * irho -
                           0.25f * (bxl * bxl + bxh * bxh + byl * byl + byh * byh + bzl * bzl +
                                    bzh * bzh);
        const float pres = eint * b.gm1;
        if (b.source == 1) {
          const float ax_ = accel_x(b, uy);
          const float ay_ = accel_y(b, ux, y);
          ux += h * ax_;
          uy += h * ay_;
        }
        Wo.f[0][a] = rho;
        Wo.f[1][a] = ux;
        Wo.f[2][a] = uy;
        Wo.f[3][a] = uz;
        Wo.f[4][a] = pres;
        Wo.f[5][a] = bxl;
        Wo.f[6][a] = byl;
        Wo.f[7][a] = bzl;
        if (Cfl) {
          // CFL time of the new state for cells the next ghost fill leaves
          // untouched (interior rows/planes); the ghost cells are reduced
          // after the fill
          const bool ghost = (y < b.m[1]) || (y >= ny - b.m[1]) ||
                             (b.m[2] > 0 && (r < b.m[2] || r >= nz - b.m[2] - 1));
          if (!ghost) {
            // same rounding as the full CFL pass (no contraction), so an
            // incrementally reduced dt equals it bit for bit
#pragma clang fp contract(off)
            const float b2 = fmaxf(bxl * bxl, bxh * bxh) + fmaxf(byl * byl, byh * byh) +
                             fmaxf(bzl * bzl, bzh * bzh);
            const float cs = sqrtf((b.gamma * pres + b2) / rho);
            const float t = fminf(b.cfl * b.d[0] / fabsf(ux + cs),
                                  fminf(b.cfl * b.d[1] / fabsf(uy + cs),
                                        b.cfl * b.d[2] / fabsf(uz + cs)));
            cfl_t = fminf(cfl_t, t);
          }
        }
      }
    }
    __syncthreads();
    if (more) {
      store3(r + 2);
      storez(r + 2);
      storep(r + 1);
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        fz_lo[k] = fz_hi[k];
        fz_hi[k] = fz_nx[k];
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) wb[k] = wb_nx[k];
      bz_hi = bz_nx;
    }
    __syncthreads();
  }
}

}  // namespace tiled
}  // namespace nmhd
