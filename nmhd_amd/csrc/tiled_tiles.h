// tiled_tiles.h -- the two stage kernels of the production path as tile
// functions, shared by the single-box launches (kernels_tiled.hip) and the
// persistent multi-box launches (kernels_batch.hip).
//
//  faces_tile   all three face fluxes of every cell of a TX x TY column tile
//               over planes [z0, zend), one HLLD per face plus the tile's
//               upper boundary faces.  The x and y stencils (cells i-2..i+1)
//               are read from an LDS ring of two planes that the workgroup
//               stages cooperatively; the z stencil (planes k-2..k+1) lives
//               in a per-thread register queue filled as the march passes
//               each plane.  Differences the hydro fluxes in the tile (LDS
//               for x and y, registers for z) and writes the hydro
//               divergence of every cell (5 floats) and the six B fluxes the
//               CT needs (6 floats).
//  update_tile  per cell of the tile: the 24 upwinded edge EMFs of the
//               cell's six faces, the conserved update, primitive
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
#include <type_traits>

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

// Results that the next kernel reads are stored non-temporally: written
// once per step, they would otherwise displace the halo planes that
// neighbouring tiles re-read from L2 (1-3% faster per kernel, measured).
__device__ __forceinline__ void put_nt(float *dst, float v) { __builtin_nontemporal_store(v, dst); }
// Likewise for values that only the loading thread reads (no halo reuse).
__device__ __forceinline__ float get_nt(const float *src) { return __builtin_nontemporal_load(src); }

// The element at byte offset boff (32-bit) from a workgroup-uniform base:
// loads and stores through it take the scalar-base + 32-bit vector-offset
// form, with no 64-bit address arithmetic per lane.  The offsets used are
// in-plane (an XY plane holds fewer than 2^30 cells, checked at creation).
template <class T>
__device__ __forceinline__ T *at_byte(T *base, unsigned boff) {
  using C = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return reinterpret_cast<T *>(reinterpret_cast<C *>(base) + boff);
}

// A per-lane offset made opaque to the optimiser for the current plane (an
// empty asm operand, no instruction): otherwise the loop-invariant sum of
// an array base and the offset is hoisted out of the plane loop as a 64-bit
// per-lane pointer per array, which costs registers and 64-bit adds.
__device__ __forceinline__ unsigned opaque(unsigned v) {
  asm volatile("" : "+v"(v));
  return v;
}

__device__ __forceinline__ void recon(int mode, float x0, float x1, float x2, float x3, float &l,
                                      float &r) {
  if (mode == 1) {
    l = x1;
    r = x2;
  } else {
    recon_minbee2_med3(x0, x1, x2, x3, l, r);
  }
}

// Reconstruction of a cell-centred B component, the average of its cell's two
// faces, from the four stencil cells' face sums a + b instead of the
// averages 0.5 (a + b): every step of the reconstruction (differences,
// median of three, the half-slope multiply-add) commutes exactly with the
// scaling by 2, so halving the two face states gives the same bits as
// reconstructing the averages, with 4 multiplies fewer per variable.
__device__ __forceinline__ void recon_bsum(int mode, float s0, float s1, float s2, float s3,
                                           float &l, float &r) {
  recon(mode, s0, s1, s2, s3, l, r);
  l *= 0.5f;
  r *= 0.5f;
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

// What the face kernel hands to the update kernel, in the 21-array flux
// storage of a stage: the hydro divergence of every cell (5 arrays, in the
// slots of the x-sweep's hydro fluxes) and the six face fluxes of B that
// the CT edge EMFs need.
struct StageBuf {
  float *pace[5];  // P_rho, P_mx, P_my, P_mz, P_E
  float *fb[6];    // x-face F(B_y), x-face F(B_z), y-face F(B_z), y-face F(B_x),
                   // z-face F(B_x), z-face F(B_y)
};

__host__ __device__ inline StageBuf stage_buf(const FluxSet &F) {
  StageBuf s;
  for (int k = 0; k < 5; ++k) s.pace[k] = F.dir[0].f[k];
  s.fb[0] = F.dir[0].f[5];
  s.fb[1] = F.dir[0].f[6];
  s.fb[2] = F.dir[1].f[5];
  s.fb[3] = F.dir[1].f[6];
  s.fb[4] = F.dir[2].f[5];
  s.fb[5] = F.dir[2].f[6];
  return s;
}

template <int TX, int TY>
struct FacesShape {
  static_assert(TX * TY == kTileThreads, "tiles have 256 threads");
  static_assert(TX <= 64 && TY <= 64, "a boundary face row or column fits one wave");
  static constexpr int LX = TX + 4, LY = TY + 4, LN = LX * LY;
  static constexpr int NS = (LN + kTileThreads - 1) / kTileThreads;  // slots per thread
  // LDS (floats): two staged planes of the 8 fields, then the hydro fluxes
  // of the plane's x-faces [5][TY][TX+1] and y-faces [5][TY+1][TX]
  static constexpr int kPlanes = 2 * 8 * LN;
  static constexpr int kFx = kPlanes;
  static constexpr int kFy = kFx + 5 * TY * (TX + 1);
  // then the global in-plane offset of every staged slot (int), kept in LDS
  // rather than in registers (read once per plane by the slot's own thread)
  static constexpr int kGxy = kFy + 5 * (TY + 1) * TX;
  static constexpr int kLds = kGxy + LN;
  // waves per SIMD the LDS allows (4 workgroups of 4 waves per CU when 4
  // copies fit the 160 KiB, else 3): the kernels' occupancy target
  static constexpr int kWavesPerSimd = 4 * kLds * 4 <= 160 * 1024 ? 4 : 3;
};

// Face fluxes of a TX x TY column tile over planes [z0, zend), and the
// hydro divergence of its cells.  Per plane p the workgroup computes the
// x-faces i in [x0, x0+TX] and the y-faces j in [y0, y0+TY] (the upper
// boundary faces of the tile, i.e. one face column and one face row more
// than it owns, spread over the first TY+TX threads), keeps their hydro
// fluxes in LDS and differences them; the z-faces come from the register
// queue, one plane of faces beyond zend.  Only the tile's own faces write
// their B fluxes to memory.  Recon is the box's reconstruction (b.recon: 0
// minbee2, 1 piecewise constant) as a compile-time constant, so that the
// scheme costs no per-variable selects; faces_tile_any dispatches on it.
template <int TX, int TY, int Recon>
__device__ __forceinline__ void faces_tile(const Box &b, const FieldPtrs &W, const FluxSet &Fs,
                                           float h, int x0, int y0, int z0, int zend,
                                           float *lds) {
  using S = FacesShape<TX, TY>;
  constexpr int NT = kTileThreads, LX = S::LX, LN = S::LN, NS = S::NS;
  const StageBuf SB = stage_buf(Fs);
  float *const FXL = lds + S::kFx;
  float *const FYL = lds + S::kFy;
  const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
  const int nx = b.n[0], ny = b.n[1], nz = b.n[2];
  const int plane = nx * ny;
  const int x = x0 + tx, y = y0 + ty;
  const bool valid = (x < nx) && (y < ny);
  constexpr int mode = Recon;
  const bool src = b.source == 1;
  const float idx = b.inv_d[0], idy = b.inv_d[1], idz = b.inv_d[2];

  // global in-plane byte offset of every LDS slot this thread stages (each
  // thread writes and reads only its own slots)
  unsigned *const GXY = reinterpret_cast<unsigned *>(lds + S::kGxy);
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int li = tid + s * NT;
    const int lx = li % LX, ly = li / LX;
    if (li < LN) GXY[li] = (unsigned)(pmod(y0 - 2 + ly, ny) * nx + pmod(x0 - 2 + lx, nx)) * 4u;
  }
  // this thread's own cell: in-plane byte offset
  const unsigned own = (unsigned)(y * nx + x) * 4u;
  // start of plane p (wrapped) in an array
  auto plane_of = [&](auto *arr, int p) { return arr + (size_t)wrap(p, nz) * plane; };
  float stage[NS][8];
  auto load_plane = [&](int p) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      if (tid + s * NT < LN) {
        const unsigned g = opaque(GXY[tid + s * NT]);
#pragma unroll
        for (int f = 0; f < 8; ++f) stage[s][f] = *at_byte(plane_of(W.f[f], p), g);
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

  // lower x-face of the cell at LDS index c of plane buffer cb (nb: next plane)
  auto face_x = [&](int cb, int nb, int c) -> Flux7 {
    const float *R = lds + (cb * 8 + 0) * LN, *VX = lds + (cb * 8 + 1) * LN;
    const float *VY = lds + (cb * 8 + 2) * LN, *VZ = lds + (cb * 8 + 3) * LN;
    const float *PR = lds + (cb * 8 + 4) * LN, *BX = lds + (cb * 8 + 5) * LN;
    const float *BY = lds + (cb * 8 + 6) * LN, *BZ = lds + (cb * 8 + 7) * LN;
    const float *BZN = lds + (nb * 8 + 7) * LN;
    float L[7], Rr[7];
    const int s0 = c - 2, s1 = c - 1, s2 = c, s3 = c + 1;
    recon(mode, R[s0], R[s1], R[s2], R[s3], L[0], Rr[0]);
    recon(mode, VX[s0], VX[s1], VX[s2], VX[s3], L[1], Rr[1]);
    recon(mode, VY[s0], VY[s1], VY[s2], VY[s3], L[2], Rr[2]);
    recon(mode, VZ[s0], VZ[s1], VZ[s2], VZ[s3], L[3], Rr[3]);
    recon(mode, PR[s0], PR[s1], PR[s2], PR[s3], L[4], Rr[4]);
    recon_bsum(mode, BY[s0] + BY[s0 + LX], BY[s1] + BY[s1 + LX], BY[s2] + BY[s2 + LX],
               BY[s3] + BY[s3 + LX], L[5], Rr[5]);
    recon_bsum(mode, BZ[s0] + BZN[s0], BZ[s1] + BZN[s1], BZ[s2] + BZN[s2], BZ[s3] + BZN[s3],
               L[6], Rr[6]);
    if (src) {
      const float mbv = 0.5f * h * accel_x(b, 0.5f * (VY[s1] + VY[s2]));
      L[1] += mbv;
      Rr[1] += mbv;
    }
    return hlld(b.gamma, b.inv_gm1, BX[c], L[0], Rr[0], L[1], Rr[1], L[2], Rr[2], L[3], Rr[3],
                L[4], Rr[4], L[5], Rr[5], L[6], Rr[6]);
  };
  // lower y-face of the cell at LDS index c; yrow = the face's array row
  auto face_y = [&](int cb, int nb, int c, int yrow) -> Flux7 {
    const float *R = lds + (cb * 8 + 0) * LN, *VX = lds + (cb * 8 + 1) * LN;
    const float *VY = lds + (cb * 8 + 2) * LN, *VZ = lds + (cb * 8 + 3) * LN;
    const float *PR = lds + (cb * 8 + 4) * LN, *BX = lds + (cb * 8 + 5) * LN;
    const float *BY = lds + (cb * 8 + 6) * LN, *BZ = lds + (cb * 8 + 7) * LN;
    const float *BZN = lds + (nb * 8 + 7) * LN;
    float L[7], Rr[7];
    const int s0 = c - 2 * LX, s1 = c - LX, s2 = c, s3 = c + LX;
    recon(mode, R[s0], R[s1], R[s2], R[s3], L[0], Rr[0]);
    recon(mode, VY[s0], VY[s1], VY[s2], VY[s3], L[1], Rr[1]);
    recon(mode, VZ[s0], VZ[s1], VZ[s2], VZ[s3], L[2], Rr[2]);
    recon(mode, VX[s0], VX[s1], VX[s2], VX[s3], L[3], Rr[3]);
    recon(mode, PR[s0], PR[s1], PR[s2], PR[s3], L[4], Rr[4]);
    recon_bsum(mode, BZ[s0] + BZN[s0], BZ[s1] + BZN[s1], BZ[s2] + BZN[s2], BZ[s3] + BZN[s3],
               L[5], Rr[5]);
    recon_bsum(mode, BX[s0] + BX[s0 + 1], BX[s1] + BX[s1 + 1], BX[s2] + BX[s2 + 1],
               BX[s3] + BX[s3 + 1], L[6], Rr[6]);
    if (src) {
      const float mbv = 0.5f * h * accel_y_face(b, 0.5f * (VX[s1] + VX[s2]), yrow);
      L[1] += mbv;
      Rr[1] += mbv;
    }
    return hlld(b.gamma, b.inv_gm1, BY[c], L[0], Rr[0], L[1], Rr[1], L[2], Rr[2], L[3], Rr[3],
                L[4], Rr[4], L[5], Rr[5], L[6], Rr[6]);
  };
  // the lower y-face (isy) or x-face of the cell at LDS index c, with the
  // stencil direction chosen per lane: the same expressions, in the same
  // order, as face_y / face_x, so the two give the same bits.  The five
  // hydro variables are reconstructed in the fixed order (rho, vx, vy, vz, p)
  // and the velocities then permuted into (normal, t1, t2); of the
  // transverse B averages, the one along z (B_z of this and the next plane)
  // is common to both and the other (B_y along y for an x-face, B_x along x
  // for a y-face) is read through a per-lane pointer.
  auto face_xy = [&](int cb, int nb, int c, bool isy, int yrow) -> Flux7 {
    const float *R = lds + (cb * 8 + 0) * LN, *VX = lds + (cb * 8 + 1) * LN;
    const float *VY = lds + (cb * 8 + 2) * LN, *VZ = lds + (cb * 8 + 3) * LN;
    const float *PR = lds + (cb * 8 + 4) * LN, *BX = lds + (cb * 8 + 5) * LN;
    const float *BY = lds + (cb * 8 + 6) * LN, *BZ = lds + (cb * 8 + 7) * LN;
    const float *BZN = lds + (nb * 8 + 7) * LN;
    const float *OA = isy ? BX : BY, *OB = isy ? BX + 1 : BY + LX;
    const int st = isy ? LX : 1;
    const int s0 = c - 2 * st, s1 = c - st, s2 = c, s3 = c + st;
    float L[7], Rr[7], lx, rx, ly, ry, lz, rz, lo, ro, lc, rc;
    recon(mode, R[s0], R[s1], R[s2], R[s3], L[0], Rr[0]);
    recon(mode, VX[s0], VX[s1], VX[s2], VX[s3], lx, rx);
    recon(mode, VY[s0], VY[s1], VY[s2], VY[s3], ly, ry);
    recon(mode, VZ[s0], VZ[s1], VZ[s2], VZ[s3], lz, rz);
    recon(mode, PR[s0], PR[s1], PR[s2], PR[s3], L[4], Rr[4]);
    recon_bsum(mode, OA[s0] + OB[s0], OA[s1] + OB[s1], OA[s2] + OB[s2], OA[s3] + OB[s3], lo, ro);
    recon_bsum(mode, BZ[s0] + BZN[s0], BZ[s1] + BZN[s1], BZ[s2] + BZN[s2], BZ[s3] + BZN[s3], lc,
               rc);
    // x-face: (vx, vy, vz), (B_y, B_z); y-face: (vy, vz, vx), (B_z, B_x)
    L[1] = isy ? ly : lx, Rr[1] = isy ? ry : rx;
    L[2] = isy ? lz : ly, Rr[2] = isy ? rz : ry;
    L[3] = isy ? lx : lz, Rr[3] = isy ? rx : rz;
    L[5] = isy ? lc : lo, Rr[5] = isy ? rc : ro;
    L[6] = isy ? lo : lc, Rr[6] = isy ? ro : rc;
    if (src) {
      const float *VA = isy ? VX : VY;
      const float va = 0.5f * (VA[s1] + VA[s2]);
      const float mbv = 0.5f * h * (isy ? accel_y_face(b, va, yrow) : accel_x(b, va));
      L[1] += mbv;
      Rr[1] += mbv;
    }
    return hlld(b.gamma, b.inv_gm1, isy ? BY[c] : BX[c], L[0], Rr[0], L[1], Rr[1], L[2], Rr[2],
                L[3], Rr[3], L[4], Rr[4], L[5], Rr[5], L[6], Rr[6]);
  };
  // hydro fluxes in the conserved order (rho, m_x, m_y, m_z, E)
  auto put_x = [&](int row, int col, const Flux7 &r) {
    float *d = FXL + row * (TX + 1) + col;
    constexpr int st = TY * (TX + 1);
    d[0] = r.rho;
    d[st] = r.mn;
    d[2 * st] = r.mt1;
    d[3 * st] = r.mt2;
    d[4 * st] = r.e;
  };
  auto put_y = [&](int row, int col, const Flux7 &r) {
    float *d = FYL + row * TX + col;
    constexpr int st = (TY + 1) * TX;
    d[0] = r.rho;
    d[st] = r.mt2;
    d[2 * st] = r.mn;
    d[3 * st] = r.mt1;
    d[4 * st] = r.e;
  };

  // own cell in the LDS tile
  const int c = (ty + 2) * LX + tx + 2;
  // z-stencil queue: q[s] holds plane p-3+s: rho, vz, vx, vy, p, Bx_avg, By_avg
  float q[4][7];
  // limited z-slope of plane p-2 (carried from the previous plane: it is the
  // left state's slope of this plane's z-face and the right state's of the
  // previous one)
  float zs_prev[7] = {};
  float bz_prev = 0.0f;  // B_z on the lower z-face of plane p-1
  float bz_cur = 0.0f;
  // divergence pipeline: x+y part of planes p-1, p-2; z-flux of the lower
  // face of plane p-2
  float pxy1[5] = {}, pxy2[5] = {}, fzl[5] = {};

  const int zs = z0 - 2, zlast = zend + 1;
  load_plane(zs);
  store_plane(zs & 1);
  load_plane(zs + 1);
  store_plane((zs + 1) & 1);
  __syncthreads();

  for (int p = zs; p <= zlast; ++p) {
    const int cb = p & 1, nb = (p + 1) & 1;
    const bool prefetch = (p + 2 <= zlast);
    const unsigned ownp = opaque(own);
    if constexpr (TX + TY <= 64) {
      // the wave that also evaluates this plane's boundary faces sets the
      // workgroup's pace at the plane barrier: it issues first
      if (__builtin_amdgcn_readfirstlane(tid >> 6) == ((p + x0 / TX + y0 / TY) & 3))
        __builtin_amdgcn_s_setprio(1);
      else
        __builtin_amdgcn_s_setprio(0);
    }
    if (prefetch) load_plane(p + 2);
    {
      const float *P0 = lds + cb * 8 * LN;
      // push plane p into the z queue
#pragma unroll
      for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int v = 0; v < 7; ++v) q[s][v] = q[s + 1][v];
      q[3][0] = P0[c];
      q[3][1] = P0[3 * LN + c];
      q[3][2] = P0[1 * LN + c];
      q[3][3] = P0[2 * LN + c];
      q[3][4] = P0[4 * LN + c];
      q[3][5] = 0.5f * (P0[5 * LN + c] + P0[5 * LN + c + 1]);
      q[3][6] = 0.5f * (P0[6 * LN + c] + P0[6 * LN + c + LX]);
      bz_prev = bz_cur;
      bz_cur = P0[7 * LN + c];
    }
    const bool xy_plane = p >= z0 && p < zend;
    if (xy_plane) {
      const int yrow = y < ny ? y : y - ny;  // the face row y == ny is face 0
      if (y < ny && x <= nx) {
        const Flux7 r = face_x(cb, nb, c);
        put_x(ty, tx, r);
        if (valid) {
          put_nt(at_byte(plane_of(SB.fb[0], p), ownp), r.bt1);  // B_y
          put_nt(at_byte(plane_of(SB.fb[1], p), ownp), r.bt2);  // B_z
        }
      }
      if (x < nx && y <= ny) {
        const Flux7 r = face_y(cb, nb, c, yrow);
        put_y(ty, tx, r);
        if (valid) {
          put_nt(at_byte(plane_of(SB.fb[2], p), ownp), r.bt1);  // B_z
          put_nt(at_byte(plane_of(SB.fb[3], p), ownp), r.bt2);  // B_x
        }
      }
      // the tile's upper boundary faces, the y-faces of row y0+TY and the
      // x-faces of column x0+TX, on a wave that rotates with the plane and
      // starts at an offset given by the tile's position, so that the SIMDs
      // of a CU share the extra HLLD evaluations evenly also when the
      // workgroups on it march in step: when both fit one wave (TX + TY <=
      // 64), in one pass (lanes [0, TX) the y-faces, [TX, TX+TY) the
      // x-faces), else on two waves
      const int lane = tid & 63, wave = tid >> 6;
      const int rot = p + x0 / TX + y0 / TY;
      if constexpr (TX + TY <= 64) {
        if (wave == (rot & 3) && lane < TX + TY) {
          const bool isy = lane < TX;
          const int row = isy ? TY : lane - TX, col = isy ? lane : TX;
          const int yy = y0 + row;
          if (isy ? (x0 + lane < nx && yy <= ny) : (yy < ny && x0 + TX <= nx)) {
            const Flux7 r = face_xy(cb, nb, (row + 2) * LX + col + 2, isy, yy < ny ? yy : yy - ny);
            if (isy)
              put_y(TY, lane, r);
            else
              put_x(row, TX, r);
          }
        }
      } else {
        if (wave == ((2 * rot) & 3) && lane < TY) {
          const int yy = y0 + lane;
          if (yy < ny && x0 + TX <= nx) put_x(lane, TX, face_x(cb, nb, (lane + 2) * LX + TX + 2));
        } else if (wave == ((2 * rot + 1) & 3) && lane < TX) {
          const int yy = y0 + TY;
          if (x0 + lane < nx && yy <= ny)
            put_y(TY, lane, face_y(cb, nb, (TY + 2) * LX + lane + 2, yy < ny ? yy : yy - ny));
        }
      }
    }
    // z-face below plane p-1 (one plane beyond the chunk: the upper face of
    // its last plane)
    float fzn[5] = {};
    // limited slope of plane p-1, minmod(x_{p-1} - x_{p-2}, x_p - x_{p-1})
    float zs_cur[7];
#pragma unroll
    for (int v = 0; v < 7; ++v) zs_cur[v] = minmod_med3(q[2][v] - q[1][v], q[3][v] - q[2][v]);
    if (valid && p - 1 >= z0 && p - 1 <= zend) {
      float L[7], Rr[7];
#pragma unroll
      for (int v = 0; v < 7; ++v) {
        if (mode == 1) {
          L[v] = q[1][v];
          Rr[v] = q[2][v];
        } else {
          L[v] = q[1][v] + 0.5f * zs_prev[v];
          Rr[v] = q[2][v] - 0.5f * zs_cur[v];
        }
      }
      // the z acceleration is zero: the face predictor adds nothing
      const Flux7 r = hlld(b.gamma, b.inv_gm1, bz_prev, L[0], Rr[0], L[1], Rr[1], L[2], Rr[2],
                           L[3], Rr[3], L[4], Rr[4], L[5], Rr[5], L[6], Rr[6]);
      fzn[0] = r.rho;
      fzn[1] = r.mt1;
      fzn[2] = r.mt2;
      fzn[3] = r.mn;
      fzn[4] = r.e;
      if (p - 1 < zend) {
        put_nt(at_byte(plane_of(SB.fb[4], p - 1), ownp), r.bt1);  // B_x
        put_nt(at_byte(plane_of(SB.fb[5], p - 1), ownp), r.bt2);  // B_y
      }
    }
    __syncthreads();
    // x + y divergence of plane p (X, Y order)
    float pxy[5] = {};
    if (valid && xy_plane) {
      constexpr int sx = TY * (TX + 1), sy = (TY + 1) * TX;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const float *fx = FXL + k * sx + ty * (TX + 1) + tx;
        const float *fy = FYL + k * sy + ty * TX + tx;
        float sacc = (fx[0] - fx[1]) * idx;
        sacc += (fy[0] - fy[TX]) * idy;
        pxy[k] = sacc;
      }
    }
    // plane p-2 complete: its lower z-face (fzl) and upper z-face (fzn)
    if (valid && p - 2 >= z0 && p - 2 < zend) {
#pragma unroll
      for (int k = 0; k < 5; ++k)
        put_nt(at_byte(plane_of(SB.pace[k], p - 2), ownp), pxy2[k] + (fzl[k] - fzn[k]) * idz);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      pxy2[k] = pxy1[k];
      pxy1[k] = pxy[k];
      fzl[k] = fzn[k];
    }
#pragma unroll
    for (int v = 0; v < 7; ++v) zs_prev[v] = zs_cur[v];
    if (prefetch) store_plane(cb);  // plane p is no longer needed
    __syncthreads();
  }
}

// faces_tile for the box's reconstruction (a uniform branch: two copies of
// the march, each without scheme selects).
template <int TX, int TY>
__device__ __forceinline__ void faces_tile_any(const Box &b, const FieldPtrs &W,
                                               const FluxSet &Fs, float h, int x0, int y0,
                                               int z0, int zend, float *lds) {
  if (b.recon == 1)
    faces_tile<TX, TY, 1>(b, W, Fs, h, x0, y0, z0, zend, lds);
  else
    faces_tile<TX, TY, 0>(b, W, Fs, h, x0, y0, z0, zend, lds);
}

template <int TX, int TY>
struct UpdateShape {
  static_assert(TX * TY == kTileThreads, "tiles have 256 threads");
  static constexpr int RX = TX + 2, RN = RX * (TY + 2);
  static constexpr int NR = (RN + kTileThreads - 1) / kTileThreads;  // ring slots per thread
  // LDS layout (floats): S3[5][3][RN] | SZ[2][2][RN] | SP[4][RN]
  static constexpr int kS3 = 0, kSZ = 15 * RN, kSP = 19 * RN;
  static constexpr int kLds = 23 * RN;
};

// The cell update as a z march: a TX x TY column tile plus a one-cell ring
// is staged in LDS, one plane ahead (prefetched into registers while the
// current plane is computed), so every neighbour access of the EMFs is an
// LDS read and each HBM value is fetched once per tile.
//   S3 (planes r-1, r, r+1): v_x, v_y, v_z, x-face F(B_z), y-face F(B_z);
//   SZ (faces r, r+1): z-face F(B_x), F(B_y);
//   SP (plane r): x-face F(B_y), y-face F(B_x), base B_x, B_y (upper faces).
// The hydro divergence and the base cell values are the thread's own and
// stay in registers.  cfl_t is lowered to the CFL time of every interior
// cell the thread updates (stage 2 with Cfl).
template <int TX, int TY, bool PaceOnly, bool Cfl>
__device__ __forceinline__ void update_tile(const Box &b, const FieldPtrs &Win, const FluxSet &Fs,
                                            const FieldPtrs &Wb, const FieldPtrs &Wo, float h,
                                            int x0, int y0, int z0, int zend, float *lds,
                                            float &cfl_t) {
  using S = UpdateShape<TX, TY>;
  constexpr int NT = kTileThreads, RX = S::RX, RN = S::RN, NR = S::NR;
  const StageBuf SB = stage_buf(Fs);
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
  const float *a3[5] = {Win.f[1], Win.f[2], Win.f[3], SB.fb[1], SB.fb[2]};
  const float *az[2] = {SB.fb[4], SB.fb[5]};
  const float *ap[4] = {SB.fb[0], SB.fb[3], Wb.f[5], Wb.f[6]};
  // staging registers
  float r3[NR][5], rz[NR][2], rp[NR][4];
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
        for (int k = 0; k < 4; ++k) rp[s][k] = ap[k][base + gxy[s]];
  };
  auto storep = [&]() {
#pragma unroll
    for (int s = 0; s < NR; ++s)
      if (tid + s * NT < RN)
#pragma unroll
        for (int k = 0; k < 4; ++k) SP[k * RN + tid + s * NT] = rp[s][k];
  };
  // own-column registers: hydro divergence and base cell values of plane r.
  // Values the rings already hold are read from LDS instead of memory: the
  // base B_x, B_y (SP) always, the base velocities (S3) when the stage input
  // is the base state (the first stage); the base B_z of plane r is the
  // upper B_z of the previous plane.
  const int own = valid ? y * nx + x : 0;
  const bool in_is_base = Win.f[1] == Wb.f[1] && Win.f[2] == Wb.f[2] && Win.f[3] == Wb.f[3];
  float pc[5], pc_nx[5], wb[5] = {}, wb_nx[5] = {}, bz_lo = 0.0f, bz_hi = 0.0f, bz_nx = 0.0f;
  auto load_own = [&](int p, float *dpc, float *dwb) {
    const int base = wrap(p, nz) * plane + own;
#pragma unroll
    for (int k = 0; k < 5; ++k) dpc[k] = get_nt(&SB.pace[k][base]);
    dwb[0] = get_nt(&Wb.f[0][base]);
    dwb[4] = get_nt(&Wb.f[4][base]);
    if (!in_is_base) {
#pragma unroll
      for (int k = 1; k < 4; ++k) dwb[k] = get_nt(&Wb.f[k][base]);
    }
  };

  // prologue: planes z0-1, z0, z0+1 of the 3-ring, faces z0, z0+1, plane z0
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
  storep();
  load_own(z0, pc, wb);
  bz_lo = Wb.f[7][wrap(z0, nz) * plane + own];
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
      load_own(r + 1, pc_nx, wb_nx);
      bz_nx = get_nt(&Wb.f[7][wrap(r + 2, nz) * plane + own]);
    }
    const int sm = (r + 2) % 3, s0 = (r + 3) % 3, sp = (r + 4) % 3;  // planes r-1, r, r+1
    const int f0 = r & 1, f1 = (r + 1) & 1;                         // faces r, r+1
    const float *VX[3] = {S3 + (0 * 3 + sm) * RN, S3 + (0 * 3 + s0) * RN, S3 + (0 * 3 + sp) * RN};
    const float *VY[3] = {S3 + (1 * 3 + sm) * RN, S3 + (1 * 3 + s0) * RN, S3 + (1 * 3 + sp) * RN};
    const float *VZ[3] = {S3 + (2 * 3 + sm) * RN, S3 + (2 * 3 + s0) * RN, S3 + (2 * 3 + sp) * RN};
    const float *FX6[3] = {S3 + (3 * 3 + sm) * RN, S3 + (3 * 3 + s0) * RN,
                           S3 + (3 * 3 + sp) * RN};
    const float *FY5[3] = {S3 + (4 * 3 + sm) * RN, S3 + (4 * 3 + s0) * RN,
                           S3 + (4 * 3 + sp) * RN};
    const float *FZ5[2] = {SZ + (0 * 2 + f0) * RN, SZ + (0 * 2 + f1) * RN};
    const float *FZ6[2] = {SZ + (1 * 2 + f0) * RN, SZ + (1 * 2 + f1) * RN};
    const float *FX5 = SP, *FY6 = SP + RN, *BXB = SP + 2 * RN, *BYB = SP + 3 * RN;

    if (valid) {
      const int a = r * plane + own;
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
          EXY[u][v] = upwind(sy, FX5[e - RX], FX5[e]);
          EYX[u][v] = upwind(sx, FY6[e - 1], FY6[e]);
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
      const float idx_ = idx, idy_ = idy, idz_ = idz;
      float pbx[2], pby[2], pbz[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        pbx[u] = 0.5f * ((EXY[u][1] - EXY[u][0]) * idy_ + (EXZ[u][1] - EXZ[u][0]) * idz_) -
                 0.5f * (EYX[u][1] - EYX[u][0]) * idy_ - 0.5f * (EZX[u][1] - EZX[u][0]) * idz_;
        pby[u] = -0.5f * (EXY[1][u] - EXY[0][u]) * idx_ +
                 0.5f * ((EYZ[u][1] - EYZ[u][0]) * idz_ + (EYX[1][u] - EYX[0][u]) * idx_) -
                 0.5f * (EZY[u][1] - EZY[u][0]) * idz_;
        pbz[u] = -0.5f * (EXZ[1][u] - EXZ[0][u]) * idx_ - 0.5f * (EYZ[1][u] - EYZ[0][u]) * idy_ +
                 0.5f * ((EZX[1][u] - EZX[0][u]) * idx_ + (EZY[1][u] - EZY[0][u]) * idy_);
      }
      if (PaceOnly) {
#pragma unroll
        for (int k = 0; k < 5; ++k) put_nt(&Wo.f[k][a], pc[k]);
        put_nt(&Wo.f[5][a], pbx[0]);
        put_nt(&Wo.f[6][a], pby[0]);
        put_nt(&Wo.f[7][a], pbz[0]);
      } else {
        const float rho0 = wb[0];
        const float vx0 = in_is_base ? VX[1][c] : wb[1];
        const float vy0 = in_is_base ? VY[1][c] : wb[2];
        const float vz0 = in_is_base ? VZ[1][c] : wb[3];
        const float bxl0 = BXB[c], bxh0 = BXB[c + 1];
        const float byl0 = BYB[c], byh0 = BYB[c + RX];
        const float bzl0 = bz_lo, bzh0 = bz_hi;
        const float e0 = wb[4] * b.inv_gm1 + 0.5f * rho0 * (vx0 * vx0 + vy0 * vy0 + vz0 * vz0) +
                         0.25f * (bxl0 * bxl0 + bxh0 * bxh0 + byl0 * byl0 + byh0 * byh0 +
                                  bzl0 * bzl0 + bzh0 * bzh0);
        const float rho = rho0 + h * pc[0];
        const float mx = rho0 * vx0 + h * pc[1];
        const float my = rho0 * vy0 + h * pc[2];
        const float mz = rho0 * vz0 + h * pc[3];
        const float e = e0 + h * pc[4];
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
        put_nt(&Wo.f[0][a], rho);
        put_nt(&Wo.f[1][a], ux);
        put_nt(&Wo.f[2][a], uy);
        put_nt(&Wo.f[3][a], uz);
        put_nt(&Wo.f[4][a], pres);
        put_nt(&Wo.f[5][a], bxl);
        put_nt(&Wo.f[6][a], byl);
        put_nt(&Wo.f[7][a], bzl);
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
            const float t = fminf(b.cfl * b.d[0] / cfl_speed(b, ux, cs),
                                  fminf(b.cfl * b.d[1] / cfl_speed(b, uy, cs),
                                        b.cfl * b.d[2] / cfl_speed(b, uz, cs)));
            cfl_t = fminf(cfl_t, t);
          }
        }
      }
    }
    __syncthreads();
    if (more) {
      store3(r + 2);
      storez(r + 2);
      storep();
#pragma unroll
      for (int k = 0; k < 5; ++k) pc[k] = pc_nx[k];
#pragma unroll
      for (int k = 0; k < 5; ++k) wb[k] = wb_nx[k];
      bz_lo = bz_hi;
      bz_hi = bz_nx;
    }
    __syncthreads();
  }
}

}  // namespace tiled
}  // namespace nmhd
