// kernels_tiled.hip -- the production kernel path: two kernels per RK stage.
//
//  k_faces   all three face fluxes of every cell, one HLLD per face and no
//            halo recomputation.  A workgroup owns a TX x TY column tile and
//            marches along z over a chunk of planes.  The x and y stencils
//            (cells i-2..i+1) are read from an LDS ring of two planes that
//            the workgroup stages cooperatively (each value read from HBM
//            once per tile, halo re-reads come from L2); the z stencil
//            (planes k-2..k+1) lives in a per-thread register queue filled as
//            the march passes each plane.  Writes the seven face fluxes of
//            each direction (21 floats per cell).
//  k_update  per cell: flux divergence, the 24 upwinded edge EMFs of the
//            cell's six faces, the conserved update, primitive recovery and
//            the shearing-box kick; in stage 2 it also reduces the CFL time
//            of the new state (interior cells) with a wave shuffle min, a
//            block min and one atomic per block.
//
// Reference correspondence (INL = library/mhd_solver.inl): calc_flux_kernel
// (INL:760-816) for the three sweeps -> k_faces; add_flux_kernel x3
// (INL:851-897), corner_magnet_flux (INL:819-840), to_conserved
// (INL:990-1007), add_pace / update (INL:1029-1071), to_primitive
// (INL:1009-1027), external_acceleration_kernel (INL:1073-1084) and
// cfl_condition_kernel (INL:714-739) -> k_update.
#include "nmhd_kernels.h"

namespace nmhd {
namespace tiled {

__device__ __forceinline__ float stage_h(const StepState *st, int hsel) {
  return hsel == H_HALF ? st->h : (hsel == H_FULL ? st->dt : st->h_diag);
}

// offset (in elements) of coordinate c+o relative to c along an axis of
// size n with periodic wrap, times the axis stride
__device__ __forceinline__ int woff(int c, int o, int n, int stride) {
  return (wrap(c + o, n) - c) * stride;
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

// ------------------------------------------------------------- k_faces

template <int TX, int TY>
__global__ void __launch_bounds__(TX *TY) k_faces(Box b, FieldPtrs W, FluxSet F,
                                                  const StepState *st, int hsel, int zc, int gx,
                                                  int gy) {
  if (!st->active) return;
  constexpr int NT = TX * TY;
  constexpr int LX = TX + 4, LY = TY + 4, LN = LX * LY;
  constexpr int NS = (LN + NT - 1) / NT;  // LDS slots staged per thread
  __shared__ float P[2][8][LN];

  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
  const int nx = b.n[0], ny = b.n[1], nz = b.n[2];
  const int plane = nx * ny;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int x0 = (lb % gx) * TX, y0 = ((lb / gx) % gy) * TY, z0 = (lb / gx / gy) * zc;
  const int x = x0 + tx, y = y0 + ty;
  const bool valid = (x < nx) && (y < ny);
  const int zend = min(z0 + zc, nz);
  const float h = stage_h(st, hsel);
  const int mode = b.recon;
  const bool src = b.source == 1;

  // global in-plane offset of every LDS slot this thread stages
  int gxy[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int li = tid + s * NT;
    const int lx = li % LX, ly = li / LX;
    gxy[s] = (li < LN) ? wrap(y0 - 2 + ly, ny) * nx + wrap(x0 - 2 + lx, nx) : 0;
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
        for (int f = 0; f < 8; ++f) P[buf][f][li] = stage[s][f];
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
    const float *R = P[cb][0];
    const float *VX = P[cb][1];
    const float *VY = P[cb][2];
    const float *VZ = P[cb][3];
    const float *PR = P[cb][4];
    const float *BX = P[cb][5];
    const float *BY = P[cb][6];
    const float *BZ = P[cb][7];
    const float *BZN = P[nb][7];

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

// ------------------------------------------------------------ k_update

__device__ __forceinline__ float upwind(float s, float f0, float f1) {
  return s > 0.0f ? f0 : (s < 0.0f ? f1 : 0.5f * (f0 + f1));
}

__device__ __forceinline__ float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}

template <bool PaceOnly, bool Cfl>
__global__ void __launch_bounds__(256) k_update(Box b, FieldPtrs Win, FluxSet F, FieldPtrs Wb,
                                                FieldPtrs Wo, StepState *st, int hsel, int gx,
                                                int gy) {
  if (!st->active) return;
  const int nx = b.n[0], ny = b.n[1], nz = b.n[2];
  const int plane = nx * ny;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int x = (lb % gx) * blockDim.x + threadIdx.x;
  const int y = ((lb / gx) % gy) * blockDim.y + threadIdx.y;
  const int z = lb / gx / gy;
  const bool valid = x < nx && y < ny;
  float cfl_t = 3.0e38f;
  if (valid) {
    const int a = (z * ny + y) * nx + x;
    const int xm = woff(x, -1, nx, 1), xp = woff(x, 1, nx, 1);
    const int ym = woff(y, -1, ny, nx), yp = woff(y, 1, ny, nx);
    const int zm = woff(z, -1, nz, plane), zp = woff(z, 1, nz, plane);
    const float *vx = Win.f[1], *vy = Win.f[2], *vz = Win.f[3];

    // divergence of rho, m, E (X, Y, Z order)
    float Pc[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      float s = (F.dir[0].f[k][a] - F.dir[0].f[k][a + xp]) * b.inv_d[0];
      s += (F.dir[1].f[k][a] - F.dir[1].f[k][a + yp]) * b.inv_d[1];
      s += (F.dir[2].f[k][a] - F.dir[2].f[k][a + zp]) * b.inv_d[2];
      Pc[k] = s;
    }

    // winds around the cell: w[di+1][dj+1] in the xy plane, etc.
    // Z-edges (x-face i+u, y-face j+v), u,v in {0,1}
    float vx_xy[3][3], vy_xy[3][3];
    const int ox[3] = {xm, 0, xp}, oy[3] = {ym, 0, yp}, oz[3] = {zm, 0, zp};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        vx_xy[i][j] = vx[a + ox[i] + oy[j]];
        vy_xy[i][j] = vy[a + ox[i] + oy[j]];
      }
    float vx_xz[3][3], vz_xz[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        vx_xz[i][k] = (k == 1) ? vx_xy[i][1] : vx[a + ox[i] + oz[k]];
        vz_xz[i][k] = vz[a + ox[i] + oz[k]];
      }
    float vy_yz[3][3], vz_yz[3][3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        vy_yz[j][k] = (k == 1) ? vy_xy[1][j] : vy[a + oy[j] + oz[k]];
        vz_yz[j][k] = (j == 1) ? vz_xz[1][k] : vz[a + oy[j] + oz[k]];
      }

    // edge EMFs; index u (0: lower, 1: upper) of the edge along each axis.
    // E_XY[u][v]: x-flux of B_y at x-face i+u, upwinded in y across y-face j+v
    // E_YX[u][v]: y-flux of B_x at y-face j+v, upwinded in x across x-face i+u
    float EXY[2][2], EYX[2][2], EXZ[2][2], EZX[2][2], EYZ[2][2], EZY[2][2];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        // winds at cells (i+u-1 .. i+u) x (j+v-1 .. j+v)
        const float sy = vy_xy[u + 1][v + 1] + vy_xy[u + 1][v] + vy_xy[u][v + 1] + vy_xy[u][v];
        const float sx = vx_xy[u + 1][v + 1] + vx_xy[u][v + 1] + vx_xy[u + 1][v] + vx_xy[u][v];
        const int ax = a + (u ? xp : 0), ay = a + (v ? yp : 0);
        EXY[u][v] = upwind(sy, F.dir[0].f[5][ax + (v ? 0 : ym)], F.dir[0].f[5][ax + (v ? yp : 0)]);
        EYX[u][v] = upwind(sx, F.dir[1].f[6][ay + (u ? 0 : xm)], F.dir[1].f[6][ay + (u ? xp : 0)]);
        // x-face i+u, z-face k+v
        const float sz = vz_xz[u + 1][v + 1] + vz_xz[u + 1][v] + vz_xz[u][v + 1] + vz_xz[u][v];
        const float sx2 = vx_xz[u + 1][v + 1] + vx_xz[u][v + 1] + vx_xz[u + 1][v] + vx_xz[u][v];
        const int az = a + (v ? zp : 0);
        EXZ[u][v] = upwind(sz, F.dir[0].f[6][ax + (v ? 0 : zm)], F.dir[0].f[6][ax + (v ? zp : 0)]);
        EZX[u][v] = upwind(sx2, F.dir[2].f[5][az + (u ? 0 : xm)], F.dir[2].f[5][az + (u ? xp : 0)]);
        // y-face j+u, z-face k+v
        const float sz2 = vz_yz[u + 1][v + 1] + vz_yz[u + 1][v] + vz_yz[u][v + 1] + vz_yz[u][v];
        const float sy2 = vy_yz[u + 1][v + 1] + vy_yz[u][v + 1] + vy_yz[u + 1][v] + vy_yz[u][v];
        const int ay2 = a + (u ? yp : 0);
        EYZ[u][v] = upwind(sz2, F.dir[1].f[5][ay2 + (v ? 0 : zm)], F.dir[1].f[5][ay2 + (v ? zp : 0)]);
        EZY[u][v] = upwind(sy2, F.dir[2].f[6][az + (u ? 0 : ym)], F.dir[2].f[6][az + (u ? yp : 0)]);
      }
    // face right-hand sides at the lower (u = 0) and upper (u = 1) faces
    const float idx = b.inv_d[0], idy = b.inv_d[1], idz = b.inv_d[2];
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
      const float h = stage_h(st, hsel);
      const float rho0 = Wb.f[0][a];
      const float vx0 = Wb.f[1][a], vy0 = Wb.f[2][a], vz0 = Wb.f[3][a];
      const float bxl0 = Wb.f[5][a], bxh0 = Wb.f[5][a + xp];
      const float byl0 = Wb.f[6][a], byh0 = Wb.f[6][a + yp];
      const float bzl0 = Wb.f[7][a], bzh0 = Wb.f[7][a + zp];
      const float e0 = Wb.f[4][a] * b.inv_gm1 +
                       0.5f * rho0 * (vx0 * vx0 + vy0 * vy0 + vz0 * vz0) +
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
      const float pr = eint * b.gm1;
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
      Wo.f[4][a] = pr;
      Wo.f[5][a] = bxl;
      Wo.f[6][a] = byl;
      Wo.f[7][a] = bzl;
      if (Cfl) {
        // CFL time of the new state for cells the next ghost fill leaves
        // untouched (interior rows/planes); the ghost cells are reduced by
        // k_cfl_ghosts after the fill
        const bool ghost = (y < b.m[1]) || (y >= ny - b.m[1]) ||
                           (b.m[2] > 0 && (z < b.m[2] || z >= nz - b.m[2] - 1));
        if (!ghost) {
          // same rounding as k_cfl, so an incrementally reduced dt equals a
          // full CFL pass over the same state bit for bit
#pragma clang fp contract(off)
          const float b2 = fmaxf(bxl * bxl, bxh * bxh) + fmaxf(byl * byl, byh * byh) +
                           fmaxf(bzl * bzl, bzh * bzh);
          const float cs = sqrtf((b.gamma * pr + b2) / rho);
          cfl_t = fminf(b.cfl * b.d[0] / fabsf(ux + cs),
                        fminf(b.cfl * b.d[1] / fabsf(uy + cs), b.cfl * b.d[2] / fabsf(uz + cs)));
        }
      }
    }
  }
  if (Cfl && !PaceOnly) {
    __shared__ float red[4];
    const float m = wave_min(cfl_t);
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    if ((tid & 63) == 0) red[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
      float r = red[0];
      for (int i = 1; i < (int)((blockDim.x * blockDim.y) >> 6); ++i) r = fminf(r, red[i]);
      atomicMin(&st->cfl_next_bits, __float_as_uint(r));
    }
  }
}

// ------------------------------------------------------- k_update_march
//
// The cell update as a z march: a TX x TY column tile plus a one-cell ring
// is staged in LDS rings, one plane ahead (prefetched into registers while
// the current plane is computed), so every neighbour access of the EMFs and
// the divergence is an LDS read and each HBM value is fetched once per tile.
//   planes r-1, r, r+1 : v_x, v_y, v_z, x-flux of B_z, y-flux of B_y... see
//   the array list below; faces r, r+1 : z-fluxes of B_x, B_y; plane r: the
//   hydro x/y fluxes, the x-flux of B_y, the y-flux of B_x and the base B_x,
//   B_y (upper faces).  The z-flux hydro components and the base cell values
//   are the thread's own and stay in registers.
template <int TX, int TY, bool PaceOnly, bool Cfl>
__global__ void __launch_bounds__(TX *TY) k_update_march(Box b, FieldPtrs Win, FluxSet F,
                                                         FieldPtrs Wb, FieldPtrs Wo,
                                                         StepState *st, int hsel, int zc,
                                                         int gx, int gy) {
  if (!st->active) return;
  constexpr int NT = TX * TY;
  constexpr int RX = TX + 2, RN = RX * (TY + 2);
  constexpr int NR = (RN + NT - 1) / NT;  // ring slots per thread
  // 3-plane rings (slot p % 3): vx, vy, vz, Fx6 (x-flux of B_z), Fy5 (y-flux of B_z)
  __shared__ float S3[5][3][RN];
  // 2-face ring (slot p & 1): Fz5 (z-flux of B_x), Fz6 (z-flux of B_y)
  __shared__ float SZ[2][2][RN];
  // 2-plane ring (slot p & 1): Fx0..4, Fx5 (B_y), Fy0..4, Fy6 (B_x), base Bx, By
  __shared__ float SP[2][14][RN];

  const int tid = threadIdx.y * TX + threadIdx.x;
  const int nx = b.n[0], ny = b.n[1], nz = b.n[2];
  const int plane = nx * ny;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int x0 = (lb % gx) * TX, y0 = ((lb / gx) % gy) * TY, z0 = (lb / gx / gy) * zc;
  const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
  const bool valid = x < nx && y < ny;
  const int zend = min(z0 + zc, nz);
  const float h = stage_h(st, hsel);

  int gxy[NR];
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    const int li = tid + s * NT;
    gxy[s] = (li < RN) ? wrap(y0 - 1 + li / RX, ny) * nx + wrap(x0 - 1 + li % RX, nx) : 0;
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
        for (int k = 0; k < 5; ++k) S3[k][slot][tid + s * NT] = r3[s][k];
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
        for (int k = 0; k < 2; ++k) SZ[k][p & 1][tid + s * NT] = rz[s][k];
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
        for (int k = 0; k < 14; ++k) SP[p & 1][k][tid + s * NT] = rp[s][k];
  };
  // own-column registers: z-flux hydro at faces r, r+1 and base cell values
  const int own = x < nx && y < ny ? y * nx + x : 0;
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

  const int c = (threadIdx.y + 1) * RX + threadIdx.x + 1;  // own slot in the ring tile
  const float idx = b.inv_d[0], idy = b.inv_d[1], idz = b.inv_d[2];
  float cfl_t = 3.0e38f;

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
    const float *VX[3] = {S3[0][sm], S3[0][s0], S3[0][sp]};
    const float *VY[3] = {S3[1][sm], S3[1][s0], S3[1][sp]};
    const float *VZ[3] = {S3[2][sm], S3[2][s0], S3[2][sp]};
    const float *FX6[3] = {S3[3][sm], S3[3][s0], S3[3][sp]};
    const float *FY5[3] = {S3[4][sm], S3[4][s0], S3[4][sp]};
    const float *FZ5[2] = {SZ[0][f0], SZ[0][f1]};
    const float *FZ6[2] = {SZ[1][f0], SZ[1][f1]};
    const float *const *PP = nullptr;
    (void)PP;
    const float *FXk[6], *FYk[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      FXk[k] = SP[pr_][k];
      FYk[k] = SP[pr_][6 + k];
    }
    const float *BXB = SP[pr_][12], *BYB = SP[pr_][13];

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
      // edge EMFs (see k_update for the index conventions)
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
          const bool ghost = (y < b.m[1]) || (y >= ny - b.m[1]) ||
                             (b.m[2] > 0 && (r < b.m[2] || r >= nz - b.m[2] - 1));
          if (!ghost) {
            // same rounding as k_cfl (no contraction)
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
  if (Cfl && !PaceOnly) {
    __shared__ float red[NT / 64];
    const float m = wave_min(cfl_t);
    if ((tid & 63) == 0) red[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
      float rr = red[0];
      for (int i = 1; i < NT / 64; ++i) rr = fminf(rr, red[i]);
      atomicMin(&st->cfl_next_bits, __float_as_uint(rr));
    }
  }
}

}  // namespace tiled

// ------------------------------------------------------------- launchers

int tiled_zchunk(const Box &b) {
  // enough workgroups to fill the chip, z chunks of at least 8 planes
  const int nz = b.n[2];
  int zc = 32;
  while (zc > 8 && (long)((b.n[0] + 63) / 64) * ((b.n[1] + 3) / 4) * ((nz + zc - 1) / zc) < 2048)
    zc /= 2;
  return zc;
}

void launch_faces(hipStream_t s, const Box &b, FieldPtrs W, const FluxSet &F, const StepState *st,
                  int hsel) {
  const int zc = tiled_zchunk(b);
  constexpr int TX = 64, TY = 4;
  const int gx = (b.n[0] + TX - 1) / TX, gy = (b.n[1] + TY - 1) / TY, gz = (b.n[2] + zc - 1) / zc;
  hipLaunchKernelGGL((tiled::k_faces<TX, TY>), dim3(gx * gy * gz), dim3(TX, TY), 0, s, b, W, F, st,
                     hsel, zc, gx, gy);
}

void launch_update_tiled(hipStream_t s, const Box &b, FieldPtrs Win, const FluxSet &F,
                         FieldPtrs Wbase, FieldPtrs Wout, StepState *st, int hsel, bool pace_only,
                         bool cfl) {
  {
    constexpr int TX = 32, TY = 8;
    const int zc = tiled_zchunk(b);
    const int gx = (b.n[0] + TX - 1) / TX, gy = (b.n[1] + TY - 1) / TY,
              gz = (b.n[2] + zc - 1) / zc;
    const dim3 grid(gx * gy * gz), blk(TX, TY);
    if (pace_only)
      hipLaunchKernelGGL((tiled::k_update_march<TX, TY, true, false>), grid, blk, 0, s, b, Win, F,
                         Wbase, Wout, st, hsel, zc, gx, gy);
    else if (cfl)
      hipLaunchKernelGGL((tiled::k_update_march<TX, TY, false, true>), grid, blk, 0, s, b, Win, F,
                         Wbase, Wout, st, hsel, zc, gx, gy);
    else
      hipLaunchKernelGGL((tiled::k_update_march<TX, TY, false, false>), grid, blk, 0, s, b, Win,
                         F, Wbase, Wout, st, hsel, zc, gx, gy);
    return;
  }
  const int gx = (b.n[0] + 63) / 64, gy = (b.n[1] + 3) / 4;
  const dim3 grid(gx * gy * b.n[2]), blk(64, 4);
  if (pace_only)
    hipLaunchKernelGGL((tiled::k_update<true, false>), grid, blk, 0, s, b, Win, F, Wbase, Wout, st,
                       hsel, gx, gy);
  else if (cfl)
    hipLaunchKernelGGL((tiled::k_update<false, true>), grid, blk, 0, s, b, Win, F, Wbase, Wout, st,
                       hsel, gx, gy);
  else
    hipLaunchKernelGGL((tiled::k_update<false, false>), grid, blk, 0, s, b, Win, F, Wbase, Wout, st,
                       hsel, gx, gy);
}

}  // namespace nmhd
