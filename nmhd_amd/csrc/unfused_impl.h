// unfused_impl.h -- the straightforward kernel path: one face-flux kernel per
// sweep direction writing the seven face fluxes to HBM, then one update
// kernel per stage that forms the whole right-hand side (flux divergence of
// the three sweeps + constrained-transport EMFs), the conserved update, the
// primitive recovery and the source kick.
//
// Instantiated twice: kernels_unfused.hip (fast arithmetic: hardware
// reciprocal / square root, FMA contraction) and kernels_strict.hip (built
// with -ffp-contract=off; IEEE division and square root, every expression in
// the reference's order and the reference's FP64-promoted HLLD literals), the
// latter being the verification path that reproduces the CPU oracle bit for
// bit.
//
// Reference correspondence (INL = library/mhd_solver.inl):
//   k_flux<D>    calc_flux_kernel<EX> (INL:760-816)
//   k_update     clear_pace (INL:741-755) + add_flux_kernel x3 (INL:851-897,
//                corner_magnet_flux INL:819-840) + to_conserved
//                (INL:990-1007) + add_pace / update (INL:1029-1071) +
//                to_primitive (INL:1009-1027) + external_acceleration_kernel
//                (INL:1073-1084), one pass over the cells.
// Every kernel covers the whole array, ghost cells included, with periodic
// wrap on every axis, as the reference's CRYSTAL_MAP kernels do.
#pragma once

#include "nmhd_kernels.h"

namespace nmhd {
namespace unfused {

struct Idx {
  int v[3];
};

__device__ __forceinline__ long addr_i(const Box &b, Idx i) {
  return addr(b, wrap(i.v[0], b.n[0]), wrap(i.v[1], b.n[1]), wrap(i.v[2], b.n[2]));
}
__device__ __forceinline__ Idx shift(Idx i, int d, int s) {
  i.v[d] += s;
  return i;
}

__device__ __forceinline__ float stage_h(const StepState *st, int hsel) {
  return hsel == H_HALF ? st->h : (hsel == H_FULL ? st->dt : st->h_diag);
}

// (a - b) / d: a multiply by 1/d on the fast path, a division when strict.
template <bool Strict>
__device__ __forceinline__ float over_d(float x, float d, float inv_d) {
  if constexpr (Strict)
    return x / d;
  else
    return x * inv_d;
}

// Face fluxes across the lower D-face of every cell.
template <int D, bool Strict>
__global__ void __launch_bounds__(256) k_flux(Box b, FieldPtrs W, FluxOut F, const StepState *st,
                                              int hsel) {
  if (!st->active) return;
  constexpr int EY = (D + 1) % 3, EZ = (D + 2) % 3;
  const long a = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= b.ncell) return;
  Idx c;
  c.v[0] = (int)(a % b.n[0]);
  c.v[1] = (int)((a / b.n[0]) % b.n[1]);
  c.v[2] = (int)(a / b.n[0] / b.n[1]);
  long s[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) s[k] = addr_i(b, shift(c, D, k - 2));
  float q[7][4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    q[0][k] = W.f[0][s[k]];
    q[1][k] = W.f[1 + D][s[k]];
    q[2][k] = W.f[1 + EY][s[k]];
    q[3][k] = W.f[1 + EZ][s[k]];
    q[4][k] = W.f[4][s[k]];
    const Idx ck = shift(c, D, k - 2);
    q[5][k] = 0.5f * (W.f[5 + EY][s[k]] + W.f[5 + EY][addr_i(b, shift(ck, EY, 1))]);
    q[6][k] = 0.5f * (W.f[5 + EZ][s[k]] + W.f[5 + EZ][addr_i(b, shift(ck, EZ, 1))]);
  }
  float L[7], R[7];
#pragma unroll
  for (int v = 0; v < 7; ++v) {
    if (b.recon == 1) {
      L[v] = q[v][1];
      R[v] = q[v][2];
    } else {
      if constexpr (Strict)
        recon_minbee2(q[v][0], q[v][1], q[v][2], q[v][3], L[v], R[v]);
      else
        recon_minbee2_med3(q[v][0], q[v][1], q[v][2], q[v][3], L[v], R[v]);
    }
  }
  if (b.source == 1 && D != 2) {
    // face predictor v_n += 0.5 h a_n, a from the two cells' mean velocity at
    // the centre of the face's own index (INL:793-797), or at the face
    // itself with NMHD_FIX_FACE_POSITION
    const float h = stage_h(st, hsel);
    float acc;
    if (D == 0)
      acc = accel_x(b, 0.5f * (W.f[2][s[1]] + W.f[2][s[2]]));
    else
      acc = accel_y_face(b, 0.5f * (W.f[1][s[1]] + W.f[1][s[2]]), c.v[1]);
    const float mbv = 0.5f * h * acc;
    L[1] += mbv;
    R[1] += mbv;
  }
  Flux7 r;
  if constexpr (Strict)
    r = hlld_strict(b.gamma, W.f[5 + D][a], L[0], R[0], L[1], R[1], L[2], R[2], L[3], R[3],
                    L[4], R[4], L[5], R[5], L[6], R[6]);
  else
    r = hlld(b.gamma, b.inv_gm1, W.f[5 + D][a], L[0], R[0], L[1], R[1], L[2], R[2], L[3], R[3],
             L[4], R[4], L[5], R[5], L[6], R[6]);
  F.f[0][a] = r.rho;
  F.f[1 + D][a] = r.mn;
  F.f[1 + EY][a] = r.mt1;
  F.f[1 + EZ][a] = r.mt2;
  F.f[4][a] = r.e;
  F.f[5][a] = r.bt1;
  F.f[6][a] = r.bt2;
}

// corner_magnet_flux (INL:819-840): the D-flux of B_E at the edge whose
// upper cells along D and E are at index e, upwinded along E by the sign of
// the four surrounding v_E.
template <int D, int E>
__device__ __forceinline__ float edge_flux(const Box &b, const FieldPtrs &W, const FluxSet &F,
                                           Idx e) {
  constexpr int comp = (E == (D + 1) % 3) ? 5 : 6;
  const float *w = W.f[1 + E];
  const Idx eE = shift(e, E, -1);
  const float sum = w[addr_i(b, e)] + w[addr_i(b, eE)] + w[addr_i(b, shift(e, D, -1))] +
                    w[addr_i(b, shift(eE, D, -1))];
  const float f0 = F.dir[D].f[comp][addr_i(b, eE)];
  const float f1 = F.dir[D].f[comp][addr_i(b, e)];
  return sum > 0.0f ? f0 : (sum < 0.0f ? f1 : 0.5f * (f0 + f1));
}

// Right-hand side of the face B_A on the lower A-face of cell c, summed over
// the sweeps in X, Y, Z order with the reference's 0.5 factor (INL:869-896).
template <int A, bool Strict>
__device__ __forceinline__ float face_pace(const Box &b, const FieldPtrs &W, const FluxSet &F,
                                           Idx c) {
  float p = 0.0f;
#pragma unroll
  for (int D = 0; D < 3; ++D) {
    if (D == A) {
      // "ox": += 0.5 * (diff along next(D) / d + diff along prev(D) / d)
      constexpr int NY = (A + 1) % 3, NZ = (A + 2) % 3;
      const float dy_ = edge_flux<A, NY>(b, W, F, shift(c, NY, 1)) - edge_flux<A, NY>(b, W, F, c);
      const float dz_ = edge_flux<A, NZ>(b, W, F, shift(c, NZ, 1)) - edge_flux<A, NZ>(b, W, F, c);
      p += 0.5f * (over_d<Strict>(dy_, b.d[NY], b.inv_d[NY]) +
                   over_d<Strict>(dz_, b.d[NZ], b.inv_d[NZ]));
    } else {
      // "oy"/"oz" of sweep D: -= 0.5 * (E(c + e_D) - E(c)) / d_D
      float diff;
      if (D == 0)
        diff = edge_flux<0, A>(b, W, F, shift(c, 0, 1)) - edge_flux<0, A>(b, W, F, c);
      else if (D == 1)
        diff = edge_flux<1, A>(b, W, F, shift(c, 1, 1)) - edge_flux<1, A>(b, W, F, c);
      else
        diff = edge_flux<2, A>(b, W, F, shift(c, 2, 1)) - edge_flux<2, A>(b, W, F, c);
      p -= over_d<Strict>(0.5f * diff, b.d[D], b.inv_d[D]);
    }
  }
  return p;
}

template <bool PaceOnly, bool Strict>
__global__ void __launch_bounds__(256) k_update(Box b, FieldPtrs Win, FluxSet F, FieldPtrs Wb,
                                                FieldPtrs Wo, const StepState *st, int hsel) {
  if (!st->active) return;
  const long a = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= b.ncell) return;
  Idx c;
  c.v[0] = (int)(a % b.n[0]);
  c.v[1] = (int)((a / b.n[0]) % b.n[1]);
  c.v[2] = (int)(a / b.n[0] / b.n[1]);
  // flux divergence of rho, m, E (X, Y, Z sweep order)
  float P[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int D = 0; D < 3; ++D) {
    const long r = addr_i(b, shift(c, D, 1));
#pragma unroll
    for (int k = 0; k < 5; ++k)
      P[k] += over_d<Strict>(F.dir[D].f[k][a] - F.dir[D].f[k][r], b.d[D], b.inv_d[D]);
  }
  const float pbx = face_pace<0, Strict>(b, Win, F, c);
  const float pby = face_pace<1, Strict>(b, Win, F, c);
  const float pbz = face_pace<2, Strict>(b, Win, F, c);
  if (PaceOnly) {
#pragma unroll
    for (int k = 0; k < 5; ++k) Wo.f[k][a] = P[k];
    Wo.f[5][a] = pbx;
    Wo.f[6][a] = pby;
    Wo.f[7][a] = pbz;
    return;
  }
  const float h = stage_h(st, hsel);
  // upper faces of the cell, needed for the face-averaged magnetic energy
  const Idx cx = shift(c, 0, 1), cy = shift(c, 1, 1), cz = shift(c, 2, 1);
  const long ax = addr_i(b, cx), ay = addr_i(b, cy), az = addr_i(b, cz);
  const float pbx_h = face_pace<0, Strict>(b, Win, F, cx);
  const float pby_h = face_pace<1, Strict>(b, Win, F, cy);
  const float pbz_h = face_pace<2, Strict>(b, Win, F, cz);

  // conserved base from Wb (to_conserved)
  const float rho0 = Wb.f[0][a];
  const float vx0 = Wb.f[1][a], vy0 = Wb.f[2][a], vz0 = Wb.f[3][a];
  const float bxl0 = Wb.f[5][a], bxh0 = Wb.f[5][ax];
  const float byl0 = Wb.f[6][a], byh0 = Wb.f[6][ay];
  const float bzl0 = Wb.f[7][a], bzh0 = Wb.f[7][az];
  float eth0;
  if constexpr (Strict)
    eth0 = Wb.f[4][a] / b.gm1;
  else
    eth0 = Wb.f[4][a] * b.inv_gm1;
  const float e0 = eth0 + 0.5f * rho0 * (vx0 * vx0 + vy0 * vy0 + vz0 * vz0) +
                   0.25f * (bxl0 * bxl0 + bxh0 * bxh0 + byl0 * byl0 + byh0 * byh0 +
                            bzl0 * bzl0 + bzh0 * bzh0);
  // update (add_pace / update_kernel)
  const float mx0 = rho0 * vx0, my0 = rho0 * vy0, mz0 = rho0 * vz0;
  const float rho = rho0 + h * P[0];
  const float mx = mx0 + h * P[1];
  const float my = my0 + h * P[2];
  const float mz = mz0 + h * P[3];
  const float e = e0 + h * P[4];
  const float bxl = bxl0 + h * pbx, bxh = bxh0 + h * pbx_h;
  const float byl = byl0 + h * pby, byh = byh0 + h * pby_h;
  const float bzl = bzl0 + h * pbz, bzh = bzh0 + h * pbz_h;
  // to_primitive
  float vx, vy, vz, kin;
  if constexpr (Strict) {
    vx = mx / rho;
    vy = my / rho;
    vz = mz / rho;
    kin = 0.5f * (mx * mx + my * my + mz * mz) / rho;
  } else {
    const float irho = frcp(rho);
    vx = mx * irho;
    vy = my * irho;
    vz = mz * irho;
    kin = 0.5f * (mx * mx + my * my + mz * mz) * irho;
  }
  const float eint =
      e - kin - 0.25f * (bxl * bxl + bxh * bxh + byl * byl + byh * byh + bzl * bzl + bzh * bzh);
  const float p = eint * b.gm1;
  if (b.source == 1) {
    const float axx = accel_x(b, vy);
    const float ayy = accel_y(b, vx, c.v[1]);
    vx += h * axx;
    vy += h * ayy;
  }
  Wo.f[0][a] = rho;
  Wo.f[1][a] = vx;
  Wo.f[2][a] = vy;
  Wo.f[3][a] = vz;
  Wo.f[4][a] = p;
  Wo.f[5][a] = bxl;
  Wo.f[6][a] = byl;
  Wo.f[7][a] = bzl;
}

template <bool Strict>
void launch_flux_t(hipStream_t s, int dir, const Box &b, FieldPtrs W, FluxOut F,
                   const StepState *st, int hsel) {
  const dim3 g(grid_for(b.ncell)), blk(kBlock);
  if (dir == 0)
    hipLaunchKernelGGL((k_flux<0, Strict>), g, blk, 0, s, b, W, F, st, hsel);
  else if (dir == 1)
    hipLaunchKernelGGL((k_flux<1, Strict>), g, blk, 0, s, b, W, F, st, hsel);
  else
    hipLaunchKernelGGL((k_flux<2, Strict>), g, blk, 0, s, b, W, F, st, hsel);
}

template <bool Strict>
void launch_update_t(hipStream_t s, const Box &b, FieldPtrs Win, const FluxSet &F,
                     FieldPtrs Wbase, FieldPtrs Wout, const StepState *st, int hsel,
                     bool pace_only) {
  const dim3 g(grid_for(b.ncell)), blk(kBlock);
  if (pace_only)
    hipLaunchKernelGGL((k_update<true, Strict>), g, blk, 0, s, b, Win, F, Wbase, Wout, st, hsel);
  else
    hipLaunchKernelGGL((k_update<false, Strict>), g, blk, 0, s, b, Win, F, Wbase, Wout, st, hsel);
}

}  // namespace unfused
}  // namespace nmhd
