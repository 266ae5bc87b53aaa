// kernels_diag.hip -- on-device run diagnostics: the quantities the
// reference's monitor hooks compute on the host after copying whole arrays
// back every step (application/kepler_disk_demo/mhd_solver.curb:107-130:
// <B_x^2>, <B_y^2> over every face of the array), plus conservation totals and
// the divergence of B over the interior.
//
// Two passes, both deterministic (no atomics): every block reduces its cells
// into one record of partial sums (FP64, as the reference accumulates in
// PreciseReal) and a per-block max; one block then folds the records in a
// fixed order.
#include "nmhd_kernels.h"

namespace nmhd {

namespace {

constexpr int kDiagBlock = 256;

__device__ __forceinline__ double warp_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double warp_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Block-reduce the kDiagVals values of every thread (sums, and a max in the
// last slot) into out[blockIdx.x * kDiagVals ...].
__device__ void block_fold(double (&v)[kDiagVals], double *out) {
  __shared__ double red[kDiagBlock / 64][kDiagVals];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int k = 0; k < kDiagVals; ++k) {
    const double r = (k == kDiagVals - 1) ? warp_max(v[k]) : warp_sum(v[k]);
    if (lane == 0) red[wid][k] = r;
  }
  __syncthreads();
  if (threadIdx.x < kDiagVals) {
    const int k = threadIdx.x;
    double r = red[0][k];
    for (int w = 1; w < kDiagBlock / 64; ++w)
      r = (k == kDiagVals - 1) ? fmax(r, red[w][k]) : r + red[w][k];
    out[(long)blockIdx.x * kDiagVals + k] = r;
  }
}

// Slots: 0-2 sum of B_d^2 over every face of the array; 3-10 interior sums of
// rho, m_x, m_y, m_z, E (face-averaged B^2), B_x, B_y, B_z; 11 max over the
// interior of |div B| * min(d).
__global__ void __launch_bounds__(kDiagBlock) k_diag_partial(Box b, FieldPtrs W, double *part) {
  double v[kDiagVals];
  for (int k = 0; k < kDiagVals; ++k) v[k] = 0.0;
  const int nx = b.n[0], ny = b.n[1], nz = b.n[2];
  const float dmin = fminf(b.d[0], fminf(b.d[1], b.d[2]));
  for (long a = (long)blockIdx.x * blockDim.x + threadIdx.x; a < b.ncell;
       a += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    cell_coords(b, a, x, y, z);
    const float bx = W.f[5][a], by = W.f[6][a], bz = W.f[7][a];
    // the monitor squares in the field's precision and sums in FP64
    v[0] += (double)(bx * bx);
    v[1] += (double)(by * by);
    v[2] += (double)(bz * bz);
    const bool interior = y >= b.m[1] && y < ny - b.m[1] && z >= b.m[2] && z < nz - b.m[2];
    if (!interior) continue;
    const double rho = W.f[0][a], vx = W.f[1][a], vy = W.f[2][a], vz = W.f[3][a];
    const double p = W.f[4][a];
    const double bxh = W.f[5][addr(b, wrap(x + 1, nx), y, z)];
    const double byh = W.f[6][addr(b, x, wrap(y + 1, ny), z)];
    const double bzh = W.f[7][addr(b, x, y, wrap(z + 1, nz))];
    v[3] += rho;
    v[4] += rho * vx;
    v[5] += rho * vy;
    v[6] += rho * vz;
    v[7] += p / ((double)b.gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz) +
            0.25 * ((double)bx * bx + bxh * bxh + (double)by * by + byh * byh +
                    (double)bz * bz + bzh * bzh);
    v[8] += bx;
    v[9] += by;
    v[10] += bz;
    const double div = (bxh - bx) / b.d[0] + (byh - by) / b.d[1] + (bzh - bz) / b.d[2];
    v[11] = fmax(v[11], fabs(div) * dmin);
  }
  block_fold(v, part);
}

__global__ void __launch_bounds__(kDiagBlock) k_diag_final(const double *part, int nblocks,
                                                           double *out) {
  double v[kDiagVals];
  for (int k = 0; k < kDiagVals; ++k) v[k] = 0.0;
  for (int i = threadIdx.x; i < nblocks; i += blockDim.x)
    for (int k = 0; k < kDiagVals; ++k)
      v[k] = (k == kDiagVals - 1) ? fmax(v[k], part[(long)i * kDiagVals + k])
                                  : v[k] + part[(long)i * kDiagVals + k];
  block_fold(v, out);
}

}  // namespace

int diag_blocks(const Box &b) {
  const long want = (b.ncell + kDiagBlock - 1) / kDiagBlock;
  return (int)(want < 1024 ? want : 1024);
}

void launch_diag(hipStream_t s, const Box &b, FieldPtrs W, double *part, double *out) {
  const int nb = diag_blocks(b);
  hipLaunchKernelGGL(k_diag_partial, dim3(nb), dim3(kDiagBlock), 0, s, b, W, part);
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(kDiagBlock), 0, s, part, nb, out);
}

}  // namespace nmhd
