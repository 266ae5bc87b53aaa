// ghost_fill.h -- the ghost fill of one cell and the CFL time of one ghost
// cell, shared by the single-box kernels (kernels_common.hip) and the batched
// ones (kernels_batch.hip), so that a box's ghosts are filled bit for bit the
// same way in both.
//
// Reference correspondence (INL = library/mhd_solver.inl):
//   ghost_y_cell   communicate_gather_kernel_y + communicate_scatter_kernel_y
//                  (INL:90-286) with relative_displacement (INL:288-295)
//   ghost_z_cell   the Z halo copies of communicate() (INL:446-708)
#pragma once

#include "nmhd_kernels.h"

namespace nmhd {

// relative_displacement (INL:288-295) with D0 = 0 and
// relvel = -q*Omega*kLocalExtentY (application/kepler_disk_demo/
// mhd_solver.curb:46-52), in FP64 as the reference evaluates it.
__host__ __device__ inline void displacement(const Box &b, double t, float *normalized,
                                             int *n, float *minor) {
  const float relvel = -b.qshear * b.omega * b.local_ext[1];
  const double d0 = 0.0 + t * (double)relvel;
  const int fold = (int)floor(d0 / (double)b.ext[0]);
  const double d1 = d0 - (double)((float)fold * b.ext[0]);
  *normalized = (float)d1;
  *n = (int)floor(d1 / (double)b.d[0]);
  *minor = (float)(d1 / (double)b.d[0] - *n);
}

// One (x, row, z) of the M-row Y ghost slab, a in [0, nx * M * nz): the
// lower ghost row j takes the top interior row j shifted by +n along X, the
// upper ghost row takes the bottom interior row shifted by -n-1, v_X offset
// by -/+ relvel; the first upper ghost By face (the outer radial face) is
// kept.  With fix_interp the shifted values are interpolated between the
// cells at +n and +n+1 (-n-1 and -n) with the gather kernel's weights
// (INL:100-110).  t: the box's committed time.
__device__ __forceinline__ void ghost_y_cell(const Box &b, const FieldPtrs &W, double t, long a) {
#pragma clang fp contract(off)
  const int nx = b.n[0], ny = b.n[1], M = b.m[1];
  // 32-bit index arithmetic: the slab holds fewer than 2^31 cells
  const unsigned u = (unsigned)a, row = u / (unsigned)nx;
  const int sx = (int)(u - row * (unsigned)nx);
  const int sz = (int)(row / (unsigned)M);
  const int sy = (int)(row - (unsigned)sz * (unsigned)M);
  int n = 0;
  float dr = 0.0f, nd, relvel = 0.0f;
  int ix = sx, dx = sx, ix1 = sx, dx1 = sx;
  if (b.shear_y) {
    displacement(b, t, &nd, &n, &dr);
    relvel = -b.qshear * b.omega * b.local_ext[1];
    n %= nx;  // same shift, and no int overflow in sx + n for huge t
    ix = ((sx + n) % nx + nx) % nx;
    dx = ((sx - n - 1) % nx + nx) % nx;
    ix1 = ix + 1 < nx ? ix + 1 : 0;
    dx1 = dx + 1 < nx ? dx + 1 : 0;
  }
  const long src_inc = addr(b, ix, ny - 2 * M + sy, sz);
  const long src_dec = addr(b, dx, M + sy, sz);
  const long src_inc1 = addr(b, ix1, ny - 2 * M + sy, sz);
  const long src_dec1 = addr(b, dx1, M + sy, sz);
  const long dst_inc = addr(b, sx, sy, sz);
  const long dst_dec = addr(b, sx, ny - M + sy, sz);
  const bool interp = b.shear_y && b.fix_interp;
  for (int f = 0; f < 8; ++f) {
    float vi = W.f[f][src_inc];
    float vd = W.f[f][src_dec];
    // the second source cells (the reference's val_inc1 / val_dec1)
    const float vi1 = interp ? W.f[f][src_inc1] : vi;
    const float vd1 = interp ? W.f[f][src_dec1] : vd;
    if (b.shear_y) {
      // (1-d)*v + d*v, exactly as the gather kernel writes it (INL:108-110);
      // with fix_interp (1-d)*v0 + d*v1 and d*v0 + (1-d)*v1
      float bi = (1.0f - dr) * vi + dr * vi1;
      float bd = dr * vd + (1.0f - dr) * vd1;
      if (f == 1) {
        bi = bi - relvel;
        bd = bd + relvel;
      }
      vi = bi;
      vd = bd;
    }
    W.f[f][dst_inc] = vi;
    if (!(f == 6 && sy == 0)) W.f[f][dst_dec] = vd;
  }
}

// Z ghost planes, a in [0, nx * ny * M): [0, M) <- [nz-2M, nz-M),
// [nz-M, nz) <- [M, 2M).
__device__ __forceinline__ void ghost_z_cell(const Box &b, const FieldPtrs &W, long a) {
  const long unit = (long)b.n[0] * b.n[1] * b.m[2];
  for (int f = 0; f < 8; ++f) {
    float *p = W.f[f];
    p[a] = p[b.ncell - 2 * unit + a];
    p[b.ncell - unit + a] = p[unit + a];
  }
}

// Cells a ghost fill rewrites, for the CFL reduction: slab 0, the Y ghost
// rows (nx * 2My * nz cells); slab 1, the Z ghost planes plus the last
// interior plane, whose upper B_z face is a ghost face (nx * ny * (2Mz + 1)).
__device__ __forceinline__ long ghost_cfl_count(const Box &b, int slab) {
  const int M = slab == 0 ? b.m[1] : b.m[2];
  if (M == 0) return 0;
  return slab == 0 ? (long)b.n[0] * 2 * M * b.n[2] : (long)b.n[0] * b.n[1] * (2 * M + 1);
}

// CFL time of the i-th of those cells.
__device__ __forceinline__ float ghost_cfl_cell(const Box &b, const FieldPtrs &W, int slab,
                                                long i) {
  const int nx = b.n[0], ny = b.n[1], nz = b.n[2];
  const int M = slab == 0 ? b.m[1] : b.m[2];
  // 32-bit index arithmetic: a box holds fewer than 2^31 cells
  const unsigned u = (unsigned)i, row = u / (unsigned)nx;
  const int x = (int)(u - row * (unsigned)nx);
  int y, z;
  if (slab == 0) {
    const unsigned m2 = 2u * (unsigned)M;
    z = (int)(row / m2);
    const int r = (int)(row - (unsigned)z * m2);
    y = r < M ? r : ny - 2 * M + r;
  } else {
    const int r = (int)(row / (unsigned)ny);
    y = (int)(row - (unsigned)r * (unsigned)ny);
    z = r < M ? r : nz - 2 * M - 1 + r;
  }
  return cell_cfl(b, W, addr(b, x, y, z), x, y, z);
}

}  // namespace nmhd
