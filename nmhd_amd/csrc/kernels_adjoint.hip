// kernels_adjoint.hip -- the hand-written reverse (adjoint) of one RK stage
// of the periodic step, U_out = cons(W_base) + h P(W_in), W_out = prim(U_out),
// with dt frozen.  Given lam_out = dL/dW_out it produces dL/dW_base and
// dL/dW_in.  No reference counterpart: the reference has no derivatives; the
// forward it reverses is proceed()'s stage (library/mhd_solver.inl:
// 990-1071 for cons/prim/update, 760-897 for the fluxes and CT, HLLD
// library/hlld.inl:7-284).
//
// Kernels of one stage reverse (all gathers: every thread writes only its
// own cell, no atomics, results independent of scheduling):
//   k_adj_prim    adjoint of prim(U) and of cons(W_base): gU (5 per cell)
//                 and lam_base; and the adjoints of the 6 CT edge-EMF families
//                 (3 distinct values per cell: lamE_x, lamE_y, lamE_z) from the
//                 face-B adjoints of the cell and its lower neighbours, which
//                 the thread recomputes rather than reading them back
//   k_adj_faces   per face of the three directions (blockIdx.y): adjoint of
//                 the 7 fluxes (divergence + upwinded EMFs, with the forward's
//                 upwind choice recomputed), then the HLLD reverse sweep
//                 (adjoint_hlld.h) -> adjoints of the 7 left/right face states
//                 and of the normal field
//   k_adj_gather  per cell: the reconstruction adjoints of the three
//                 directions, gathered from the 4 faces whose stencils hold
//                 the cell (their state adjoints and the minmod branches the
//                 face reverse recorded) -> lam_in of rho, v, p and of the
//                 cell-averaged B
//   k_adj_bfinish face B: cell-average and normal-field adjoints -> lam_in B
#include "adjoint_hlld.h"
#include "nmhd_kernels.h"

namespace nmhd {

namespace {

__device__ __forceinline__ long cell_index(const Box &b, int x, int y, int z) {
  return ((long)z * b.n[1] + y) * b.n[0] + x;
}

// neighbour of (x, y, z) shifted by s along axis d (periodic)
__device__ __forceinline__ long nb(const Box &b, int x, int y, int z, int d, int s) {
  if (d == 0) x = wrap(x + s, b.n[0]);
  if (d == 1) y = wrap(y + s, b.n[1]);
  if (d == 2) z = wrap(z + s, b.n[2]);
  return cell_index(b, x, y, z);
}

// The thread's cell; 32-bit unsigned index arithmetic (a box holds fewer
// than 2^31 cells).
__device__ __forceinline__ bool cell_of(const Box &b, int &x, int &y, int &z, long &a) {
  a = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= b.ncell) return false;
  const unsigned u = (unsigned)a, n0 = (unsigned)b.n[0], n1 = (unsigned)b.n[1];
  const unsigned row = u / n0;
  x = (int)(u - row * n0);
  z = (int)(row / n1);
  y = (int)(row - (unsigned)z * n1);
  return true;
}

// weights of the two candidates of the upwind choice s > 0 ? f0 : s < 0 ? f1 : avg
__device__ __forceinline__ float w_lo(float s) { return s > 0.0f ? 1.0f : (s < 0.0f ? 0.0f : 0.5f); }
__device__ __forceinline__ float w_hi(float s) { return s > 0.0f ? 0.0f : (s < 0.0f ? 1.0f : 0.5f); }

// The forward's sum of four cell velocities around an edge (same order as the
// stage kernels, so that the sign decision is the same).  The edge lies on
// the lower p-face and lower q-face of cell (x,y,z) (p < q axes); comp is
// the velocity component summed.
__device__ __forceinline__ float edge_sum(const Box &b, const float *v, int x, int y, int z,
                                          int p, int q, int comp) {
  const long c = cell_index(b, x, y, z);
  const long cp = nb(b, x, y, z, p, -1);
  const long cq = nb(b, x, y, z, q, -1);
  int xp = x, yp = y, zp = z;
  if (p == 0) xp = wrap(x - 1, b.n[0]);
  if (p == 1) yp = wrap(y - 1, b.n[1]);
  const long cpq = nb(b, xp, yp, zp, q, -1);
  // winds summed along the first edge axis first when comp is the second
  // edge axis' component, as the stage kernels do
  if (comp == q) return v[c] + v[cq] + v[cp] + v[cpq];
  return v[c] + v[cp] + v[cq] + v[cpq];
}

// ------------------------------------------------------------- k_adj_prim

// Adjoint of the face value B_d at cell c (its lower d-face) from the stage
// output: its own adjoint, and through the pressure of the two cells that
// share the face (c and its lower d-neighbour m).
__device__ __forceinline__ float face_b_adj(const Box &b, const FieldPtrs &Wout,
                                            const FieldPtrs &lam, int d, long c, long m) {
  return lam.f[5 + d][c] - b.gm1 * 0.5f * Wout.f[5 + d][c] * (lam.f[4][c] + lam.f[4][m]);
}

// lamE_z = adjoint of the z-edge EMF from the x-sweep (= minus that of the
// y-sweep's), at the edge on the lower x- and y-faces of the cell; likewise
// lamE_y (x-sweep / z-sweep pair on the lower x- and z-faces) and lamE_x
// (y-sweep / z-sweep pair on the lower y- and z-faces), scaled by h.
template <bool AccBase>
__global__ void __launch_bounds__(256) k_adj_prim(Box b, FieldPtrs Wout, FieldPtrs Wbase,
                                                  FieldPtrs lam, float *gU0, float *gU1,
                                                  float *gU2, float *gU3, float *gU4, float h,
                                                  float *lEx, float *lEy, float *lEz,
                                                  FieldPtrs lam_base) {
  int x, y, z;
  long a;
  if (!cell_of(b, x, y, z, a)) return;
  const float gm1 = b.gm1, ig = b.inv_gm1;
  const float rho = Wout.f[0][a];
  const float v[3] = {Wout.f[1][a], Wout.f[2][a], Wout.f[3][a]};
  const float lp = lam.f[4][a];
  const float irho = frcp(rho);
  const float gE = gm1 * lp;
  float gm[3];
  float grho = lam.f[0][a] + gm1 * lp * 0.5f * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int d = 0; d < 3; ++d) {
    gm[d] = lam.f[1 + d][a] * irho - gm1 * lp * v[d];
    grho -= lam.f[1 + d][a] * v[d] * irho;
  }
  const long xm = nb(b, x, y, z, 0, -1), ym = nb(b, x, y, z, 1, -1), zm = nb(b, x, y, z, 2, -1);
  const long mm[3] = {xm, ym, zm};
  float gB[3];
  for (int d = 0; d < 3; ++d) gB[d] = face_b_adj(b, Wout, lam, d, a, mm[d]);
  gU0[a] = grho;
  gU1[a] = gm[0];
  gU2[a] = gm[1];
  gU3[a] = gm[2];
  gU4[a] = gE;
  // the edge-EMF adjoints: face-B adjoints of the lower x-, y- and
  // z-neighbours, each through its own lower neighbour along its face axis
  {
    const float idx = b.inv_d[0], idy = b.inv_d[1], idz = b.inv_d[2];
    const float gB0_ym = face_b_adj(b, Wout, lam, 0, ym, nb(b, x, wrap(y - 1, b.n[1]), z, 0, -1));
    const float gB0_zm = face_b_adj(b, Wout, lam, 0, zm, nb(b, x, y, wrap(z - 1, b.n[2]), 0, -1));
    const float gB1_xm = face_b_adj(b, Wout, lam, 1, xm, nb(b, wrap(x - 1, b.n[0]), y, z, 1, -1));
    const float gB1_zm = face_b_adj(b, Wout, lam, 1, zm, nb(b, x, y, wrap(z - 1, b.n[2]), 1, -1));
    const float gB2_xm = face_b_adj(b, Wout, lam, 2, xm, nb(b, wrap(x - 1, b.n[0]), y, z, 2, -1));
    const float gB2_ym = face_b_adj(b, Wout, lam, 2, ym, nb(b, x, wrap(y - 1, b.n[1]), z, 2, -1));
    lEz[a] = h * (0.5f * idy * (gB0_ym - gB[0]) + 0.5f * idx * (gB[1] - gB1_xm));
    lEy[a] = h * (0.5f * idz * (gB0_zm - gB[0]) + 0.5f * idx * (gB[2] - gB2_xm));
    lEx[a] = h * (0.5f * idz * (gB1_zm - gB[1]) + 0.5f * idy * (gB[2] - gB2_ym));
  }
  // cons(W_base): rho, m = rho v, E = p/(g-1) + rho v^2/2 + (B_lo^2 + B_hi^2)/4
  const float rho0 = Wbase.f[0][a];
  const float v0[3] = {Wbase.f[1][a], Wbase.f[2][a], Wbase.f[3][a]};
  float out[8];
  out[0] = grho + 0.5f * gE * (v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2]);
  for (int d = 0; d < 3; ++d) {
    out[0] += gm[d] * v0[d];
    out[1 + d] = gm[d] * rho0 + gE * rho0 * v0[d];
  }
  out[4] = gE * ig;
  for (int d = 0; d < 3; ++d)
    out[5 + d] = gB[d] + 0.5f * Wbase.f[5 + d][a] * (gE + gm1 * lam.f[4][mm[d]]);
  for (int f = 0; f < 8; ++f) lam_base.f[f][a] = AccBase ? lam_base.f[f][a] + out[f] : out[f];
}

// ------------------------------------------------------------- k_adj_face
//
// Face = lower D-face of cell (x,y,z).  Flux components in the forward's
// order (rho, m_D, m_D+1, m_D+2, E, B_D+1, B_D+2).
//
// The face reverse hands the gather the adjoints of its two reconstructed
// states, aL[k] and aR[k] (variable k), and the limiter branches the
// forward took for them, 4 bits per variable in one word: bit 4k (4k+1)
// when the left state's limited slope took its lower (upper) difference,
// bit 4k+2 (4k+3) likewise for the right state's; no bit for a zero slope.
// The gather then spreads the adjoints over the stencil cells without
// reading the stencils again.
struct AdjFaceOut {
  float *aL[7];
  float *aR[7];
  unsigned *br;
  float *abn;
};

// Derivative of minmod(a, b) with respect to a and b on the branch the
// forward took (ties and sign changes as minmod takes them).
__device__ __forceinline__ void mm_grad(float a, float b, float &da, float &db) {
  const bool nz = !(a * b <= 0.0f);
  const bool pick_a = fabsf(a) < fabsf(b);
  da = (nz && pick_a) ? 1.0f : 0.0f;
  db = (nz && !pick_a) ? 1.0f : 0.0f;
}

template <int D>
__device__ __forceinline__ void adj_face(const Box &b, const FieldPtrs &W, const float *gU0,
                                         const float *gU1, const float *gU2, const float *gU3,
                                         const float *gU4, const float *lEx, const float *lEy,
                                         const float *lEz, float h, const AdjFaceOut &out,
                                         int x, int y, int z, long a) {
  constexpr int D1 = (D + 1) % 3, D2 = (D + 2) % 3;
  const float *gU[5] = {gU0, gU1, gU2, gU3, gU4};
  const long m1 = nb(b, x, y, z, D, -1);
  const float id = b.inv_d[D];
  adj::HOut<float> aF;
  aF.rho = h * (gU[0][a] - gU[0][m1]) * id;
  aF.mn = h * (gU[1 + D][a] - gU[1 + D][m1]) * id;
  aF.mt1 = h * (gU[1 + D1][a] - gU[1 + D1][m1]) * id;
  aF.mt2 = h * (gU[1 + D2][a] - gU[1 + D2][m1]) * id;
  aF.e = h * (gU[4][a] - gU[4][m1]) * id;
  // B fluxes through the upwinded edge EMFs.  For the x-sweep: F(B_y) feeds
  // the z-edges on y-faces j (as the upper candidate) and j+1 (lower one),
  // upwinded by the y-velocity; F(B_z) the y-edges on z-faces k and k+1.
  const float *v[3] = {W.f[1], W.f[2], W.f[3]};
  {
    // component bt1 = B_D1 flux, used by edges along D2 on D1-faces
    // component bt2 = B_D2 flux, used by edges along D1 on D2-faces
    auto edge_adj = [&](int q, int xx, int yy, int zz) -> float {
      // lam of the edge on the lower D-face and lower q-face of cell (xx,yy,zz),
      // with the sign of the D-sweep's member of the pair
      const long e = cell_index(b, xx, yy, zz);
      const int other = 3 - D - q;  // the edge's axis
      const float *arr = other == 0 ? lEx : (other == 1 ? lEy : lEz);
      // the x-sweep (or, for x-edges, the y-sweep) member carries +lam,
      // the other sweep's member -lam
      const int first = D < q ? D : q;
      return D == first ? arr[e] : -arr[e];
    };
    auto sum_at = [&](int q, int xx, int yy, int zz) -> float {
      // upwind sum of the q-velocity around the edge on the lower D- and
      // q-faces of cell (xx,yy,zz), in the forward's order
      const int p = D < q ? D : q, r = D < q ? q : D;
      return edge_sum(b, v[q], xx, yy, zz, p, r, q);
    };
    int xs[3] = {x, y, z};
    // bt1: flux of B_D1, edges on D1-faces j (this face's own row) and j+1
    {
      const int q = D1;
      int up[3] = {xs[0], xs[1], xs[2]};
      up[q] = wrap(up[q] + 1, b.n[q]);
      const float s0 = sum_at(q, xs[0], xs[1], xs[2]);
      const float s1 = sum_at(q, up[0], up[1], up[2]);
      aF.bt1 = edge_adj(q, xs[0], xs[1], xs[2]) * w_hi(s0) +
               edge_adj(q, up[0], up[1], up[2]) * w_lo(s1);
    }
    {
      const int q = D2;
      int up[3] = {xs[0], xs[1], xs[2]};
      up[q] = wrap(up[q] + 1, b.n[q]);
      const float s0 = sum_at(q, xs[0], xs[1], xs[2]);
      const float s1 = sum_at(q, up[0], up[1], up[2]);
      aF.bt2 = edge_adj(q, xs[0], xs[1], xs[2]) * w_hi(s0) +
               edge_adj(q, up[0], up[1], up[2]) * w_lo(s1);
    }
  }
  // recompute the face states (same reconstruction as the forward)
  // stencil cells c-2..c+1 along D, with their coordinates
  long cs[4];
  int cc[4][3];
  for (int s = 0; s < 4; ++s) {
    cc[s][0] = x, cc[s][1] = y, cc[s][2] = z;
    cc[s][D] = wrap(cc[s][D] + s - 2, b.n[D]);
    cs[s] = cell_index(b, cc[s][0], cc[s][1], cc[s][2]);
  }
  // cell-centred B_t of stencil cell s: the average of its two t-faces
  auto bavg = [&](int t, int s) -> float {
    return 0.5f * (W.f[5 + t][cs[s]] + W.f[5 + t][nb(b, cc[s][0], cc[s][1], cc[s][2], t, 1)]);
  };
  float q[7][4];
  for (int s = 0; s < 4; ++s) {
    q[0][s] = W.f[0][cs[s]];
    q[1][s] = W.f[1 + D][cs[s]];
    q[2][s] = W.f[1 + D1][cs[s]];
    q[3][s] = W.f[1 + D2][cs[s]];
    q[4][s] = W.f[4][cs[s]];
    q[5][s] = bavg(D1, s);
    q[6][s] = bavg(D2, s);
  }
  float Lq[7], Rq[7];
  for (int k = 0; k < 7; ++k) {
    if (b.recon == 1) {
      Lq[k] = q[k][1];
      Rq[k] = q[k][2];
    } else {
      recon_minbee2_med3(q[k][0], q[k][1], q[k][2], q[k][3], Lq[k], Rq[k]);
    }
  }
  // the limiter branches of L = x1 + m(x1 - x0, x2 - x1)/2 and
  // R = x2 - m(x2 - x1, x3 - x2)/2 over the stencil x0..x3 (cells c-2..c+1),
  // for the gather
  unsigned br = 0u;
  if (b.recon != 1) {
    for (int k = 0; k < 7; ++k) {
      float da, db, da2, db2;
      mm_grad(q[k][1] - q[k][0], q[k][2] - q[k][1], da, db);
      mm_grad(q[k][2] - q[k][1], q[k][3] - q[k][2], da2, db2);
      br |= ((da != 0.0f ? 1u : 0u) | (db != 0.0f ? 2u : 0u) | (da2 != 0.0f ? 4u : 0u) |
             (db2 != 0.0f ? 8u : 0u)) << (4 * k);
    }
  }
  adj::HIn<float> in;
  in.bx = W.f[5 + D][a];
  in.rL = Lq[0], in.rR = Rq[0], in.unL = Lq[1], in.unR = Rq[1], in.u1L = Lq[2], in.u1R = Rq[2];
  in.u2L = Lq[3], in.u2R = Rq[3], in.pL = Lq[4], in.pR = Rq[4], in.b1L = Lq[5], in.b1R = Rq[5];
  in.b2L = Lq[6], in.b2R = Rq[6];
  adj::HTape<float> tape;
  adj::HOut<float> F;
  adj::hlld_forward(b.gamma, b.inv_gm1, in, tape, F);
  adj::HIn<float> g;
  adj::hlld_reverse(b.gamma, b.inv_gm1, in, tape, aF, g);
  const float gl[7] = {g.rL, g.unL, g.u1L, g.u2L, g.pL, g.b1L, g.b2L};
  const float gr[7] = {g.rR, g.unR, g.u1R, g.u2R, g.pR, g.b1R, g.b2R};
  for (int k = 0; k < 7; ++k) {
    out.aL[k][a] = gl[k];
    out.aR[k][a] = gr[k];
  }
  out.br[a] = br;
  out.abn[a] = g.bx;
}

// All faces of the three directions in one launch: blockIdx.y is the
// direction (a uniform branch per workgroup).
struct AdjFaceOut3 {
  AdjFaceOut d[3];
};

__global__ void __launch_bounds__(256) k_adj_faces(Box b, FieldPtrs W, const float *gU0,
                                                   const float *gU1, const float *gU2,
                                                   const float *gU3, const float *gU4,
                                                   const float *lEx, const float *lEy,
                                                   const float *lEz, float h, AdjFaceOut3 out) {
  int x, y, z;
  long a;
  if (!cell_of(b, x, y, z, a)) return;
  if (blockIdx.y == 0)
    adj_face<0>(b, W, gU0, gU1, gU2, gU3, gU4, lEx, lEy, lEz, h, out.d[0], x, y, z, a);
  else if (blockIdx.y == 1)
    adj_face<1>(b, W, gU0, gU1, gU2, gU3, gU4, lEx, lEy, lEz, h, out.d[1], x, y, z, a);
  else
    adj_face<2>(b, W, gU0, gU1, gU2, gU3, gU4, lEx, lEy, lEz, h, out.d[2], x, y, z, a);
}

// ------------------------------------------------------------ k_adj_gather
//
// The reconstruction reverse of direction D at cell (x,y,z): res[0..6] =
// adjoints of rho, v_D, v_D+1, v_D+2, p, Bbar_D+1, Bbar_D+2 of the cell,
// gathered from the four faces whose stencils hold it: faces c+2, c+1, c
// and c-1 along D, where it is stencil cell 0, 1, 2 and 3.  A face's left
// state L = x1 + m/2 and right state R = x2 - m'/2 (m the limited slope of
// x0, x1, x2, m' that of x1, x2, x3) pass their adjoints aL, aR to the
// stencil cells through the branches the forward took:
//   cell 0: -aL/2 [m took x1 - x0]
//   cell 1:  aL + aL/2 ([m took x1 - x0] - [m took x2 - x1]) + aR/2 [m' took x2 - x1]
//   cell 2:  aR + aL/2 [m took x2 - x1] - aR/2 ([m' took x2 - x1] - [m' took x3 - x2])
//   cell 3: -aR/2 [m' took x3 - x2]
template <int D>
__device__ __forceinline__ void adj_gather(const Box &b, const AdjFaceOut &fa, int x, int y,
                                           int z, float (&res)[7]) {
  const long f0 = nb(b, x, y, z, D, 2), f1 = nb(b, x, y, z, D, 1), f3 = nb(b, x, y, z, D, -1);
  const long f2 = cell_index(b, x, y, z);
  const unsigned b0 = fa.br[f0], b1 = fa.br[f1], b2 = fa.br[f2], b3 = fa.br[f3];
  auto bit = [](unsigned w, int i) { return ((w >> i) & 1u) ? 1.0f : 0.0f; };
  for (int k = 0; k < 7; ++k) {
    const int s = 4 * k;
    const float l0 = fa.aL[k][f0], l1 = fa.aL[k][f1], r1 = fa.aR[k][f1];
    const float l2 = fa.aL[k][f2], r2 = fa.aR[k][f2], r3 = fa.aR[k][f3];
    const float c0 = -0.5f * l0 * bit(b0, s);
    const float c1 = l1 + 0.5f * l1 * (bit(b1, s) - bit(b1, s + 1)) + 0.5f * r1 * bit(b1, s + 2);
    const float c2 =
        r2 + 0.5f * l2 * bit(b2, s + 1) - 0.5f * r2 * (bit(b2, s + 2) - bit(b2, s + 3));
    const float c3 = -0.5f * r3 * bit(b3, s + 3);
    res[k] = c0 + c1 + c2 + c3;
  }
}

// lam_in of rho, v, p from the three directions' reconstruction reverses
// (written, or added to with accumulate), and the cell-average B adjoints
// lbb[t] (written).
__global__ void __launch_bounds__(256) k_adj_gather(Box b, AdjFaceOut3 fa, FieldPtrs lam_in,
                                                    int accumulate, float *lbb0, float *lbb1,
                                                    float *lbb2) {
  int x, y, z;
  long a;
  if (!cell_of(b, x, y, z, a)) return;
  float rho = 0.0f, v[3] = {0.0f, 0.0f, 0.0f}, p = 0.0f, bb[3] = {0.0f, 0.0f, 0.0f};
  float res[7];
  adj_gather<0>(b, fa.d[0], x, y, z, res);  // X: (rho, vx, vy, vz, p, Bby, Bbz)
  rho += res[0], v[0] += res[1], v[1] += res[2], v[2] += res[3], p += res[4];
  bb[1] += res[5], bb[2] += res[6];
  adj_gather<1>(b, fa.d[1], x, y, z, res);  // Y: (rho, vy, vz, vx, p, Bbz, Bbx)
  rho += res[0], v[1] += res[1], v[2] += res[2], v[0] += res[3], p += res[4];
  bb[2] += res[5], bb[0] += res[6];
  adj_gather<2>(b, fa.d[2], x, y, z, res);  // Z: (rho, vz, vx, vy, p, Bbx, Bby)
  rho += res[0], v[2] += res[1], v[0] += res[2], v[1] += res[3], p += res[4];
  bb[0] += res[5], bb[1] += res[6];
  const float out[5] = {rho, v[0], v[1], v[2], p};
  for (int f = 0; f < 5; ++f) lam_in.f[f][a] = accumulate ? lam_in.f[f][a] + out[f] : out[f];
  lbb0[a] = bb[0];
  lbb1[a] = bb[1];
  lbb2[a] = bb[2];
}

// Face B_t of lam_in: half of each adjacent cell's cell-average adjoint plus
// the normal-field adjoint of the t-sweep's face.
__global__ void __launch_bounds__(256) k_adj_bfinish(Box b, const float *lbb0, const float *lbb1,
                                                     const float *lbb2, const float *abn0,
                                                     const float *abn1, const float *abn2,
                                                     FieldPtrs lam_in, int accumulate) {
  int x, y, z;
  long a;
  if (!cell_of(b, x, y, z, a)) return;
  const float *lbb[3] = {lbb0, lbb1, lbb2};
  const float *abn[3] = {abn0, abn1, abn2};
  for (int t = 0; t < 3; ++t) {
    const long m = nb(b, x, y, z, t, -1);
    const float v = 0.5f * (lbb[t][a] + lbb[t][m]) + abn[t][a];
    lam_in.f[5 + t][a] = accumulate ? lam_in.f[5 + t][a] + v : v;
  }
}

// Step record of a stage with a frozen dt.
__global__ void k_adj_set_dt(StepState *st, float dt) {
  st->dt = dt;
  st->h = dt / 2;
  st->h_diag = dt;
  st->active = 1;
  st->finished = 0;
}

}  // namespace

// ------------------------------------------------------------- launchers

void launch_adj_set_dt(hipStream_t s, StepState *st, float dt) {
  hipLaunchKernelGGL(k_adj_set_dt, dim3(1), dim3(1), 0, s, st, dt);
}

void launch_adj_stage(hipStream_t s, const Box &b, const AdjStageArgs &A) {
  const dim3 g(grid_for(b.ncell)), blk(kBlock);
  if (A.acc_base)
    hipLaunchKernelGGL(k_adj_prim<true>, g, blk, 0, s, b, A.Wout, A.Wbase, A.lam_out, A.gU[0],
                       A.gU[1], A.gU[2], A.gU[3], A.gU[4], A.h, A.lE[0], A.lE[1], A.lE[2],
                       A.lam_base);
  else
    hipLaunchKernelGGL(k_adj_prim<false>, g, blk, 0, s, b, A.Wout, A.Wbase, A.lam_out, A.gU[0],
                       A.gU[1], A.gU[2], A.gU[3], A.gU[4], A.h, A.lE[0], A.lE[1], A.lE[2],
                       A.lam_base);
  AdjFaceOut3 fo;
  for (int d = 0; d < 3; ++d) {
    for (int k = 0; k < 7; ++k) {
      fo.d[d].aL[k] = A.face[d * 15 + k];
      fo.d[d].aR[k] = A.face[d * 15 + 7 + k];
    }
    fo.d[d].br = reinterpret_cast<unsigned *>(A.face[d * 15 + 14]);
    fo.d[d].abn = A.abn[d];
  }
  const dim3 g3(grid_for(b.ncell), 3);
  hipLaunchKernelGGL(k_adj_faces, g3, blk, 0, s, b, A.Win, A.gU[0], A.gU[1], A.gU[2], A.gU[3],
                     A.gU[4], A.lE[0], A.lE[1], A.lE[2], A.h, fo);
  hipLaunchKernelGGL(k_adj_gather, g, blk, 0, s, b, fo, A.lam_in, A.acc_in ? 1 : 0, A.lbb[0],
                     A.lbb[1], A.lbb[2]);
  hipLaunchKernelGGL(k_adj_bfinish, g, blk, 0, s, b, A.lbb[0], A.lbb[1], A.lbb[2], A.abn[0],
                     A.abn[1], A.abn[2], A.lam_in, A.acc_in ? 1 : 0);
}

}  // namespace nmhd
