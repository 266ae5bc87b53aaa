// nmhd_device.h -- device building blocks shared by every kernel path:
// box parameters, periodic addressing, reconstruction, the HLLD Riemann
// solver and the shearing-box acceleration.  All arithmetic is FP32.
//
// Reference correspondence (paths relative to the reference's src/ tree,
// INL = library/mhd_solver.inl, HLLD = library/hlld.inl,
// KEP = application/kepler_disk_demo/mhd_solver.curb):
//   recon_minbee2       INL:68-84   (absolute_smaller + interpolate)
//   hlld                HLLD:7-284  (Miyoshi-Kusano HLLD, fast-speed bounds)
//   accel_*             KEP:64-81   (Coriolis + tidal acceleration)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nmhd {

// Run-time box description passed by value to every kernel.
struct Box {
  int n[3];         // array size per axis incl. ghosts
  int m[3];         // ghost layers per side
  long ncell;       // n0*n1*n2
  float d[3];       // spacing
  float inv_d[3];   // 1/spacing (rounded once)
  float ext[3];     // array extent incl. ghosts
  float local_ext[3];
  float gamma;
  float inv_gm1;    // 1/(gamma-1)
  float gm1;        // gamma-1
  float cfl;
  int source;       // 0 none, 1 shearing box
  float omega, qshear;
  int shear_y;
  int recon;        // 0 minbee2, 1 piecewiseconst1
  int fix_cfl;      // CFL speed |v| + c instead of |v + c|
  int fix_interp;   // shear fill interpolates between the two source cells
  float face_off;   // radial face position in the predictor: y + face_off
                    // cells (0.5: the reference's cell centre; 0: the face)
};

// Device-side step record: dt is computed, clipped and consumed on the GPU so
// that many steps can be enqueued without host round trips.
struct StepState {
  double t;          // committed time (GlobalSetup::current_time)
  double t_next;     // time after the step in flight
  float dt;          // dt of the step in flight
  float h;           // first-stage step dt/2
  float h_diag;      // step used by the rhs diagnostic
  unsigned cfl_bits; // float bits of the running CFL minimum (atomic min)
  int finished;      // an earlier step reached goal_time
  int active;        // the step in flight executes
  int generation;
  int steps_done;
  int goal_reached;
  unsigned cfl_next_bits; // CFL minimum of the next state, reduced by the
                          // stage-2 epilogue over the cells a ghost fill
                          // does not touch
};

// Periodic wrap valid for offsets within [-2n, 2n) (RCo::calc_addr with
// CRYSTAL_PERIODIC_*, crystal.h:31-50).
__device__ __forceinline__ int wrap(int i, int n) {
  if (i < 0) i += n;
  if (i < 0) i += n;
  if (i >= n) i -= n;
  if (i >= n) i -= n;
  return i;
}

__device__ __forceinline__ long addr(const Box &b, int x, int y, int z) {
  return ((long)z * b.n[1] + y) * b.n[0] + x;
}

__device__ __forceinline__ float sq(float x) { return x * x; }

// Workgroups are dispatched to the 8 XCDs round-robin (block b -> XCD b % 8),
// and each XCD has its own L2.  Map the hardware block id to a logical tile id
// so that XCD k processes the contiguous range of tiles
// [k*(N/8) + min(k, N%8), ...): neighbouring tiles, which re-read each
// other's halos, then share an L2.  The map is a bijection on [0, N).
__device__ __forceinline__ int xcd_logical_block(int b, int N) {
  const int k = b & 7, q = N >> 3, r = N & 7;
  return k * q + (k < r ? k : r) + (b >> 3);
}

// Division and square root used on the hot path: the hardware reciprocal and
// square root (1 ulp) instead of the IEEE correctly rounded sequences.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fdiv(float a, float b) { return a * frcp(b); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// minmod of two slopes: 0 on a sign change, else the one of smaller modulus.
__device__ __forceinline__ float minmod(float a, float b) {
  return a * b <= 0.0f ? 0.0f : (fabsf(a) < fabsf(b) ? a : b);
}

// MUSCL-minmod face states from the 4-cell stencil x0 x1 | x2 x3.
__device__ __forceinline__ void recon_minbee2(float x0, float x1, float x2, float x3,
                                              float &left, float &right) {
  const float d12 = x2 - x1;
  left = x1 + 0.5f * minmod(x1 - x0, d12);
  right = x2 - 0.5f * minmod(d12, x3 - x2);
}

// The same limiter as one median-of-three instruction: med3(a, b, 0) is 0
// when a and b differ in sign and otherwise the one nearer 0, i.e. minmod.
// It differs from minmod() only where a*b underflows to 0 (minmod: 0) and
// in the sign of a zero result; the fast paths use it.
__device__ __forceinline__ float minmod_med3(float a, float b) {
  return __builtin_amdgcn_fmed3f(a, b, 0.0f);
}

__device__ __forceinline__ void recon_minbee2_med3(float x0, float x1, float x2, float x3,
                                                   float &left, float &right) {
  const float d12 = x2 - x1;
  left = x1 + 0.5f * minmod_med3(x1 - x0, d12);
  right = x2 - 0.5f * minmod_med3(d12, x3 - x2);
}

// HLLD flux across a face with normal field bx (Miyoshi & Kusano 2005) with
// the outer signal speeds min/max(vL, vR) -/+ sqrt(2) max(cfL, cfR), the
// 1e-10 guard in the single-star denominators and signBx = 0 for bx = 0.
// Inputs are the left/right states (rho, vn, vt1, vt2, p, Bt1, Bt2); outputs
// the fluxes of rho, m_n, m_t1, m_t2, E, B_t1, B_t2.
struct Flux7 {
  float rho, mn, mt1, mt2, e, bt1, bt2;
};

__device__ __forceinline__ Flux7 hlld(float gamma, float inv_gm1, float bx,
                                      float rL, float rR, float unL, float unR,
                                      float ut1L, float ut1R, float ut2L, float ut2R,
                                      float pL, float pR, float b1L, float b1R,
                                      float b2L, float b2R) {
  const float TINY = 1.0e-10f;
  const float abx = fabsf(bx);
  // sign(bx), 0 for bx = 0 (a sign-bit copy and one select)
  const float sgn = bx == 0.0f ? 0.0f : copysignf(1.0f, bx);
  const float bx2 = bx * bx;

  // outer states: pt = p + B^2/2 (the energy E = p/(gamma-1) + (rho v^2 +
  // B^2)/2 is formed below, on the side the fan region takes)
  const float B2L = bx2 + b1L * b1L + b2L * b2L;
  const float B2R = bx2 + b1R * b1R + b2R * b2R;
  const float ptL = pL + 0.5f * B2L;
  const float ptR = pR + 0.5f * B2R;

  // fast magnetosonic speeds and outer signal speeds
  const float gpL = gamma * pL, gpR = gamma * pR;
  const float gpbL = gpL + B2L, gpbR = gpR + B2R;
  const float bx2x4 = 4.0f * bx2;
  // cf^2 = a / (2 rho); sqrt(2) max(cfL, cfR) = sqrt(a / rho) of the side
  // with the larger a / rho (rho > 0): one reciprocal and one root
  const float aL = gpbL + fsqrt(gpbL * gpbL - gpL * bx2x4);
  const float aR = gpbR + fsqrt(gpbR * gpbR - gpR * bx2x4);
  const bool lmax = aL * rR >= aR * rL;
  const float cmax = fsqrt(fdiv(lmax ? aL : aR, lmax ? rL : rR));
  // min / max of the normal velocities by one compare (no NaN handling)
  const bool lslow = unL < unR;
  const float SL = (lslow ? unL : unR) - cmax;
  const float SR = (lslow ? unR : unL) + cmax;

  // contact speed and star total pressure
  const float sLu = SL - unL, sRu = SR - unR;
  const float den = sRu * rR - sLu * rL;
  const float iden = frcp(den);
  const float SM = (sRu * rR * unR - sLu * rL * unL - ptR + ptL) * iden;
  const float pts = (sRu * rR * ptL - sLu * rL * ptR + rL * rR * sRu * sLu * (unR - unL)) * iden;

  // single-star states (both sides: the double-star states need them)
  const float sLm = SL - SM, sRm = SR - SM;
  const float isLm = frcp(sLm), isRm = frcp(sRm);
  const float rLs = rL * sLu * isLm;
  const float rRs = rR * sRu * isRm;
  const float dL = rL * sLu * sLm - bx2 + TINY;
  const float dR = rR * sRu * sRm - bx2 + TINY;
  const float idL = frcp(dL), idR = frcp(dR);
  const float kvL = bx * (SM - unL) * idL;
  const float kvR = bx * (SM - unR) * idR;
  const float kbL = (rL * sLu * sLu - bx2) * idL;
  const float kbR = (rR * sRu * sRu - bx2) * idR;
  const float u1Ls = ut1L - b1L * kvL, u2Ls = ut2L - b2L * kvL;
  const float u1Rs = ut1R - b1R * kvR, u2Rs = ut2R - b2R * kvR;
  const float b1Ls = b1L * kbL, b2Ls = b2L * kbL;
  const float b1Rs = b1R * kbR, b2Rs = b2R * kbR;

  // double-star (rotational) states; sqrt(rho*) and 1/sqrt(rho*) from one
  // reciprocal square root each
  const float isrL = __builtin_amdgcn_rsqf(rLs), isrR = __builtin_amdgcn_rsqf(rRs);
  const float srL = rLs * isrL, srR = rRs * isrR;
  const float f = frcp(srL + srR);
  const float u1ss = f * (srL * u1Ls + srR * u1Rs + (b1Rs - b1Ls) * sgn);
  const float u2ss = f * (srL * u2Ls + srR * u2Rs + (b2Rs - b2Ls) * sgn);
  const float srLR = srL * srR;
  const float b1ss = f * (srL * b1Rs + srR * b1Ls + srLR * (u1Rs - u1Ls) * sgn);
  const float b2ss = f * (srL * b2Rs + srR * b2Ls + srLR * (u2Rs - u2Ls) * sgn);
  const float vbss = SM * bx + u1ss * b1ss + u2ss * b2ss;

  // Select the fan region; every region's flux is F_K + sum of jumps.
  const bool left = SM >= 0.0f;  // regions 1-3 use the left side
  // outer and single-star state on the chosen side
  const float r = left ? rL : rR, un = left ? unL : unR;
  const float u1 = left ? ut1L : ut1R, u2 = left ? ut2L : ut2R;
  const float b1 = left ? b1L : b1R, b2 = left ? b2L : b2R;
  const float p = left ? pL : pR, B2 = left ? B2L : B2R, pt = left ? ptL : ptR;
  const float S = left ? SL : SR, sKu = left ? sLu : sRu, isKm = left ? isLm : isRm;
  const float rs = left ? rLs : rRs;
  const float u1s = left ? u1Ls : u1Rs, u2s = left ? u2Ls : u2Rs;
  const float b1s = left ? b1Ls : b1Rs, b2s = left ? b2Ls : b2Rs;
  // the chosen side's energy chain: E, E* and E** (the other side's is not
  // needed by any region)
  const float v2 = un * un + u1 * u1 + u2 * u2;
  const float kin = r * v2 + B2;
  const float e = p * inv_gm1 + 0.5f * kin;
  const float vb = un * bx + u1 * b1 + u2 * b2;
  const float vbs = SM * bx + u1s * b1s + u2s * b2s;
  const float es = (sKu * e - pt * un + pts * SM + bx * (vb - vbs)) * isKm;
  // S*_K = SM -+ |Bx| / sqrt(rho*_K); E**_K = E*_K -+ sqrt(rho*_K) (v*.B* -
  // v**.B**) sgn(Bx), - on the left, + on the right
  const float sabx = left ? -abx : abx, ssr = left ? -srL : srR;
  const float Ss = SM + sabx * (left ? isrL : isrR);
  const float ess = es + ssr * (vbs - vbss) * sgn;

  Flux7 F;
  F.rho = r * un;
  F.mn = r * un * un + pt - bx2;
  F.mt1 = r * u1 * un - bx * b1;
  F.mt2 = r * u2 * un - bx * b2;
  F.e = (e + pt) * un - bx * vb;
  F.bt1 = b1 * un - bx * u1;
  F.bt2 = b2 * un - bx * u2;

  // fan region: outer state, single star (star) or double star (star and
  // dstar).  Left of the contact: star = !(SL > 0), dstar = !(S*_L >= 0);
  // right: star = !(SR < 0) = !(-SR > 0), dstar = !(S*_R < 0).  Written as
  // one float select and one compare, and the two dstar tests merged by
  // mask operations (dB ^ (left & (dA ^ dB)) = left ? dA : dB), so that the
  // compiler keeps them as lane masks instead of integers per lane (10
  // fewer VALU instructions per face, same result for every input)
  const float sw = left ? SL : -SR;
  const bool star = !(sw > 0.0f);
  const bool dA = !(0.0f <= Ss), dB = !(Ss < 0.0f);
  const bool dstar = dB ^ (left & (dA ^ dB));
  if (star) {
    // F_K + S (U*_K - U_K)
    F.rho += S * (rs - r);
    F.mn += S * (rs * SM - r * un);
    F.mt1 += S * (rs * u1s - r * u1);
    F.mt2 += S * (rs * u2s - r * u2);
    F.e += S * (es - e);
    F.bt1 += S * (b1s - b1);
    F.bt2 += S * (b2s - b2);
    if (dstar) {
      // + S*_K (U**_K - U*_K)
      F.mt1 += Ss * rs * (u1ss - u1s);
      F.mt2 += Ss * rs * (u2ss - u2s);
      F.e += Ss * (ess - es);
      F.bt1 += Ss * (b1ss - b1s);
      F.bt2 += Ss * (b2ss - b2s);
    }
  }
  return F;
}

// The same HLLD solve with the reference's rounding (HLLD:7-284): IEEE
// division and square root, each expression in the reference's order, and the
// double literals (0.5, 1.0, 2.0) that promote the energies, the fast speeds,
// the contact speed, the star pressure and f = 1/(sqrt(rho*_L)+sqrt(rho*_R))
// to FP64 before rounding back to FP32.  Used by the strict verification path
// (compiled without FMA contraction), which reproduces the CPU oracle bit for
// bit.
__device__ __forceinline__ Flux7 hlld_strict(float gamma, float Bx, float rL, float rR,
                                             float uL, float uR, float vL, float vR,
                                             float wL, float wR, float pL, float pR,
                                             float yL, float yR, float zL, float zR) {
  const float TINY = 1.0e-10;
  float sgn = 0;
  if (fabsf(Bx) > 0) sgn = Bx / fabsf(Bx);
  const float mxL = rL * uL, myL = rL * vL, mzL = rL * wL;
  const float mxR = rR * uR, myR = rR * vR, mzR = rR * wR;
  const float B2L = sq(Bx) + sq(yL) + sq(zL);
  const float v2L = sq(uL) + sq(vL) + sq(wL);
  const float eL = pL / (gamma - 1) + 0.5 * rL * v2L + 0.5 * B2L;
  const float B2R = sq(Bx) + sq(yR) + sq(zR);
  const float v2R = sq(uR) + sq(vR) + sq(wR);
  const float eR = pR / (gamma - 1) + 0.5 * rR * v2R + 0.5 * B2R;
  const float gpl = gamma * pL, gpr = gamma * pR;
  const float gpbl = gpl + B2L, gpbr = gpr + B2R;
  const float fsig = 1.41421356f;  // sqrtf(2.0)
  const float cfl = sqrtf((gpbl + sqrtf(sq(gpbl) - 4 * gpl * sq(Bx))) / (2.0 * rL));
  const float cfr = sqrtf((gpbr + sqrtf(sq(gpbr) - 4 * gpr * sq(Bx))) / (2.0 * rR));
  const float cfmax = fsig * fmaxf(cfl, cfr);
  const float SL = fminf(uL, uR) - cfmax;
  const float SR = fmaxf(uL, uR) + cfmax;
  const float ptL = pL + 0.5 * B2L;
  const float ptR = pR + 0.5 * B2R;
  float SM = (SR - uR) * mxR - (SL - uL) * mxL - ptR + ptL;
  SM *= 1.0 / ((SR - uR) * rR - (SL - uL) * rL);
  float pts = (SR - uR) * rR * ptL - (SL - uL) * rL * ptR;
  pts += rL * rR * (SR - uR) * (SL - uL) * (uR - uL);
  pts *= 1.0 / ((SR - uR) * rR - (SL - uL) * rL);
  const float B2x = Bx * Bx;
  const float dLs = rL * (SL - uL) / (SL - SM);
  const float vLs = vL - Bx * yL * (SM - uL) / (rL * (SL - uL) * (SL - SM) - B2x + TINY);
  const float wLs = wL - Bx * zL * (SM - uL) / (rL * (SL - uL) * (SL - SM) - B2x + TINY);
  const float yLs = yL * (rL * sq(SL - uL) - B2x) / (rL * (SL - uL) * (SL - SM) - B2x + TINY);
  const float zLs = zL * (rL * sq(SL - uL) - B2x) / (rL * (SL - uL) * (SL - SM) - B2x + TINY);
  const float dRs = rR * (SR - uR) / (SR - SM);
  const float vRs = vR - Bx * yR * (SM - uR) / (rR * (SR - uR) * (SR - SM) - B2x + TINY);
  const float wRs = wR - Bx * zR * (SM - uR) / (rR * (SR - uR) * (SR - SM) - B2x + TINY);
  const float yRs = yR * (rR * sq(SR - uR) - B2x) / (rR * (SR - uR) * (SR - SM) - B2x + TINY);
  const float zRs = zR * (rR * sq(SR - uR) - B2x) / (rR * (SR - uR) * (SR - SM) - B2x + TINY);
  const float vbL = uL * Bx + vL * yL + wL * zL;
  const float vbLs = SM * Bx + vLs * yLs + wLs * zLs;
  const float eLs = ((SL - uL) * eL - ptL * uL + pts * SM + Bx * (vbL - vbLs)) / (SL - SM);
  const float vbR = uR * Bx + vR * yR + wR * zR;
  const float vbRs = SM * Bx + vRs * yRs + wRs * zRs;
  const float eRs = ((SR - uR) * eR - ptR * uR + pts * SM + Bx * (vbR - vbRs)) / (SR - SM);
  const float SLs = SM - fabsf(Bx) / sqrtf(dLs);
  const float SRs = SM + fabsf(Bx) / sqrtf(dRs);
  const float f = 1.0 / (sqrtf(dLs) + sqrtf(dRs));
  const float vss = f * (sqrtf(dLs) * vLs + sqrtf(dRs) * vRs + (yRs - yLs) * sgn);
  const float wss = f * (sqrtf(dLs) * wLs + sqrtf(dRs) * wRs + (zRs - zLs) * sgn);
  const float yss = f * (sqrtf(dLs) * yRs + sqrtf(dRs) * yLs + sqrtf(dLs * dRs) * (vRs - vLs) * sgn);
  const float zss = f * (sqrtf(dLs) * zRs + sqrtf(dRs) * zLs + sqrtf(dLs * dRs) * (wRs - wLs) * sgn);
  const float vbss = SM * Bx + vss * yss + wss * zss;
  const float eLss = eLs - sqrtf(dLs) * (vbLs - vbss) * sgn;
  const float eRss = eRs + sqrtf(dRs) * (vbRs - vbss) * sgn;
  // outer fluxes
  const float FdL = rL * uL, FxL = mxL * uL + ptL - B2x, FyL = myL * uL - Bx * yL;
  const float FzL = mzL * uL - Bx * zL, FeL = eL * uL + ptL * uL - Bx * vbL;
  const float FbyL = yL * uL - Bx * vL, FbzL = zL * uL - Bx * wL;
  const float FdR = rR * uR, FxR = mxR * uR + ptR - B2x, FyR = myR * uR - Bx * yR;
  const float FzR = mzR * uR - Bx * zR, FeR = eR * uR + ptR * uR - Bx * vbR;
  const float FbyR = yR * uR - Bx * vR, FbzR = zR * uR - Bx * wR;
  // star momenta
  const float mxLs = dLs * SM, myLs = dLs * vLs, mzLs = dLs * wLs;
  const float myLss = dLs * vss, mzLss = dLs * wss;
  const float mxRs = dRs * SM, myRs = dRs * vRs, mzRs = dRs * wRs;
  const float myRss = dRs * vss, mzRss = dRs * wss;
  Flux7 F;
  if (SL > 0) {
    F = {FdL, FxL, FyL, FzL, FeL, FbyL, FbzL};
  } else if (SL <= 0 && 0 <= SLs) {
    F = {FdL + SL * dLs - SL * rL, FxL + SL * mxLs - SL * mxL, FyL + SL * myLs - SL * myL,
         FzL + SL * mzLs - SL * mzL, FeL + SL * eLs - SL * eL, FbyL + SL * yLs - SL * yL,
         FbzL + SL * zLs - SL * zL};
  } else if (SLs <= 0 && 0 <= SM) {
    F = {FdL + SLs * dLs - (SLs - SL) * dLs - SL * rL,
         FxL + SLs * mxLs - (SLs - SL) * mxLs - SL * mxL,
         FyL + SLs * myLss - (SLs - SL) * myLs - SL * myL,
         FzL + SLs * mzLss - (SLs - SL) * mzLs - SL * mzL,
         FeL + SLs * eLss - (SLs - SL) * eLs - SL * eL,
         FbyL + SLs * yss - (SLs - SL) * yLs - SL * yL,
         FbzL + SLs * zss - (SLs - SL) * zLs - SL * zL};
  } else if (SM <= 0 && 0 <= SRs) {
    F = {FdR + SRs * dRs - (SRs - SR) * dRs - SR * rR,
         FxR + SRs * mxRs - (SRs - SR) * mxRs - SR * mxR,
         FyR + SRs * myRss - (SRs - SR) * myRs - SR * myR,
         FzR + SRs * mzRss - (SRs - SR) * mzRs - SR * mzR,
         FeR + SRs * eRss - (SRs - SR) * eRs - SR * eR,
         FbyR + SRs * yss - (SRs - SR) * yRs - SR * yR,
         FbzR + SRs * zss - (SRs - SR) * zRs - SR * zR};
  } else if (SRs <= 0 && 0 <= SR) {
    F = {FdR + SR * dRs - SR * rR, FxR + SR * mxRs - SR * mxR, FyR + SR * myRs - SR * myR,
         FzR + SR * mzRs - SR * mzR, FeR + SR * eRs - SR * eR, FbyR + SR * yRs - SR * yR,
         FbzR + SR * zRs - SR * zR};
  } else {
    F = {FdR, FxR, FyR, FzR, FeR, FbyR, FbzR};
  }
  return F;
}

// Radial (Y) position y + off cells from the array's lower edge, relative
// to the box centre: corner + d*(y + off) with corner = -ext/2.
__device__ __forceinline__ float radial_position_at(const Box &b, int y, float off) {
  return -b.ext[1] * 0.5f + b.d[1] * ((float)y + off);
}
// Centre of row y.  Every staggered position gets the same +0.5
// (crystal.h:86-104), so the reference evaluates faces there too.
__device__ __forceinline__ float radial_position(const Box &b, int y) {
  return radial_position_at(b, y, 0.5f);
}

// Shearing-box acceleration (KEP:78-80): a_X = -2 Omega v_Y,
// a_Y = 2 Omega v_X + 2 q Omega^2 Y, a_Z = 0.
__device__ __forceinline__ float accel_x(const Box &b, float vY) {
  return -2.0f * b.omega * vY;
}
__device__ __forceinline__ float accel_y(const Box &b, float vX, int y) {
  return 2.0f * b.omega * vX + 2.0f * b.qshear * (b.omega * b.omega) * radial_position(b, y);
}
// The same on the radial face below cell y, for the face predictor: at
// y + 0.5 cells as the reference evaluates it, or at the face itself
// (face_off = 0, NMHD_FIX_FACE_POSITION).
__device__ __forceinline__ float accel_y_face(const Box &b, float vX, int y) {
  return 2.0f * b.omega * vX +
         2.0f * b.qshear * (b.omega * b.omega) * radial_position_at(b, y, b.face_off);
}
// CFL signal speed along one axis: |v + c| as the reference computes it
// (INL:715-720), or |v| + c (NMHD_FIX_CFL_SPEED; c > 0).
__device__ __forceinline__ float cfl_speed(const Box &b, float v, float c) {
  return fabsf((b.fix_cfl ? fabsf(v) : v) + c);
}

}  // namespace nmhd
