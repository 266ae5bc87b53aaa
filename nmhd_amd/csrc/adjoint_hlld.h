// adjoint_hlld.h -- reverse-mode derivative of the HLLD flux, written by
// hand: the forward solve is recomputed with every intermediate kept in a
// register "tape", the fan region and side it took are kept, and the reverse
// sweep walks the intermediates backwards (the branches of the forward solve
// are followed, min/max pass the derivative to the argument they selected).
//
// The forward here is the same Miyoshi-Kusano HLLD as hlld() in
// nmhd_device.h (reference: library/hlld.inl:7-284), written with plain
// division and square root and templated on the scalar type, so that the
// host build (double) can check the reverse sweep against finite
// differences and the device build (float) runs it in the adjoint kernels.
//
// Inputs (15): bx, rL, rR, unL, unR, u1L, u1R, u2L, u2R, pL, pR, b1L, b1R,
// b2L, b2R; outputs (7): fluxes of rho, m_n, m_t1, m_t2, E, B_t1, B_t2.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace nmhd {
namespace adj {

// Division, reciprocal and square root of the tape and the reverse sweep: on
// the device the hardware reciprocal and square root (1 ulp), the arithmetic
// of the production HLLD, instead of the IEEE sequences (a tenth of the
// instructions); on the host (the double build that checks the reverse
// sweep against finite differences) the IEEE operations.
__device__ inline float qdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ inline float qrcp(float a) { return __builtin_amdgcn_rcpf(a); }
__device__ inline float qsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__host__ inline float qdiv(float a, float b) { return a / b; }
__host__ inline float qrcp(float a) { return 1.0f / a; }
__host__ inline float qsqrt(float x) { return std::sqrt(x); }
__host__ __device__ inline double qdiv(double a, double b) { return a / b; }
__host__ __device__ inline double qrcp(double a) { return 1.0 / a; }
__host__ __device__ inline double qsqrt(double x) { return std::sqrt(x); }

template <class T>
struct HIn {
  T bx, rL, rR, unL, unR, u1L, u1R, u2L, u2R, pL, pR, b1L, b1R, b2L, b2R;
};

template <class T>
struct HOut {
  T rho, mn, mt1, mt2, e, bt1, bt2;
};

// Every intermediate of the forward solve that the reverse sweep reads.
template <class T>
struct HTape {
  T bx2, abx, sgn;
  T v2L, v2R, eL, eR, ptL, ptR, gpL, gpR, gpbL, gpbR, sdL, sdR, cf2L, cf2R, cfL, cfR;
  T SL, SR, sLu, sRu, den, iden, SMn, SM, ptsn, pts, sLm, sRm, isLm, isRm;
  T rLs, rRs, dL, dR, idL, idR, kvL, kvR, kbLn, kbRn, kbL, kbR;
  T u1Ls, u2Ls, u1Rs, u2Rs, b1Ls, b2Ls, b1Rs, b2Rs;
  T vbL, vbR, vbLs, vbRs, eLsn, eRsn, eLs, eRs;
  T srL, srR, SLs, SRs, f, u1ssn, u2ssn, u1ss, u2ss, srLR, b1ssn, b2ssn, b1ss, b2ss, vbss;
  T eLss, eRss;
  bool useL_c, minL, maxL, left;
  int region;  // 0 outer state, 1 single star, 2 double star
};

template <class T>
__host__ __device__ inline void hlld_forward(T gamma, T ig, const HIn<T> &in, HTape<T> &t,
                                             HOut<T> &F) {
  using std::sqrt;
  const T TINY = T(1.0e-10);
  const T SQ2 = T(1.41421356);
  const T bx = in.bx;
  t.bx2 = bx * bx;
  t.abx = bx < T(0) ? -bx : bx;
  t.sgn = bx > T(0) ? T(1) : (bx < T(0) ? T(-1) : T(0));
  const T B2L = t.bx2 + in.b1L * in.b1L + in.b2L * in.b2L;
  const T B2R = t.bx2 + in.b1R * in.b1R + in.b2R * in.b2R;
  t.v2L = in.unL * in.unL + in.u1L * in.u1L + in.u2L * in.u2L;
  t.v2R = in.unR * in.unR + in.u1R * in.u1R + in.u2R * in.u2R;
  t.eL = in.pL * ig + T(0.5) * in.rL * t.v2L + T(0.5) * B2L;
  t.eR = in.pR * ig + T(0.5) * in.rR * t.v2R + T(0.5) * B2R;
  t.ptL = in.pL + T(0.5) * B2L;
  t.ptR = in.pR + T(0.5) * B2R;
  t.gpL = gamma * in.pL;
  t.gpR = gamma * in.pR;
  t.gpbL = t.gpL + B2L;
  t.gpbR = t.gpR + B2R;
  const T discL = t.gpbL * t.gpbL - T(4) * t.gpL * t.bx2;
  const T discR = t.gpbR * t.gpbR - T(4) * t.gpR * t.bx2;
  t.sdL = qsqrt(discL > T(0) ? discL : T(0));
  t.sdR = qsqrt(discR > T(0) ? discR : T(0));
  t.cf2L = qdiv(t.gpbL + t.sdL, T(2) * in.rL);
  t.cf2R = qdiv(t.gpbR + t.sdR, T(2) * in.rR);
  t.cfL = qsqrt(t.cf2L);
  t.cfR = qsqrt(t.cf2R);
  t.useL_c = t.cfL >= t.cfR;
  const T cmax = SQ2 * (t.useL_c ? t.cfL : t.cfR);
  t.minL = in.unL <= in.unR;
  t.maxL = in.unL >= in.unR;
  t.SL = (t.minL ? in.unL : in.unR) - cmax;
  t.SR = (t.maxL ? in.unL : in.unR) + cmax;
  t.sLu = t.SL - in.unL;
  t.sRu = t.SR - in.unR;
  t.den = t.sRu * in.rR - t.sLu * in.rL;
  t.iden = qrcp(t.den);
  t.SMn = t.sRu * in.rR * in.unR - t.sLu * in.rL * in.unL - t.ptR + t.ptL;
  t.SM = t.SMn * t.iden;
  t.ptsn = t.sRu * in.rR * t.ptL - t.sLu * in.rL * t.ptR +
           in.rL * in.rR * t.sRu * t.sLu * (in.unR - in.unL);
  t.pts = t.ptsn * t.iden;
  t.sLm = t.SL - t.SM;
  t.sRm = t.SR - t.SM;
  t.isLm = qrcp(t.sLm);
  t.isRm = qrcp(t.sRm);
  t.rLs = in.rL * t.sLu * t.isLm;
  t.rRs = in.rR * t.sRu * t.isRm;
  t.dL = in.rL * t.sLu * t.sLm - t.bx2 + TINY;
  t.dR = in.rR * t.sRu * t.sRm - t.bx2 + TINY;
  t.idL = qrcp(t.dL);
  t.idR = qrcp(t.dR);
  t.kvL = bx * (t.SM - in.unL) * t.idL;
  t.kvR = bx * (t.SM - in.unR) * t.idR;
  t.kbLn = in.rL * t.sLu * t.sLu - t.bx2;
  t.kbRn = in.rR * t.sRu * t.sRu - t.bx2;
  t.kbL = t.kbLn * t.idL;
  t.kbR = t.kbRn * t.idR;
  t.u1Ls = in.u1L - in.b1L * t.kvL;
  t.u2Ls = in.u2L - in.b2L * t.kvL;
  t.u1Rs = in.u1R - in.b1R * t.kvR;
  t.u2Rs = in.u2R - in.b2R * t.kvR;
  t.b1Ls = in.b1L * t.kbL;
  t.b2Ls = in.b2L * t.kbL;
  t.b1Rs = in.b1R * t.kbR;
  t.b2Rs = in.b2R * t.kbR;
  t.vbL = in.unL * bx + in.u1L * in.b1L + in.u2L * in.b2L;
  t.vbR = in.unR * bx + in.u1R * in.b1R + in.u2R * in.b2R;
  t.vbLs = t.SM * bx + t.u1Ls * t.b1Ls + t.u2Ls * t.b2Ls;
  t.vbRs = t.SM * bx + t.u1Rs * t.b1Rs + t.u2Rs * t.b2Rs;
  t.eLsn = t.sLu * t.eL - t.ptL * in.unL + t.pts * t.SM + bx * (t.vbL - t.vbLs);
  t.eRsn = t.sRu * t.eR - t.ptR * in.unR + t.pts * t.SM + bx * (t.vbR - t.vbRs);
  t.eLs = t.eLsn * t.isLm;
  t.eRs = t.eRsn * t.isRm;
  t.srL = qsqrt(t.rLs);
  t.srR = qsqrt(t.rRs);
  t.SLs = t.SM - qdiv(t.abx, t.srL);
  t.SRs = t.SM + qdiv(t.abx, t.srR);
  t.f = qrcp(t.srL + t.srR);
  t.u1ssn = t.srL * t.u1Ls + t.srR * t.u1Rs + (t.b1Rs - t.b1Ls) * t.sgn;
  t.u2ssn = t.srL * t.u2Ls + t.srR * t.u2Rs + (t.b2Rs - t.b2Ls) * t.sgn;
  t.u1ss = t.f * t.u1ssn;
  t.u2ss = t.f * t.u2ssn;
  t.srLR = qsqrt(t.rLs * t.rRs);
  t.b1ssn = t.srL * t.b1Rs + t.srR * t.b1Ls + t.srLR * (t.u1Rs - t.u1Ls) * t.sgn;
  t.b2ssn = t.srL * t.b2Rs + t.srR * t.b2Ls + t.srLR * (t.u2Rs - t.u2Ls) * t.sgn;
  t.b1ss = t.f * t.b1ssn;
  t.b2ss = t.f * t.b2ssn;
  t.vbss = t.SM * bx + t.u1ss * t.b1ss + t.u2ss * t.b2ss;
  t.eLss = t.eLs - t.srL * (t.vbLs - t.vbss) * t.sgn;
  t.eRss = t.eRs + t.srR * (t.vbRs - t.vbss) * t.sgn;

  t.left = t.SM >= T(0);
  if (t.left)
    t.region = t.SL > T(0) ? 0 : (T(0) <= t.SLs ? 1 : 2);
  else
    t.region = t.SR < T(0) ? 0 : (t.SRs < T(0) ? 1 : 2);
  const bool L = t.left;
  const T r = L ? in.rL : in.rR, un = L ? in.unL : in.unR;
  const T u1 = L ? in.u1L : in.u1R, u2 = L ? in.u2L : in.u2R;
  const T b1 = L ? in.b1L : in.b1R, b2 = L ? in.b2L : in.b2R;
  const T e = L ? t.eL : t.eR, pt = L ? t.ptL : t.ptR, vb = L ? t.vbL : t.vbR;
  const T S = L ? t.SL : t.SR, Ss = L ? t.SLs : t.SRs, rs = L ? t.rLs : t.rRs;
  const T u1s = L ? t.u1Ls : t.u1Rs, u2s = L ? t.u2Ls : t.u2Rs;
  const T b1s = L ? t.b1Ls : t.b1Rs, b2s = L ? t.b2Ls : t.b2Rs;
  const T es = L ? t.eLs : t.eRs, ess = L ? t.eLss : t.eRss;
  F.rho = r * un;
  F.mn = r * un * un + pt - t.bx2;
  F.mt1 = r * u1 * un - bx * b1;
  F.mt2 = r * u2 * un - bx * b2;
  F.e = (e + pt) * un - bx * vb;
  F.bt1 = b1 * un - bx * u1;
  F.bt2 = b2 * un - bx * u2;
  if (t.region >= 1) {
    F.rho += S * (rs - r);
    F.mn += S * (rs * t.SM - r * un);
    F.mt1 += S * (rs * u1s - r * u1);
    F.mt2 += S * (rs * u2s - r * u2);
    F.e += S * (es - e);
    F.bt1 += S * (b1s - b1);
    F.bt2 += S * (b2s - b2);
    if (t.region == 2) {
      F.mt1 += Ss * rs * (t.u1ss - u1s);
      F.mt2 += Ss * rs * (t.u2ss - u2s);
      F.e += Ss * (ess - es);
      F.bt1 += Ss * (t.b1ss - b1s);
      F.bt2 += Ss * (t.b2ss - b2s);
    }
  }
}

// Reverse sweep: a (the adjoint of the 15 inputs) = J^T aF, with J the
// Jacobian of the forward solve on the branches recorded in the tape.
template <class T>
__host__ __device__ inline void hlld_reverse(T gamma, T ig, const HIn<T> &in, const HTape<T> &t,
                                             const HOut<T> &aF, HIn<T> &a) {
  const T SQ2 = T(1.41421356);
  const T bx = in.bx;
  const T Z = T(0);
  // adjoints of the intermediates (zero-initialised)
  T a_bx = Z, a_bx2 = Z, a_abx = Z;
  T a_rL = Z, a_rR = Z, a_unL = Z, a_unR = Z, a_u1L = Z, a_u1R = Z, a_u2L = Z, a_u2R = Z;
  T a_pL = Z, a_pR = Z, a_b1L = Z, a_b1R = Z, a_b2L = Z, a_b2R = Z;
  T a_eL = Z, a_eR = Z, a_ptL = Z, a_ptR = Z, a_vbL = Z, a_vbR = Z;
  T a_SL = Z, a_SR = Z, a_SLs = Z, a_SRs = Z, a_rLs = Z, a_rRs = Z;
  T a_u1Ls = Z, a_u2Ls = Z, a_u1Rs = Z, a_u2Rs = Z, a_b1Ls = Z, a_b2Ls = Z, a_b1Rs = Z,
    a_b2Rs = Z;
  T a_eLs = Z, a_eRs = Z, a_eLss = Z, a_eRss = Z;
  T a_u1ss = Z, a_u2ss = Z, a_b1ss = Z, a_b2ss = Z, a_SM = Z;

  // ---- flux of the selected region (side K)
  {
    const bool L = t.left;
    const bool r1 = t.region >= 1, r2 = t.region == 2;
    const T r = L ? in.rL : in.rR, un = L ? in.unL : in.unR;
    const T u1 = L ? in.u1L : in.u1R, u2 = L ? in.u2L : in.u2R;
    const T b1 = L ? in.b1L : in.b1R, b2 = L ? in.b2L : in.b2R;
    const T e = L ? t.eL : t.eR, pt = L ? t.ptL : t.ptR, vb = L ? t.vbL : t.vbR;
    const T S = L ? t.SL : t.SR, Ss = L ? t.SLs : t.SRs, rs = L ? t.rLs : t.rRs;
    const T u1s = L ? t.u1Ls : t.u1Rs, u2s = L ? t.u2Ls : t.u2Rs;
    const T b1s = L ? t.b1Ls : t.b1Rs, b2s = L ? t.b2Ls : t.b2Rs;
    const T es = L ? t.eLs : t.eRs, ess = L ? t.eLss : t.eRss;
    T ar = Z, aun = Z, au1 = Z, au2 = Z, ab1 = Z, ab2 = Z, ae = Z, apt = Z, avb = Z;
    T aS = Z, aSs = Z, ars = Z, au1s = Z, au2s = Z, ab1s = Z, ab2s = Z, aes = Z, aess = Z;
    // rho
    ar += aF.rho * un;
    aun += aF.rho * r;
    // m_n
    ar += aF.mn * un * un;
    aun += aF.mn * T(2) * r * un;
    apt += aF.mn;
    a_bx2 -= aF.mn;
    // m_t1, m_t2
    ar += aF.mt1 * u1 * un + aF.mt2 * u2 * un;
    au1 += aF.mt1 * r * un;
    au2 += aF.mt2 * r * un;
    aun += aF.mt1 * r * u1 + aF.mt2 * r * u2;
    a_bx -= aF.mt1 * b1 + aF.mt2 * b2;
    ab1 -= aF.mt1 * bx;
    ab2 -= aF.mt2 * bx;
    // E
    ae += aF.e * un;
    apt += aF.e * un;
    aun += aF.e * (e + pt);
    a_bx -= aF.e * vb;
    avb -= aF.e * bx;
    // B_t1, B_t2
    ab1 += aF.bt1 * un;
    ab2 += aF.bt2 * un;
    aun += aF.bt1 * b1 + aF.bt2 * b2;
    a_bx -= aF.bt1 * u1 + aF.bt2 * u2;
    au1 -= aF.bt1 * bx;
    au2 -= aF.bt2 * bx;
    if (r1) {
      aS += aF.rho * (rs - r) + aF.mn * (rs * t.SM - r * un) + aF.mt1 * (rs * u1s - r * u1) +
            aF.mt2 * (rs * u2s - r * u2) + aF.e * (es - e) + aF.bt1 * (b1s - b1) +
            aF.bt2 * (b2s - b2);
      ars += S * (aF.rho + aF.mn * t.SM + aF.mt1 * u1s + aF.mt2 * u2s);
      ar -= S * (aF.rho + aF.mn * un + aF.mt1 * u1 + aF.mt2 * u2);
      aun -= S * aF.mn * r;
      a_SM += S * aF.mn * rs;
      au1s += S * aF.mt1 * rs;
      au2s += S * aF.mt2 * rs;
      au1 -= S * aF.mt1 * r;
      au2 -= S * aF.mt2 * r;
      aes += S * aF.e;
      ae -= S * aF.e;
      ab1s += S * aF.bt1;
      ab1 -= S * aF.bt1;
      ab2s += S * aF.bt2;
      ab2 -= S * aF.bt2;
    }
    if (r2) {
      aSs += aF.mt1 * rs * (t.u1ss - u1s) + aF.mt2 * rs * (t.u2ss - u2s) + aF.e * (ess - es) +
             aF.bt1 * (t.b1ss - b1s) + aF.bt2 * (t.b2ss - b2s);
      ars += Ss * (aF.mt1 * (t.u1ss - u1s) + aF.mt2 * (t.u2ss - u2s));
      a_u1ss += Ss * rs * aF.mt1;
      a_u2ss += Ss * rs * aF.mt2;
      au1s -= Ss * rs * aF.mt1;
      au2s -= Ss * rs * aF.mt2;
      aess += Ss * aF.e;
      aes -= Ss * aF.e;
      a_b1ss += Ss * aF.bt1;
      a_b2ss += Ss * aF.bt2;
      ab1s -= Ss * aF.bt1;
      ab2s -= Ss * aF.bt2;
    }
    if (L) {
      a_rL += ar, a_unL += aun, a_u1L += au1, a_u2L += au2, a_b1L += ab1, a_b2L += ab2;
      a_eL += ae, a_ptL += apt, a_vbL += avb, a_SL += aS, a_SLs += aSs, a_rLs += ars;
      a_u1Ls += au1s, a_u2Ls += au2s, a_b1Ls += ab1s, a_b2Ls += ab2s, a_eLs += aes,
          a_eLss += aess;
    } else {
      a_rR += ar, a_unR += aun, a_u1R += au1, a_u2R += au2, a_b1R += ab1, a_b2R += ab2;
      a_eR += ae, a_ptR += apt, a_vbR += avb, a_SR += aS, a_SRs += aSs, a_rRs += ars;
      a_u1Rs += au1s, a_u2Rs += au2s, a_b1Rs += ab1s, a_b2Rs += ab2s, a_eRs += aes,
          a_eRss += aess;
    }
  }

  // ---- double-star states
  T a_vbss = Z, a_srL = Z, a_srR = Z, a_f = Z, a_srLR = Z, a_vbLs = Z, a_vbRs = Z;
  // eLss = eLs - srL (vbLs - vbss) sgn;  eRss = eRs + srR (vbRs - vbss) sgn
  a_eLs += a_eLss;
  a_srL -= a_eLss * (t.vbLs - t.vbss) * t.sgn;
  a_vbLs -= a_eLss * t.srL * t.sgn;
  a_vbss += a_eLss * t.srL * t.sgn;
  a_eRs += a_eRss;
  a_srR += a_eRss * (t.vbRs - t.vbss) * t.sgn;
  a_vbRs += a_eRss * t.srR * t.sgn;
  a_vbss -= a_eRss * t.srR * t.sgn;
  // vbss = SM bx + u1ss b1ss + u2ss b2ss
  a_SM += a_vbss * bx;
  a_bx += a_vbss * t.SM;
  a_u1ss += a_vbss * t.b1ss;
  a_b1ss += a_vbss * t.u1ss;
  a_u2ss += a_vbss * t.b2ss;
  a_b2ss += a_vbss * t.u2ss;
  // b1ss = f b1ssn, b1ssn = srL b1Rs + srR b1Ls + srLR (u1Rs - u1Ls) sgn (and b2ss)
  {
    const T g1 = a_b1ss * t.f, g2 = a_b2ss * t.f;
    a_f += a_b1ss * t.b1ssn + a_b2ss * t.b2ssn;
    a_srL += g1 * t.b1Rs + g2 * t.b2Rs;
    a_b1Rs += g1 * t.srL;
    a_b2Rs += g2 * t.srL;
    a_srR += g1 * t.b1Ls + g2 * t.b2Ls;
    a_b1Ls += g1 * t.srR;
    a_b2Ls += g2 * t.srR;
    a_srLR += (g1 * (t.u1Rs - t.u1Ls) + g2 * (t.u2Rs - t.u2Ls)) * t.sgn;
    a_u1Rs += g1 * t.srLR * t.sgn;
    a_u1Ls -= g1 * t.srLR * t.sgn;
    a_u2Rs += g2 * t.srLR * t.sgn;
    a_u2Ls -= g2 * t.srLR * t.sgn;
  }
  // srLR = sqrt(rLs rRs)
  if (t.srLR > Z) {
    const T g = qdiv(a_srLR, T(2) * t.srLR);
    a_rLs += g * t.rRs;
    a_rRs += g * t.rLs;
  }
  // u1ss = f u1ssn, u1ssn = srL u1Ls + srR u1Rs + (b1Rs - b1Ls) sgn (and u2ss)
  {
    const T g1 = a_u1ss * t.f, g2 = a_u2ss * t.f;
    a_f += a_u1ss * t.u1ssn + a_u2ss * t.u2ssn;
    a_srL += g1 * t.u1Ls + g2 * t.u2Ls;
    a_u1Ls += g1 * t.srL;
    a_u2Ls += g2 * t.srL;
    a_srR += g1 * t.u1Rs + g2 * t.u2Rs;
    a_u1Rs += g1 * t.srR;
    a_u2Rs += g2 * t.srR;
    a_b1Rs += g1 * t.sgn;
    a_b1Ls -= g1 * t.sgn;
    a_b2Rs += g2 * t.sgn;
    a_b2Ls -= g2 * t.sgn;
  }
  // f = 1/(srL + srR)
  a_srL -= a_f * t.f * t.f;
  a_srR -= a_f * t.f * t.f;
  // SLs = SM - abx/srL; SRs = SM + abx/srR
  a_SM += a_SLs + a_SRs;
  a_abx += -qdiv(a_SLs, t.srL) + qdiv(a_SRs, t.srR);
  a_srL += qdiv(a_SLs * t.abx, t.srL * t.srL);
  a_srR -= qdiv(a_SRs * t.abx, t.srR * t.srR);
  // srL = sqrt(rLs)
  a_rLs += qdiv(a_srL, T(2) * t.srL);
  a_rRs += qdiv(a_srR, T(2) * t.srR);

  // ---- single-star states
  T a_sLu = Z, a_sRu = Z, a_isLm = Z, a_isRm = Z, a_pts = Z;
  {  // eLs = eLsn isLm
    const T g = a_eLs * t.isLm;
    a_isLm += a_eLs * t.eLsn;
    a_sLu += g * t.eL;
    a_eL += g * t.sLu;
    a_ptL -= g * in.unL;
    a_unL -= g * t.ptL;
    a_pts += g * t.SM;
    a_SM += g * t.pts;
    a_bx += g * (t.vbL - t.vbLs);
    a_vbL += g * bx;
    a_vbLs -= g * bx;
  }
  {  // eRs = eRsn isRm
    const T g = a_eRs * t.isRm;
    a_isRm += a_eRs * t.eRsn;
    a_sRu += g * t.eR;
    a_eR += g * t.sRu;
    a_ptR -= g * in.unR;
    a_unR -= g * t.ptR;
    a_pts += g * t.SM;
    a_SM += g * t.pts;
    a_bx += g * (t.vbR - t.vbRs);
    a_vbR += g * bx;
    a_vbRs -= g * bx;
  }
  // vbLs = SM bx + u1Ls b1Ls + u2Ls b2Ls (and R)
  a_SM += (a_vbLs + a_vbRs) * bx;
  a_bx += (a_vbLs + a_vbRs) * t.SM;
  a_u1Ls += a_vbLs * t.b1Ls;
  a_b1Ls += a_vbLs * t.u1Ls;
  a_u2Ls += a_vbLs * t.b2Ls;
  a_b2Ls += a_vbLs * t.u2Ls;
  a_u1Rs += a_vbRs * t.b1Rs;
  a_b1Rs += a_vbRs * t.u1Rs;
  a_u2Rs += a_vbRs * t.b2Rs;
  a_b2Rs += a_vbRs * t.u2Rs;
  // vbL = unL bx + u1L b1L + u2L b2L (and R)
  a_unL += a_vbL * bx;
  a_bx += a_vbL * in.unL + a_vbR * in.unR;
  a_u1L += a_vbL * in.b1L;
  a_b1L += a_vbL * in.u1L;
  a_u2L += a_vbL * in.b2L;
  a_b2L += a_vbL * in.u2L;
  a_unR += a_vbR * bx;
  a_u1R += a_vbR * in.b1R;
  a_b1R += a_vbR * in.u1R;
  a_u2R += a_vbR * in.b2R;
  a_b2R += a_vbR * in.u2R;
  // b1Ls = b1L kbL, b2Ls = b2L kbL; u1Ls = u1L - b1L kvL, u2Ls = u2L - b2L kvL (and R)
  const T a_kbL = a_b1Ls * in.b1L + a_b2Ls * in.b2L;
  const T a_kbR = a_b1Rs * in.b1R + a_b2Rs * in.b2R;
  a_b1L += a_b1Ls * t.kbL;
  a_b2L += a_b2Ls * t.kbL;
  a_b1R += a_b1Rs * t.kbR;
  a_b2R += a_b2Rs * t.kbR;
  const T a_kvL = -(a_u1Ls * in.b1L + a_u2Ls * in.b2L);
  const T a_kvR = -(a_u1Rs * in.b1R + a_u2Rs * in.b2R);
  a_u1L += a_u1Ls;
  a_u2L += a_u2Ls;
  a_b1L -= a_u1Ls * t.kvL;
  a_b2L -= a_u2Ls * t.kvL;
  a_u1R += a_u1Rs;
  a_u2R += a_u2Rs;
  a_b1R -= a_u1Rs * t.kvR;
  a_b2R -= a_u2Rs * t.kvR;
  T a_idL = Z, a_idR = Z, a_sLm = Z, a_sRm = Z;
  {  // kbL = kbLn idL, kbLn = rL sLu^2 - bx2
    const T g = a_kbL * t.idL;
    a_idL += a_kbL * t.kbLn;
    a_rL += g * t.sLu * t.sLu;
    a_sLu += g * T(2) * in.rL * t.sLu;
    a_bx2 -= g;
  }
  {
    const T g = a_kbR * t.idR;
    a_idR += a_kbR * t.kbRn;
    a_rR += g * t.sRu * t.sRu;
    a_sRu += g * T(2) * in.rR * t.sRu;
    a_bx2 -= g;
  }
  // kvL = bx (SM - unL) idL
  a_bx += a_kvL * (t.SM - in.unL) * t.idL + a_kvR * (t.SM - in.unR) * t.idR;
  a_SM += a_kvL * bx * t.idL + a_kvR * bx * t.idR;
  a_unL -= a_kvL * bx * t.idL;
  a_unR -= a_kvR * bx * t.idR;
  a_idL += a_kvL * bx * (t.SM - in.unL);
  a_idR += a_kvR * bx * (t.SM - in.unR);
  {  // idL = 1/dL, dL = rL sLu sLm - bx2 + TINY
    const T gL = -a_idL * t.idL * t.idL, gR = -a_idR * t.idR * t.idR;
    a_rL += gL * t.sLu * t.sLm;
    a_sLu += gL * in.rL * t.sLm;
    a_sLm += gL * in.rL * t.sLu;
    a_rR += gR * t.sRu * t.sRm;
    a_sRu += gR * in.rR * t.sRm;
    a_sRm += gR * in.rR * t.sRu;
    a_bx2 -= gL + gR;
  }
  // rLs = rL sLu isLm
  a_rL += a_rLs * t.sLu * t.isLm;
  a_sLu += a_rLs * in.rL * t.isLm;
  a_isLm += a_rLs * in.rL * t.sLu;
  a_rR += a_rRs * t.sRu * t.isRm;
  a_sRu += a_rRs * in.rR * t.isRm;
  a_isRm += a_rRs * in.rR * t.sRu;
  // isLm = 1/sLm
  a_sLm -= a_isLm * t.isLm * t.isLm;
  a_sRm -= a_isRm * t.isRm * t.isRm;
  // sLm = SL - SM
  a_SL += a_sLm;
  a_SR += a_sRm;
  a_SM -= a_sLm + a_sRm;
  // pts = ptsn iden
  T a_iden = a_pts * t.ptsn;
  {
    const T g = a_pts * t.iden;
    const T du = in.unR - in.unL;
    a_sRu += g * (in.rR * t.ptL + in.rL * in.rR * t.sLu * du);
    a_rR += g * (t.sRu * t.ptL + in.rL * t.sRu * t.sLu * du);
    a_ptL += g * t.sRu * in.rR;
    a_sLu += g * (-in.rL * t.ptR + in.rL * in.rR * t.sRu * du);
    a_rL += g * (-t.sLu * t.ptR + in.rR * t.sRu * t.sLu * du);
    a_ptR -= g * t.sLu * in.rL;
    a_unR += g * in.rL * in.rR * t.sRu * t.sLu;
    a_unL -= g * in.rL * in.rR * t.sRu * t.sLu;
  }
  // SM = SMn iden, SMn = sRu rR unR - sLu rL unL - ptR + ptL
  a_iden += a_SM * t.SMn;
  {
    const T g = a_SM * t.iden;
    a_sRu += g * in.rR * in.unR;
    a_rR += g * t.sRu * in.unR;
    a_unR += g * t.sRu * in.rR;
    a_sLu -= g * in.rL * in.unL;
    a_rL -= g * t.sLu * in.unL;
    a_unL -= g * t.sLu * in.rL;
    a_ptR -= g;
    a_ptL += g;
  }
  {  // iden = 1/den, den = sRu rR - sLu rL
    const T g = -a_iden * t.iden * t.iden;
    a_sRu += g * in.rR;
    a_rR += g * t.sRu;
    a_sLu -= g * in.rL;
    a_rL -= g * t.sLu;
  }
  // sLu = SL - unL, sRu = SR - unR
  a_SL += a_sLu;
  a_unL -= a_sLu;
  a_SR += a_sRu;
  a_unR -= a_sRu;
  // SL = min(unL, unR) - cmax, SR = max(unL, unR) + cmax
  const T a_cmax = a_SR - a_SL;
  if (t.minL)
    a_unL += a_SL;
  else
    a_unR += a_SL;
  if (t.maxL)
    a_unL += a_SR;
  else
    a_unR += a_SR;
  T a_cfL = Z, a_cfR = Z;
  if (t.useL_c)
    a_cfL += SQ2 * a_cmax;
  else
    a_cfR += SQ2 * a_cmax;
  T a_gpL = Z, a_gpR = Z, a_B2L = Z, a_B2R = Z;
  {  // cfL = sqrt(cf2L), cf2L = (gpbL + sdL)/(2 rL), sdL = sqrt(gpbL^2 - 4 gpL bx2)
    const T a_cf2L = t.cfL > Z ? qdiv(a_cfL, T(2) * t.cfL) : Z;
    const T a_cf2R = t.cfR > Z ? qdiv(a_cfR, T(2) * t.cfR) : Z;
    T a_gpbL = qdiv(a_cf2L, T(2) * in.rL), a_gpbR = qdiv(a_cf2R, T(2) * in.rR);
    const T a_sdL = a_gpbL, a_sdR = a_gpbR;
    a_rL -= qdiv(a_cf2L * t.cf2L, in.rL);
    a_rR -= qdiv(a_cf2R * t.cf2R, in.rR);
    const T a_discL = t.sdL > Z ? qdiv(a_sdL, T(2) * t.sdL) : Z;
    const T a_discR = t.sdR > Z ? qdiv(a_sdR, T(2) * t.sdR) : Z;
    a_gpbL += a_discL * T(2) * t.gpbL;
    a_gpbR += a_discR * T(2) * t.gpbR;
    a_gpL -= a_discL * T(4) * t.bx2;
    a_gpR -= a_discR * T(4) * t.bx2;
    a_bx2 -= a_discL * T(4) * t.gpL + a_discR * T(4) * t.gpR;
    a_gpL += a_gpbL;
    a_gpR += a_gpbR;
    a_B2L += a_gpbL;
    a_B2R += a_gpbR;
  }
  a_pL += gamma * a_gpL;
  a_pR += gamma * a_gpR;
  // ptL = pL + B2L/2
  a_pL += a_ptL;
  a_pR += a_ptR;
  a_B2L += T(0.5) * a_ptL;
  a_B2R += T(0.5) * a_ptR;
  // eL = pL ig + rL v2L/2 + B2L/2
  a_pL += a_eL * ig;
  a_pR += a_eR * ig;
  a_rL += T(0.5) * a_eL * t.v2L;
  a_rR += T(0.5) * a_eR * t.v2R;
  const T a_v2L = T(0.5) * a_eL * in.rL, a_v2R = T(0.5) * a_eR * in.rR;
  a_B2L += T(0.5) * a_eL;
  a_B2R += T(0.5) * a_eR;
  // v2L = unL^2 + u1L^2 + u2L^2
  a_unL += T(2) * a_v2L * in.unL;
  a_u1L += T(2) * a_v2L * in.u1L;
  a_u2L += T(2) * a_v2L * in.u2L;
  a_unR += T(2) * a_v2R * in.unR;
  a_u1R += T(2) * a_v2R * in.u1R;
  a_u2R += T(2) * a_v2R * in.u2R;
  // B2L = bx2 + b1L^2 + b2L^2
  a_bx2 += a_B2L + a_B2R;
  a_b1L += T(2) * a_B2L * in.b1L;
  a_b2L += T(2) * a_B2L * in.b2L;
  a_b1R += T(2) * a_B2R * in.b1R;
  a_b2R += T(2) * a_B2R * in.b2R;
  // abx = |bx|, bx2 = bx^2
  a_bx += a_abx * t.sgn + T(2) * a_bx2 * bx;

  a.bx = a_bx;
  a.rL = a_rL;
  a.rR = a_rR;
  a.unL = a_unL;
  a.unR = a_unR;
  a.u1L = a_u1L;
  a.u1R = a_u1R;
  a.u2L = a_u2L;
  a.u2R = a_u2R;
  a.pL = a_pL;
  a.pR = a_pR;
  a.b1L = a_b1L;
  a.b1R = a_b1R;
  a.b2L = a_b2L;
  a.b2R = a_b2R;
}

}  // namespace adj
}  // namespace nmhd
