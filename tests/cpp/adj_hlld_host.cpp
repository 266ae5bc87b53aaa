// Host (double precision) build of the hand-written HLLD reverse sweep, for
// the finite-difference checks in tests/test_adjoint_cpu.py.
#include "adjoint_hlld.h"

using nmhd::adj::HIn;
using nmhd::adj::HOut;
using nmhd::adj::HTape;

extern "C" {

// in: n x 15, out: n x 7 (fluxes), region: n (side * 3 + region)
void adj_hlld_forward(long n, double gamma, const double *in, double *out, int *region) {
  const double ig = 1.0 / (gamma - 1.0);
  for (long i = 0; i < n; ++i) {
    HIn<double> x;
    double *px = &x.bx;
    for (int k = 0; k < 15; ++k) px[k] = in[i * 15 + k];
    HTape<double> t;
    HOut<double> F;
    nmhd::adj::hlld_forward(gamma, ig, x, t, F);
    const double *pf = &F.rho;
    for (int k = 0; k < 7; ++k) out[i * 7 + k] = pf[k];
    if (region) region[i] = (t.left ? 0 : 3) + t.region;
  }
}

// aout: n x 7 adjoints of the fluxes -> ain: n x 15 adjoints of the inputs
void adj_hlld_reverse(long n, double gamma, const double *in, const double *aout, double *ain) {
  const double ig = 1.0 / (gamma - 1.0);
  for (long i = 0; i < n; ++i) {
    HIn<double> x, a;
    double *px = &x.bx;
    for (int k = 0; k < 15; ++k) px[k] = in[i * 15 + k];
    HTape<double> t;
    HOut<double> F, aF;
    nmhd::adj::hlld_forward(gamma, ig, x, t, F);
    double *pa = &aF.rho;
    for (int k = 0; k < 7; ++k) pa[k] = aout[i * 7 + k];
    nmhd::adj::hlld_reverse(gamma, ig, x, t, aF, a);
    const double *pr = &a.bx;
    for (int k = 0; k < 15; ++k) ain[i * 15 + k] = pr[k];
  }
}

}  // extern "C"
