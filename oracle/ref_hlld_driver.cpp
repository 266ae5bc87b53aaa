// ref_hlld_driver.cpp -- compiles the reference's own HLLD Riemann solver
// (library/hlld.inl, read in place from the reference tree) for the host, so
// the oracle's restatement of it can be checked bit for bit.
//
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ by oracle/build.py when
// the reference tree is present; never linked into the product.
//
// hlld.inl is a `__host__ __device__` function that expects the includer to
// provide Real, kGamma, sq, absR and sqrtR (mhd_solver.inl:758 includes it
// after simulation_config.h and library/math.h).  hipcc supplies the CUDA
// qualifiers; library/math.h is the reference's own header; Real is the
// single-precision type of simulation_config_default.h:11-13, and kGamma is a
// float variable so one build serves every adiabatic index.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>

#define USE_SINGLE_PRECISION
typedef float Real;
#include "library/math.h"

static Real g_ref_gamma = Real(5.0f / 3.0f);
#define kGamma g_ref_gamma

// The reference ran hlld_flux as CUDA device code, where fabs/fmin/fmax of a
// float resolve to the float overloads.  In host C++ the unqualified names
// find only the C double versions, which would promote
// `absR(Bx)/sqrtR(dens_L_s)` (hlld.inl:167,170) to FP64.  Bring the float
// overloads into scope so the host build rounds as the device code did.
using std::fabs;
using std::fmax;
using std::fmin;

#include "library/hlld.inl"

extern "C" void ref_hlld_batch(size_t n, float gamma, float *const in[15],
                               float *out[7]) {
  g_ref_gamma = gamma;
  for (size_t i = 0; i < n; ++i) {
    hlld_flux(in[0][i], in[1][i], in[2][i], in[3][i], in[4][i], in[5][i],
              in[6][i], in[7][i], in[8][i], in[9][i], in[10][i], in[11][i],
              in[12][i], in[13][i], in[14][i], out[0][i], out[1][i], out[2][i],
              out[3][i], out[4][i], out[5][i], out[6][i]);
  }
}
