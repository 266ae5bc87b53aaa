// kernels_strict.hip -- the verification instantiation of the unfused path
// (see unfused_impl.h).  This translation unit is compiled with
// -ffp-contract=off so that no multiply-add is fused; together with IEEE
// division / square root and the reference's FP64-promoted HLLD literals it
// reproduces the CPU restatement of the reference step bit for bit.
#include "unfused_impl.h"

namespace nmhd {

void launch_flux_strict(hipStream_t s, int dir, const Box &b, FieldPtrs W, FluxOut F,
                        const StepState *st, int hsel) {
  unfused::launch_flux_t<true>(s, dir, b, W, F, st, hsel);
}

void launch_update_strict(hipStream_t s, const Box &b, FieldPtrs Win, const FluxSet &F,
                          FieldPtrs Wbase, FieldPtrs Wout, const StepState *st, int hsel,
                          bool pace_only) {
  unfused::launch_update_t<true>(s, b, Win, F, Wbase, Wout, st, hsel, pace_only);
}

__global__ void k_hlld_batch_strict(long n, float gamma, FluxIn in, FluxOut out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Flux7 F = hlld_strict(gamma, in.f[0][i], in.f[1][i], in.f[2][i], in.f[3][i], in.f[4][i],
                        in.f[5][i], in.f[6][i], in.f[7][i], in.f[8][i], in.f[9][i], in.f[10][i],
                        in.f[11][i], in.f[12][i], in.f[13][i], in.f[14][i]);
  out.f[0][i] = F.rho;
  out.f[1][i] = F.mn;
  out.f[2][i] = F.mt1;
  out.f[3][i] = F.mt2;
  out.f[4][i] = F.e;
  out.f[5][i] = F.bt1;
  out.f[6][i] = F.bt2;
}

void launch_hlld_batch_strict(hipStream_t s, long n, float gamma, FluxIn in, FluxOut out) {
  hipLaunchKernelGGL(k_hlld_batch_strict, dim3(grid_for(n)), dim3(kBlock), 0, s, n, gamma, in,
                     out);
}

}  // namespace nmhd
