// kernels_unfused.hip -- fast-arithmetic instantiation of the unfused path
// (see unfused_impl.h).
#include "unfused_impl.h"

namespace nmhd {

void launch_flux(hipStream_t s, int dir, const Box &b, FieldPtrs W, FluxOut F,
                 const StepState *st, int hsel) {
  unfused::launch_flux_t<false>(s, dir, b, W, F, st, hsel);
}

void launch_update(hipStream_t s, const Box &b, FieldPtrs Win, const FluxSet &F,
                   FieldPtrs Wbase, FieldPtrs Wout, const StepState *st, int hsel,
                   bool pace_only) {
  unfused::launch_update_t<false>(s, b, Win, F, Wbase, Wout, st, hsel, pace_only);
}

}  // namespace nmhd
