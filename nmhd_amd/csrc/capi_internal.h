// capi_internal.h -- descriptor checks and derived constants shared by the
// single-box (capi.hip) and multi-box (capi_batch.hip) entry points.
#pragma once

#include <string>

#include "../../include/nmhd.h"
#include "nmhd_kernels.h"

namespace nmhd {

// Validate a box descriptor; on failure `why` names the offending field.
bool desc_valid(const nmhd_desc *d, std::string &why);

// Derived constants, in FP32 exactly as the reference's constant
// expressions (simulation_config_default.h:79-103).
void derive_box(const nmhd_desc *d, Box &b, nmhd_geometry &g);

// The frozen-dt step W_out = S(W_in; dt) of a periodic box with the
// production stage kernels, enqueued on s: st receives dt (no other field
// of it is read), Wstar holds the midpoint stage state (unused for
// NMHD_EULER1), and of F only the 11 arrays the face kernel hands to the
// update kernel (tiled::stage_buf) are used.
void enqueue_frozen_step(hipStream_t s, const Box &b, int integrator, FieldPtrs Win,
                         FieldPtrs Wstar, FieldPtrs Wout, const FluxSet &F, StepState *st,
                         float dt);

// The message nmhd_adj_last_error(NULL) returns (failures of calls that
// have no nmhd_adj object).
void set_adj_global_error(const std::string &msg);

}  // namespace nmhd
