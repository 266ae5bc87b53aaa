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

}  // namespace nmhd
