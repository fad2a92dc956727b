// Job planning shared by the forward-transform handle and the DWTForward53/97
// drop-in classes.
#pragma once
#include <cstdint>
#include <vector>

#include "../tc_kernels.h"

namespace tc {
void plan_forward_levels(int w, int h, int levels, int first, uint64_t in_off, uint32_t in_pitch, uint64_t plane,
                         const uint64_t work[2], int TX, int TY, std::vector<std::vector<tc_fdwt_job>>& out);
}  // namespace tc
