// nsd_stage.hpp -- launchers of the time-step kernels, instantiated per
// transform length (compile-time lines) and launch variant in nsd_stage.hip.
#pragma once

#include "nsd_kernels.hpp"

namespace nsd {

// number of launch variants per length (variant 0 is the default shape)
constexpr int kStageVariants = 5;

// k_row in ROW_STAGE mode for rows of length L (power of two, 16..8192),
// `members` batch members along grid y
template <typename T>
void launch_row_stage(const RowArgs<T>& a, int L, int items, int members, int variant,
                      hipStream_t st);
// k_col in COL_STAGE (scheme from a.phys.scheme) or COL_PRIME mode, length L
template <typename T>
void launch_col_fused(const ColArgs<T>& a, int L, int ncols, int members, int variant,
                      hipStream_t st);
// 1 if `variant` exists for length L and fits `maxlds` bytes of LDS
template <typename T>
int stage_variant_valid(int L, int variant, int maxlds);
// raise the dynamic-LDS limit of every instantiation above
template <typename T>
void stage_lds_limits(int bytes);

}  // namespace nsd
