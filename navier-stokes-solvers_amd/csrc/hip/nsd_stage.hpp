// nsd_stage.hpp -- launchers of the time-step kernels, instantiated per
// transform length (compile-time lines) in nsd_stage.hip.
#pragma once

#include "nsd_kernels.hpp"

namespace nsd {

// k_row in ROW_STAGE mode for rows of length L (power of two, 16..8192),
// `members` batch members along grid y
template <typename T>
void launch_row_stage(const RowArgs<T>& a, int L, int items, int members, hipStream_t st);
// k_col in COL_STAGE (scheme from a.phys.scheme) or COL_PRIME mode, length L
template <typename T>
void launch_col_fused(const ColArgs<T>& a, int L, int ncols, int members, hipStream_t st);
// raise the dynamic-LDS limit of every instantiation above
template <typename T>
void stage_lds_limits(int bytes);

}  // namespace nsd
