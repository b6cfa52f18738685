// nsd_stage.hpp -- launchers of the time-step kernels, instantiated per
// transform length (compile-time lines) and launch variant in nsd_stage.hip.
#pragma once

#include "nsd_kernels.hpp"

namespace nsd {

// number of launch variants per length (variant 0 is the default shape)
constexpr int kStageVariants = 5;

// k_row in ROW_STAGE mode for rows of length L (power of two, 16..8192),
// `members` batch members along grid y
// (errctl: the error controller's instantiation, launch variant 0)
template <typename T>
void launch_row_stage(const RowArgs<T>& a, int L, int items, int members, int variant,
                      hipStream_t st, bool errctl = false);
// k_col in COL_STAGE (scheme from a.phys.scheme) or COL_PRIME mode, length L
template <typename T>
void launch_col_fused(const ColArgs<T>& a, int L, int ncols, int members, int variant,
                      hipStream_t st, bool errctl = false);
// workgroups of the default (variant 0) column launch over ncols columns
template <typename T>
int col_fused_blocks(int L, int ncols);
// 1 if `variant` exists for length L and fits `maxlds` bytes of LDS
template <typename T>
int stage_variant_valid(int L, int variant, int maxlds);
// 1 if the stage kernels of `variant` at length L read (k_row) and write
// (k_col) the mixed fields in a layout of their own (stage_layout)
template <typename T>
int stage_variant_pairs(int L, int variant);
// raise the dynamic-LDS limit of every instantiation above
template <typename T>
void stage_lds_limits(int bytes);

}  // namespace nsd
