// nsd_stage.hip -- time-step kernels instantiated for every supported
// transform length, so that all LDS addresses and strides are compile-time
// constants.  Compiled once per precision with -DNSD_T=float / double.
#include "nsd_stage.hpp"

#ifndef NSD_T
#error "NSD_T must be float or double"
#endif

namespace nsd {

namespace {

template <int L>
constexpr int nt_for() {
    return L / P > 256 ? L / P : 256;
}

template <typename T, int LOG>
void row_stage_L(const RowArgs<T>& a, int items, int members, hipStream_t st) {
    constexpr int L = 1 << LOG;
    constexpr int NT = nt_for<L>();
    constexpr int lines = NT * P / L;
    const int blocks = (items + lines - 1) / lines;
    const size_t lds = (size_t)lines * line_regions<T, L>() * (L + L / 16) * sizeof(cx<T>);
    hipLaunchKernelGGL((k_row<T, NT, ROW_STAGE, L>), dim3(blocks, members), dim3(NT), lds, st, a);
}

template <typename T, int LOG>
void col_fused_L(const ColArgs<T>& a, int ncols, int members, hipStream_t st) {
    constexpr int L = 1 << LOG;
    constexpr int NT = nt_for<L>();
    constexpr int lines = NT * P / L;
    const int blocks = (ncols + lines - 1) / lines;
    const size_t lds = (size_t)lines * line_regions<T, L>() * (L + L / 16) * sizeof(cx<T>);
    if (a.mode == COL_PRIME)
        hipLaunchKernelGGL((k_col<T, NT, COL_PRIME, 0, L>), dim3(blocks, members), dim3(NT), lds, st,
                           a);
    else if (a.phys.scheme == 0)
        hipLaunchKernelGGL((k_col<T, NT, COL_STAGE, 0, L>), dim3(blocks, members), dim3(NT), lds,
                           st, a);
    else
        hipLaunchKernelGGL((k_col<T, NT, COL_STAGE, 1, L>), dim3(blocks, members), dim3(NT), lds,
                           st, a);
}

template <typename K>
void lds_attr(K kernel, int bytes) {
    (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              bytes);
}

template <typename T, int LOG>
void limits_L(int bytes) {
    constexpr int L = 1 << LOG;
    constexpr int NT = nt_for<L>();
    lds_attr(k_row<T, NT, ROW_STAGE, L>, bytes);
    lds_attr(k_col<T, NT, COL_PRIME, 0, L>, bytes);
    lds_attr(k_col<T, NT, COL_STAGE, 0, L>, bytes);
    lds_attr(k_col<T, NT, COL_STAGE, 1, L>, bytes);
}

int log2i(int L) {
    int l = 0;
    while ((1 << l) < L) ++l;
    return l;
}

}  // namespace

#define NSD_FOR_LOGS(X) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13)

template <typename T>
void launch_row_stage(const RowArgs<T>& a, int L, int items, int members, hipStream_t st) {
    switch (log2i(L)) {
#define X(n) \
    case n: row_stage_L<T, n>(a, items, members, st); break;
        NSD_FOR_LOGS(X)
#undef X
    default: break;
    }
}

template <typename T>
void launch_col_fused(const ColArgs<T>& a, int L, int ncols, int members, hipStream_t st) {
    switch (log2i(L)) {
#define X(n) \
    case n: col_fused_L<T, n>(a, ncols, members, st); break;
        NSD_FOR_LOGS(X)
#undef X
    default: break;
    }
}

template <typename T>
void stage_lds_limits(int bytes) {
#define X(n) limits_L<T, n>(bytes);
    NSD_FOR_LOGS(X)
#undef X
}

template void launch_row_stage<NSD_T>(const RowArgs<NSD_T>&, int, int, int, hipStream_t);
template void launch_col_fused<NSD_T>(const ColArgs<NSD_T>&, int, int, int, hipStream_t);
template void stage_lds_limits<NSD_T>(int);

}  // namespace nsd
