// nsd_stage.hip -- time-step kernels instantiated for every supported
// transform length, so that all LDS addresses and strides are compile-time
// constants.  Compiled once per precision with -DNSD_T=float / double.
//
// Lengths 64..1024 (the ensemble sizes) also get launch variants that the
// autotuner chooses between: workgroup size (lines per workgroup) and whether
// the inverse transforms run in lock-step pairs.
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

// variant V of length L: threads per workgroup and pairing (-1 automatic)
template <int L, int V>
constexpr int var_nt() {
    return V <= 1 ? nt_for<L>() : (V <= 3 ? 2 * nt_for<L>() : nt_for<L>() / 2);
}
template <int V>
constexpr int var_pr() {
    return (V == 1 || V == 3) ? 0 : -1;
}
template <int L, int V>
constexpr bool var_ok() {
    constexpr int NT = var_nt<L, V>();
    return V == 0 || (L >= 64 && L <= 1024 && NT <= 1024 && NT >= L / P && NT >= 64);
}

// the stage kernels' LDS (FP64 lines of 4096 take a second line region)
template <typename T, int L, int V>
size_t stage_lds() {
    constexpr int NT = var_nt<L, V>();
    constexpr int lines = NT * P / L;
    constexpr int PR = var_pr<V>();
    return (size_t)lines * stage_regions<T, L, NT, PR>() * region_elems<T, L, NT, PR>() *
           sizeof(cx<T>);
}

// ERR: the error controller's instantiations (launch variant 0 only)
template <typename T, int LOG, int V, bool ERR = false>
void row_stage_LV(const RowArgs<T>& a, int items, int members, hipStream_t st) {
    constexpr int L = 1 << LOG;
    if constexpr (var_ok<L, V>()) {
        constexpr int NT = var_nt<L, V>();
        constexpr int PR = var_pr<V>();
        constexpr int lines = NT * P / L;
        const int blocks = (items + lines - 1) / lines;
        const size_t lds = stage_lds<T, L, V>();
        hipLaunchKernelGGL((k_row<T, NT, ROW_STAGE, L, PR, ERR>), dim3(blocks, members), dim3(NT),
                           lds, st, a);
    }
}

template <typename T, int LOG, int V, bool ERR = false>
void col_fused_LV(const ColArgs<T>& a, int ncols, int members, hipStream_t st) {
    constexpr int L = 1 << LOG;
    if constexpr (var_ok<L, V>()) {
        constexpr int NT = var_nt<L, V>();
        constexpr int PR = var_pr<V>();
        constexpr int lines = NT * P / L;
        const int blocks = (ncols + lines - 1) / lines;
        const dim3 grid(blocks, members);
        const size_t lds = stage_lds<T, L, V>();
        if (a.mode == COL_PRIME)
            hipLaunchKernelGGL((k_col<T, NT, COL_PRIME, 0, L, PR>), grid, dim3(NT), lds, st, a);
        else if (a.phys.scheme == 0)
            hipLaunchKernelGGL((k_col<T, NT, COL_STAGE, 0, L, PR, ERR>), grid, dim3(NT), lds, st,
                               a);
        else
            hipLaunchKernelGGL((k_col<T, NT, COL_STAGE, 1, L, PR, ERR>), grid, dim3(NT), lds, st,
                               a);
    }
}

template <typename K>
void lds_attr(K kernel, int bytes) {
    (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              bytes);
}

template <typename T, int LOG, int V>
void limits_LV(int bytes) {
    constexpr int L = 1 << LOG;
    if constexpr (var_ok<L, V>()) {
        constexpr int NT = var_nt<L, V>();
        constexpr int PR = var_pr<V>();
        lds_attr(k_row<T, NT, ROW_STAGE, L, PR>, bytes);
        lds_attr(k_col<T, NT, COL_PRIME, 0, L, PR>, bytes);
        lds_attr(k_col<T, NT, COL_STAGE, 0, L, PR>, bytes);
        lds_attr(k_col<T, NT, COL_STAGE, 1, L, PR>, bytes);
        if constexpr (V == 0) {
            lds_attr(k_row<T, NT, ROW_STAGE, L, PR, true>, bytes);
            lds_attr(k_col<T, NT, COL_STAGE, 0, L, PR, true>, bytes);
            lds_attr(k_col<T, NT, COL_STAGE, 1, L, PR, true>, bytes);
        }
    }
}

template <typename T, int LOG, int V>
int var_pairs() {
    constexpr int L = 1 << LOG;
    if constexpr (var_ok<L, V>()) return stage_layout<T, L, var_nt<L, V>(), var_pr<V>()>() ? 1 : 0;
    return 0;
}

template <typename T, int LOG, int V>
int var_valid(int maxlds) {
    constexpr int L = 1 << LOG;
    if constexpr (var_ok<L, V>())
        return stage_lds<T, L, V>() <= (size_t)maxlds ? 1 : 0;
    return 0;
}

int log2i(int L) {
    int l = 0;
    while ((1 << l) < L) ++l;
    return l;
}

}  // namespace

#define NSD_FOR_LOGS(X) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13)
#define NSD_FOR_VARS(X, n) X(n, 0) X(n, 1) X(n, 2) X(n, 3) X(n, 4)

template <typename T>
void launch_row_stage(const RowArgs<T>& a, int L, int items, int members, int variant,
                      hipStream_t st, bool errctl) {
    const int lg = log2i(L);
    if (errctl) {
#define E_(n) \
    if (lg == n) return row_stage_LV<T, n, 0, true>(a, items, members, st);
        NSD_FOR_LOGS(E_)
#undef E_
        return;
    }
#define V_(n, v) \
    if (lg == n && variant == v) return row_stage_LV<T, n, v>(a, items, members, st);
#define X(n) NSD_FOR_VARS(V_, n)
    NSD_FOR_LOGS(X)
#undef X
#undef V_
}

template <typename T>
void launch_col_fused(const ColArgs<T>& a, int L, int ncols, int members, int variant,
                      hipStream_t st, bool errctl) {
    const int lg = log2i(L);
    if (errctl && a.mode == COL_STAGE) {
#define E_(n) \
    if (lg == n) return col_fused_LV<T, n, 0, true>(a, ncols, members, st);
        NSD_FOR_LOGS(E_)
#undef E_
        return;
    }
#define V_(n, v) \
    if (lg == n && variant == v) return col_fused_LV<T, n, v>(a, ncols, members, st);
#define X(n) NSD_FOR_VARS(V_, n)
    NSD_FOR_LOGS(X)
#undef X
#undef V_
}

template <typename T>
int stage_variant_valid(int L, int variant, int maxlds) {
    const int lg = log2i(L);
#define V_(n, v) \
    if (lg == n && variant == v) return var_valid<T, n, v>(maxlds);
#define X(n) NSD_FOR_VARS(V_, n)
    NSD_FOR_LOGS(X)
#undef X
#undef V_
    return 0;
}

template <typename T>
int stage_variant_pairs(int L, int variant) {
    const int lg = log2i(L);
#define V_(n, v) \
    if (lg == n && variant == v) return var_pairs<T, n, v>();
#define X(n) NSD_FOR_VARS(V_, n)
    NSD_FOR_LOGS(X)
#undef X
#undef V_
    return 0;
}

template <typename T>
void stage_lds_limits(int bytes) {
#define V_(n, v) limits_LV<T, n, v>(bytes);
#define X(n) NSD_FOR_VARS(V_, n)
    NSD_FOR_LOGS(X)
#undef X
#undef V_
}

template void launch_row_stage<NSD_T>(const RowArgs<NSD_T>&, int, int, int, int, hipStream_t,
                                      bool);
template void launch_col_fused<NSD_T>(const ColArgs<NSD_T>&, int, int, int, int, hipStream_t,
                                      bool);

// workgroups of the default column launch over ncols columns of length L
// (the number of partials the error controller's last stage leaves)
template <typename T>
int col_fused_blocks(int L, int ncols) {
    const int NT = L / P > 256 ? L / P : 256;
    const int lines = NT * P / L;
    return (ncols + lines - 1) / lines;
}
template int col_fused_blocks<NSD_T>(int, int);
template int stage_variant_valid<NSD_T>(int, int, int);
template int stage_variant_pairs<NSD_T>(int, int);
template void stage_lds_limits<NSD_T>(int);

}  // namespace nsd
