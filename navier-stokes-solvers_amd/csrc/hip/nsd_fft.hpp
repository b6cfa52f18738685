// nsd_fft.hpp -- workgroup-level Stockham FFT building blocks for gfx950.
//
// A "line" is one 1-D transform of length L (power of two, 16..8192) held by
// TPL = L/P threads, each thread owning P complex values (P = NSD_POINTS,
// 8 by default).  Between passes the line lives in LDS.  Thread j of a line
// always reads, and after the last pass always holds, the elements
// n = j + s*TPL (s = 0..P-1): the "standard pattern".  Because the first pass reads and the last pass
// writes that same pattern, pre- and post-processing (spectral derivatives,
// masks, stage updates, pointwise products) are fused into the registers
// between transforms without any extra data movement.
//
// Pass plan: floor(log_P L) radix-P passes, then one radix-2/4/8 pass when
// L is not a power of P.  Pass with radix R and stride Ns (product of the
// previous radices): butterfly b reads b + r*L/R, multiplies by
// W_{Ns R}^{(b mod Ns) r}, does an R-point DFT, and writes (b - b mod Ns)*R
// + b mod Ns + r*Ns (Stockham autosort, natural-order output).
// Twiddles come from a table W_L^m = exp(-2 pi i m / L), m < L, computed on
// the host in FP64 with exact integer angle reduction (FP32 tables are the
// FP64 values rounded once); a radix-R butterfly loads W, W^2, W^4, W^8 and
// forms the other powers by complex products.
#pragma once

#include <hip/hip_runtime.h>

#ifndef NSD_POINTS
#define NSD_POINTS 8
#endif

namespace nsd {

constexpr int P = NSD_POINTS;  // complex values per thread per line
static_assert(P == 8 || P == 16, "points per thread must be 8 or 16");

template <typename T>
struct cx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> mk(T a, T b) {
    cx<T> r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cx<T> operator*(T s, cx<T> a) { return mk(s * a.x, s * a.y); }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {
    return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T>
__device__ __forceinline__ cx<T> conjg(cx<T> a) { return mk(a.x, -a.y); }
// multiply by i*s
template <typename T>
__device__ __forceinline__ cx<T> mul_i(cx<T> a, T s) { return mk(-s * a.y, s * a.x); }
// multiply by DIR*i (DIR = -1: -i, forward; +1: +i, inverse)
template <int DIR, typename T>
__device__ __forceinline__ cx<T> rot90(cx<T> a) {
    return DIR < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}

// LDS padding: one spare element every 16 keeps the strided Stockham
// writes of the first passes off a single bank group.
__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }

template <int DIR, typename T>
__device__ __forceinline__ void dft2(cx<T>& a, cx<T>& b) {
    cx<T> t = a - b;
    a = a + b;
    b = t;
}

template <int DIR, typename T>
__device__ __forceinline__ void dft4(cx<T>& a0, cx<T>& a1, cx<T>& a2, cx<T>& a3) {
    cx<T> s02 = a0 + a2, d02 = a0 - a2;
    cx<T> s13 = a1 + a3, d13 = rot90<DIR>(a1 - a3);
    a0 = s02 + s13;
    a2 = s02 - s13;
    a1 = d02 + d13;
    a3 = d02 - d13;
}

template <typename T>
struct Consts {
    static constexpr T c8 = T(0.92387953251128675612818318939678828682241662586);  // cos(pi/8)
    static constexpr T s8 = T(0.38268343236508977172845998403039886676134456248);  // sin(pi/8)
    static constexpr T r2 = T(0.70710678118654752440084436210484903928483593768);  // sqrt(1/2)
};

// multiply by W^e, W = exp(DIR * 2 pi i / 16)
template <int DIR, int E, typename T>
__device__ __forceinline__ cx<T> w16(cx<T> a) {
    constexpr T c8 = Consts<T>::c8, s8 = Consts<T>::s8, r2 = Consts<T>::r2;
    constexpr T sg = DIR < 0 ? T(-1) : T(1);
    if constexpr (E == 0) {
        return a;
    } else if constexpr (E == 1) {
        return cmul(a, mk(c8, sg * s8));
    } else if constexpr (E == 2) {
        return cmul(a, mk(r2, sg * r2));
    } else if constexpr (E == 3) {
        return cmul(a, mk(s8, sg * c8));
    } else if constexpr (E == 4) {
        return rot90<DIR>(a);
    } else if constexpr (E == 6) {
        return cmul(a, mk(-r2, sg * r2));
    } else {  // E == 9
        return cmul(a, mk(-c8, -sg * s8));
    }
}

// 8-point DFT (natural order in and out), v0..v7 references.
template <int DIR, typename T>
__device__ __forceinline__ void dft8(cx<T>& v0, cx<T>& v1, cx<T>& v2, cx<T>& v3, cx<T>& v4,
                                     cx<T>& v5, cx<T>& v6, cx<T>& v7) {
    // r = 2a + b: 4-point DFTs over a for b = 0,1
    dft4<DIR>(v0, v2, v4, v6);
    dft4<DIR>(v1, v3, v5, v7);
    // twiddle W8^{b k1} for b = 1: k1 = 0..3 -> W16^{0,2,4,6}
    v3 = w16<DIR, 2>(v3);
    v5 = w16<DIR, 4>(v5);
    v7 = w16<DIR, 6>(v7);
    // 2-point over b; output k = k1 + 4 k2
    cx<T> o0 = v0 + v1, o4 = v0 - v1;
    cx<T> o1 = v2 + v3, o5 = v2 - v3;
    cx<T> o2 = v4 + v5, o6 = v4 - v5;
    cx<T> o3 = v6 + v7, o7 = v6 - v7;
    v0 = o0; v1 = o1; v2 = o2; v3 = o3; v4 = o4; v5 = o5; v6 = o6; v7 = o7;
}

// 16-point DFT on v[base + stride*r], r = 0..15 (natural order in and out).
template <int DIR, typename T>
__device__ __forceinline__ void dft16(cx<T>* v) {
    // r = 4a + b, k = k1 + 4 k2
#pragma unroll
    for (int b = 0; b < 4; ++b) dft4<DIR>(v[b], v[b + 4], v[b + 8], v[b + 12]);
    // v[b + 4 k1] now holds y[b][k1]; twiddle by W16^{b k1}
    v[5] = w16<DIR, 1>(v[5]);
    v[9] = w16<DIR, 2>(v[9]);
    v[13] = w16<DIR, 3>(v[13]);
    v[6] = w16<DIR, 2>(v[6]);
    v[10] = w16<DIR, 4>(v[10]);
    v[14] = w16<DIR, 6>(v[14]);
    v[7] = w16<DIR, 3>(v[7]);
    v[11] = w16<DIR, 6>(v[11]);
    v[15] = w16<DIR, 9>(v[15]);
    // 4-point DFTs over b for each k1: inputs v[4k1 + b], outputs X[k1 + 4 k2]
    cx<T> o[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        cx<T> a0 = v[4 * k1 + 0], a1 = v[4 * k1 + 1], a2 = v[4 * k1 + 2], a3 = v[4 * k1 + 3];
        dft4<DIR>(a0, a1, a2, a3);
        o[k1] = a0;
        o[k1 + 4] = a1;
        o[k1 + 8] = a2;
        o[k1 + 12] = a3;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = o[i];
}

// Line geometry shared by every kernel.
struct Line {
    int L;    // transform length
    int TPL;  // threads per line = L/P
    int nP;   // number of radix-P passes
    int rem;  // trailing radix (1, 2, 4 or 8)
};

__host__ __device__ inline Line make_line(int L) {
    Line g;
    g.L = L;
    g.TPL = L / P;
    int n = 0, r = L;
    while (r >= P && (r % P) == 0) {
        r /= P;
        ++n;
    }
    g.nP = n;
    g.rem = r;
    return g;
}

// threads per workgroup for lines of length L (a workgroup holds whole lines)
__host__ __device__ constexpr int threads_for(int L) { return L / P > 256 ? L / P : 256; }

// One Stockham pass of radix R on the thread's P registers (Q = P/R
// butterflies per thread).  If `last`, the outputs stay in v (standard
// pattern); otherwise they are written to the line's LDS region.
template <int DIR, int R, typename T>
__device__ __forceinline__ void stockham_pass(cx<T> (&v)[P], int j, const Line& g, int Ns,
                                              const cx<T>* __restrict__ tw, cx<T>* lds,
                                              bool last) {
    constexpr int Q = P / R;
    const int tstride = g.L / (Ns * R);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int b = j + q * g.TPL;
        const int k = b & (Ns - 1);
        if (Ns > 1) {
            // powers W^r of W = W_{Ns R}^k: W, W^2, W^4, W^8 from the table,
            // the others as products (at most three roundings each)
            const int step = k * tstride;
            cx<T> w[R];
            w[1] = tw[step];
            if constexpr (R > 2) w[2] = tw[2 * step];
            if constexpr (R > 4) w[4] = tw[4 * step];
            if constexpr (R > 8) w[8] = tw[8 * step];
            if constexpr (R > 2) w[3] = cmul(w[1], w[2]);
            if constexpr (R > 4) {
                w[5] = cmul(w[1], w[4]);
                w[6] = cmul(w[2], w[4]);
                w[7] = cmul(w[3], w[4]);
            }
            if constexpr (R > 8) {
#pragma unroll
                for (int r = 9; r < 16; ++r) w[r] = cmul(w[r - 8], w[8]);
            }
#pragma unroll
            for (int r = 1; r < R; ++r)
                v[q + r * Q] = DIR < 0 ? cmul(v[q + r * Q], w[r]) : cmulc(v[q + r * Q], w[r]);
        }
        if constexpr (R == 16) {
            dft16<DIR>(v);
        } else if constexpr (R == 8) {
            dft8<DIR>(v[q], v[q + Q], v[q + 2 * Q], v[q + 3 * Q], v[q + 4 * Q], v[q + 5 * Q],
                      v[q + 6 * Q], v[q + 7 * Q]);
        } else if constexpr (R == 4) {
            dft4<DIR>(v[q], v[q + Q], v[q + 2 * Q], v[q + 3 * Q]);
        } else {
            dft2<DIR>(v[q], v[q + Q]);
        }
        if (!last) {
            const int base = (b - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lds[pad(base + r * Ns)] = v[q + r * Q];
        }
    }
}

template <typename T>
__device__ __forceinline__ void lds_read_std(cx<T> (&v)[P], const cx<T>* lds, int j, int TPL) {
#pragma unroll
    for (int s = 0; s < P; ++s) v[s] = lds[pad(j + s * TPL)];
}

template <typename T>
__device__ __forceinline__ void lds_write_std(const cx<T> (&v)[P], cx<T>* lds, int j, int TPL) {
#pragma unroll
    for (int s = 0; s < P; ++s) lds[pad(j + s * TPL)] = v[s];
}

// Full line transform: v (standard pattern, in registers) -> v (standard
// pattern, natural-order result).  Unnormalised.  DIR = -1 forward
// (exp(-i...)), +1 inverse.  Every thread of the workgroup must call it
// (it contains workgroup barriers); the line's LDS region must be free on
// entry (a barrier is taken first) and is free again on return.
template <int DIR, typename T>
__device__ __forceinline__ void line_fft(cx<T> (&v)[P], cx<T>* lds, int j, const Line& g,
                                         const cx<T>* __restrict__ tw) {
    int Ns = 1;
    const int passes = g.nP + (g.rem > 1 ? 1 : 0);
    for (int p = 0; p < passes; ++p) {
        const bool last = (p == passes - 1);
        if (p > 0) {
            lds_read_std(v, lds, j, g.TPL);
        }
        __syncthreads();  // reads of the previous pass (or earlier users) done
        if (p < g.nP) {
            stockham_pass<DIR, P>(v, j, g, Ns, tw, lds, last);
            Ns *= P;
        } else if (g.rem == 8) {
            stockham_pass<DIR, 8>(v, j, g, Ns, tw, lds, last);
        } else if (g.rem == 4) {
            stockham_pass<DIR, 4>(v, j, g, Ns, tw, lds, last);
        } else {
            stockham_pass<DIR, 2>(v, j, g, Ns, tw, lds, last);
        }
        if (!last) __syncthreads();
    }
}

// ------------------------------------------------------------------------
// Compile-time lines: the same Stockham plan with L, the radix of every
// pass and every stride known to the compiler, so that LDS addresses and
// twiddle indices fold into immediate offsets.  Used by the time-step
// kernels, which run at a fixed size for the whole simulation.

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
constexpr int ipowc(int b, int e) { return e == 0 ? 1 : b * ipowc(b, e - 1); }

template <int L>
struct CLine {
    static constexpr int TPL = L / P;
    static constexpr int nP = ilog2c(L) / ilog2c(P);
    static constexpr int rem = L / ipowc(P, nP);
    static constexpr int passes = nP + (rem > 1 ? 1 : 0);
    static constexpr int radix(int p) { return p < nP ? P : rem; }
    static constexpr int stride(int p) { return p == 0 ? 1 : stride(p - 1) * radix(p - 1); }
};

// LDS offset of element i under the one-in-sixteen padding
constexpr int padc(int i) { return i + (i >> 4); }

// Twiddle sources of the compile-time passes.  A radix-R butterfly of pass
// PASS needs W^(k 2^e), e = 0 .. log2(R) - 1, for k = its index mod Ns.
// TwTable reads them from the table in global memory.  TwRegs holds W^k of
// every full-radix pass in registers, loaded once per thread, and forms the
// powers by repeated squaring: no global load then sits inside those passes,
// so they never wait for loads issued ahead of the transform (the vector
// memory counter retires loads in order).  Squaring adds a few ulp to the
// higher powers (W^4k is three roundings from the table value).  The short
// closing pass of a length that is not a power of P (several butterflies
// per thread) still reads the table.
template <typename T>
struct TwTable {
    const cx<T>* __restrict__ tw;
    template <int L, int PASS>
    __device__ __forceinline__ cx<T> get(int, int e, int k) const {
        constexpr int tstride = L / (CLine<L>::stride(PASS) * CLine<L>::radix(PASS));
        return tw[(k << e) * tstride];
    }
};

template <int L, typename T>
struct TwRegs {
    const cx<T>* __restrict__ tab;
    cx<T> w[CLine<L>::passes];  // W^k of each full-radix pass
    template <int L2, int PASS>
    __device__ __forceinline__ cx<T> get(int q, int e, int k) const {
        if constexpr (CLine<L>::radix(PASS) == P) {
            cx<T> r = w[PASS];
            for (int i = 0; i < e; ++i) r = cmul(r, r);
            return r;
        } else {
            return TwTable<T>{tab}.template get<L, PASS>(q, e, k);
        }
    }
};

template <int L, int PASS, typename T>
__device__ __forceinline__ void load_tw_regs(TwRegs<L, T>& t, int j, const cx<T>* __restrict__ tw) {
    if constexpr (PASS == 0) t.tab = tw;
    if constexpr (PASS < CLine<L>::passes) {
        using G = CLine<L>;
        constexpr int Ns = G::stride(PASS);
        if constexpr (Ns > 1 && G::radix(PASS) == P)
            t.w[PASS] = tw[(j & (Ns - 1)) * (L / (Ns * P))];
        load_tw_regs<L, PASS + 1>(t, j, tw);
    }
}

template <int L, typename T>
__device__ __forceinline__ void lds_read_c(cx<T> (&v)[P], const cx<T>* lds, int j) {
    constexpr int TPL = CLine<L>::TPL;
    if constexpr (TPL % 16 == 0) {
        const cx<T>* b = lds + pad(j);
#pragma unroll
        for (int s = 0; s < P; ++s) v[s] = b[padc(s * TPL)];
    } else {
#pragma unroll
        for (int s = 0; s < P; ++s) v[s] = lds[pad(j + s * TPL)];
    }
}

template <int L, typename T>
__device__ __forceinline__ void lds_write_c(const cx<T> (&v)[P], cx<T>* lds, int j) {
    constexpr int TPL = CLine<L>::TPL;
    if constexpr (TPL % 16 == 0) {
        cx<T>* b = lds + pad(j);
#pragma unroll
        for (int s = 0; s < P; ++s) b[padc(s * TPL)] = v[s];
    } else {
#pragma unroll
        for (int s = 0; s < P; ++s) lds[pad(j + s * TPL)] = v[s];
    }
}

template <int DIR, int L, int PASS, typename T, typename TW>
__device__ __forceinline__ void cpass(cx<T> (&v)[P], int j, const TW& tw, cx<T>* lds) {
    using G = CLine<L>;
    constexpr int R = G::radix(PASS);
    constexpr int Ns = G::stride(PASS);
    constexpr bool last = (PASS == G::passes - 1);
    constexpr int Q = P / R;
    constexpr int TPL = G::TPL;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int b = j + q * TPL;
        const int k = b & (Ns - 1);
        if constexpr (Ns > 1) {
            cx<T> w[R];
            w[1] = tw.template get<L, PASS>(q, 0, k);
            if constexpr (R > 2) w[2] = tw.template get<L, PASS>(q, 1, k);
            if constexpr (R > 4) w[4] = tw.template get<L, PASS>(q, 2, k);
            if constexpr (R > 8) w[8] = tw.template get<L, PASS>(q, 3, k);
            if constexpr (R > 2) w[3] = cmul(w[1], w[2]);
            if constexpr (R > 4) {
                w[5] = cmul(w[1], w[4]);
                w[6] = cmul(w[2], w[4]);
                w[7] = cmul(w[3], w[4]);
            }
            if constexpr (R > 8) {
#pragma unroll
                for (int r = 9; r < 16; ++r) w[r] = cmul(w[r - 8], w[8]);
            }
#pragma unroll
            for (int r = 1; r < R; ++r)
                v[q + r * Q] = DIR < 0 ? cmul(v[q + r * Q], w[r]) : cmulc(v[q + r * Q], w[r]);
        }
        if constexpr (R == 16) {
            dft16<DIR>(v);
        } else if constexpr (R == 8) {
            dft8<DIR>(v[q], v[q + Q], v[q + 2 * Q], v[q + 3 * Q], v[q + 4 * Q], v[q + 5 * Q],
                      v[q + 6 * Q], v[q + 7 * Q]);
        } else if constexpr (R == 4) {
            dft4<DIR>(v[q], v[q + Q], v[q + 2 * Q], v[q + 3 * Q]);
        } else if constexpr (R == 2) {
            dft2<DIR>(v[q], v[q + Q]);
        }
        if constexpr (!last) {
            const int base = (b - k) * R + k;
            if constexpr (Ns % 16 == 0) {
                cx<T>* o = lds + pad(base);
#pragma unroll
                for (int r = 0; r < R; ++r) o[r * (Ns + Ns / 16)] = v[q + r * Q];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) lds[pad(base + r * Ns)] = v[q + r * Q];
            }
        }
    }
}

// Butterfly of thread j in pass PASS: j itself, except in the stride-8 pass
// of 4096-point FP32 lines.  There thread j = 8 a + k writes its eight
// outputs at 68 a + k + 8 r (one-in-sixteen padding), so the two values of a
// in a 16-lane write phase sit 8 banks apart and half their banks collide;
// swapping bits 3 and 4 of j pairs a with a + 2 (136 elements, 16 banks
// apart) and the phase covers 32 banks once.  The pass reads its inputs by
// the same index, contiguous 32-lane groups either way; twiddles depend on
// k = j mod 8 only.
template <typename T, int L, int PASS>
__device__ __forceinline__ int pass_lane(int j) {
    if constexpr (sizeof(T) == 4 && L == 4096 && CLine<L>::stride(PASS) == 8 &&
                  PASS + 1 < CLine<L>::passes)
        return j ^ ((((j >> 3) ^ (j >> 4)) & 1) * 0x18);
    else
        return j;
}

template <int DIR, int L, int PASS, typename T, typename TW>
__device__ __forceinline__ void cpasses(cx<T> (&v)[P], cx<T>* lds, int j, const TW& tw) {
    if constexpr (PASS < CLine<L>::passes) {
        const int jp = pass_lane<T, L, PASS>(j);
        if constexpr (PASS > 0) lds_read_c<L>(v, lds, jp);
        __syncthreads();
        cpass<DIR, L, PASS>(v, jp, tw, lds);
        if constexpr (PASS + 1 < CLine<L>::passes) __syncthreads();
        cpasses<DIR, L, PASS + 1>(v, lds, j, tw);
    }
}

// Compile-time-length version of line_fft (same contract).
template <int DIR, int L, typename T>
__device__ __forceinline__ void line_fft_c(cx<T> (&v)[P], cx<T>* lds, int j,
                                           const cx<T>* __restrict__ tw) {
    cpasses<DIR, L, 0>(v, lds, j, TwTable<T>{tw});
}
// the same with register-resident twiddles
template <int DIR, int L, typename T>
__device__ __forceinline__ void line_fft_c(cx<T> (&v)[P], cx<T>* lds, int j,
                                           const TwRegs<L, T>& tw) {
    cpasses<DIR, L, 0>(v, lds, j, tw);
}

// Two independent transforms of compile-time length L run in lock-step: one
// barrier pair per pass serves both, the twiddles are computed once, and the
// two dependency chains interleave.  lds1/lds2 are two separate line regions.
template <int DIR, int L, int PASS, typename T, typename TW>
__device__ __forceinline__ void cpass2(cx<T> (&v1)[P], cx<T> (&v2)[P], int j, const TW& tw,
                                       cx<T>* lds1, cx<T>* lds2) {
    using G = CLine<L>;
    constexpr int R = G::radix(PASS);
    constexpr int Ns = G::stride(PASS);
    constexpr bool last = (PASS == G::passes - 1);
    constexpr int Q = P / R;
    constexpr int TPL = G::TPL;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int b = j + q * TPL;
        const int k = b & (Ns - 1);
        if constexpr (Ns > 1) {
            cx<T> w[R];
            w[1] = tw.template get<L, PASS>(q, 0, k);
            if constexpr (R > 2) w[2] = tw.template get<L, PASS>(q, 1, k);
            if constexpr (R > 4) w[4] = tw.template get<L, PASS>(q, 2, k);
            if constexpr (R > 8) w[8] = tw.template get<L, PASS>(q, 3, k);
            if constexpr (R > 2) w[3] = cmul(w[1], w[2]);
            if constexpr (R > 4) {
                w[5] = cmul(w[1], w[4]);
                w[6] = cmul(w[2], w[4]);
                w[7] = cmul(w[3], w[4]);
            }
            if constexpr (R > 8) {
#pragma unroll
                for (int r = 9; r < 16; ++r) w[r] = cmul(w[r - 8], w[8]);
            }
#pragma unroll
            for (int r = 1; r < R; ++r) {
                v1[q + r * Q] = DIR < 0 ? cmul(v1[q + r * Q], w[r]) : cmulc(v1[q + r * Q], w[r]);
                v2[q + r * Q] = DIR < 0 ? cmul(v2[q + r * Q], w[r]) : cmulc(v2[q + r * Q], w[r]);
            }
        }
        if constexpr (R == 16) {
            dft16<DIR>(v1);
            dft16<DIR>(v2);
        } else if constexpr (R == 8) {
            dft8<DIR>(v1[q], v1[q + Q], v1[q + 2 * Q], v1[q + 3 * Q], v1[q + 4 * Q], v1[q + 5 * Q],
                      v1[q + 6 * Q], v1[q + 7 * Q]);
            dft8<DIR>(v2[q], v2[q + Q], v2[q + 2 * Q], v2[q + 3 * Q], v2[q + 4 * Q], v2[q + 5 * Q],
                      v2[q + 6 * Q], v2[q + 7 * Q]);
        } else if constexpr (R == 4) {
            dft4<DIR>(v1[q], v1[q + Q], v1[q + 2 * Q], v1[q + 3 * Q]);
            dft4<DIR>(v2[q], v2[q + Q], v2[q + 2 * Q], v2[q + 3 * Q]);
        } else if constexpr (R == 2) {
            dft2<DIR>(v1[q], v1[q + Q]);
            dft2<DIR>(v2[q], v2[q + Q]);
        }
        if constexpr (!last) {
            const int base = (b - k) * R + k;
            if constexpr (Ns % 16 == 0) {
                cx<T>* o1 = lds1 + pad(base);
                cx<T>* o2 = lds2 + pad(base);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    o1[r * (Ns + Ns / 16)] = v1[q + r * Q];
                    o2[r * (Ns + Ns / 16)] = v2[q + r * Q];
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    lds1[pad(base + r * Ns)] = v1[q + r * Q];
                    lds2[pad(base + r * Ns)] = v2[q + r * Q];
                }
            }
        }
    }
}

template <int DIR, int L, int PASS, typename T, typename TW>
__device__ __forceinline__ void cpasses2(cx<T> (&v1)[P], cx<T> (&v2)[P], cx<T>* lds1,
                                         cx<T>* lds2, int j, const TW& tw) {
    if constexpr (PASS < CLine<L>::passes) {
        if constexpr (PASS > 0) {
            lds_read_c<L>(v1, lds1, pass_lane<T, L, PASS>(j));
            lds_read_c<L>(v2, lds2, pass_lane<T, L, PASS>(j));
        }
        __syncthreads();
        cpass2<DIR, L, PASS>(v1, v2, pass_lane<T, L, PASS>(j), tw, lds1, lds2);
        if constexpr (PASS + 1 < CLine<L>::passes) __syncthreads();
        cpasses2<DIR, L, PASS + 1>(v1, v2, lds1, lds2, j, tw);
    }
}

// Two transforms in lock-step (same contract as line_fft_c for each region).
template <int DIR, int L, typename T>
__device__ __forceinline__ void line_fft2_c(cx<T> (&v1)[P], cx<T> (&v2)[P], cx<T>* lds1,
                                            cx<T>* lds2, int j, const cx<T>* __restrict__ tw) {
    cpasses2<DIR, L, 0>(v1, v2, lds1, lds2, j, TwTable<T>{tw});
}
template <int DIR, int L, typename T>
__device__ __forceinline__ void line_fft2_c(cx<T> (&v1)[P], cx<T> (&v2)[P], cx<T>* lds1,
                                            cx<T>* lds2, int j, const TwRegs<L, T>& tw) {
    cpasses2<DIR, L, 0>(v1, v2, lds1, lds2, j, tw);
}

// ------------------------------------------------------------------------
// 4096-point lines on 512 threads as 8 x 512 (wave-split transforms).
//
// Thread j = 64 w + l is lane l of wave w.  With n = n2 + 512 n1 and
// k = k1 + 8 k2 the transform splits into 512 eight-point DFTs across the
// waves and eight 512-point DFTs, one per wave:
//   X[k1 + 8 k2] = sum_n2 W512^(n2 k2) W4096^(n2 k1) sum_n1 x[n2 + 512 n1] W8^(n1 k1).
// Only the step between the two crosses waves.  The 512-point transform of
// wave w runs on row w of the line region (512 elements, padded one in
// sixteen) in three radix-8 Stockham passes.  Its exchanges
// stay inside the wave, whose LDS operations complete in order, so they need
// a compiler fence and no workgroup barrier: two barriers per transform
// remain instead of seven, and the cross-wave step reads and writes rows and
// columns at unit stride.
//
// ws_a: standard pattern in (thread j holds x[j + 512 s]), wave-strided
//       pattern out (lane l of wave w holds X[w + 8 l + 512 s]).
// ws_b: wave-strided pattern in, standard pattern out: the 512-point DFTs of
//       y[w + 8 m] first, then the cross-wave step with twiddle W4096^(j w).
// A pointwise product between ws_a and ws_b transforms does not see the
// pattern, which is how the row stage uses them.

// Row w of a line region starts at 576 w: 512 elements with the
// one-in-sixteen padding take 544, and a stride of 576 elements (4608 bytes
// at FP32, a multiple of 512) lets the compiler pair a thread's accesses to
// two rows into one ds_read2st64 / ds_write2st64.  Single 64-bit reads of
// consecutive padded elements lose a cycle to a bank conflict wherever a
// 32-lane phase crosses a padding gap; the paired forms do not.
constexpr int WS_ROW = 576;
constexpr int WS_REGION = 8 * WS_ROW;  // elements of a wave-split line region

template <typename T>
struct TwWs {
    cx<T> wj;   // W4096^j: the cross-wave step
    cx<T> w8;   // W64^(l mod 8): local pass of stride 8
    cx<T> w64;  // W512^l: local pass of stride 64
};

// tw is the 4096-point table W4096^m
template <typename T>
__device__ __forceinline__ void load_tw_ws(TwWs<T>& t, int j, const cx<T>* __restrict__ tw) {
    const int l = j & 63;
    t.wj = tw[j];
    t.w8 = tw[(l & 7) << 6];
    t.w64 = tw[l << 3];
}

// W^1 .. W^7 from W by squaring and products (as in cpass)
template <typename T>
__device__ __forceinline__ void tw_powers8(cx<T> (&w)[8], cx<T> b) {
    w[1] = b;
    w[2] = cmul(b, b);
    w[4] = cmul(w[2], w[2]);
    w[3] = cmul(w[1], w[2]);
    w[5] = cmul(w[1], w[4]);
    w[6] = cmul(w[2], w[4]);
    w[7] = cmul(w[3], w[4]);
}

template <int DIR, typename T>
__device__ __forceinline__ void tw_apply8(cx<T> (&v)[P], const cx<T> (&w)[8]) {
#pragma unroll
    for (int r = 1; r < 8; ++r) v[r] = DIR < 0 ? cmul(v[r], w[r]) : cmulc(v[r], w[r]);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft8v(cx<T> (&v)[P]) {
    dft8<DIR>(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
}

// A wave's LDS operations complete in the order issued: between its writes
// and its reads of its own row only the compiler has to be kept from
// reordering them.
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Butterfly of lane l in the stride-8 local pass: bits 3 and 4 swapped where
// they differ.  Butterfly b = 8 a + k writes 64 a + k + 8 r, padded 68 a + k +
// 8 r + r / 2, so a and a + 1 in one 16-lane write phase sit 8 banks apart;
// a and a + 2 sit 16 apart and the phase covers the 32 banks once.
__device__ __forceinline__ int ws_lane8(int l) { return l ^ ((((l >> 3) ^ (l >> 4)) & 1) * 0x18); }

// Positions inside a wave-split row.  Each exchange (a write phase and the
// read phase after it) may lay the row out its own way.
// - Padded rows (SWZ false, 8-byte elements): every exchange keeps the
//   one-in-sixteen padding.  It spreads the stride-8 writes of the local
//   passes, and the compiler pairs the unit-stride reads into ds_read2 forms,
//   which padding gaps do not slow.
// - 16-byte rows (SWZ, single 128-bit accesses): a write serves eight lanes
//   per cycle on element banks mod 8, and a unit-stride read of padded
//   elements is slowed by the gaps (probe: twice the LDS cycles).  The
//   exchanges at unit stride on both sides (cross-wave step, pass 1 to 2 with
//   writes 64 a + k + 8 r, k = lane mod 8) go unpadded; pass 0 to 1, whose
//   writes 8 l + r are stride 8, puts element c at c ^ ((c >> 3) & 7), which
//   spreads them mod 8 and only permutes aligned groups of eight for the
//   reads.
// ws_pos(c + 64 s) = ws_pos(c) + ws_off(s) in every layout, so a lane's
// reads take compile-time offsets.
template <bool SWZ>
__device__ __forceinline__ int ws_pos(int c) { return SWZ ? c : padc(c); }
template <bool SWZ>
constexpr int ws_off(int s) { return SWZ ? 64 * s : 68 * s; }
__device__ __forceinline__ int ws_swz(int c) { return c ^ ((c >> 3) & 7); }

// NF rows (1 or 2) of the wave's 512-point transform, lane l holding elements
// l + 64 s of each row in and out.  row[f] is this wave's row of field f; its
// contents may be overwritten.
template <int DIR, int NF, bool SWZ = false, typename T>
__device__ __forceinline__ void ws_local(cx<T> (*const (&v)[NF])[P], cx<T>* const (&row)[NF],
                                         int l, const TwWs<T>& tw) {
    static_assert(!SWZ || sizeof(T) == 8, "16-byte row layouts");
    // pass 0 (stride 1, no twiddles): butterfly l writes 8 l + r (padded: one
    // 16-element padding group; 16-byte rows: r ^ (l mod 8) inside the
    // lane's block of eight)
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        dft8v<DIR>(*v[f]);
        cx<T>* o = row[f] + (SWZ ? 8 * l : padc(8 * l));
#pragma unroll
        for (int r = 0; r < 8; ++r) o[SWZ ? (r ^ (l & 7)) : r] = (*v[f])[r];
    }
    wave_lds_fence();
    // pass 1 (stride 8); padded rows swap bits 3 and 4 of the lane (ws_lane8)
    {
        const int b = SWZ ? l : ws_lane8(l);
        const int k = b & 7;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const cx<T>* i = row[f] + (SWZ ? ws_swz(b) : padc(b));
#pragma unroll
            for (int s = 0; s < 8; ++s) (*v[f])[s] = i[ws_off<SWZ>(s)];
        }
        wave_lds_fence();
        cx<T> w[8];
        tw_powers8(w, tw.w8);
        // outputs at 64 a + k + 8 r (padded: 68 a + k + 8 r + r / 2)
        const int ob = SWZ ? (b - k) * 8 + k : padc((b - k) * 8) + k;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            tw_apply8<DIR>(*v[f], w);
            dft8v<DIR>(*v[f]);
#pragma unroll
            for (int r = 0; r < 8; ++r) row[f][SWZ ? ob + 8 * r : ob + 8 * r + (r >> 1)] = (*v[f])[r];
        }
    }
    wave_lds_fence();
    // pass 2 (stride 64, last): butterfly l; the result stays in registers
    {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const cx<T>* i = row[f] + ws_pos<SWZ>(l);
#pragma unroll
            for (int s = 0; s < 8; ++s) (*v[f])[s] = i[ws_off<SWZ>(s)];
        }
        wave_lds_fence();
        cx<T> w[8];
        tw_powers8(w, tw.w64);
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            tw_apply8<DIR>(*v[f], w);
            dft8v<DIR>(*v[f]);
        }
    }
}

// NF 4096-point transforms (regions lds[f]), standard pattern in,
// wave-strided pattern out.  Every thread must be free to write the regions
// on entry (the caller's barrier).  On return each wave still works in its
// own rows: other waves may write the regions only after a workgroup barrier.
template <int DIR, int NF, bool SWZ = false, typename T>
__device__ __forceinline__ void ws_a(cx<T> (*const (&v)[NF])[P], cx<T>* const (&lds)[NF], int j,
                                     const TwWs<T>& tw) {
    cx<T> w[8];
    tw_powers8(w, tw.wj);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        dft8v<DIR>(*v[f]);
        tw_apply8<DIR>(*v[f], w);
        cx<T>* o = lds[f] + ws_pos<SWZ>(j);  // u[k1] of column j to row k1
#pragma unroll
        for (int r = 0; r < 8; ++r) o[r * WS_ROW] = (*v[f])[r];
    }
    __syncthreads();
    const int wv = j >> 6, l = j & 63;
    cx<T>* row[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        row[f] = lds[f] + wv * WS_ROW;
        const cx<T>* i = row[f] + ws_pos<SWZ>(l);
#pragma unroll
        for (int s = 0; s < 8; ++s) (*v[f])[s] = i[ws_off<SWZ>(s)];
    }
    wave_lds_fence();
    ws_local<DIR, NF, SWZ>(v, row, l, tw);
}

// NF 4096-point transforms, wave-strided pattern in, standard pattern out.
// On entry only the wave's own rows of the regions have to be free (as they
// are right after ws_a on the same regions); on return other threads may
// still be reading the regions.
template <int DIR, int NF, bool SWZ = false, typename T>
__device__ __forceinline__ void ws_b(cx<T> (*const (&v)[NF])[P], cx<T>* const (&lds)[NF], int j,
                                     const TwWs<T>& tw) {
    const int wv = j >> 6, l = j & 63;
    cx<T>* row[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) row[f] = lds[f] + wv * WS_ROW;
    ws_local<DIR, NF, SWZ>(v, row, l, tw);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        cx<T>* o = row[f] + ws_pos<SWZ>(l);  // Y_w[l + 64 s] to row w
#pragma unroll
        for (int s = 0; s < 8; ++s) o[ws_off<SWZ>(s)] = (*v[f])[s];
    }
    __syncthreads();
    cx<T> w[8];
    tw_powers8(w, tw.wj);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const cx<T>* i = lds[f] + ws_pos<SWZ>(j);  // column j of every row
#pragma unroll
        for (int r = 0; r < 8; ++r) (*v[f])[r] = i[r * WS_ROW];
        tw_apply8<DIR>(*v[f], w);
        dft8v<DIR>(*v[f]);
    }
}

// Wave-strided pattern to standard pattern through region tr.  Wave w writes
// element c = l + 64 s of its transform at 544 w + (c + 4 w) mod 512 (its own
// row: the skew by 4 w puts the eight rows of one column in different banks,
// and each 16-lane write phase stays on 16 consecutive elements mod 512);
// after a barrier thread j reads X[j + 512 s] from row j mod 8, column
// j / 8 + 64 s, 32 lanes on 32 distinct banks pairs.  On entry the wave's own
// row of tr must be free; on return other threads may still read tr.
template <int NF, typename T>
__device__ __forceinline__ void ws_transpose(cx<T> (*const (&v)[NF])[P], cx<T>* const (&tr)[NF],
                                             int j) {
    const int wv = j >> 6, l = j & 63;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        cx<T>* o = tr[f] + wv * WS_ROW;
#pragma unroll
        for (int s = 0; s < 8; ++s) o[(l + 4 * wv + 64 * s) & 511] = (*v[f])[s];
    }
    __syncthreads();
    const int rw = j & 7, c0 = j >> 3;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const cx<T>* i = tr[f] + rw * WS_ROW;
#pragma unroll
        for (int s = 0; s < 8; ++s) (*v[f])[s] = i[(c0 + 4 * rw + 64 * s) & 511];
    }
}

// NF 4096-point transforms, standard pattern in and out: ws_a on regions
// rows, then the transposing exchange through regions tr (which may be the
// same regions: each wave writes only its own row there).  Entry condition
// as for ws_a on rows and, when tr differs, every thread free to write tr;
// on return other threads may still read tr and work in their rows.
template <int DIR, int NF, bool SWZ = false, typename T>
__device__ __forceinline__ void ws_c(cx<T> (*const (&v)[NF])[P], cx<T>* const (&rows)[NF],
                                     cx<T>* const (&tr)[NF], int j, const TwWs<T>& tw) {
    ws_a<DIR, NF, SWZ>(v, rows, j, tw);
    ws_transpose<NF>(v, tr, j);
}

// XCD-aware block order: MI355X dispatches workgroup b to XCD b % 8; this
// gives each XCD a contiguous range of logical tiles so that neighbouring
// tiles (which share cache lines of the transposed intermediates) meet in
// the same L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    if (nb & 7) return b;
    return (b & 7) * (nb >> 3) + (b >> 3);
}

// Column groups on one XCD: blocks b, b + 8, ..., b + 8 (G - 1) (the same
// XCD, one after the other in its dispatch order) get columns G q .. G q +
// G - 1, group q = 8 (b / 8 G) + b % 8, so that G columns that share cache
// lines are read or written through one L2 at about the same time.  A
// bijection on [0, nb): the blocks past the last multiple of 8 G keep their
// index.
template <int G>
__device__ __forceinline__ int xcd_group_remap(int b, int nb) {
    static_assert(G == 1 || G == 2 || G == 4, "groups of 1, 2 or 4 columns");
    if (G == 1 || b >= (nb & ~(8 * G - 1))) return b;
    const int x = b & 7, idx = b >> 3;
    return ((idx / G) * 8 + x) * G + idx % G;
}

}  // namespace nsd
