// nsd_closure.hip -- the learned subgrid closure: a 3x3 CNN of four GELU
// hidden layers of 32 channels and a linear output layer (1 -> 32 -> 32 ->
// 32 -> 32 -> 1 channels, periodic padding, 28,353 parameters) evaluated and
// trained with bf16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation).
//
// Layout: activations are NHWC bf16 ([batch][y][x][channel]); weights
// [cout][tap][cin] with tap = 3 ky + kx for the offset (ky - 1, kx - 1);
// FP32 master weights, bias, gradients and Adam moments.
//
// Every convolution is a GEMM over pixels: out[p][n] = sum_k A[p][k] B[k][n].
//   forward:        A[p][(tap, ci)] = X[p + off(tap)][ci],  B = W[n][tap][ci]
//   input gradient: A[p][(tap, co)] = dZ[p - off(tap)][co], B = W[co][tap][n]
//   weight grad.:   dW[co][(tap, ci)] = sum_p dZ[p][co] X[p + off(tap)][ci]
// A wave computes 16 pixels x 16 channels per MFMA (v_mfma_f32_16x16x32_bf16,
// depth 32: one 3x3 tap of 32 channels); the B operand (at most 288 x 32
// bf16) sits in LDS, transposed so each lane's 8 k-values are one 16-byte
// read, and each lane's A values are one 16-byte load of 8 channels.  A workgroup stages B once and walks TILES tiles of 64
// pixels, so the staging is paid once per 64 x TILES pixels.  Grids that are
// multiples of 64 use the LDS-tiled kernels instead: k_conv_tile2 (forward
// layers with 32 input channels: persistent, the next patch streamed into LDS
// while the current one is computed), k_conv_tile (the input layer and the
// input gradients) and k_conv_dw_tile / k_conv_dw_tile_in (weight
// gradients); k_conv and k_conv_dw serve the smaller grids.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ns_closure.h"
#include "ns_device.h"
#include "nsd_closure_link.hpp"

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

constexpr int kHidden = 32;  // channels of the three hidden layers
constexpr int kLayers = 5;                  // four GELU hidden layers, linear output
constexpr int kHiddenLayers = kLayers - 1;

__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((unsigned)h << 16); }
// round to nearest even: gfx950's v_cvt_pk_bf16_f32 (NaN stays NaN)
__device__ __forceinline__ bf16 f2bf(float f) { return __builtin_bit_cast(bf16, (__bf16)f); }

// GELU, tanh form: g(z) = z/2 (1 + tanh(u)), u = c (z + a z^3), c = sqrt(2/pi),
// a = 0.044715.  Since (1 + tanh u)/2 = sigma(2u) = 1/(1 + e^(-2u)), it is
// evaluated as z sigma(2u) with the hardware exponential and reciprocal
// (a few ulp of FP32; the results are stored as bf16).  g'(z) = s + z s (1 - s)
// 2 c (1 + 3 a z^2) with s = sigma(2u).
__device__ __forceinline__ float sig2u(float z) {
    const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
    return __builtin_amdgcn_rcpf(1.f + __expf(-2.f * u));
}
__device__ __forceinline__ float gelu(float z) { return z * sig2u(z); }
__device__ __forceinline__ float gelu_d(float z) {
    const float sg = sig2u(z);
    return sg + z * sg * (1.f - sg) * (2.f * 0.7978845608028654f) * (1.f + 3.f * 0.044715f * z * z);
}

// pixel p (row-major in an n x n grid, n a power of two) shifted by tap's
// offset times s, periodic
__device__ __forceinline__ int shift(int p, int tap, int s, int lg, int n) {
    const int y = p >> lg, x = p & (n - 1);
    const int dy = (tap / 3 - 1) * s, dx = (tap % 3 - 1) * s;
    return (((y + dy) & (n - 1)) << lg) | ((x + dx) & (n - 1));
}

enum ConvMode { CONV_FWD_GELU = 0, CONV_FWD_LAST = 1, CONV_BWD_DATA = 2 };

struct ConvArgs {
    int n, lg, npix;
    const bf16* in;     // A source: [batch][npix][KC]
    const bf16* w;      // layer weights, bf16 [cout][9][cin]
    const float* bias;  // [NN] (forward)
    bf16* z;            // forward: pre-activations [batch][npix][NN] (null: not kept)
    bf16* act;          // forward: GELU(z)
    float* y;           // last layer: [batch][npix]
    float yscale;       // last layer: output multiplied by this
    const bf16* zprev;  // input gradient: pre-activations of the layer below [batch][npix][NN]
    bf16* dzprev;       // input gradient: dL/dz of the layer below
};

// KC = channels of the A source, NN = output channels of this GEMM, TILES =
// 64-pixel tiles per workgroup (npix must be a multiple of 64 TILES)
template <int KC, int NN, int MODE, int TILES>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
    constexpr int KP = (9 * KC + 31) / 32 * 32;  // GEMM depth, padded to the MFMA depth
    constexpr int NP = (NN + 15) / 16 * 16;      // GEMM width, padded
    __shared__ __align__(16) bf16 Bt[NP * KP];   // B transposed: Bt[n][k]
    const int tid = threadIdx.x;
    for (int e = tid; e < NP * KP; e += 256) {
        const int nn = e / KP, k = e % KP;
        bf16 v = 0;
        if (nn < NN && k < 9 * KC) {
            const int tap = k / KC, kc = k % KC;
            v = MODE == CONV_BWD_DATA ? a.w[(kc * 9 + tap) * NN + nn] : a.w[(nn * 9 + tap) * KC + kc];
        }
        Bt[e] = v;
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, kq = lane >> 4;
    const size_t boff = (size_t)blockIdx.y * a.npix;
    const int s = MODE == CONV_BWD_DATA ? -1 : 1;
#pragma unroll 1
    for (int t = 0; t < TILES; ++t) {
        const int p0 = ((blockIdx.x * TILES + t) * 4 + wave) * 16;
        const int p = p0 + m;
        f4 acc[NP / 16];
#pragma unroll
        for (int nb = 0; nb < NP / 16; ++nb) acc[nb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < KP / 32; ++c) {
            const int k0 = c * 32 + 8 * kq;  // A lane l: row l % 16, k = 8 (l / 16) + i
            s8 av;
            if constexpr (KC % 32 == 0) {
                const int tap = k0 / KC, kc = k0 % KC;
                const bf16* src = a.in + (boff + shift(p, tap, s, a.lg, a.n)) * KC + kc;
                av = *reinterpret_cast<const s8*>(src);
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int k = k0 + i;
                    short v = 0;
                    if (k < 9 * KC) {
                        const int tap = k / KC, kc = k % KC;
                        v = (short)a.in[(boff + shift(p, tap, s, a.lg, a.n)) * KC + kc];
                    }
                    av[i] = v;
                }
            }
#pragma unroll
            for (int nb = 0; nb < NP / 16; ++nb) {
                const s8 bv = *reinterpret_cast<const s8*>(&Bt[(nb * 16 + m) * KP + k0]);
                if constexpr (MODE == CONV_FWD_LAST)  // D[channel][pixel]
                    acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf8, bv), __builtin_bit_cast(bf8, av), acc[nb], 0, 0, 0);
                else  // D[pixel][channel]
                    acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf8, av), __builtin_bit_cast(bf8, bv), acc[nb], 0, 0, 0);
            }
        }
        if constexpr (MODE == CONV_FWD_LAST) {
            // D[row = channel 4 kq + i][col = pixel p0 + (lane & 15)]: channel 0
            // sits in lanes 0..15, one pixel each
            if (kq == 0) a.y[boff + p] = (acc[0][0] + a.bias[0]) * a.yscale;
        } else {
            // D[row = pixel p0 + 4 kq + i][col = channel nb 16 + (lane & 15)]
#pragma unroll
            for (int nb = 0; nb < NP / 16; ++nb) {
                const int nn = nb * 16 + m;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const size_t q = boff + p0 + 4 * kq + i;
                    const float v = acc[nb][i];
                    if constexpr (MODE == CONV_FWD_GELU) {
                        if (nn < NN) {
                            const float zz = v + a.bias[nn];
                            if (a.z) a.z[q * NN + nn] = f2bf(zz);
                            a.act[q * NN + nn] = f2bf(gelu(zz));
                        }
                    } else {
                        if (nn < NN)
                            a.dzprev[q * NN + nn] = f2bf(v * gelu_d(bf2f(a.zprev[q * NN + nn])));
                    }
                }
            }
        }
    }  // tiles
}

// grid and launch of k_conv: 8 tiles per workgroup when the grid allows it
template <int KC, int NN, int MODE>
void launch_conv(const ConvArgs& a, int batch, hipStream_t st) {
    if (a.npix % (64 * 8) == 0)
        hipLaunchKernelGGL((k_conv<KC, NN, MODE, 8>), dim3(a.npix / (64 * 8), batch), dim3(256), 0,
                           st, a);
    else
        hipLaunchKernelGGL((k_conv<KC, NN, MODE, 1>), dim3(a.npix / 64, batch), dim3(256), 0, st,
                           a);
}

// Layers from an LDS-staged input patch (forward, and the input gradient,
// whose A operand is dZ shifted the other way).  A workgroup of 512
// threads computes a tile of 8 rows x 64 pixels of one member: it stages the
// tile's input with its one-pixel periodic halo (10 x 66 pixels, all KC
// channels) in LDS with coalesced 16-byte loads issued together, holds its
// part of the transposed weights in registers, and then each wave (one tile
// row, four blocks of 16 pixels) reads every tap's A operand from LDS.
// k_conv reloads each pixel's neighbourhood from global memory for each tap,
// a chain of dependent loads per 16-pixel tile.  Same GEMM, lane layouts and
// epilogue as k_conv; n must be a multiple of 64.
constexpr int kTileRows = 8, kTileCols = 64;

template <int KC, int NN, int MODE>
__global__ __launch_bounds__(512) void k_conv_tile(ConvArgs a) {
    static_assert(kTileRows * 64 == 512, "one wave per tile row");
    constexpr int S = MODE == CONV_BWD_DATA ? -1 : 1;  // tap offsets applied to the input
    constexpr int KP = (9 * KC + 31) / 32 * 32;
    constexpr int NP = (NN + 15) / 16 * 16;
    constexpr int PR = kTileRows + 2, PC = kTileCols + 2;
    __shared__ __align__(16) bf16 Bt[NP * KP];
    __shared__ __align__(16) bf16 Xs[PR * PC * KC];
    const int tid = threadIdx.x;
    const int tiles_x = a.n / kTileCols;
    const int r0 = (blockIdx.x / tiles_x) * kTileRows, c0 = (blockIdx.x % tiles_x) * kTileCols;
    const size_t boff = (size_t)blockIdx.y * a.npix;
    const int nm = a.n - 1;
    // input patch: Xs[(r PC + c) KC + ch] = in[(r0 + r - 1, c0 + c - 1) mod n][ch]
    if constexpr (KC % 8 == 0) {
        constexpr int CH = KC / 8;  // 16-byte pieces per pixel
#pragma unroll 4
        for (int e = tid; e < PR * PC * CH; e += 512) {
            const int pix = e / CH, q = e - pix * CH;
            const int r = pix / PC, c = pix - r * PC;
            const int gy = (r0 + r - 1) & nm, gx = (c0 + c - 1) & nm;
            const s8 v = *reinterpret_cast<const s8*>(
                a.in + (boff + ((size_t)gy << a.lg) + gx) * KC + 8 * q);
            *reinterpret_cast<s8*>(&Xs[pix * KC + 8 * q]) = v;
        }
    } else {
        for (int e = tid; e < PR * PC * KC; e += 512) {
            const int pix = e / KC, q = e - pix * KC;
            const int r = pix / PC, c = pix - r * PC;
            const int gy = (r0 + r - 1) & nm, gx = (c0 + c - 1) & nm;
            Xs[e] = a.in[(boff + ((size_t)gy << a.lg) + gx) * KC + q];
        }
    }
    for (int e = tid; e < NP * KP; e += 512) {
        const int nn = e / KP, k = e % KP;
        bf16 v = 0;
        if (nn < NN && k < 9 * KC) {
            const int tap = k / KC, kc = k % KC;
            v = MODE == CONV_BWD_DATA ? a.w[(kc * 9 + tap) * NN + nn] : a.w[(nn * 9 + tap) * KC + kc];
        }
        Bt[e] = v;
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, kq = lane >> 4;
    // this lane's B operands for the whole tile
    s8 bv[KP / 32][NP / 16];
#pragma unroll
    for (int c = 0; c < KP / 32; ++c)
#pragma unroll
        for (int nb = 0; nb < NP / 16; ++nb)
            bv[c][nb] = *reinterpret_cast<const s8*>(&Bt[(nb * 16 + m) * KP + c * 32 + 8 * kq]);
    const int lr = wave;  // tile row of this wave
#pragma unroll 1
    for (int blk = 0; blk < kTileCols / 16; ++blk) {
        const int lc = blk * 16;
        f4 acc[NP / 16];
#pragma unroll
        for (int nb = 0; nb < NP / 16; ++nb) acc[nb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < KP / 32; ++c) {
            const int k0 = c * 32 + 8 * kq;  // A lane l: row l % 16, k = 8 (l / 16) + i
            s8 av;
            if constexpr (KC % 32 == 0) {
                const int tap = k0 / KC, kc = k0 % KC;
                const int pr = lr + 1 + S * (tap / 3 - 1), pc = lc + m + 1 + S * (tap % 3 - 1);
                av = *reinterpret_cast<const s8*>(&Xs[(pr * PC + pc) * KC + kc]);
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int k = k0 + i;
                    short v = 0;
                    if (k < 9 * KC) {
                        const int tap = k / KC, kc = k % KC;
                        const int pr = lr + 1 + S * (tap / 3 - 1),
                                  pc = lc + m + 1 + S * (tap % 3 - 1);
                        v = (short)Xs[(pr * PC + pc) * KC + kc];
                    }
                    av[i] = v;
                }
            }
#pragma unroll
            for (int nb = 0; nb < NP / 16; ++nb) {
                if constexpr (MODE == CONV_FWD_LAST)  // D[channel][pixel]
                    acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf8, bv[c][nb]), __builtin_bit_cast(bf8, av), acc[nb], 0,
                        0, 0);
                else  // D[pixel][channel]
                    acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf8, av), __builtin_bit_cast(bf8, bv[c][nb]), acc[nb], 0,
                        0, 0);
            }
        }
        const size_t q0 = boff + ((size_t)(r0 + lr) << a.lg) + c0 + lc;  // pixel of block column 0
        if constexpr (MODE == CONV_FWD_LAST) {
            if (kq == 0) a.y[q0 + m] = (acc[0][0] + a.bias[0]) * a.yscale;
        } else {
#pragma unroll
            for (int nb = 0; nb < NP / 16; ++nb) {
                const int nn = nb * 16 + m;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const size_t q = q0 + 4 * kq + i;
                    if (nn < NN) {
                        if constexpr (MODE == CONV_BWD_DATA) {
                            a.dzprev[q * NN + nn] =
                                f2bf(acc[nb][i] * gelu_d(bf2f(a.zprev[q * NN + nn])));
                        } else {
                            const float zz = acc[nb][i] + a.bias[nn];
                            if (a.z) a.z[q * NN + nn] = f2bf(zz);
                            a.act[q * NN + nn] = f2bf(gelu(zz));
                        }
                    }
                }
            }
        }
    }
}

// 32-channel layers, persistent: two workgroups per CU walk the 8 x 32-pixel
// tiles of all members.  A workgroup stages the layer's weights once (held
// in registers, as in k_conv_tile) and keeps two LDS patches of 10 x 34
// pixels: while it computes one tile from one patch, the next tile's patch
// streams into the other with global_load_lds_dwordx4 (each lane's 16 bytes
// land at the wave's LDS base + 16 x lane, which is the patch's own order of
// 16-byte pieces), so no register holds it and one barrier per tile serves
// both the reads of the old patch and the arrival of the new one.
constexpr int kT2Cols = 32;

__device__ __forceinline__ void g2lds16(const void* g, bf16* lds_wave_base) {
    __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void*)lds_wave_base, 16,
                                     0, 0);
}

template <int NN, int MODE>
__global__ __launch_bounds__(512) void k_conv_tile2(ConvArgs a, int ntiles) {
    constexpr int KC = 32;
    constexpr int S = MODE == CONV_BWD_DATA ? -1 : 1;
    constexpr int KP = 9 * KC;
    constexpr int NP = (NN + 15) / 16 * 16;
    constexpr int PR = kTileRows + 2, PC = kT2Cols + 2;
    constexpr int NPC = PR * PC * (KC / 8);  // 16-byte pieces per patch
    constexpr int PER = (NPC + 511) / 512;
    __shared__ __align__(16) bf16 Bt[NP * KP];
    // two distinct arrays (not one of two rows): the compiler's wait for the
    // patch loads in flight then covers reads of the patch being loaded only
    __shared__ __align__(16) bf16 X0[PR * PC * KC];
    __shared__ __align__(16) bf16 X1[PR * PC * KC];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_x = a.n / kT2Cols, per_mem = tiles_x * (a.n / kTileRows);
    const int nm = a.n - 1;
    for (int e = tid; e < NP * KP; e += 512) {
        const int nn = e / KP, k = e % KP;
        bf16 v = 0;
        if (nn < NN) {
            const int tap = k / KC, kc = k % KC;
            v = MODE == CONV_BWD_DATA ? a.w[(kc * 9 + tap) * NN + nn] : a.w[(nn * 9 + tap) * KC + kc];
        }
        Bt[e] = v;
    }
    // patch of tile t into buffer b: piece e = pixel (r PC + c), channel group e % 4
    auto issue = [&](int t, bf16* Xd) {
        const int mem = t / per_mem, rem = t - mem * per_mem;
        const int r0 = (rem / tiles_x) * kTileRows, c0 = (rem % tiles_x) * kT2Cols;
        const size_t boff = (size_t)mem * a.npix;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e0 = 512 * i + 64 * wave;  // this wave's first piece
            if (e0 < NPC) {
                const int e = e0 + lane;
                if (e < NPC) {
                    const int pix = e >> 2, q = e & 3;
                    const int r = pix / PC, c = pix - r * PC;
                    const int gy = (r0 + r - 1) & nm, gx = (c0 + c - 1) & nm;
                    g2lds16(a.in + (boff + ((size_t)gy << a.lg) + gx) * KC + 8 * q, Xd + 8 * e0);
                }
            }
        }
    };
    int t = blockIdx.x;
    if (t < ntiles) issue(t, X0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int m = lane & 15, kq = lane >> 4;
    s8 bv[KP / 32][NP / 16];
#pragma unroll
    for (int c = 0; c < KP / 32; ++c)
#pragma unroll
        for (int nb = 0; nb < NP / 16; ++nb)
            bv[c][nb] = *reinterpret_cast<const s8*>(&Bt[(nb * 16 + m) * KP + c * 32 + 8 * kq]);
    const int lr = wave;
    // biases in registers: a global load in the loop would wait for the
    // patch loads in flight (the vector memory counter retires in order)
    // Both forward modes compute D[channel][pixel] (weights as the A
    // operand): a lane then holds channels 4 kq .. 4 kq + 3 of one pixel, which
    // leave as one 8-byte store.
    f4 bias4[NP / 16];
#pragma unroll
    for (int nb = 0; nb < NP / 16; ++nb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int nn = nb * 16 + 4 * kq + i;
            bias4[nb][i] = MODE == CONV_BWD_DATA || nn >= NN ? 0.f : a.bias[nn];
        }
    // one tile from patch X while the next tile's patch streams into Xn
    auto step = [&](int t, const bf16* X, bf16* Xn) {
        if (t + (int)gridDim.x < ntiles) issue(t + gridDim.x, Xn);
        const int mem = t / per_mem, rem = t - mem * per_mem;
        const int r0 = (rem / tiles_x) * kTileRows, c0 = (rem % tiles_x) * kT2Cols;
        const size_t boff = (size_t)mem * a.npix;
#pragma unroll 1
        for (int blk = 0; blk < kT2Cols / 16; ++blk) {
            const int lc = blk * 16;
            f4 acc[NP / 16];
#pragma unroll
            for (int nb = 0; nb < NP / 16; ++nb) acc[nb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < KP / 32; ++c) {  // one tap of 32 channels per MFMA
                const int tap = c, kc = 8 * kq;
                const int pr = lr + 1 + S * (tap / 3 - 1), pc = lc + m + 1 + S * (tap % 3 - 1);
                const s8 av = *reinterpret_cast<const s8*>(&X[(pr * PC + pc) * KC + kc]);
#pragma unroll
                for (int nb = 0; nb < NP / 16; ++nb) {
                    if constexpr (MODE != CONV_BWD_DATA)  // D[channel][pixel]
                        acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf8, bv[c][nb]), __builtin_bit_cast(bf8, av),
                            acc[nb], 0, 0, 0);
                    else  // D[pixel][channel]
                        acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf8, av), __builtin_bit_cast(bf8, bv[c][nb]),
                            acc[nb], 0, 0, 0);
                }
            }
            const size_t q0 = boff + ((size_t)(r0 + lr) << a.lg) + c0 + lc;
            if constexpr (MODE == CONV_FWD_LAST) {
                // D[row = channel 4 kq + i][col = pixel m]: channel 0 in lanes 0..15
                if (kq == 0) a.y[q0 + m] = (acc[0][0] + bias4[0][0]) * a.yscale;
            } else if constexpr (MODE == CONV_FWD_GELU) {
                const size_t q = q0 + m;  // this lane's pixel
#pragma unroll
                for (int nb = 0; nb < NP / 16; ++nb) {
                    const int cb = nb * 16 + 4 * kq;  // its first channel
                    if (cb < NN) {
                        float zz[4], g[4];
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            zz[i] = acc[nb][i] + bias4[nb][i];
                            g[i] = gelu(zz[i]);
                        }
                        auto pack = [](float x0, float x1) {
                            return (unsigned)f2bf(x0) | ((unsigned)f2bf(x1) << 16);
                        };
                        if (a.z)
                            *reinterpret_cast<uint2*>(&a.z[q * NN + cb]) =
                                uint2{pack(zz[0], zz[1]), pack(zz[2], zz[3])};
                        *reinterpret_cast<uint2*>(&a.act[q * NN + cb]) =
                            uint2{pack(g[0], g[1]), pack(g[2], g[3])};
                    }
                }
            } else {
#pragma unroll
                for (int nb = 0; nb < NP / 16; ++nb) {
                    const int nn = nb * 16 + m;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const size_t q = q0 + 4 * kq + i;
                        if (nn < NN) {
                            if constexpr (MODE == CONV_BWD_DATA) {
                                a.dzprev[q * NN + nn] =
                                    f2bf(acc[nb][i] * gelu_d(bf2f(a.zprev[q * NN + nn])));
                            }
                        }
                    }
                }
            }
        }
        // the next patch has arrived and every read of this one is done
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    };
    while (t < ntiles) {
        step(t, X0, X1);
        t += gridDim.x;
        if (t >= ntiles) break;
        step(t, X1, X0);
        t += gridDim.x;
    }
}

int conv_tile2_workgroups() {
    static int wg = 0;
    if (!wg) {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
            wg = 2 * ncu;
        if (wg <= 0) wg = 512;
    }
    return wg;
}

// the tiled kernels where the grid is a multiple of their tiles (the input
// gradient stays on k_conv_tile: its epilogue loads the layer below's
// pre-activations, which would wait behind the patch in flight)
template <int KC, int NN, int MODE>
void launch_conv_lds(const ConvArgs& a, int batch, hipStream_t st) {
    if constexpr (KC == 32 && MODE != CONV_BWD_DATA) {
        if (a.n % kT2Cols == 0 && a.n % kTileRows == 0) {
            const int ntiles = (a.n / kT2Cols) * (a.n / kTileRows) * batch;
            hipLaunchKernelGGL((k_conv_tile2<NN, MODE>),
                               dim3(std::min(ntiles, conv_tile2_workgroups())), dim3(512), 0, st,
                               a, ntiles);
            return;
        }
    }
    if (a.n % kTileCols == 0)
        hipLaunchKernelGGL((k_conv_tile<KC, NN, MODE>),
                           dim3((a.n / kTileCols) * (a.n / kTileRows), batch), dim3(512), 0, st, a);
    else
        launch_conv<KC, NN, MODE>(a, batch, st);
}

// dW[co][(tap, ci)] += sum_q dZ[q][co] X[q + off(tap)][ci] over this
// workgroup's pixel range (all batch members flattened); dbias likewise.
struct DwArgs {
    int n, lg, npix, total;  // total = batch * npix
    int chunk;               // pixels per workgroup (multiple of 16)
    const bf16* x;           // layer input [total][CIN]
    const bf16* dz;          // [total][COUT]
    float* dw;               // [COUT][9][CIN]
    float* db;               // [COUT]
};

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv_dw(DwArgs a) {
    constexpr int MT = (9 * CIN + 15) / 16, NT = (COUT + 15) / 16, T = MT * NT;
    constexpr int TW = (T + 3) / 4;  // tiles per wave
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, kq = lane >> 4;
    f4 acc[TW];
#pragma unroll
    for (int t = 0; t < TW; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
    float dbs[NT];  // bias gradient: wave 0, lanes 0..15, channels 16 j + lane
#pragma unroll
    for (int j = 0; j < NT; ++j) dbs[j] = 0.f;
    const int q_begin = blockIdx.x * a.chunk;
    const int q_end = min(q_begin + a.chunk, a.total);
    for (int q0 = q_begin; q0 < q_end; q0 += 16) {
#pragma unroll
        for (int t = 0; t < TW; ++t) {
            const int tile = wave + 4 * t;
            if (tile >= T) continue;
            const int mb = tile / NT, nb = tile % NT;
            const int mrow = mb * 16 + r;          // (tap, ci)
            const int co = nb * 16 + r;
            s4 av, bv;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int q = q0 + 4 * kq + i;
                short xa = 0, dzv = 0;
                if (q < q_end) {
                    const int b = q / a.npix, p = q - b * a.npix;
                    if (mrow < 9 * CIN) {
                        const int tap = mrow / CIN, ci = mrow % CIN;
                        xa = (short)a.x[((size_t)b * a.npix + shift(p, tap, 1, a.lg, a.n)) * CIN + ci];
                    }
                    if (co < COUT) dzv = (short)a.dz[(size_t)q * COUT + co];
                }
                av[i] = xa;
                bv[i] = dzv;
            }
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
        }
        if (wave == 0 && kq == 0) {
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const int co = 16 * j + r;
                if (co < COUT)
                    for (int e = 0; e < 16 && q0 + e < q_end; ++e)
                        dbs[j] += bf2f(a.dz[(size_t)(q0 + e) * COUT + co]);
            }
        }
    }
    if (wave == 0 && kq == 0) {
#pragma unroll
        for (int j = 0; j < NT; ++j)
            if (16 * j + r < COUT) atomicAdd(&a.db[16 * j + r], dbs[j]);
    }
#pragma unroll
    for (int t = 0; t < TW; ++t) {
        const int tile = wave + 4 * t;
        if (tile >= T) continue;
        const int mb = tile / NT, nb = tile % NT;
        const int co = nb * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int mrow = mb * 16 + 4 * kq + i;
            if (mrow < 9 * CIN && co < COUT) atomicAdd(&a.dw[co * 9 * CIN + mrow], acc[t][i]);
        }
    }
}

// Weight gradient of a 32-channel-input layer from LDS-staged tiles (n a
// multiple of 64).  dW^T[(tap, ci)][co] = sum_p X[p + off(tap)][ci] dZ[p][co]
// is a GEMM of depth "pixels".  A workgroup of six waves walks `per_wg` tiles
// of 4 rows x 64 pixels of one member; per tile it stages
//   Xl[ci][yy][xx]: the input with its one-pixel periodic halo, channel-major
//                   (tile row y at yy = y + 1, column x at xx = x + 8), and
//   Dl[co][y][x]:   dZ, channel-major (zero rows up to 16 channels),
// so that 8 consecutive pixels of one channel are 16 contiguous bytes.  Wave
// (ky, h) owns the three taps (ky, kx = 0..2) of channels 16 h .. 16 h + 15:
// per 32 pixels a lane reads one 16-byte chunk of its row (plus one dword on
// each side), forms the three column shifts with byte-align operations, and
// issues 3 x NT MFMAs (v_mfma_f32_16x16x32_bf16).  Partial sums, and the
// bias gradient, go to part[workgroup][...]; k_dw_reduce adds them up in a
// fixed order.
constexpr int kDwRows = 4, kDwCols = 64;
constexpr int kDwXs = 80;                     // LDS row stride of Xl (elements)
constexpr int kDwXc = (kDwRows + 2) * kDwXs;  // LDS channel stride of Xl
constexpr int kDwMaxWg = 1024;                // workgroups of one launch, at most

struct DwTileArgs {
    int n, lg, npix;
    int tiles;         // members x (n / 4) x (n / 64)
    int per_wg;        // tiles per workgroup
    const bf16* x;     // layer input [member][npix][32]
    const bf16* dz;    // [member][npix][COUT]
    float* part;       // [gridDim.x][9 x 32 x COUT + COUT]
};

__device__ __forceinline__ unsigned align2(unsigned hi, unsigned lo) {
    return __builtin_amdgcn_alignbyte(hi, lo, 2);
}

template <int COUT>
__global__ __launch_bounds__(384) void k_conv_dw_tile(DwTileArgs a) {
    constexpr int CIN = 32;
    constexpr int NT = (COUT + 15) / 16;
    constexpr int DC = NT * 16;
    constexpr int NW = 9 * CIN * COUT;
    __shared__ __align__(16) bf16 Xl[CIN * kDwXc];
    __shared__ __align__(16) bf16 Dl[DC * kDwRows * kDwCols];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ky = wave >> 1, h = wave & 1;
    const int r = lane & 15, g = lane >> 4;
    f4 acc[3][NT];
#pragma unroll
    for (int kx = 0; kx < 3; ++kx)
#pragma unroll
        for (int nb = 0; nb < NT; ++nb) acc[kx][nb] = f4{0.f, 0.f, 0.f, 0.f};
    float dbs = 0.f;  // bias gradient: thread 8 co + seg sums a quarter row of Dl[co]
    if constexpr (COUT < DC) {
        for (int e = COUT * kDwRows * kDwCols + tid; e < DC * kDwRows * kDwCols; e += 384)
            Dl[e] = 0;
    }
    const int tx_n = a.n / kDwCols, ty_n = a.n / kDwRows;
    const int t_begin = blockIdx.x * a.per_wg;
    const int t_end = min(t_begin + a.per_wg, a.tiles);
    for (int t = t_begin; t < t_end; ++t) {
        const int mem = t / (tx_n * ty_n), rem = t - mem * (tx_n * ty_n);
        const int y0 = (rem / tx_n) * kDwRows, x0 = (rem % tx_n) * kDwCols;
        const bf16* xb = a.x + (size_t)mem * a.npix * CIN;
        const bf16* db = a.dz + (size_t)mem * a.npix * COUT;
        __syncthreads();  // the previous tile's reads are done
        {   // interior: 6 rows x 16 quads of pixels x 4 groups of 8 channels
            const int cg = tid & 3, qd = (tid >> 2) & 15, yy = tid >> 6;
            const int gy = (y0 + yy - 1) & (a.n - 1);
            const bf16* src = xb + ((size_t)(gy << a.lg) + x0 + 4 * qd) * CIN + 8 * cg;
            const s8 v0 = *reinterpret_cast<const s8*>(src);
            const s8 v1 = *reinterpret_cast<const s8*>(src + CIN);
            const s8 v2 = *reinterpret_cast<const s8*>(src + 2 * CIN);
            const s8 v3 = *reinterpret_cast<const s8*>(src + 3 * CIN);
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const s4 q = {v0[c], v1[c], v2[c], v3[c]};
                *reinterpret_cast<s4*>(&Xl[(8 * cg + c) * kDwXc + yy * kDwXs + 8 + 4 * qd]) = q;
            }
        }
        {   // halo columns x = -1 and x = 64: 6 rows x 2 sides x 32 channels
            const int side = tid & 1, ci = (tid >> 1) & 31, yy = tid >> 6;
            const int gy = (y0 + yy - 1) & (a.n - 1);
            const int gx = (side ? x0 + kDwCols : x0 - 1) & (a.n - 1);
            Xl[ci * kDwXc + yy * kDwXs + (side ? 8 + kDwCols : 7)] =
                xb[((size_t)(gy << a.lg) + gx) * CIN + ci];
        }
        if constexpr (COUT == 32) {  // 4 rows x 16 quads x 4 groups of 8 channels
            if (tid < 256) {
                const int cg = tid & 3, qd = (tid >> 2) & 15, yr = tid >> 6;
                const bf16* src =
                    db + ((size_t)((y0 + yr) << a.lg) + x0 + 4 * qd) * COUT + 8 * cg;
                const s8 v0 = *reinterpret_cast<const s8*>(src);
                const s8 v1 = *reinterpret_cast<const s8*>(src + COUT);
                const s8 v2 = *reinterpret_cast<const s8*>(src + 2 * COUT);
                const s8 v3 = *reinterpret_cast<const s8*>(src + 3 * COUT);
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const s4 q = {v0[c], v1[c], v2[c], v3[c]};
                    *reinterpret_cast<s4*>(
                        &Dl[(8 * cg + c) * (kDwRows * kDwCols) + yr * kDwCols + 4 * qd]) = q;
                }
            }
        } else {  // one channel: rows of 64 contiguous values
            static_assert(COUT == 1, "weight-gradient tiles for 32 or 1 output channels");
            if (tid < 32) {
                const int yr = tid >> 3, seg = tid & 7;
                *reinterpret_cast<s8*>(&Dl[yr * kDwCols + 8 * seg]) = *reinterpret_cast<const s8*>(
                    db + ((size_t)((y0 + yr) << a.lg) + x0 + 8 * seg));
            }
        }
        __syncthreads();
        if (tid < 8 * COUT) {  // bias: channel tid / 8, pixels 32 (tid % 8) .. + 31
            const bf16* d = &Dl[(tid >> 3) * (kDwRows * kDwCols) + 32 * (tid & 7)];
#pragma unroll
            for (int e = 0; e < 32; ++e) dbs += bf2f(d[e]);
        }
#pragma unroll
        for (int yr = 0; yr < kDwRows; ++yr) {
#pragma unroll
            for (int hx = 0; hx < 2; ++hx) {
                const int xs = 32 * hx + 8 * g;  // this lane's first pixel in the tile row
                const bf16* xrow = &Xl[(16 * h + r) * kDwXc + (yr + ky) * kDwXs + xs];
                const unsigned lo = *reinterpret_cast<const unsigned*>(xrow + 6);
                const uint4 mid = *reinterpret_cast<const uint4*>(xrow + 8);
                const unsigned hi = *reinterpret_cast<const unsigned*>(xrow + 16);
                const unsigned u1 = align2(mid.y, mid.x), u2 = align2(mid.z, mid.y),
                               u3 = align2(mid.w, mid.z);
                const uint4 av[3] = {uint4{align2(mid.x, lo), u1, u2, u3}, mid,
                                     uint4{u1, u2, u3, align2(hi, mid.w)}};
#pragma unroll
                for (int nb = 0; nb < NT; ++nb) {
                    const uint4 bv = *reinterpret_cast<const uint4*>(
                        &Dl[(16 * nb + r) * (kDwRows * kDwCols) + yr * kDwCols + xs]);
#pragma unroll
                    for (int kx = 0; kx < 3; ++kx)
                        acc[kx][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf8, av[kx]), __builtin_bit_cast(bf8, bv),
                            acc[kx][nb], 0, 0, 0);
                }
            }
        }
    }
    // D[row = channel 16 h + 4 g + i][col = output channel 16 nb + r] of tap (ky, kx)
    float* out = a.part + (size_t)blockIdx.x * (NW + COUT);
#pragma unroll
    for (int kx = 0; kx < 3; ++kx)
#pragma unroll
        for (int nb = 0; nb < NT; ++nb) {
            const int co = 16 * nb + r;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (co < COUT) out[(co * 9 + 3 * ky + kx) * CIN + 16 * h + 4 * g + i] = acc[kx][nb][i];
        }
    if (tid < 8 * COUT) {  // eight neighbouring lanes hold one channel's parts
        dbs += __shfl_xor(dbs, 1);
        dbs += __shfl_xor(dbs, 2);
        dbs += __shfl_xor(dbs, 4);
        if ((tid & 7) == 0) out[NW + (tid >> 3)] = dbs;
    }
}

// The same for the one-channel input layer: dW^T[tap][co] = sum_p X[p +
// off(tap)] dZ[p][co], nine taps padded to one 16-row MFMA tile.  Four waves,
// one per tile row; lane row r < 9 is tap (r / 3, r % 3) and picks its column
// shift from the three byte-aligned forms.  Each wave writes its own partial
// row (part[4 workgroup + wave][9 x 32 + 32]).
template <int COUT>
__global__ __launch_bounds__(256) void k_conv_dw_tile_in(DwTileArgs a) {
    static_assert(COUT == 32, "input-layer tiles for 32 output channels");
    constexpr int NT = COUT / 16;
    constexpr int NW = 9 * COUT;
    __shared__ __align__(16) bf16 Xl[kDwXc];
    __shared__ __align__(16) bf16 Dl[COUT * kDwRows * kDwCols];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int tky = r / 3, tkx = r - 3 * (r / 3);  // tap of this lane's A row (r < 9)
    f4 acc[NT];
#pragma unroll
    for (int nb = 0; nb < NT; ++nb) acc[nb] = f4{0.f, 0.f, 0.f, 0.f};
    float dbs = 0.f;
    const int tx_n = a.n / kDwCols, ty_n = a.n / kDwRows;
    const int t_begin = blockIdx.x * a.per_wg;
    const int t_end = min(t_begin + a.per_wg, a.tiles);
    for (int t = t_begin; t < t_end; ++t) {
        const int mem = t / (tx_n * ty_n), rem = t - mem * (tx_n * ty_n);
        const int y0 = (rem / tx_n) * kDwRows, x0 = (rem % tx_n) * kDwCols;
        const bf16* xb = a.x + (size_t)mem * a.npix;
        const bf16* db = a.dz + (size_t)mem * a.npix * COUT;
        __syncthreads();
        if (tid < 48) {  // interior: 6 rows x 8 chunks of 8 pixels
            const int yy = tid >> 3, seg = tid & 7;
            const int gy = (y0 + yy - 1) & (a.n - 1);
            *reinterpret_cast<s8*>(&Xl[yy * kDwXs + 8 + 8 * seg]) =
                *reinterpret_cast<const s8*>(xb + ((size_t)(gy << a.lg) + x0 + 8 * seg));
        } else if (tid < 60) {  // halo columns: 6 rows x 2 sides
            const int e = tid - 48, yy = e >> 1, side = e & 1;
            const int gy = (y0 + yy - 1) & (a.n - 1);
            const int gx = (side ? x0 + kDwCols : x0 - 1) & (a.n - 1);
            Xl[yy * kDwXs + (side ? 8 + kDwCols : 7)] = xb[(size_t)(gy << a.lg) + gx];
        }
        {   // dZ: 4 rows x 16 quads x 4 groups of 8 channels
            const int cg = tid & 3, qd = (tid >> 2) & 15, yr = tid >> 6;
            const bf16* src = db + ((size_t)((y0 + yr) << a.lg) + x0 + 4 * qd) * COUT + 8 * cg;
            const s8 v0 = *reinterpret_cast<const s8*>(src);
            const s8 v1 = *reinterpret_cast<const s8*>(src + COUT);
            const s8 v2 = *reinterpret_cast<const s8*>(src + 2 * COUT);
            const s8 v3 = *reinterpret_cast<const s8*>(src + 3 * COUT);
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const s4 q = {v0[c], v1[c], v2[c], v3[c]};
                *reinterpret_cast<s4*>(
                    &Dl[(8 * cg + c) * (kDwRows * kDwCols) + yr * kDwCols + 4 * qd]) = q;
            }
        }
        __syncthreads();
        {   // bias: channel tid / 8, pixels 32 (tid % 8) .. + 31
            const bf16* d = &Dl[(tid >> 3) * (kDwRows * kDwCols) + 32 * (tid & 7)];
#pragma unroll
            for (int e = 0; e < 32; ++e) dbs += bf2f(d[e]);
        }
        const int yr = wave;
#pragma unroll
        for (int hx = 0; hx < 2; ++hx) {
            const int xs = 32 * hx + 8 * g;
            uint4 av = uint4{0u, 0u, 0u, 0u};
            if (r < 9) {
                const bf16* xrow = &Xl[(yr + tky) * kDwXs + xs];
                const unsigned lo = *reinterpret_cast<const unsigned*>(xrow + 6);
                const uint4 mid = *reinterpret_cast<const uint4*>(xrow + 8);
                const unsigned hi = *reinterpret_cast<const unsigned*>(xrow + 16);
                const unsigned u1 = align2(mid.y, mid.x), u2 = align2(mid.z, mid.y),
                               u3 = align2(mid.w, mid.z);
                av = tkx == 1 ? mid
                     : tkx == 0 ? uint4{align2(mid.x, lo), u1, u2, u3}
                                : uint4{u1, u2, u3, align2(hi, mid.w)};
            }
#pragma unroll
            for (int nb = 0; nb < NT; ++nb) {
                const uint4 bv = *reinterpret_cast<const uint4*>(
                    &Dl[(16 * nb + r) * (kDwRows * kDwCols) + yr * kDwCols + xs]);
                acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    __builtin_bit_cast(bf8, av), __builtin_bit_cast(bf8, bv), acc[nb], 0, 0, 0);
            }
        }
    }
    // D[row = tap 4 g + i][col = output channel 16 nb + r]
    float* out = a.part + ((size_t)blockIdx.x * 4 + wave) * (NW + COUT);
#pragma unroll
    for (int nb = 0; nb < NT; ++nb)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (4 * g + i < 9) out[(16 * nb + r) * 9 + 4 * g + i] = acc[nb][i];
    dbs += __shfl_xor(dbs, 1);
    dbs += __shfl_xor(dbs, 2);
    dbs += __shfl_xor(dbs, 4);
    // this wave's bias channels are 8 wave .. 8 wave + 7; the others are zero in its row
    if ((lane & 7) == 0) out[NW + (tid >> 3)] = dbs;
    if (lane < COUT && (lane >> 3) != wave) out[NW + lane] = 0.f;
}

// grad[i] += sum over workgroups w of part[w][i]: 64 outputs per block; wave
// q sums the workgroups w = q mod 4 in order, and the four sums are added in
// a fixed order (the result does not depend on timing)
__global__ __launch_bounds__(256) void k_dw_reduce(int nwg, int stride, const float* part,
                                                   float* dw, int nw, float* db) {
    __shared__ float red[4][64];
    const int o = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + o;
    float s = 0.f;
    if (i < stride) {
#pragma unroll 8
        for (int w = q; w < nwg; w += 4) s += part[(size_t)w * stride + i];
    }
    red[q][o] = s;
    __syncthreads();
    if (q == 0 && i < stride) {
        const float t = (red[0][o] + red[1][o]) + (red[2][o] + red[3][o]);
        if (i < nw) dw[i] += t;
        else db[i - nw] += t;
    }
}

// mean squared error and its gradient dL/dy = 2 (y - t) / total, in bf16
__global__ void k_mse(int total, const float* y, const float* t, bf16* dy, double* loss) {
    __shared__ double red[4];
    double s = 0.0;
    const float sc = 2.f / (float)total;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
        const float d = y[q] - t[q];
        s += (double)d * d;
        dy[q] = f2bf(sc * d);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (red[0] + red[1] + red[2] + red[3]) / total);
}

// minibatch from the data set: x (scaled, bf16) and target (scaled, FP32)
__global__ void k_gather_batch(int npix, const int* idx, const float* ds_x, const float* ds_t,
                               float xs, float ts, bf16* x0, float* t) {
    const int b = blockIdx.y;
    const size_t src = (size_t)idx[b] * npix, dst = (size_t)b * npix;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
        x0[dst + p] = f2bf(xs * ds_x[src + p]);
        t[dst + p] = ts * ds_t[src + p];
    }
}

// sums of squares of the data set's inputs and targets (for the scales)
__global__ void k_sumsq(size_t count, const float* x, const float* t, double* out) {
    __shared__ double red[2][4];
    double sx = 0.0, st = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x) {
        sx += (double)x[i] * x[i];
        st += (double)t[i] * t[i];
    }
    for (int o = 32; o > 0; o >>= 1) {
        sx += __shfl_xor(sx, o, 64);
        st += __shfl_xor(st, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = sx;
        red[1][threadIdx.x >> 6] = st;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&out[0], red[0][0] + red[0][1] + red[0][2] + red[0][3]);
        atomicAdd(&out[1], red[1][0] + red[1][1] + red[1][2] + red[1][3]);
    }
}

__global__ void k_to_bf16(size_t count, const float* src, float scale, bf16* dst) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = f2bf(scale * src[i]);
}

// Adam on the FP32 master parameters; refreshes the bf16 weight copies
__global__ void k_adam(int count, float* w, const float* g, float* m, float* v, bf16* wb,
                       const int* is_weight, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (is_weight[i] >= 0) wb[is_weight[i]] = f2bf(wi);
    }
}

}  // namespace

// ----------------------------------------------------------------- closure

struct NsdClosure {
    int n = 0, lg = 0, npix = 0, max_batch = 0;
    int cin[kLayers] = {1, kHidden, kHidden, kHidden, kHidden};
    int cout[kLayers] = {kHidden, kHidden, kHidden, kHidden, 1};
    size_t woff[kLayers] = {}, boff[kLayers] = {}, bwoff[kLayers] = {};
    size_t nparam = 0, nweight = 0;
    float* prm = nullptr;    // FP32 master: per layer W then b
    float* grad = nullptr;
    float* adm = nullptr;    // Adam first moments
    float* adv = nullptr;    // Adam second moments
    int* widx = nullptr;     // parameter -> bf16 weight index (-1 for biases)
    bf16* wb = nullptr;      // bf16 weights, per layer
    bf16* x0 = nullptr;      // input [B][npix][1]
    bf16* z[kHiddenLayers] = {};    // hidden pre-activations [B][npix][32]
    bf16* act[kHiddenLayers] = {};  // hidden activations
    float* y = nullptr;      // output [B][npix]
    bf16* dz[kHiddenLayers] = {};   // gradients of the hidden pre-activations
    bf16* dy = nullptr;      // gradient of the output [B][npix]
    float* xin = nullptr;    // FP32 staging [B][npix]
    float* tin = nullptr;
    double* loss = nullptr;
    unsigned long long adam_t = 0;
    float in_scale = 1.f, out_scale = 1.f;  // net input = in_scale x; tendency = out_scale net
    float* ds_x = nullptr;   // data set: inputs [count][npix]
    float* ds_t = nullptr;   // data set: targets [count][npix]
    int ds_count = 0;
    int* idx = nullptr;      // minibatch indices
    double* sums = nullptr;
    float* dwpart = nullptr;  // weight-gradient partial sums per workgroup
    int dw_wg = 0;            // their rows
    hipStream_t stream = nullptr;
    std::string err;
};

namespace {

int ck(NsdClosure* c, hipError_t e, const char* what) {
    if (e == hipSuccess) return 0;
    c->err = std::string(what) + ": " + hipGetErrorString(e);
    return 5;  // SOLVER_MEM_ERROR + 1 would collide; SolverStatus has 0..4
}
#define CKC(call) \
    do { if (ck(c, (call), #call)) return 3; } while (0)

unsigned long long splitmix(unsigned long long& s) {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// the layers on `st`, output (times yscale) into y; keep_z keeps the
// pre-activations for a backward pass
int launch_forward_to(NsdClosure* c, int batch, float* y, float yscale, hipStream_t st,
                      bool keep_z) {
    ConvArgs a;
    memset(&a, 0, sizeof(a));
    a.n = c->n;
    a.lg = c->lg;
    a.npix = c->npix;
    a.yscale = yscale;
    for (int l = 0; l < kLayers; ++l) {
        a.in = l == 0 ? c->x0 : c->act[l - 1];
        a.w = c->wb + c->bwoff[l];
        a.bias = c->prm + c->boff[l];
        if (l < kHiddenLayers) {
            a.z = keep_z ? c->z[l] : nullptr;
            a.act = c->act[l];
        } else {
            a.y = y;
        }
        if (l == 0) launch_conv_lds<1, kHidden, CONV_FWD_GELU>(a, batch, st);
        else if (l < kHiddenLayers) launch_conv_lds<kHidden, kHidden, CONV_FWD_GELU>(a, batch, st);
        else launch_conv_lds<kHidden, 1, CONV_FWD_LAST>(a, batch, st);
    }
    return ck(c, hipGetLastError(), "closure forward") ? 3 : 0;
}

int launch_forward(NsdClosure* c, int batch) {
    return launch_forward_to(c, batch, c->y, 1.f, c->stream, true);
}

int adam_step(NsdClosure* c, float lr) {
    const float b1 = 0.9f, b2 = 0.999f;
    c->adam_t++;
    const float c1 = 1.f - powf(b1, (float)c->adam_t), c2 = 1.f - powf(b2, (float)c->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(64), dim3(256), 0, c->stream, (int)c->nparam, c->prm, c->grad,
                       c->adm, c->adv, c->wb, c->widx, lr, b1, b2, 1e-8f, c1, c2);
    return ck(c, hipGetLastError(), "adam") ? 3 : 0;
}

int launch_backward(NsdClosure* c, int batch) {
    const int total = batch * c->npix;
    CKC(hipMemsetAsync(c->grad, 0, sizeof(float) * c->nparam, c->stream));
    // dW, db of layer l from its input and dz_l; then dz_{l-1} = (W_l^T dz_l) * gelu'(z_{l-1})
    DwArgs d;
    memset(&d, 0, sizeof(d));
    d.n = c->n;
    d.lg = c->lg;
    d.npix = c->npix;
    d.total = total;
    d.chunk = 1024;
    const int nblk = (total + d.chunk - 1) / d.chunk;
    ConvArgs a;
    memset(&a, 0, sizeof(a));
    a.n = c->n;
    a.lg = c->lg;
    a.npix = c->npix;
    for (int l = kLayers - 1; l >= 0; --l) {
        const bf16* dzl = l == kHiddenLayers ? c->dy : c->dz[l];
        d.x = l == 0 ? c->x0 : c->act[l - 1];
        d.dz = dzl;
        d.dw = c->grad + c->woff[l];
        d.db = c->grad + c->boff[l];
        if (l > 0 && c->dwpart) {
            DwTileArgs e;
            e.n = c->n;
            e.lg = c->lg;
            e.npix = c->npix;
            e.tiles = batch * (c->n / kDwRows) * (c->n / kDwCols);
            const int per = (e.tiles + c->dw_wg - 1) / c->dw_wg;
            e.per_wg = std::max(per, 8);
            const int nwg = (e.tiles + e.per_wg - 1) / e.per_wg;
            e.x = d.x;
            e.dz = d.dz;
            e.part = c->dwpart;
            const int co = l == kHiddenLayers ? 1 : kHidden, nw = 9 * kHidden * co;
            if (co == 1)
                hipLaunchKernelGGL((k_conv_dw_tile<1>), dim3(nwg), dim3(384), 0, c->stream, e);
            else
                hipLaunchKernelGGL((k_conv_dw_tile<kHidden>), dim3(nwg), dim3(384), 0, c->stream, e);
            hipLaunchKernelGGL(k_dw_reduce, dim3((nw + co + 63) / 64), dim3(256), 0, c->stream,
                               nwg, nw + co, c->dwpart, d.dw, nw, d.db);
        } else if (l == 0 && c->dwpart) {
            DwTileArgs e;
            e.n = c->n;
            e.lg = c->lg;
            e.npix = c->npix;
            e.tiles = batch * (c->n / kDwRows) * (c->n / kDwCols);
            const int per = (e.tiles + c->dw_wg - 1) / c->dw_wg;
            e.per_wg = std::max(per, 8);
            const int nwg = (e.tiles + e.per_wg - 1) / e.per_wg;
            e.x = d.x;
            e.dz = d.dz;
            e.part = c->dwpart;
            const int nw = 9 * kHidden;
            hipLaunchKernelGGL((k_conv_dw_tile_in<kHidden>), dim3(nwg), dim3(256), 0, c->stream, e);
            hipLaunchKernelGGL(k_dw_reduce, dim3((nw + kHidden + 63) / 64), dim3(256), 0,
                               c->stream, 4 * nwg, nw + kHidden, c->dwpart, d.dw, nw, d.db);
        } else if (l == kHiddenLayers)
            hipLaunchKernelGGL((k_conv_dw<kHidden, 1>), dim3(nblk), dim3(256), 0, c->stream, d);
        else if (l > 0)
            hipLaunchKernelGGL((k_conv_dw<kHidden, kHidden>), dim3(nblk), dim3(256), 0, c->stream, d);
        else
            hipLaunchKernelGGL((k_conv_dw<1, kHidden>), dim3(nblk), dim3(256), 0, c->stream, d);
        if (l == 0) break;
        a.in = dzl;
        a.w = c->wb + c->bwoff[l];
        a.zprev = c->z[l - 1];
        a.dzprev = c->dz[l - 1];
        if (l == kHiddenLayers) launch_conv_lds<1, kHidden, CONV_BWD_DATA>(a, batch, c->stream);
        else launch_conv_lds<kHidden, kHidden, CONV_BWD_DATA>(a, batch, c->stream);
    }
    return ck(c, hipGetLastError(), "closure backward") ? 3 : 0;
}

int stage_input(NsdClosure* c, const float* x_host, int batch) {
    const size_t cnt = (size_t)batch * c->npix;
    CKC(hipMemcpyAsync(c->xin, x_host, sizeof(float) * cnt, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_to_bf16, dim3(256), dim3(256), 0, c->stream, cnt, c->xin, 1.f, c->x0);
    return 0;
}

int refresh_bf16(NsdClosure* c) {
    std::vector<float> h(c->nparam);
    CKC(hipMemcpy(h.data(), c->prm, sizeof(float) * c->nparam, hipMemcpyDeviceToHost));
    std::vector<int> wi(c->nparam);
    CKC(hipMemcpy(wi.data(), c->widx, sizeof(int) * c->nparam, hipMemcpyDeviceToHost));
    std::vector<bf16> wb(c->nweight);
    for (size_t i = 0; i < c->nparam; ++i)
        if (wi[i] >= 0) {
            unsigned u;
            memcpy(&u, &h[i], 4);
            u += 0x7fffu + ((u >> 16) & 1u);
            wb[wi[i]] = (bf16)(u >> 16);
        }
    CKC(hipMemcpy(c->wb, wb.data(), sizeof(bf16) * c->nweight, hipMemcpyHostToDevice));
    return 0;
}

}  // namespace

extern "C" {

void nsd_closure_destroy(NsdClosure* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void* bufs[] = {c->prm, c->grad, c->adm, c->adv, c->widx, c->wb, c->x0, c->y, c->dy,
                    c->xin, c->tin, c->loss, c->ds_x, c->ds_t, c->idx, c->sums, c->dwpart};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (int i = 0; i < kHiddenLayers; ++i)
        for (void* b : {(void*)c->z[i], (void*)c->act[i], (void*)c->dz[i]})
            if (b) (void)hipFree(b);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

NsdClosure* nsd_closure_create(int n, int max_batch, unsigned long long seed, char* err,
                               size_t errlen) {
    auto fail = [&](const std::string& m, NsdClosure* c) -> NsdClosure* {
        if (err && errlen) snprintf(err, errlen, "%s", m.c_str());
        nsd_closure_destroy(c);
        return nullptr;
    };
    if (n < 16 || n > 4096 || (n & (n - 1)) || max_batch < 1)
        return fail("closure grid must be a power of two in [16, 4096], batch >= 1", nullptr);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no HIP device", nullptr);
    NsdClosure* c = new NsdClosure();
    c->n = n;
    while ((1 << c->lg) < n) ++c->lg;
    c->npix = n * n;
    c->max_batch = max_batch;
    size_t off = 0, wo = 0;
    for (int l = 0; l < kLayers; ++l) {
        c->woff[l] = off;
        off += (size_t)c->cout[l] * 9 * c->cin[l];
        c->boff[l] = off;
        off += c->cout[l];
        c->bwoff[l] = wo;
        wo += (size_t)c->cout[l] * 9 * c->cin[l];
    }
    c->nparam = off;
    c->nweight = wo;
    const size_t P = (size_t)max_batch * c->npix;
    auto al = [&](void** p, size_t bytes) { return hipMalloc(p, bytes) == hipSuccess; };
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && al((void**)&c->prm, 4 * off) && al((void**)&c->grad, 4 * off) &&
         al((void**)&c->adm, 4 * off) && al((void**)&c->adv, 4 * off) &&
         al((void**)&c->widx, 4 * off) && al((void**)&c->wb, 2 * wo) &&
         al((void**)&c->x0, 2 * P) && al((void**)&c->y, 4 * P) && al((void**)&c->dy, 2 * P) &&
         al((void**)&c->xin, 4 * P) && al((void**)&c->tin, 4 * P) &&
         al((void**)&c->loss, sizeof(double)) && al((void**)&c->idx, sizeof(int) * max_batch) &&
         al((void**)&c->sums, 2 * sizeof(double));
    for (int i = 0; i < kHiddenLayers && ok; ++i)
        ok = al((void**)&c->z[i], 2 * P * kHidden) && al((void**)&c->act[i], 2 * P * kHidden) &&
             al((void**)&c->dz[i], 2 * P * kHidden);
    if (ok && n % kDwCols == 0) {
        const long tiles = (long)max_batch * (n / kDwRows) * (n / kDwCols);
        c->dw_wg = (int)std::min<long>(kDwMaxWg, (tiles + 7) / 8);
        ok = al((void**)&c->dwpart, sizeof(float) * (size_t)c->dw_wg * (9 * kHidden * kHidden + kHidden));
    }
    if (!ok) return fail("closure: device allocation failed", c);
    // He-uniform initial weights from a SplitMix64 stream, zero biases
    std::vector<float> h(off, 0.f);
    std::vector<int> wi(off, -1);
    unsigned long long s = seed ? seed : 0x5eedull;
    for (int l = 0; l < kLayers; ++l) {
        const size_t cnt = (size_t)c->cout[l] * 9 * c->cin[l];
        const double lim = std::sqrt(6.0 / (9.0 * c->cin[l]));
        for (size_t i = 0; i < cnt; ++i) {
            const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
            h[c->woff[l] + i] = (float)((2.0 * u - 1.0) * lim);
            wi[c->woff[l] + i] = (int)(c->bwoff[l] + i);
        }
    }
    if (hipMemcpy(c->prm, h.data(), 4 * off, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->widx, wi.data(), 4 * off, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->adm, 0, 4 * off) != hipSuccess || hipMemset(c->adv, 0, 4 * off) != hipSuccess ||
        refresh_bf16(c))
        return fail("closure: initialisation failed", c);
    return c;
}

const char* nsd_closure_error(const NsdClosure* c) { return c ? c->err.c_str() : "null closure"; }

int nsd_closure_param_count(const NsdClosure* c) { return c ? (int)c->nparam : 0; }

int nsd_closure_params(NsdClosure* c, float* params, int to_device) {
    if (!c || !params) return 4;
    if (to_device) {
        CKC(hipMemcpy(c->prm, params, 4 * c->nparam, hipMemcpyHostToDevice));
        return refresh_bf16(c);
    }
    CKC(hipMemcpy(params, c->prm, 4 * c->nparam, hipMemcpyDeviceToHost));
    return 0;
}

int nsd_closure_forward(NsdClosure* c, const float* x, int batch, float* y) {
    if (!c || !x || !y || batch < 1 || batch > c->max_batch) return 4;
    // inference: the pre-activations are not kept
    if (stage_input(c, x, batch) || launch_forward_to(c, batch, c->y, 1.f, c->stream, false))
        return 3;
    CKC(hipMemcpyAsync(y, c->y, sizeof(float) * batch * c->npix, hipMemcpyDeviceToHost, c->stream));
    CKC(hipStreamSynchronize(c->stream));
    return 0;
}

int nsd_closure_gradient(NsdClosure* c, const float* x, const float* t, int batch, double* loss,
                         float* grad) {
    if (!c || !x || !t || batch < 1 || batch > c->max_batch) return 4;
    const int total = batch * c->npix;
    if (stage_input(c, x, batch) || launch_forward(c, batch)) return 3;
    CKC(hipMemcpyAsync(c->tin, t, sizeof(float) * total, hipMemcpyHostToDevice, c->stream));
    CKC(hipMemsetAsync(c->loss, 0, sizeof(double), c->stream));
    hipLaunchKernelGGL(k_mse, dim3(256), dim3(256), 0, c->stream, total, c->y, c->tin, c->dy,
                       c->loss);
    if (launch_backward(c, batch)) return 3;
    if (loss) CKC(hipMemcpyAsync(loss, c->loss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (grad)
        CKC(hipMemcpyAsync(grad, c->grad, 4 * c->nparam, hipMemcpyDeviceToHost, c->stream));
    CKC(hipStreamSynchronize(c->stream));
    return 0;
}


int nsd_closure_set_scales(NsdClosure* c, double in_scale, double out_scale) {
    if (!c || !(in_scale > 0.0) || !(out_scale > 0.0)) return 4;
    c->in_scale = (float)in_scale;
    c->out_scale = (float)out_scale;
    return 0;
}

int nsd_closure_get_scales(const NsdClosure* c, double* in_scale, double* out_scale) {
    if (!c) return 4;
    if (in_scale) *in_scale = c->in_scale;
    if (out_scale) *out_scale = c->out_scale;
    return 0;
}

int nsd_closure_dataset_reserve(NsdClosure* c, int count) {
    if (!c || count < 1) return 4;
    if (c->ds_x) (void)hipFree(c->ds_x);
    if (c->ds_t) (void)hipFree(c->ds_t);
    c->ds_x = c->ds_t = nullptr;
    c->ds_count = 0;
    const size_t b = sizeof(float) * (size_t)count * c->npix;
    CKC(hipMalloc(&c->ds_x, b));
    CKC(hipMalloc(&c->ds_t, b));
    c->ds_count = count;
    return 0;
}

int nsd_closure_dataset_slot(NsdClosure* c, int index, float** x, float** t) {
    if (!c || index < 0 || index >= c->ds_count || !x || !t) return 4;
    *x = c->ds_x + (size_t)index * c->npix;
    *t = c->ds_t + (size_t)index * c->npix;
    return 0;
}

int nsd_closure_dataset_fill(NsdClosure* c, int index, const float* x, const float* t) {
    float *dx, *dt;
    if (nsd_closure_dataset_slot(c, index, &dx, &dt)) return 4;
    CKC(hipMemcpy(dx, x, sizeof(float) * c->npix, hipMemcpyHostToDevice));
    CKC(hipMemcpy(dt, t, sizeof(float) * c->npix, hipMemcpyHostToDevice));
    return 0;
}

int nsd_closure_dataset_read(NsdClosure* c, int index, float* x, float* t) {
    float *dx, *dt;
    if (!x || !t || nsd_closure_dataset_slot(c, index, &dx, &dt)) return 4;
    CKC(hipMemcpy(x, dx, sizeof(float) * c->npix, hipMemcpyDeviceToHost));
    CKC(hipMemcpy(t, dt, sizeof(float) * c->npix, hipMemcpyDeviceToHost));
    return 0;
}

int nsd_closure_train(NsdClosure* c, int count, int epochs, int batch, double lr,
                      unsigned long long seed, double* loss_first, double* loss_last,
                      double* loss_valid) {
    if (!c || count < 2 || count > c->ds_count || epochs < 1 || batch < 1 ||
        batch > c->max_batch || !(lr > 0.0))
        return 4;
    const size_t cnt = (size_t)count * c->npix;
    // scales: unit mean square for the network's input and output
    CKC(hipMemsetAsync(c->sums, 0, 2 * sizeof(double), c->stream));
    hipLaunchKernelGGL(k_sumsq, dim3(1024), dim3(256), 0, c->stream, cnt, c->ds_x, c->ds_t, c->sums);
    double sums[2];
    CKC(hipMemcpyAsync(sums, c->sums, sizeof(sums), hipMemcpyDeviceToHost, c->stream));
    CKC(hipStreamSynchronize(c->stream));
    if (!(sums[0] > 0.0) || !(sums[1] > 0.0)) {
        c->err = "closure data set is empty or zero";
        return 4;
    }
    c->in_scale = (float)std::sqrt(cnt / sums[0]);
    c->out_scale = (float)std::sqrt(sums[1] / cnt);
    const float ts = 1.f / c->out_scale;
    // every tenth pair is held out for validation (spread over members and times)
    std::vector<int> order, valid;
    for (int i = 0; i < count; ++i) (i % 10 == 9 ? valid : order).push_back(i);
    if (valid.empty()) {
        valid.push_back(order.back());
        order.pop_back();
    }
    const int ntrain = (int)order.size();
    unsigned long long s = seed ? seed : 0xc105eull;
    const dim3 ggrid((c->npix + 255) / 256 < 64 ? (c->npix + 255) / 256 : 64, batch);
    // learning rate: cosine decay from lr to lr / 20 over all training steps
    const long total_steps = (long)epochs * (ntrain / batch);
    long step = 0;
    auto rate = [&]() {
        const double f = total_steps > 1 ? (double)step / (double)(total_steps - 1) : 0.0;
        return lr * (0.05 + 0.475 * (1.0 + std::cos(3.14159265358979323846 * f)));
    };
    auto run_batch = [&](const int* ids, int nb, bool learn, double* loss_acc) -> int {
        CKC(hipMemcpyAsync(c->idx, ids, sizeof(int) * nb, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_gather_batch, dim3(ggrid.x, nb), dim3(256), 0, c->stream, c->npix,
                           c->idx, c->ds_x, c->ds_t, c->in_scale, ts, c->x0, c->tin);
        if (launch_forward(c, nb)) return 3;
        CKC(hipMemsetAsync(c->loss, 0, sizeof(double), c->stream));
        hipLaunchKernelGGL(k_mse, dim3(256), dim3(256), 0, c->stream, nb * c->npix, c->y, c->tin,
                           c->dy, c->loss);
        if (learn && (launch_backward(c, nb) || adam_step(c, (float)rate()))) return 3;
        if (learn) ++step;
        double l = 0.0;
        CKC(hipMemcpyAsync(&l, c->loss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        CKC(hipStreamSynchronize(c->stream));
        *loss_acc += l * nb;
        return 0;
    };
    for (int e = 0; e < epochs; ++e) {
        for (int i = ntrain - 1; i > 0; --i) {  // Fisher-Yates with SplitMix64
            const int j = (int)(splitmix(s) % (unsigned long long)(i + 1));
            std::swap(order[i], order[j]);
        }
        double acc = 0.0;
        for (int b0 = 0; b0 + batch <= ntrain; b0 += batch)
            if (run_batch(&order[b0], batch, true, &acc)) return 3;
        const double mean = acc / (double)(ntrain / batch * batch);
        if (e == 0 && loss_first) *loss_first = mean;
        if (loss_last) *loss_last = mean;
    }
    double vacc = 0.0;
    int done = 0;
    for (size_t b0 = 0; b0 < valid.size(); b0 += batch) {
        const int nb = (int)std::min<size_t>(batch, valid.size() - b0);
        if (run_batch(&valid[b0], nb, false, &vacc)) return 3;
        done += nb;
    }
    if (loss_valid) *loss_valid = vacc / done;
    return 0;
}

// tendencies (FP32 [batch][n][n], device) of the vorticities x (same shape),
// queued on `stream` (capturable; uses the network's own work buffers)
int nsd_closure_eval_batch(NsdClosure* c, const float* x, float* y, int batch, void* stream) {
    if (!c || !x || !y || batch < 1 || batch > c->max_batch) return 4;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_to_bf16, dim3(256), dim3(256), 0, st, (size_t)batch * c->npix, x,
                       c->in_scale, c->x0);
    return launch_forward_to(c, batch, y, c->out_scale, st, false);
}

int nsd_closure_eval_device(NsdClosure* c, const float* x, float* y, void* stream) {
    return nsd_closure_eval_batch(c, x, y, 1, stream);
}

int nsd_closure_max_batch(const NsdClosure* c) { return c ? c->max_batch : 0; }

int nsd_closure_size(const NsdClosure* c) { return c ? c->n : 0; }

}  // extern "C"

namespace nsd {

unsigned short* closure_input(NsdClosure* c, float* in_scale) {
    *in_scale = c->in_scale;
    return c->x0;
}

int closure_eval_staged(NsdClosure* c, float* y, int batch, void* stream) {
    if (!c || !y || batch < 1 || batch > c->max_batch) return 4;
    return launch_forward_to(c, batch, y, c->out_scale, (hipStream_t)stream, false);
}

}  // namespace nsd
