// nsd_noalloc.hip -- nsd_physical_to_spectral_noalloc: the forward transform
// of nsd_physical_to_spectral (r2c, x 1/(nx ny), 2/3 mask: the reference's
// spectral_transforms.c:9-22) on caller device arrays, with no allocation.
//
// For callers that cannot allocate (code under HIP graph capture, serving
// loops that forbid hipMalloc).  Its two kernels are the ones the context
// path launches (k_row ROW_R2C, k_col COL_FWD); what the context path keeps in
// its own buffers -- the mixed-space intermediate (kept kx columns x ny, FP64
// complex) and the two twiddle tables -- lives here in `scratch`, one
// file-scope device array per GPU.  Its size bounds the grids: nx, ny powers of
// two in [16, 1024] (1024^2 needs 5.6 MB of the 6 MiB).
//
// Every call on a GPU shares that GPU's array and overwrites it completely
// (twiddles, then every element of the intermediate) before reading it, so
// nothing needs to start at zero; calls that overlap on one GPU (two streams,
// two threads) race on it and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

#include "ns_device.h"
#include "nsd_kernels.hpp"

namespace nsd {
namespace {

constexpr size_t kScratchBytes = size_t(6) << 20;
constexpr int kMaxDevices = 64;
constexpr int kMaxLen = 1024;
// workgroup size of every line length up to kMaxLen (lines share a workgroup;
// LDS 2048 / L lines x (L + L / 16) x 16 bytes = 34 KB at every length)
constexpr int kNT = threads_for(kMaxLen);
static_assert(kNT == 256 && threads_for(16) == kNT, "one workgroup size for lengths 16..1024");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// device address of `scratch` per device, from hipGetSymbolAddress on the
// first call made on that device (the array's name is only a host-side
// stand-in and is never passed to a kernel)
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

int scratch_of_current_device(void** out, char* err, size_t errlen) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        if (err && errlen)
            snprintf(err, errlen, "no usable current HIP device (%s)", hipGetErrorString(e));
        return 3;  // SOLVER_MEM_ERROR
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            if (err && errlen)
                snprintf(err, errlen, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return 3;
        }
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return 0;
}

// W_L^m = exp(-2 pi i m / L) for m < L, with the exact quadrant reduction of
// the integer ratio that the context path uses on the host
__global__ void k_twiddles(int L, cx<double>* tw) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    const int q = (4 * m) / L, r = 4 * m - q * L;
    double c0, s0;
    if (2 * r <= L) {
        double s, c;
        sincos((kPi / 2.0) * ((double)r / (double)L), &s, &c);
        c0 = c;
        s0 = s;
    } else {
        double s, c;
        sincos((kPi / 2.0) * ((double)(L - r) / (double)L), &s, &c);
        c0 = s;
        s0 = c;
    }
    double cc, ss;
    switch (q) {
    case 0: cc = c0; ss = s0; break;
    case 1: cc = -s0; ss = c0; break;
    case 2: cc = -c0; ss = -s0; break;
    default: cc = s0; ss = -c0; break;
    }
    tw[m] = mk(cc, -ss);
}

bool pow2_len(int n) { return n >= 16 && n <= kMaxLen && (n & (n - 1)) == 0; }

size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace
}  // namespace nsd

using namespace nsd;

namespace {

template <int NT>
void row_r2c(const RowArgs<double>& r, int L, int pairs, hipStream_t st) {
    const int lines = NT * P / L;
    const int blocks = (pairs + lines - 1) / lines;
    const size_t lds = (size_t)lines * (L + L / 16) * sizeof(cx<double>);
    hipLaunchKernelGGL((k_row<double, NT, ROW_R2C, 0>), dim3(blocks), dim3(NT), lds, st, r);
}

template <int NT>
void col_fwd(const ColArgs<double>& a, int L, int ncols, hipStream_t st) {
    const int lines = NT * P / L;
    const int blocks = (ncols + lines - 1) / lines;
    const size_t lds = (size_t)lines * (L + L / 16) * sizeof(cx<double>);
    hipLaunchKernelGGL((k_col<double, NT, COL_FWD, 0, 0>), dim3(blocks), dim3(NT), lds, st, a);
}

}  // namespace

extern "C" int nsd_physical_to_spectral_noalloc(int nx, int ny, const double* field,
                                                double* field_hat, void* stream, char* err,
                                                size_t errlen) {
    if (err && errlen) err[0] = '\0';
    if (!field || !field_hat) {
        if (err && errlen) snprintf(err, errlen, "field and field_hat must be device arrays");
        return 4;  // SOLVER_INVALID_PARAM
    }
    if (!pow2_len(nx) || !pow2_len(ny)) {
        if (err && errlen)
            snprintf(err, errlen,
                     "nx = %d, ny = %d: the allocation-free transform takes powers of two in "
                     "[16, %d] (its scratch array holds %zu bytes)",
                     nx, ny, kMaxLen, kScratchBytes);
        return 4;
    }
    Geo G;
    G.nx = nx;
    G.ny = ny;
    G.nkx = nx / 2 + 1;
    G.kxmax = (2 * G.nkx) / 3;  // spectral_initialization.c:46 (quirk Q1)
    if (G.kxmax > G.nkx - 1) G.kxmax = G.nkx - 1;
    G.kymax = (2 * ny) / 6;  // spectral_initialization.c:48
    G.nkx_keep = G.kxmax + 1;
    G.nky_keep = 2 * G.kymax + 1;
    G.Lx = G.Ly = 2.0 * kPi;  // the transform and the mask do not depend on the domain
    const size_t twx_b = round256((size_t)nx * sizeof(cx<double>));
    const size_t twy_b = round256((size_t)ny * sizeof(cx<double>));
    const size_t mix_b = round256((size_t)G.nkx_keep * ny * sizeof(cx<double>));
    if (twx_b + twy_b + mix_b > kScratchBytes) {
        if (err && errlen)
            snprintf(err, errlen, "%d x %d needs %zu bytes of scratch, the array holds %zu", nx,
                     ny, twx_b + twy_b + mix_b, kScratchBytes);
        return 4;
    }
    void* base = nullptr;
    int rc = scratch_of_current_device(&base, err, errlen);
    if (rc) return rc;
    char* p = static_cast<char*>(base);
    cx<double>* twx = reinterpret_cast<cx<double>*>(p);
    cx<double>* twy = reinterpret_cast<cx<double>*>(p + twx_b);
    cx<double>* mixed = reinterpret_cast<cx<double>*>(p + twx_b + twy_b);
    hipStream_t st = static_cast<hipStream_t>(stream);

    hipLaunchKernelGGL(k_twiddles, dim3((nx + 255) / 256), dim3(256), 0, st, nx, twx);
    hipLaunchKernelGGL(k_twiddles, dim3((ny + 255) / 256), dim3(256), 0, st, ny, twy);

    RowArgs<double> r;
    memset(&r, 0, sizeof(r));
    r.geo = G;
    r.mode = ROW_R2C;
    r.tw = twx;
    r.nl = mixed;
    r.in_ref = field;
    row_r2c<kNT>(r, nx, ny / 2, st);

    ColArgs<double> a;
    memset(&a, 0, sizeof(a));
    a.geo = G;
    a.mode = COL_FWD;
    a.tw = twy;
    a.nl = mixed;
    a.out_ref = field_hat;
    a.fwd_scale = 1.0;
    col_fwd<kNT>(a, ny, G.nkx, st);

    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        if (err && errlen) snprintf(err, errlen, "launch: %s", hipGetErrorString(e));
        return 3;
    }
    return 0;
}
