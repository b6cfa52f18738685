// nsd_api.hip -- device context, small kernels and the extern "C" layer
// declared in include/ns_device.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "ns_closure.h"
#include "ns_device.h"
#include "nsd_closure_link.hpp"
#include "nsd_stage.hpp"

using namespace nsd;

namespace {

constexpr int ST_OK = 0, ST_MEM = 3, ST_PARAM = 4;

// ------------------------------------------------------------ small kernels

__global__ void k_fill(double* p, size_t n, double v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// plain streaming copy (bandwidth probe and counter calibration)
__global__ void k_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// reference layout (FP64, full H) -> compact state (kept modes)
template <typename T>
__global__ void k_gather_state(Geo G, const double* __restrict__ ref, cx<T>* __restrict__ st) {
    const size_t n = (size_t)G.nkx_keep * G.nky_keep;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(e / G.nky_keep), cy = (int)(e % G.nky_keep);
        const int jy = cy <= G.kymax ? cy : cy + (G.ny - G.nky_keep);
        st[e] = ld_ref<T>(ref, (size_t)jy * G.nkx + kx);
    }
}

// compact state -> reference layout (zeros outside the mask)
template <typename T>
__global__ void k_scatter_state(Geo G, const cx<T>* __restrict__ st, double* __restrict__ ref) {
    const size_t n = (size_t)G.nkx * G.ny;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int jy = (int)(e / G.nkx), kx = (int)(e % G.nkx);
        cx<T> v = mk(T(0), T(0));
        if (kept(G, kx, jy)) v = st[(size_t)kx * G.nky_keep + compact_y(G, jy)];
        st_ref(ref, e, v);
    }
}

// spectral_initialization.c:10-49
__global__ void k_grids(Geo G, double* x, double* y, double* kxa, double* kya, double* k2,
                        int* mask, unsigned long long* kmax_bits) {
    const size_t n = (size_t)G.nkx * G.ny;
    double m = 0.0;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int jy = (int)(e / G.nkx), i = (int)(e % G.nkx);
        const double kx = kx_of(G, i), ky = ky_of(G, jy);
        const double kk = kx * kx + ky * ky;
        if (k2) k2[e] = kk;
        if (mask) mask[e] = kept(G, i, jy) ? 1 : 0;
        m = fmax(m, sqrt(kk));
        if (jy == 0 && kxa) kxa[i] = kx;
        if (i == 0 && kya) kya[jy] = ky;
    }
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (x && t < (size_t)G.nx) x[t] = t * G.Lx / G.nx;
    if (y && t < (size_t)G.ny) y[t] = t * G.Ly / G.ny;
    __shared__ double red[4];
    block_max_to(m, kmax_bits, red);
}

// hyperviscosity.c:98-101
__global__ void k_hv_weights(Geo G, Phys P, double* w) {
    const size_t n = (size_t)G.nkx * G.ny;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int jy = (int)(e / G.nkx), i = (int)(e % G.nkx);
        const double kx = kx_of(G, i), ky = ky_of(G, jy);
        w[e] = hv_weight(P, sqrt(kx * kx + ky * ky));
    }
}

// spectral_core.c:8-34 on reference-layout arrays
__global__ void k_velocity_ref(Geo G, const double* __restrict__ w, double* __restrict__ psi,
                               double* __restrict__ u, double* __restrict__ v) {
    const size_t n = (size_t)G.nkx * G.ny;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int jy = (int)(e / G.nkx), i = (int)(e % G.nkx);
        const double kx = kx_of(G, i), ky = ky_of(G, jy);
        const double k2 = kx * kx + ky * ky;
        const cx<double> om = ld_ref<double>(w, e);
        const cx<double> ps = k2 > 1e-12 ? (-1.0 / k2) * om : mk(0.0, 0.0);
        if (psi) st_ref(psi, e, ps);
        if (u) st_ref(u, e, mul_i(ps, ky));
        if (v) st_ref(v, e, mul_i(ps, -kx));
    }
}

// spectral_transforms.c:35-43
__global__ void k_dealias_ref(Geo G, double* f) {
    const size_t n = (size_t)G.nkx * G.ny;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int jy = (int)(e / G.nkx), i = (int)(e % G.nkx);
        if (!kept(G, i, jy)) {
            f[2 * e] = 0.0;
            f[2 * e + 1] = 0.0;
        }
    }
}

// hyperviscosity.c:104-150
__global__ void k_hv_apply_ref(Geo G, Phys P, double* rhs, const double* st, const double* sfd) {
    const size_t n = (size_t)G.nkx * G.ny;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int jy = (int)(e / G.nkx), i = (int)(e % G.nkx);
        const double kx = kx_of(G, i), ky = ky_of(G, jy);
        const double km = sqrt(kx * kx + ky * ky);
        const cx<double> s = ld_ref<double>(st, e);
        cx<double> d = mk(0.0, 0.0);
        if (P.hv_on) d = d + hv_weight(P, km) * s;
        if (P.sfd_on) {
            const double sw = sfd_weight(P, km);
            if (sw > 0.0) d = d + sw * (s - ld_ref<double>(sfd, e));
        }
        st_ref(rhs, e, ld_ref<double>(rhs, e) - d);
    }
}

// hyperviscosity.c:170-177
__global__ void k_hv_post_ref(size_t n, double alpha, const double* st, double* sfd) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const cx<double> p = ld_ref<double>(sfd, e);
        st_ref(sfd, e, (1.0 - alpha) * p + alpha * ld_ref<double>(st, e));
    }
}

// spectral_initialization.c:72-95 (physical part)
__global__ void k_taylor_green(Geo G, double* u, double* v, double* w, double* p) {
    const size_t n = (size_t)G.nx * G.ny;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int jy = (int)(e / G.nx), i = (int)(e % G.nx);
        const double x = i * G.Lx / G.nx, y = jy * G.Ly / G.ny;
        const double sx = sin(2 * kPi * x / G.Lx), cxv = cos(2 * kPi * x / G.Lx);
        const double sy = sin(2 * kPi * y / G.Ly), cy = cos(2 * kPi * y / G.Ly);
        u[e] = sx * cy;
        v[e] = -cxv * sy;
        w[e] = 2.0 * (2 * kPi / G.Lx) * cxv * cy;
        p[e] = 0.0;
    }
}

// spectral_time_integration.c:14-20 and spectral_analysis.c:8-33
__global__ void k_phys_reduce(size_t n, const double* u, const double* v, const double* w,
                              DevScalars* sc) {
    double m = 0.0, e = 0.0, z = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        if (u && v) {
            const double s = u[i] * u[i] + v[i] * v[i];
            m = fmax(m, s);
            e += 0.5 * s;
        }
        if (w) z += 0.5 * w[i] * w[i];
    }
    __shared__ double red[4];
    block_max_to(m, &sc->redmax[0], red);
    block_sum_to(e, &sc->red[0], red);
    block_sum_to(z, &sc->red[1], red);
}

// spectral_analysis.c:36-66 and :69-96 (max_k from k2[last]: quirk Q7)
__global__ void k_spec_reduce(Geo G, const double* w, DevScalars* sc) {
    const size_t n = (size_t)G.nkx * G.ny;
    const double kxl = kx_of(G, G.nkx - 1), kyl = ky_of(G, G.ny - 1);
    const double k2last = kxl * kxl + kyl * kyl;
    const double max_k = sqrt(k2last);
    double tot = 0.0, high = 0.0, mx = 0.0, es = 0.0, cnt = 0.0;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int jy = (int)(e / G.nkx), i = (int)(e % G.nkx);
        const double kx = kx_of(G, i), ky = ky_of(G, jy);
        const double k2 = kx * kx + ky * ky;
        const double a = hypot(w[2 * e], w[2 * e + 1]);
        const double a2 = a * a;
        tot += a2;
        const double kn = max_k > 0.0 ? sqrt(k2) / max_k : 0.0;
        if (kn > 0.8) high += a2;
        mx = fmax(mx, a);
        if (k2 > 0.25 * k2last) {
            es += a2;
            cnt += 1.0;
        }
    }
    __shared__ double red[4];
    block_sum_to(tot, &sc->red[0], red);
    block_sum_to(high, &sc->red[1], red);
    block_max_to(mx, &sc->redmax[0], red);
    block_sum_to(es, &sc->red[2], red);
    block_sum_to(cnt, &sc->red[3], red);
}

// spectral_analysis.c:103-123: shell sums of 1/2(|u|^2+|v|^2) over stored modes
__global__ void k_spectrum(Geo G, const double* u, const double* v, double* ek, int* ck,
                           int max_k) {
    const size_t n = (size_t)G.nkx * G.ny;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int jy = (int)(e / G.nkx), i = (int)(e % G.nkx);
        const double kx = kx_of(G, i), ky = ky_of(G, jy);
        const int bin = (int)(sqrt(kx * kx + ky * ky) + 0.5);
        if (bin < max_k) {
            const double au = hypot(u[2 * e], u[2 * e + 1]), av = hypot(v[2 * e], v[2 * e + 1]);
            atomicAdd(&ek[bin], 0.5 * (au * au + av * av));
            atomicAdd(&ck[bin], 1);
        }
    }
}

// sharp spectral filter: the modes of a fine reference-layout field that
// the coarse grid keeps (its 2/3 mask), in the coarse reference layout
__global__ void k_filter_ref(Geo F, Geo Cg, const double* fine, double* coarse) {
    const size_t n = (size_t)Cg.nkx * Cg.ny;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int jy = (int)(e / Cg.nkx), i = (int)(e % Cg.nkx);
        double re = 0.0, im = 0.0;
        if (kept(Cg, i, jy)) {
            const int b = band_of(Cg, jy);
            const size_t fe = (size_t)(b >= 0 ? b : b + F.ny) * F.nkx + i;
            re = fine[2 * fe];
            im = fine[2 * fe + 1];
        }
        coarse[2 * e] = re;
        coarse[2 * e + 1] = im;
    }
}

__global__ void k_sub_d(size_t n, const double* a, const double* b, double* out) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        out[e] = a[e] - b[e];
}

__global__ void k_d2f(size_t n, const double* a, float* out) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        out[e] = (float)a[e];
}

// time, dt and forcing counter of every member
__global__ void k_set_scalars(DevScalars* sc, int nb, int which, double t, double dt,
                              unsigned long long fcount, double dt_err) {
    for (int m = threadIdx.x; m < nb; m += blockDim.x) {
        if (which & 1) {
            sc[m].t = t;
            sc[m].dt = dt;
        }
        if (which & 2) sc[m].fcount = fcount;
        if (which & 4) {  // error controller: a fresh start
            sc[m].dt_err = dt_err;
            sc[m].err = 0.0;
            sc[m].n_acc = sc[m].n_rej = 0ull;
            sc[m].pending = sc[m].redo = 0;
        }
    }
}

// Error controller, at the start of every step (one workgroup per member):
// judge the candidate the previous step left.  err = |y4 - y2| / (atol +
// rtol |y4|) from the last stage's workgroup partials, summed in index order;
// err <= 1 (or a step already at dt_min) accepts, else t goes back by the
// candidate's dt and this step restores the state kept before it (redo).
// Either way the next steps may use dt_err = dt clamp(safety err^(-1/3),
// 0.2, 5): the estimate y4 - y2 is O(dt^3).
__global__ void k_judge(DevScalars* sc, const double* part, size_t stride, Phys ph) {
    DevScalars* s = sc + blockIdx.x;
    const double* p = part + blockIdx.x * stride;
    __shared__ double re[256], rz[256];
    const int pending = s->pending;
    const int n = s->errparts;
    double e = 0.0, z = 0.0;
    if (pending)
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            e += p[2 * i];
            z += p[2 * i + 1];
        }
    re[threadIdx.x] = e;
    rz[threadIdx.x] = z;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            re[threadIdx.x] += re[threadIdx.x + o];
            rz[threadIdx.x] += rz[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (!pending) {
        s->redo = 0;
        return;
    }
    const double num = sqrt(re[0]), den = ph.err_atol + ph.err_rtol * sqrt(rz[0]);
    const double err = den > 0.0 ? num / den : (num > 0.0 ? __builtin_inf() : 0.0);
    const double fac =
        err > 0.0 ? fmin(kErrFacMax, fmax(kErrFacMin, ph.err_safety * pow(err, -1.0 / 3.0)))
                  : kErrFacMax;
    const double dt = s->dt;
    s->err = err;
    s->pending = 0;
    if (err <= 1.0 || dt <= ph.dt_lo) {
        s->n_acc++;
        s->redo = 0;
    } else {
        s->n_rej++;
        s->t -= dt;
        s->redo = 1;
    }
    s->dt_err = dt * fac;
}

// Parseval sums over the compact state: E = 1/2 sum_full |w|^2/k^2,
// Z = 1/2 sum_full |w|^2 (half-plane weights 1 at kx = 0, else 2).
// This thread's share of both sums for member blockIdx.y.
template <typename T>
__device__ void state_energy_terms(const Geo& G, const cx<T>* st, double& e, double& z) {
    const size_t n = (size_t)G.nkx_keep * G.nky_keep;
    st += blockIdx.y * n;  // member blockIdx.y of a batch
    e = 0.0;
    z = 0.0;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x) {
        const int kxi = (int)(q / G.nky_keep), cy = (int)(q % G.nky_keep);
        const int jy = cy <= G.kymax ? cy : cy + (G.ny - G.nky_keep);
        const double kx = kx_of(G, kxi), ky = ky_of(G, jy);
        const double k2 = kx * kx + ky * ky;
        const double wt = (kxi == 0 || 2 * kxi == G.nx) ? 1.0 : 2.0;
        const double a2 = (double)st[q].x * st[q].x + (double)st[q].y * st[q].y;
        z += 0.5 * wt * a2;
        if (k2 > 1e-12) e += 0.5 * wt * a2 / k2;
    }
}

// Sum of a (blockDim.x <= 256)-thread workgroup in a fixed order, valid in
// thread 0.
__device__ double block_sum_fixed(double v, double* red_lds) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red_lds[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red_lds[i];
    return s;
}

// Parseval sums in a fixed order, so that they are the same bits on every
// call: each workgroup stores its partials, part[2 (member nparts + block)
// + {0, 1}], and k_sum_parts adds them in index order.
template <typename T>
__global__ void k_state_energy_part(Geo G, const cx<T>* st, double* part) {
    double e, z;
    state_energy_terms(G, st, e, z);
    __shared__ double red[4];
    e = block_sum_fixed(e, red);
    z = block_sum_fixed(z, red);
    if (threadIdx.x == 0) {
        const size_t o = 2 * ((size_t)blockIdx.y * gridDim.x + blockIdx.x);
        part[o] = e;
        part[o + 1] = z;
    }
}

// One workgroup per member (blockIdx.x): sc[m].red[0..1] = sums of its partials.
__global__ void k_sum_parts(const double* part, int nparts, DevScalars* sc) {
    part += 2 * (size_t)blockIdx.x * nparts;
    double e = 0.0, z = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        e += part[2 * i];
        z += part[2 * i + 1];
    }
    __shared__ double red[4];
    e = block_sum_fixed(e, red);
    z = block_sum_fixed(z, red);
    if (threadIdx.x == 0) {
        sc[blockIdx.x].red[0] = e;
        sc[blockIdx.x].red[1] = z;
    }
}

__device__ __forceinline__ unsigned long long splitmix64_at(unsigned long long seed,
                                                            unsigned long long k) {
    // k-th output (k >= 1) of the SplitMix64 stream of forcing.c:11-16
    unsigned long long z = seed + k * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double uniform_at(unsigned long long seed, unsigned long long k) {
    return (double)(splitmix64_at(seed, k) >> 11) * (1.0 / 9007199254740992.0);
}

// Random-phase initial vorticity (build-defined, same definition as the oracle).
template <typename T>
__global__ void k_random_phase(Geo G, unsigned long long seed, const unsigned long long* seeds,
                               double k_peak, cx<T>* st) {
    const size_t n = (size_t)G.nkx_keep * G.nky_keep;
    st += blockIdx.y * n;  // member blockIdx.y of a batch, with its own seed
    if (seeds) seed = seeds[blockIdx.y];
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x) {
        const int kxi = (int)(q / G.nky_keep), cy = (int)(q % G.nky_keep);
        int jy = cy <= G.kymax ? cy : cy + (G.ny - G.nky_keep);
        bool conj_of = false;
        if (kxi == 0 && jy > G.ny / 2) {  // Hermitian column: mirror of (0, ny - jy)
            jy = G.ny - jy;
            conj_of = true;
        }
        const double kx = kx_of(G, kxi), ky = ky_of(G, jy);
        const double k2 = kx * kx + ky * ky;
        cx<T> r = mk(T(0), T(0));
        if (k2 > 1e-12) {
            const double km = sqrt(k2), rr = km / k_peak;
            const double Ek = km * km * km * km * exp(-rr * rr);
            const double amp = sqrt(km * Ek / kPi);
            const size_t idx = (size_t)jy * G.nkx + kxi;
            const double ph = 2.0 * kPi * uniform_at(seed, idx + 1);
            r = mk(T(amp * cos(ph)), T((conj_of ? -1.0 : 1.0) * amp * sin(ph)));
        }
        st[q] = r;
    }
}

template <typename T>
__global__ void k_scale_state(size_t n, cx<T>* st, const DevScalars* sc, double target) {
    st += blockIdx.y * n;
    const double e = sc[blockIdx.y].red[0];
    const T s = T(e > 0.0 ? sqrt(target / e) : 0.0);
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x)
        st[q] = s * st[q];
}

// Forcing draws (forcing.c:117-184 semantics through the counter form of the
// SplitMix64 stream: on call c with A active modes, mode rank a uses
// uniforms 2(cA+a)+1 and 2(cA+a)+2), then the build-defined application:
// Hermitian projection of the kx = 0 column, mask, and scaling to the
// injection rate.  One workgroup.
struct ForceArgs {
    Geo geo;
    int A;
    const int* modes;    // reference index j*nkx + i per active mode (reference order)
    const int* partner;  // rank of the (0, ny-j) partner for kx = 0 modes, else -1
    unsigned long long seed;
    int colored;
    double tau, energy_target, rate;
    double* ou;          // colored state per active rank (interleaved complex)
    double* work;        // 2 * A doubles
    int fbx, fby;
    int box_T;           // 1: box in FP32, 0: FP64
    void* fbox;
    DevScalars* sc;
    const unsigned long long* m_seed;  // per-member seeds of a batch (else seed)
    int skip_on_redo;                  // error controller: no draw on a restoring step
};

__global__ void k_force(ForceArgs a) {
    const Geo& G = a.geo;
    if (a.skip_on_redo && a.sc[blockIdx.y].redo) return;
    {  // member blockIdx.y of a batch: its own generator, OU state, box and scalars
        const int mb = blockIdx.y;
        const size_t boxn = (size_t)a.fbx * (2 * a.fby + 1);
        a.sc += mb;
        a.ou += (size_t)mb * 2 * a.A;
        a.work += (size_t)mb * 4 * a.A;
        a.fbox = (char*)a.fbox + (size_t)mb * boxn * (a.box_T ? 2 * sizeof(float) : 2 * sizeof(double));
        if (a.m_seed) a.seed = a.m_seed[mb];
    }
    const unsigned long long c = a.sc->fcount;
    const double dt = a.sc->dt;
    double alpha = 0.0, beta = 1.0;
    if (a.colored) {
        const double tau = fmax(a.tau, 1e-6);
        alpha = exp(-dt / tau);
        beta = sqrt(fmax(1.0 - alpha * alpha, 0.0));
    }
    __shared__ double red[16];
    __shared__ double s_scale;
    double part = 0.0;
    for (int r = threadIdx.x; r < a.A; r += blockDim.x) {
        const unsigned long long k = 2ull * (c * (unsigned long long)a.A + r) + 1ull;
        double u1 = uniform_at(a.seed, k), u2 = uniform_at(a.seed, k + 1);
        u1 = fmax(u1, 2.2250738585072014e-308);
        const double rad = sqrt(-2.0 * log(u1));
        const double th = 2.0 * kPi * u2;
        double vr = rad * cos(th), vi = rad * sin(th);
        if (a.colored) {
            vr = alpha * a.ou[2 * r] + beta * vr;
            vi = alpha * a.ou[2 * r + 1] + beta * vi;
        }
        a.work[2 * r] = vr;
        a.work[2 * r + 1] = vi;
        part += vr * vr + vi * vi;
    }
    part = wave_sum(part);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
        t = fmax(t, 1e-16);
        s_scale = sqrt(a.energy_target / t);
    }
    __syncthreads();
    const double sc = s_scale;
    for (int r = threadIdx.x; r < a.A; r += blockDim.x) {
        a.work[2 * r] *= sc;
        a.work[2 * r + 1] *= sc;
        if (a.colored) {
            a.ou[2 * r] = a.work[2 * r];
            a.ou[2 * r + 1] = a.work[2 * r + 1];
        }
    }
    __syncthreads();
    // Hermitian projection + mask, rate of the projected field
    double rate = 0.0;
    double* pr = a.work + 2 * a.A;  // projected values
    for (int r = threadIdx.x; r < a.A; r += blockDim.x) {
        const int idx = a.modes[r];
        const int jy = idx / G.nkx, i = idx % G.nkx;
        double fr = a.work[2 * r], fi = a.work[2 * r + 1];
        if (i == 0 || 2 * i == G.nx) {
            const int p = a.partner[r];
            if (p >= 0) {
                fr = 0.5 * (fr + a.work[2 * p]);
                fi = 0.5 * (fi - a.work[2 * p + 1]);
            } else if (jy == 0 || 2 * jy == G.ny) {
                fi = 0.0;
            }
        }
        const double kx = kx_of(G, i), ky = ky_of(G, jy);
        const double k2 = kx * kx + ky * ky;
        if (!kept(G, i, jy) || k2 <= 1e-12) fr = fi = 0.0;
        pr[2 * r] = fr;
        pr[2 * r + 1] = fi;
        const double wt = (i == 0 || 2 * i == G.nx) ? 1.0 : 2.0;
        if (k2 > 1e-12) rate += 0.5 * wt * (fr * fr + fi * fi) / k2;
    }
    rate = wave_sum(rate);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = rate;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
        s_scale = (t > 0.0 && a.rate > 0.0) ? sqrt(a.rate / t) : 0.0;
    }
    __syncthreads();
    const double s2 = s_scale;
    const int boxn = a.fbx * (2 * a.fby + 1);
    for (int e = threadIdx.x; e < boxn; e += blockDim.x) {
        if (a.box_T) reinterpret_cast<float*>(a.fbox)[2 * e] = 0.f, reinterpret_cast<float*>(a.fbox)[2 * e + 1] = 0.f;
        else reinterpret_cast<double*>(a.fbox)[2 * e] = 0.0, reinterpret_cast<double*>(a.fbox)[2 * e + 1] = 0.0;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < a.A; r += blockDim.x) {
        const int idx = a.modes[r];
        const int jy = idx / G.nkx, i = idx % G.nkx;
        const int b = band_of(G, jy);
        if (i >= a.fbx || b < -a.fby || b > a.fby) continue;
        const int e = i * (2 * a.fby + 1) + (b + a.fby);
        if (a.box_T) {
            reinterpret_cast<float*>(a.fbox)[2 * e] = float(s2 * pr[2 * r]);
            reinterpret_cast<float*>(a.fbox)[2 * e + 1] = float(s2 * pr[2 * r + 1]);
        } else {
            reinterpret_cast<double*>(a.fbox)[2 * e] = s2 * pr[2 * r];
            reinterpret_cast<double*>(a.fbox)[2 * e + 1] = s2 * pr[2 * r + 1];
        }
    }
    if (threadIdx.x == 0) a.sc->fcount = c + 1;
}

// forcing.c:117-184 on a full reference-layout output (compat entry point)
__global__ void k_force_ref(int A, const int* modes, unsigned long long seed, unsigned long long c,
                            int colored, double alpha, double beta, double energy_target,
                            double* col, double* out) {
    __shared__ double red[16];
    __shared__ double s_scale;
    double part = 0.0;
    for (int r = threadIdx.x; r < A; r += blockDim.x) {
        const unsigned long long k = 2ull * (c * (unsigned long long)A + r) + 1ull;
        double u1 = uniform_at(seed, k), u2 = uniform_at(seed, k + 1);
        u1 = fmax(u1, 2.2250738585072014e-308);
        const double rad = sqrt(-2.0 * log(u1));
        const double th = 2.0 * kPi * u2;
        double vr = rad * cos(th), vi = rad * sin(th);
        const int idx = modes[r];
        if (colored) {
            vr = alpha * col[2 * idx] + beta * vr;
            vi = alpha * col[2 * idx + 1] + beta * vi;
            col[2 * idx] = vr;
            col[2 * idx + 1] = vi;
        }
        out[2 * idx] = vr;
        out[2 * idx + 1] = vi;
        part += vr * vr + vi * vi;
    }
    part = wave_sum(part);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
        t = fmax(t, 1e-16);
        s_scale = sqrt(energy_target / t);
    }
    __syncthreads();
    for (int r = threadIdx.x; r < A; r += blockDim.x) {
        const int idx = modes[r];
        out[2 * idx] *= s_scale;
        out[2 * idx + 1] *= s_scale;
        if (colored) {
            col[2 * idx] *= s_scale;
            col[2 * idx + 1] *= s_scale;
        }
    }
}

// exp(-2 pi i m / n), exact quadrant reduction of the integer ratio
void twiddle(long m, long n, double* c, double* s) {
    m %= n;
    long q = (4 * m) / n, r = 4 * m - q * n;
    double c0, s0;
    if (2 * r <= n) {
        const double a = (kPi / 2.0) * ((double)r / (double)n);
        c0 = std::cos(a);
        s0 = std::sin(a);
    } else {
        const double a = (kPi / 2.0) * ((double)(n - r) / (double)n);
        c0 = std::sin(a);
        s0 = std::cos(a);
    }
    double cc, ss;
    switch (q) {
    case 0: cc = c0; ss = s0; break;
    case 1: cc = -s0; ss = c0; break;
    case 2: cc = -c0; ss = -s0; break;
    default: cc = s0; ss = -c0; break;
    }
    *c = cc;
    *s = -ss;
}

unsigned grid_for(size_t n, unsigned bs = 256) {
    size_t g = (n + bs - 1) / bs;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace

// ------------------------------------------------------- dispatch order

namespace {

// The staggered start of the 4096-point FP32 row stage delays workgroups
// [ncu, 2 ncu) on the assumption that the dispatcher gives every compute unit
// one workgroup before it gives any a second.  k_dispatch_probe checks that
// assumption on the device: 2 ncu workgroups, each holding the LDS of one
// 4096-point FP32 row pair (so that at most two share a CU) and resident for
// about 30 us, record the compute unit they run on, read from the HW_ID
// (CU, shader array, shader engine) and XCC_ID hardware registers.
constexpr int kProbeLds = 2 * (4096 + 4096 / 16) * 8;

__global__ __launch_bounds__(512) void k_dispatch_probe(unsigned* cu_of) {
    extern __shared__ unsigned char probe_lds[];
    if (threadIdx.x == 0) {
        // s_getreg immediates: (size - 1) << 11 | offset << 6 | register
        const unsigned hw = __builtin_amdgcn_s_getreg((31 << 11) | 4);   // HW_ID, 32 bits
        const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | 20);  // XCC_ID, bits 3:0
        probe_lds[0] = 1;
        cu_of[blockIdx.x] = (xcc << 8) | ((hw >> 8) & 0x7fu);  // CU 11:8, SH 12, SE 14:13
    }
    for (int i = 0; i < 8; ++i) __builtin_amdgcn_s_sleep(127);
}

// 1 when workgroups [0, ncu) of a launch land on ncu different compute units
// and each unit holds exactly two of the first 2 ncu; cached per device
int dispatch_breadth_first(int dev, int ncu) {
    static std::mutex mu;
    static std::map<int, int> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) return it->second;
    int ok = 0;
    unsigned* d = nullptr;
    std::vector<unsigned> h(2 * (size_t)ncu);
    if (ncu > 0 && hipMalloc(&d, h.size() * sizeof(unsigned)) == hipSuccess) {
        (void)hipFuncSetAttribute((const void*)k_dispatch_probe,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, kProbeLds);
        hipLaunchKernelGGL(k_dispatch_probe, dim3(2 * ncu), dim3(512), kProbeLds, 0, d);
        if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
            hipMemcpy(h.data(), d, h.size() * sizeof(unsigned), hipMemcpyDeviceToHost) ==
                hipSuccess) {
            std::map<unsigned, int> first, all;
            for (int b = 0; b < 2 * ncu; ++b) {
                if (b < ncu) first[h[b]]++;
                all[h[b]]++;
            }
            ok = (int)first.size() == ncu && (int)all.size() == ncu;
            for (auto& kv : all) ok = ok && kv.second == 2;
        }
        (void)hipFree(d);
    }
    cache[dev] = ok;
    return ok;
}

}  // namespace

// ------------------------------------------------------------------ context

struct NsdContext {
    Geo G;
    int prec;
    int nb = 1;                             // batch members (one field each)
    int vrow = 0, vcol = 0;                 // launch variants of the step kernels
    int maxlds = 65536;                     // LDS bytes per workgroup on this device
    int ncu = 0;                            // compute units on this device
    int stagger = 0;                        // staggered start of the 4096-point FP32 row stage
    double* m_inv_re = nullptr;             // per-member 1/Re (batches), device
    unsigned long long* m_seed = nullptr;   // per-member forcing seeds (batches), device
    unsigned long long* m_ic = nullptr;     // per-member initial-field seeds, device
    NsdClosure* closure = nullptr;          // learned subgrid tendency (FP32, single field)
    void* cl_m = nullptr;                   // closure: mixed-layout scratch
    float* cl_y = nullptr;                  // closure: physical tendency
    void* cl_src = nullptr;                 // closure: compact spectral tendency
    void* ec_prev = nullptr;                // error controller: state before the step
    void* ec_ymid = nullptr;                // error controller: midpoint solution
    double* ec_part = nullptr;              // error controller: workgroup partials
    NsdParams P;
    Phys phys;
    hipStream_t stream = nullptr;
    // the forcing draw runs beside stage 1 on its own stream (fork after
    // stage 0, join before stage 2)
    hipStream_t fstream = nullptr;
    hipEvent_t fev[2] = {nullptr, nullptr};
    size_t esz;  // bytes per complex element at the working precision
    void* state = nullptr;
    void* acc = nullptr;
    void* stagebuf = nullptr;
    void* sfd = nullptr;
    void* mixed[4] = {nullptr, nullptr, nullptr, nullptr};
    void* nl = nullptr;
    void* twx = nullptr;
    void* twy = nullptr;
    DevScalars* sc = nullptr;
    double* stH[4] = {nullptr, nullptr, nullptr, nullptr};
    double* stR[4] = {nullptr, nullptr, nullptr, nullptr};
    int* fmodes = nullptr;
    int* fpartner = nullptr;
    double* fou = nullptr;
    double* fwork = nullptr;
    void* fbox = nullptr;
    int fA = 0, fbx = 0, fby = 0;
    int prof = 0;
    std::vector<hipEvent_t> evs;
    double prof_row = 0, prof_col = 0, prof_step = 0;
    int prof_steps = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int g_adaptive = -1;
    unsigned long long gen = 0, g_gen = ~0ull;
    std::string err;
};

// profiling events per step: three around the two kernels of each stage
constexpr int kEvPerStep = 12;

#define CK(call)                                                                   \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            c->err = std::string(#call) + ": " + hipGetErrorString(e_);            \
            return ST_MEM;                                                         \
        }                                                                          \
    } while (0)

namespace {

size_t H_of(const NsdContext* c) { return (size_t)c->G.nkx * c->G.ny; }
size_t R_of(const NsdContext* c) { return (size_t)c->G.nx * c->G.ny; }
size_t C_of(const NsdContext* c) { return (size_t)c->G.nkx_keep * c->G.nky_keep; }
// mixed-field and J buffers hold an even number of kx columns (J in 2 x 2
// blocks pairs columns k and k ^ 1)
size_t M_of(const NsdContext* c) { return (size_t)((c->G.nkx_keep + 1) & ~1) * c->G.ny; }

int ensure_staging(NsdContext* c) {
    for (int i = 0; i < 4; ++i) {
        if (!c->stH[i]) CK(hipMalloc(&c->stH[i], 2 * sizeof(double) * H_of(c)));
        if (!c->stR[i]) CK(hipMalloc(&c->stR[i], sizeof(double) * R_of(c)));
    }
    return ST_OK;
}

int up(NsdContext* c, void* dst, const void* src, size_t bytes) {
    if (!src) return ST_OK;
    CK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return ST_OK;
}
int down(NsdContext* c, void* dst, const void* src, size_t bytes) {
    if (!dst) return ST_OK;
    CK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    return ST_OK;
}
int sync(NsdContext* c) {
    CK(hipStreamSynchronize(c->stream));
    return ST_OK;
}
int lastk(NsdContext* c) {
    CK(hipGetLastError());
    return ST_OK;
}

// The stage kernels' mixed fields use a layout of their own (stage_layout in
// nsd_kernels.hpp: the pair layout at FP32, row-pair tiles at FP64) when the
// row and the column instantiations this context launches both hold it
// (square 4096^2 grids); the answer follows the launch variants.
bool pairs_ok(const NsdContext* c) {
    if (c->G.nx != c->G.ny) return false;
    if (c->prec == NSD_FP32)
        return stage_variant_pairs<float>(c->G.nx, c->vrow) &&
               stage_variant_pairs<float>(c->G.ny, c->vcol);
    return stage_variant_pairs<double>(c->G.nx, c->vrow) &&
           stage_variant_pairs<double>(c->G.ny, c->vcol);
}
// J in 2 x 2 blocks: written by the FP32 pair-layout row stage
bool jblocks_ok(const NsdContext* c) { return c->prec == NSD_FP32 && pairs_ok(c); }

template <typename T>
ColArgs<T> col_args(NsdContext* c) {
    ColArgs<T> a;
    memset(&a, 0, sizeof(a));
    a.geo = c->G;
    a.phys = c->phys;
    a.fwd_scale = 1.0;
    a.tw = (const cx<T>*)c->twy;
    a.nl = (const cx<T>*)c->nl;
    for (int i = 0; i < 4; ++i) a.mixed[i] = (cx<T>*)c->mixed[i];
    a.state = (cx<T>*)c->state;
    a.acc = (cx<T>*)c->acc;
    a.stagebuf = (cx<T>*)c->stagebuf;
    a.sfd = (cx<T>*)c->sfd;
    a.fbox = (const cx<T>*)c->fbox;
    a.sc = c->sc;
    a.bm = M_of(c);
    a.bc = C_of(c);
    a.bf = (size_t)c->fbx * (2 * c->fby + 1);
    a.m_inv_re = c->m_inv_re;
    a.src = c->closure ? (const cx<T>*)c->cl_src : nullptr;
    a.pairs = pairs_ok(c) ? 1 : 0;
    a.prev = (cx<T>*)c->ec_prev;
    a.ymid = (cx<T>*)c->ec_ymid;
    a.errpart = c->ec_part;
    a.bp = 2 * (size_t)c->G.nkx_keep;
    return a;
}

template <typename T>
RowArgs<T> row_args(NsdContext* c) {
    RowArgs<T> a;
    memset(&a, 0, sizeof(a));
    a.geo = c->G;
    a.tw = (const cx<T>*)c->twx;
    for (int i = 0; i < 4; ++i) a.mixed[i] = (const cx<T>*)c->mixed[i];
    a.nl = (cx<T>*)c->nl;
    a.sc = c->sc;
    a.bm = M_of(c);
    a.pairs = pairs_ok(c) ? 1 : 0;
    return a;
}

template <typename T, int NT>
void launch_col_nt(NsdContext* c, const ColArgs<T>& a, int blocks, size_t lds) {
    const dim3 grid(blocks, a.nbat > 1 ? a.nbat : 1);
    if (a.mode == COL_FWD)
        hipLaunchKernelGGL((k_col<T, NT, COL_FWD, 0, 0>), grid, dim3(NT), lds, c->stream, a);
    else
        hipLaunchKernelGGL((k_col<T, NT, COL_INV, 0, 0>), grid, dim3(NT), lds, c->stream, a);
}

template <typename T>
void launch_col_t(NsdContext* c, const ColArgs<T>& a, int ncols, bool errctl = false) {
    const int L = c->G.ny;
    if (a.mode == COL_STAGE || a.mode == COL_PRIME) {
        launch_col_fused<T>(a, L, ncols, c->nb, c->vcol, c->stream, errctl);
        return;
    }
    const int NT = threads_for(L);
    const int lines = NT * P / L;
    const int blocks = (ncols + lines - 1) / lines;
    const size_t lds = (size_t)lines * (L + L / 16) * sizeof(cx<T>);
    if (NT == 256) launch_col_nt<T, 256>(c, a, blocks, lds);
    else if (NT == 512) launch_col_nt<T, 512>(c, a, blocks, lds);
    else launch_col_nt<T, 1024>(c, a, blocks, lds);
}

template <typename T, int NT>
void launch_row_nt(NsdContext* c, const RowArgs<T>& a, int blocks, size_t lds) {
    const dim3 grid(blocks, a.nbat > 1 ? a.nbat : 1);
    if (a.mode == ROW_C2R)
        hipLaunchKernelGGL((k_row<T, NT, ROW_C2R, 0>), grid, dim3(NT), lds, c->stream, a);
    else
        hipLaunchKernelGGL((k_row<T, NT, ROW_R2C, 0>), grid, dim3(NT), lds, c->stream, a);
}

template <typename T>
void launch_row_t(NsdContext* c, const RowArgs<T>& a, int items, bool errctl = false) {
    const int L = c->G.nx;
    if (a.mode == ROW_STAGE) {
        launch_row_stage<T>(a, L, items, c->nb, c->vrow, c->stream, errctl);
        return;
    }
    const int NT = threads_for(L);
    const int lines = NT * P / L;
    const int blocks = (items + lines - 1) / lines;
    const size_t lds = (size_t)lines * (L + L / 16) * sizeof(cx<T>);
    if (NT == 256) launch_row_nt<T, 256>(c, a, blocks, lds);
    else if (NT == 512) launch_row_nt<T, 512>(c, a, blocks, lds);
    else launch_row_nt<T, 1024>(c, a, blocks, lds);
}

template <typename K>
void lds_attr(K kernel, int bytes) {
    (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              bytes);
}

template <typename T, int NT>
void set_lds_limits_nt(int bytes) {
    lds_attr(k_col<T, NT, COL_FWD, 0, 0>, bytes);
    lds_attr(k_col<T, NT, COL_INV, 0, 0>, bytes);
    lds_attr(k_row<T, NT, ROW_C2R, 0>, bytes);
    lds_attr(k_row<T, NT, ROW_R2C, 0>, bytes);
}

// allow the large dynamic LDS footprints of the long lines (up to the
// device's per-workgroup limit); failures surface at launch instead
template <typename T>
void set_lds_limits(int bytes) {
    set_lds_limits_nt<T, 256>(bytes);
    set_lds_limits_nt<T, 512>(bytes);
    set_lds_limits_nt<T, 1024>(bytes);
    stage_lds_limits<T>(bytes);
    (void)hipGetLastError();  // attribute calls must not leave a sticky error behind
}

size_t line_lds_bytes(int L, size_t esz) {
    const int NT = threads_for(L);
    return (size_t)(NT * P / L) * (L + L / 16) * esz;
}

// --- typed operations

template <typename T>
int op_prime(NsdContext* c) {
    ColArgs<T> a = col_args<T>(c);
    a.mode = COL_PRIME;
    launch_col_t<T>(c, a, c->G.nkx_keep);
    return lastk(c);
}

// the error controller runs with adaptive steps only
bool errctl_on(const NsdContext* c, int adaptive) { return c->phys.errctl && adaptive; }

template <typename T>
void enqueue_stage(NsdContext* c, int s, int adaptive) {
    const bool ec = errctl_on(c, adaptive);
    RowArgs<T> r = row_args<T>(c);
    r.mode = ROW_STAGE;
    r.track_umax = (s == 0) ? 1 : 0;
    r.skip_on_redo = (s > 0) ? 1 : 0;
    r.stagger_lo = c->stagger ? c->ncu : 0;  // second resident workgroup of each CU
    r.stagger_hi = c->stagger ? 2 * c->ncu : 0;
    ColArgs<T> a = col_args<T>(c);
    a.mode = COL_STAGE;
    a.stage = s;
    a.phys.adaptive = adaptive;
    // profiling: three events per stage bracket exactly the two kernels
    // (before k_row, between the kernels, after k_col), so neither bracket
    // holds the forcing stream's join or another stage's event
    const bool prof = c->prof != 0;
    const size_t e0 = (size_t)c->prof_steps * kEvPerStep + 3 * s;
    if (prof) (void)hipEventRecord(c->evs[e0], c->stream);
    launch_row_t<T>(c, r, c->G.ny / 2, ec);
    if (prof) (void)hipEventRecord(c->evs[e0 + 1], c->stream);
    launch_col_t<T>(c, a, c->G.nkx_keep, ec);
    if (prof) (void)hipEventRecord(c->evs[e0 + 2], c->stream);
}

// error controller: the judge of the previous candidate (start of a step)
void enqueue_judge(NsdContext* c) {
    hipLaunchKernelGGL(k_judge, dim3(c->nb), dim3(256), 0, c->stream, c->sc,
                       (const double*)c->ec_part, 2 * (size_t)c->G.nkx_keep, c->phys);
}

// error controller, after the last step of a call: judge the last candidate
// and, if it is rejected, restore the state before it (stage-0 column pass in
// finalize mode), so that the state between calls is always an accepted one
template <typename T>
void enqueue_finalize(NsdContext* c) {
    enqueue_judge(c);
    ColArgs<T> a = col_args<T>(c);
    a.mode = COL_STAGE;
    a.stage = 0;
    a.finalize = 1;
    a.phys.adaptive = 1;
    launch_col_t<T>(c, a, c->G.nkx_keep, true);
}

template <typename T>
bool force_active(const NsdContext* c) {
    return c->phys.force_on && c->fA != 0;
}

template <typename T>
void enqueue_force(NsdContext* c, hipStream_t st) {
    if (!force_active<T>(c)) return;
    ForceArgs f;
    memset(&f, 0, sizeof(f));
    f.geo = c->G;
    f.A = c->fA;
    f.modes = c->fmodes;
    f.partner = c->fpartner;
    f.seed = c->P.force_seed ? c->P.force_seed : 0x123456789abcdefull;
    f.colored = c->P.force_colored;
    f.tau = c->P.force_tau;
    f.energy_target = c->P.force_energy_target;
    f.rate = c->P.force_rate;
    f.ou = c->fou;
    f.work = c->fwork;
    f.fbx = c->fbx;
    f.fby = c->fby;
    f.box_T = (sizeof(T) == 4) ? 1 : 0;
    f.fbox = c->fbox;
    f.sc = c->sc;
    f.m_seed = c->m_seed;
    f.skip_on_redo = c->phys.errctl ? 1 : 0;
    hipLaunchKernelGGL(k_force, dim3(1, c->nb), dim3(256), 0, st, f);
}

// learned subgrid tendency of the state at the start of the step, into the
// compact source the stage kernels add to the nonlinear term: inverse
// transform to physical space, network, forward transform (all on the
// solver's stream and private buffers, so the step stays capturable)
template <typename T>
void enqueue_closure(NsdContext* c) {
    if constexpr (std::is_same<T, float>::value) {
        const int nb = c->nb;
        ColArgs<T> a = col_args<T>(c);
        a.mode = COL_INV;
        a.nbat = nb;
        a.in_ref = 0;
        a.in0 = (const cx<T>*)c->state;
        a.nout = 1;
        a.ops[0] = OP_ID0;
        a.mixed[0] = (cx<T>*)c->cl_m;
        launch_col_t<T>(c, a, c->G.nkx_keep);
        RowArgs<T> r = row_args<T>(c);
        r.mode = ROW_C2R;
        r.nbat = nb;
        r.br = R_of(c);
        r.nfields = 1;
        r.mixed[0] = (const cx<T>*)c->cl_m;
        // the physical vorticity goes straight into the network's bf16 input
        r.out_bf16 = closure_input(c->closure, &r.out_bf16_scale);
        launch_row_t<T>(c, r, c->G.ny);
        (void)closure_eval_staged(c->closure, c->cl_y, nb, (void*)c->stream);
        RowArgs<T> r2 = row_args<T>(c);
        r2.mode = ROW_R2C;
        r2.nbat = nb;
        r2.br = R_of(c);
        r2.in_t = c->cl_y;
        r2.nl = (cx<T>*)c->cl_m;
        launch_row_t<T>(c, r2, c->G.ny / 2);
        ColArgs<T> f = col_args<T>(c);
        f.mode = COL_FWD;
        f.nbat = nb;
        f.nl = (const cx<T>*)c->cl_m;
        f.out_c = (cx<T>*)c->cl_src;
        f.fwd_scale = 1.0;
        launch_col_t<T>(c, f, c->G.nkx);
    }
}

// one RK4 step: stage 0 computes dt (adaptive) before the forcing draw,
// which needs dt for the colored-noise decay and is consumed on stage 3.
// The draw (one workgroup) runs on the side stream beside stage 1, which
// neither reads nor writes what it touches (its box, OU state, work area
// and the forcing counter).
template <typename T>
void enqueue_step(NsdContext* c, int adaptive) {
    if (errctl_on(c, adaptive)) enqueue_judge(c);
    if (c->closure) enqueue_closure<T>(c);
    const bool side = force_active<T>(c);
    for (int s = 0; s < 4; ++s) {
        if (s == 2 && side) (void)hipStreamWaitEvent(c->stream, c->fev[1], 0);
        enqueue_stage<T>(c, s, adaptive);
        if (s == 0 && side) {
            (void)hipEventRecord(c->fev[0], c->stream);
            (void)hipStreamWaitEvent(c->fstream, c->fev[0], 0);
            enqueue_force<T>(c, c->fstream);
            (void)hipEventRecord(c->fev[1], c->fstream);
        }
    }
}

template <typename T>
int op_advance(NsdContext* c, int nsteps, int adaptive) {
    const size_t last = (size_t)nsteps * kEvPerStep;
    if (c->prof) {
        const size_t need = last + 2;
        while (c->evs.size() < need) {
            hipEvent_t e;
            CK(hipEventCreate(&e));
            c->evs.push_back(e);
        }
        c->prof_steps = 0;
        CK(hipEventRecord(c->evs[last], c->stream));
    }
    for (int n = 0; n < nsteps; ++n) {
        enqueue_step<T>(c, adaptive);
        if (c->prof) c->prof_steps++;
    }
    if (errctl_on(c, adaptive) && nsteps > 0) enqueue_finalize<T>(c);
    if (c->prof) {
        CK(hipEventRecord(c->evs[last + 1], c->stream));
        CK(hipStreamSynchronize(c->stream));
        double row = 0, col = 0;
        for (int n = 0; n < nsteps; ++n) {
            for (int s = 0; s < 4; ++s) {
                const size_t e0 = (size_t)n * kEvPerStep + 3 * s;
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, c->evs[e0], c->evs[e0 + 1]);
                row += ms;
                (void)hipEventElapsedTime(&ms, c->evs[e0 + 1], c->evs[e0 + 2]);
                col += ms;
            }
        }
        float tot = 0.f;
        (void)hipEventElapsedTime(&tot, c->evs[last], c->evs[last + 1]);
        c->prof_row = row / (4.0 * nsteps);
        c->prof_col = col / (4.0 * nsteps);
        c->prof_step = tot / nsteps;
        c->prof_steps = nsteps;
    }
    return lastk(c);
}

template <typename T>
int op_advance_graph(NsdContext* c, int nsteps, int adaptive) {
    if (!c->gexec || c->g_adaptive != adaptive || c->g_gen != c->gen) {
        if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
        if (c->graph) (void)hipGraphDestroy(c->graph);
        c->gexec = nullptr;
        c->graph = nullptr;
        const int prof = c->prof;
        c->prof = 0;
        CK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        enqueue_step<T>(c, adaptive);
        hipError_t e = hipStreamEndCapture(c->stream, &c->graph);
        c->prof = prof;
        if (e != hipSuccess) {
            c->err = std::string("graph capture: ") + hipGetErrorString(e);
            return ST_MEM;
        }
        CK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
        c->g_adaptive = adaptive;
        c->g_gen = c->gen;
    }
    for (int n = 0; n < nsteps; ++n) CK(hipGraphLaunch(c->gexec, c->stream));
    if (errctl_on(c, adaptive) && nsteps > 0) {
        enqueue_finalize<T>(c);
        return lastk(c);
    }
    return ST_OK;
}

template <typename T>
int op_upload(NsdContext* c, const double* omega_hat, int member = 0) {
    if (ensure_staging(c)) return ST_MEM;
    if (up(c, c->stH[0], omega_hat, 2 * sizeof(double) * H_of(c))) return ST_MEM;
    hipLaunchKernelGGL(k_gather_state<T>, dim3(grid_for(C_of(c))), dim3(256), 0, c->stream, c->G,
                       c->stH[0], (cx<T>*)c->state + (size_t)member * C_of(c));
    if (op_prime<T>(c)) return ST_MEM;
    return sync(c);
}

template <typename T>
int op_download(NsdContext* c, double* omega_hat, void* src) {
    if (ensure_staging(c)) return ST_MEM;
    hipLaunchKernelGGL(k_scatter_state<T>, dim3(grid_for(H_of(c))), dim3(256), 0, c->stream,
                       c->G, (const cx<T>*)src, c->stH[0]);
    if (lastk(c) || down(c, omega_hat, c->stH[0], 2 * sizeof(double) * H_of(c))) return ST_MEM;
    return sync(c);
}

template <typename T>
int op_sfd_upload(NsdContext* c, const double* sfd) {
    if (ensure_staging(c)) return ST_MEM;
    if (up(c, c->stH[0], sfd, 2 * sizeof(double) * H_of(c))) return ST_MEM;
    hipLaunchKernelGGL(k_gather_state<T>, dim3(grid_for(C_of(c))), dim3(256), 0, c->stream, c->G,
                       c->stH[0], (cx<T>*)c->sfd);
    return sync(c);
}

// E and Z of every member into sc[m].red[0..1], summed in a fixed order (the
// same bits on every call).  The workgroup partials go to the J buffer, which
// holds nothing between steps: 2 doubles per workgroup of 256 kept modes, and
// the buffer has room for at least one per kept mode.
template <typename T>
void enqueue_state_energy(NsdContext* c) {
    const dim3 grid(grid_for(C_of(c)), c->nb);
    hipLaunchKernelGGL(k_state_energy_part<T>, grid, dim3(256), 0, c->stream, c->G,
                       (const cx<T>*)c->state, (double*)c->nl);
    hipLaunchKernelGGL(k_sum_parts, dim3(c->nb), dim3(256), 0, c->stream, (const double*)c->nl,
                       (int)grid.x, c->sc);
}

template <typename T>
int op_random_phase(NsdContext* c, unsigned long long seed, const unsigned long long* seeds,
                    double k_peak, double energy) {
    const dim3 grid(grid_for(C_of(c)), c->nb);
    hipLaunchKernelGGL(k_random_phase<T>, grid, dim3(256), 0, c->stream, c->G, seed, seeds,
                       k_peak, (cx<T>*)c->state);
    // summed in a fixed order: one seed gives the same field bits on every call
    enqueue_state_energy<T>(c);
    hipLaunchKernelGGL(k_scale_state<T>, grid, dim3(256), 0, c->stream, C_of(c),
                       (cx<T>*)c->state, (const DevScalars*)c->sc, energy);
    if (op_prime<T>(c)) return ST_MEM;
    return sync(c);
}

// inverse transforms of up to 4 derived fields of in0 (and in1), from the
// FP64 reference-layout staging buffers, into FP64 physical outputs
template <typename T>
int op_inverse(NsdContext* c, const double* in0, const double* in1, int nout, const int* ops,
               double* const* outs) {
    ColArgs<T> a = col_args<T>(c);
    a.mode = COL_INV;
    a.in_ref = 1;
    a.in0_ref = in0;
    a.in1_ref = in1;
    a.nout = nout;
    for (int i = 0; i < nout; ++i) a.ops[i] = ops[i];
    launch_col_t<T>(c, a, c->G.nkx_keep);
    RowArgs<T> r = row_args<T>(c);
    r.mode = ROW_C2R;
    r.nfields = nout;
    for (int i = 0; i < nout; ++i) r.out_ref[i] = outs[i];
    launch_row_t<T>(c, r, c->G.ny);
    return lastk(c);
}

template <typename T>
int op_forward(NsdContext* c, const double* in_phys, double* out_ref, double scale) {
    RowArgs<T> r = row_args<T>(c);
    r.mode = ROW_R2C;
    r.in_ref = in_phys;
    launch_row_t<T>(c, r, c->G.ny / 2);
    ColArgs<T> a = col_args<T>(c);
    a.mode = COL_FWD;
    a.out_ref = out_ref;
    a.fwd_scale = scale;
    launch_col_t<T>(c, a, c->G.nkx);
    return lastk(c);
}

template <typename T>
int op_nonlinear(NsdContext* c, const double* w_ref, const double* psi_ref, double* out_ref) {
    ColArgs<T> a = col_args<T>(c);
    a.mode = COL_INV;
    a.in_ref = 1;
    a.in0_ref = psi_ref;
    a.in1_ref = w_ref;
    a.nout = 4;
    const int ops[4] = {OP_DX0, OP_DY0, OP_DX1, OP_DY1};  // psi_x, psi_y, omega_x, omega_y
    for (int i = 0; i < 4; ++i) a.ops[i] = ops[i];
    launch_col_t<T>(c, a, c->G.nkx_keep);
    RowArgs<T> r = row_args<T>(c);
    r.mode = ROW_STAGE;
    r.track_umax = 0;
    r.pairs = 0;  // the transforms above wrote the per-field layout
    launch_row_t<T>(c, r, c->G.ny / 2);
    ColArgs<T> f = col_args<T>(c);
    f.mode = COL_FWD;
    f.jblocks = jblocks_ok(c) ? 1 : 0;  // J of a pair-layout row stage
    f.out_ref = out_ref;
    f.fwd_scale = -1.0;  // spectral_core.c:108-111
    launch_col_t<T>(c, f, c->G.nkx);
    return lastk(c);
}

template <typename T>
int op_state_derived(NsdContext* c, double* psi_hat, double* u_hat, double* v_hat, double* u,
                     double* v, double* omega) {
    if (ensure_staging(c)) return ST_MEM;
    hipLaunchKernelGGL(k_scatter_state<T>, dim3(grid_for(H_of(c))), dim3(256), 0, c->stream,
                       c->G, (const cx<T>*)c->state, c->stH[0]);
    hipLaunchKernelGGL(k_velocity_ref, dim3(grid_for(H_of(c))), dim3(256), 0, c->stream, c->G,
                       c->stH[0], c->stH[1], c->stH[2], c->stH[3]);
    const int ops[3] = {OP_U0, OP_V0, OP_ID0};
    double* outs[3] = {c->stR[0], c->stR[1], c->stR[2]};
    if (op_inverse<T>(c, c->stH[0], nullptr, 3, ops, outs)) return ST_MEM;
    // the inverse transforms reused the mixed buffers: restore the invariant
    // that they hold the derivative fields of the state
    if (op_prime<T>(c)) return ST_MEM;
    const size_t hb = 2 * sizeof(double) * H_of(c), rb = sizeof(double) * R_of(c);
    if (down(c, psi_hat, c->stH[1], hb) || down(c, u_hat, c->stH[2], hb) ||
        down(c, v_hat, c->stH[3], hb) || down(c, u, c->stR[0], rb) || down(c, v, c->stR[1], rb) ||
        down(c, omega, c->stR[2], rb))
        return ST_MEM;
    return sync(c);
}

// E and Z of every member: out[2 m], out[2 m + 1]
template <typename T>
int op_state_energy(NsdContext* c, double* out) {
    enqueue_state_energy<T>(c);
    std::vector<DevScalars> h(c->nb);
    if (lastk(c) || down(c, h.data(), c->sc, sizeof(DevScalars) * c->nb) || sync(c)) return ST_MEM;
    for (int m = 0; m < c->nb; ++m) {
        out[2 * m] = h[m].red[0];
        out[2 * m + 1] = h[m].red[1];
    }
    return ST_OK;
}

// compat calls reuse the mixed buffers; after them the invariant "mixed
// holds the derivative fields of the state" is restored by a prime pass
template <typename T>
int restore_mixed(NsdContext* c) {
    return op_prime<T>(c);
}

// ----------------------------------------------------------------- subgrid samples

// nonlinear term (spectral_core.c:46-112) of member m's state into a device
// FP64 reference-layout buffer; compute_rows = 0 reuses the J fields of a
// previous call (the row pass covers every member of a batch at once)
template <typename T>
int op_nl_member(NsdContext* c, int m, double* out_ref, int compute_rows) {
    if (compute_rows) {
        RowArgs<T> r = row_args<T>(c);
        r.mode = ROW_STAGE;
        r.track_umax = 0;
        launch_row_t<T>(c, r, c->G.ny / 2);
    }
    ColArgs<T> f = col_args<T>(c);
    f.mode = COL_FWD;
    f.nl = (const cx<T>*)c->nl + (size_t)m * M_of(c);
    f.jblocks = jblocks_ok(c) ? 1 : 0;  // J of a pair-layout row stage
    f.out_ref = out_ref;
    f.fwd_scale = -1.0;  // spectral_core.c:108-111
    launch_col_t<T>(c, f, c->G.nkx);
    return lastk(c);
}

template <typename T>
int op_scatter_member(NsdContext* c, int m, double* out_ref) {
    hipLaunchKernelGGL(k_scatter_state<T>, dim3(grid_for(H_of(c))), dim3(256), 0, c->stream, c->G,
                       (const cx<T>*)c->state + (size_t)m * C_of(c), out_ref);
    return lastk(c);
}

template <typename T>
int op_gather_prime(NsdContext* c, const double* in_ref) {
    hipLaunchKernelGGL(k_gather_state<T>, dim3(grid_for(C_of(c))), dim3(256), 0, c->stream, c->G,
                       in_ref, (cx<T>*)c->state);
    return op_prime<T>(c);
}

// ----------------------------------------------------------------- autotuner

// Tuned launch variants per (kernel, transform length, precision, batch
// members): the best launch shape depends on how many workgroups the batch
// gives the device.  Timed once per process, optionally persisted in the text
// file named by NSD_TUNE_FILE (lines "row|col L precision members variant";
// lines of another form are skipped).
using TuneKey = std::tuple<int, int, int, int>;
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tune;
bool g_tune_loaded = false;

void tune_load() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char* f = getenv("NSD_TUNE_FILE");
    if (!f) return;
    FILE* fp = fopen(f, "r");
    if (!fp) return;
    char line[128], kind[8];
    int L, prec, nb, v;
    while (fgets(line, sizeof(line), fp))
        if (sscanf(line, "%7s %d %d %d %d", kind, &L, &prec, &nb, &v) == 5)
            g_tune[TuneKey(strcmp(kind, "col") == 0 ? 1 : 0, L, prec, nb)] = v;
    fclose(fp);
}

void tune_store(int kind, int L, int prec, int nb, int v) {
    g_tune[TuneKey(kind, L, prec, nb)] = v;
    const char* f = getenv("NSD_TUNE_FILE");
    if (!f) return;
    FILE* fp = fopen(f, "a");
    if (!fp) return;
    fprintf(fp, "%s %d %d %d %d\n", kind ? "col" : "row", L, prec, nb, v);
    fclose(fp);
}

// zero every field and scalar the step reads, so that timing runs leave a
// well-defined (all-zero) state behind
int zero_fields(NsdContext* c) {
    const size_t C = C_of(c) * c->esz * c->nb, M = M_of(c) * c->esz * c->nb;
    for (void* b : {c->state, c->acc, c->stagebuf, c->sfd})
        CK(hipMemsetAsync(b, 0, C, c->stream));
    for (void* b : {c->mixed[0], c->mixed[1], c->mixed[2], c->mixed[3], c->nl})
        CK(hipMemsetAsync(b, 0, M, c->stream));
    CK(hipMemsetAsync(c->sc, 0, sizeof(DevScalars) * c->nb, c->stream));
    return sync(c);
}

template <typename T>
int op_autotune(NsdContext* c, int reps, int force) {
    if (zero_fields(c)) return ST_MEM;
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int prec = c->prec;
    for (int kind = 0; kind < 2; ++kind) {
        const int L = kind == 0 ? c->G.nx : c->G.ny;
        int best = 0;
        bool cached = false;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            tune_load();
            auto it = g_tune.find(TuneKey(kind, L, prec, c->nb));
            if (it != g_tune.end() && !force) {
                best = it->second;
                cached = true;
            }
        }
        if (!cached) {
            float best_ms = 1e30f;
            for (int v = 0; v < kStageVariants; ++v) {
                if (!stage_variant_valid<T>(L, v, c->maxlds)) continue;
                auto launch = [&]() {
                    if (kind == 0) {
                        RowArgs<T> r = row_args<T>(c);
                        r.mode = ROW_STAGE;
                        r.track_umax = 0;
                        launch_row_stage<T>(r, L, c->G.ny / 2, c->nb, v, c->stream);
                    } else {
                        ColArgs<T> a = col_args<T>(c);
                        a.mode = COL_STAGE;
                        a.stage = 1;
                        a.phys.adaptive = 0;
                        launch_col_fused<T>(a, L, c->G.nkx_keep, c->nb, v, c->stream);
                    }
                };
                launch();  // warm-up
                CK(hipEventRecord(e0, c->stream));
                for (int r = 0; r < reps; ++r) launch();
                CK(hipEventRecord(e1, c->stream));
                CK(hipEventSynchronize(e1));
                if (lastk(c)) return ST_MEM;
                float ms = 0.f;
                CK(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best_ms) {
                    best_ms = ms;
                    best = v;
                }
            }
            std::lock_guard<std::mutex> lk(g_tune_mu);
            tune_store(kind, L, prec, c->nb, best);
        }
        if (kind == 0) c->vrow = best;
        else c->vcol = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->gen++;
    return zero_fields(c);
}

// ----------------------------------------------------------------- batches

int batch_refused(NsdContext* c) {
    c->err = "compat operations need a single-member context";
    return ST_PARAM;
}

}  // namespace

// One graph for a whole ensemble: the step of every batch context, each on
// its own stream, forked from and joined back into the first context's
// stream during capture, so that one hipGraphLaunch advances all of them.
struct NsdMulti {
    std::vector<NsdContext*> cs;
    std::vector<unsigned long long> gens;
    std::vector<hipEvent_t> ev;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int adaptive = -1;
    std::string err;
};

namespace {

int multi_capture(NsdMulti* mh, int adaptive) {
    NsdContext* c0 = mh->cs[0];
    if (mh->gexec) (void)hipGraphExecDestroy(mh->gexec);
    if (mh->graph) (void)hipGraphDestroy(mh->graph);
    mh->gexec = nullptr;
    mh->graph = nullptr;
    const size_t n = mh->cs.size();
    hipError_t e = hipStreamBeginCapture(c0->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) e = hipEventRecord(mh->ev[0], c0->stream);
    for (size_t i = 1; i < n && e == hipSuccess; ++i)
        e = hipStreamWaitEvent(mh->cs[i]->stream, mh->ev[0], 0);
    for (size_t i = 0; i < n && e == hipSuccess; ++i) {
        NsdContext* c = mh->cs[i];
        const int prof = c->prof;
        c->prof = 0;
        if (c->prec == NSD_FP32) enqueue_step<float>(c, adaptive);
        else enqueue_step<double>(c, adaptive);
        c->prof = prof;
    }
    for (size_t i = 1; i < n && e == hipSuccess; ++i) {
        e = hipEventRecord(mh->ev[i], mh->cs[i]->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(c0->stream, mh->ev[i], 0);
    }
    hipError_t e2 = hipStreamEndCapture(c0->stream, &mh->graph);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&mh->gexec, mh->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        mh->err = std::string("ensemble graph capture: ") + hipGetErrorString(e);
        return ST_MEM;
    }
    mh->adaptive = adaptive;
    for (size_t i = 0; i < n; ++i) mh->gens[i] = mh->cs[i]->gen;
    return ST_OK;
}

}  // namespace

// ----------------------------------------------------------------- extern C

extern "C" {

NsdMulti* nsd_multi_create(NsdContext* const* cs, int n) {
    if (!cs || n < 1) return nullptr;
    NsdMulti* mh = new NsdMulti();
    for (int i = 0; i < n; ++i) {
        if (!cs[i] || cs[i]->prec != cs[0]->prec) {
            delete mh;
            return nullptr;
        }
        mh->cs.push_back(cs[i]);
    }
    mh->gens.assign(n, ~0ull);
    mh->ev.assign(n, nullptr);
    for (int i = 0; i < n; ++i)
        if (hipEventCreateWithFlags(&mh->ev[i], hipEventDisableTiming) != hipSuccess) {
            nsd_multi_destroy(mh);
            return nullptr;
        }
    return mh;
}

void nsd_multi_destroy(NsdMulti* mh) {
    if (!mh) return;
    for (NsdContext* c : mh->cs) (void)hipStreamSynchronize(c->stream);
    if (mh->gexec) (void)hipGraphExecDestroy(mh->gexec);
    if (mh->graph) (void)hipGraphDestroy(mh->graph);
    for (hipEvent_t e : mh->ev)
        if (e) (void)hipEventDestroy(e);
    delete mh;
}

const char* nsd_multi_error(const NsdMulti* mh) { return mh ? mh->err.c_str() : "null handle"; }

int nsd_multi_advance(NsdMulti* mh, int nsteps, int adaptive) {
    if (!mh || nsteps < 0) return ST_PARAM;
    adaptive = adaptive ? 1 : 0;
    for (NsdContext* c : mh->cs)
        if (errctl_on(c, adaptive)) {
            // its finalize pass would have to follow the whole graph on the
            // context's own stream
            mh->err = "the error controller runs on single contexts, not in multi-size graphs";
            return ST_PARAM;
        }
    bool stale = !mh->gexec || mh->adaptive != adaptive;
    for (size_t i = 0; i < mh->cs.size(); ++i) stale = stale || mh->gens[i] != mh->cs[i]->gen;
    if (stale && multi_capture(mh, adaptive)) return ST_MEM;
    for (int s = 0; s < nsteps; ++s)
        if (hipGraphLaunch(mh->gexec, mh->cs[0]->stream) != hipSuccess) {
            mh->err = "hipGraphLaunch failed";
            return ST_MEM;
        }
    return ST_OK;
}

int nsd_multi_synchronize(NsdMulti* mh) {
    if (!mh) return ST_PARAM;
    if (hipStreamSynchronize(mh->cs[0]->stream) != hipSuccess) {
        mh->err = "hipStreamSynchronize failed";
        return ST_MEM;
    }
    return ST_OK;
}

NsdContext* nsd_create(int nx, int ny, double Lx, double Ly, int precision, char* err,
                       size_t errlen) {
    return nsd_create_batch(nx, ny, Lx, Ly, precision, 1, err, errlen);
}

NsdContext* nsd_create_batch(int nx, int ny, double Lx, double Ly, int precision, int members,
                             char* err, size_t errlen) {
    auto fail = [&](const char* m) -> NsdContext* {
        if (err && errlen) snprintf(err, errlen, "%s", m);
        return nullptr;
    };
    if (members < 1 || members > 65535) return fail("members must be in [1, 65535]");
    auto pow2 = [](int n) { return n >= 16 && n <= 8192 && (n & (n - 1)) == 0; };
    if (!pow2(nx) || !pow2(ny)) return fail("nx and ny must be powers of two in [16, 8192]");
    if (!(Lx > 0.0) || !(Ly > 0.0)) return fail("Lx and Ly must be positive");
    if (precision != NSD_FP64 && precision != NSD_FP32) return fail("unknown precision");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no HIP device available");
    NsdContext* c = new NsdContext();
    c->nb = members;
    Geo& G = c->G;
    G.nx = nx;
    G.ny = ny;
    G.nkx = nx / 2 + 1;
    G.kxmax = (2 * G.nkx) / 3;  // spectral_initialization.c:46 (quirk Q1)
    if (G.kxmax > G.nkx - 1) G.kxmax = G.nkx - 1;
    G.kymax = (2 * ny) / 6;     // spectral_initialization.c:48
    G.nkx_keep = G.kxmax + 1;
    G.nky_keep = 2 * G.kymax + 1;
    G.Lx = Lx;
    G.Ly = Ly;
    c->prec = precision;
    c->esz = precision == NSD_FP32 ? 2 * sizeof(float) : 2 * sizeof(double);
    memset(&c->P, 0, sizeof(c->P));
    memset(&c->phys, 0, sizeof(c->phys));
    c->P.Re = 100.0;
    c->P.cfl_factor = 0.3;
    auto bad = [&](hipError_t e) {
        if (e != hipSuccess) {
            if (err && errlen) snprintf(err, errlen, "HIP: %s", hipGetErrorString(e));
            nsd_destroy(c);
            return true;
        }
        return false;
    };
    if (bad(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) ||
        bad(hipStreamCreateWithFlags(&c->fstream, hipStreamNonBlocking)) ||
        bad(hipEventCreateWithFlags(&c->fev[0], hipEventDisableTiming)) ||
        bad(hipEventCreateWithFlags(&c->fev[1], hipEventDisableTiming)))
        return nullptr;
    const size_t C = C_of(c) * c->esz * members, M = M_of(c) * c->esz * members;
    const size_t S = sizeof(DevScalars) * members;
    if (bad(hipMalloc(&c->state, C)) || bad(hipMalloc(&c->acc, C)) ||
        bad(hipMalloc(&c->stagebuf, C)) || bad(hipMalloc(&c->sfd, C)) ||
        bad(hipMalloc(&c->nl, M)) || bad(hipMalloc(&c->sc, S)))
        return nullptr;
    // psi_x, psi_y in one allocation and omega_x, omega_y in another, so that
    // each pair can also be held interleaved (pair layout)
    for (int i = 0; i < 4; i += 2) {
        if (bad(hipMalloc(&c->mixed[i], 2 * M))) return nullptr;
        c->mixed[i + 1] = (char*)c->mixed[i] + M;
    }
    if (bad(hipMemset(c->state, 0, C)) || bad(hipMemset(c->sfd, 0, C)) ||
        bad(hipMemset(c->sc, 0, S)))
        return nullptr;
    // twiddle tables
    for (int d = 0; d < 2; ++d) {
        const int L = d == 0 ? nx : ny;
        std::vector<double> tw(2 * (size_t)L);
        for (int m = 0; m < L; ++m) twiddle(m, L, &tw[2 * m], &tw[2 * m + 1]);
        void** dst = d == 0 ? &c->twx : &c->twy;
        if (precision == NSD_FP32) {
            std::vector<float> tf(tw.begin(), tw.end());
            if (bad(hipMalloc(dst, tf.size() * sizeof(float))) ||
                bad(hipMemcpy(*dst, tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice)))
                return nullptr;
        } else {
            if (bad(hipMalloc(dst, tw.size() * sizeof(double))) ||
                bad(hipMemcpy(*dst, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice)))
                return nullptr;
        }
    }
    int dev = 0, maxlds = 0;
    if (bad(hipGetDevice(&dev)) ||
        bad(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev)) ||
        bad(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, dev)))
        return nullptr;
    const size_t need = std::max(line_lds_bytes(nx, c->esz), line_lds_bytes(ny, c->esz));
    if (need > (size_t)maxlds) {
        if (err && errlen)
            snprintf(err, errlen, "transform length needs %zu B of LDS per workgroup, device has %d",
                     need, maxlds);
        nsd_destroy(c);
        return nullptr;
    }
    if (precision == NSD_FP32) set_lds_limits<float>(maxlds);
    else set_lds_limits<double>(maxlds);
    c->maxlds = maxlds;
    if (precision == NSD_FP32 && nx == 4096) c->stagger = dispatch_breadth_first(dev, c->ncu);
    nsd_set_params(c, &c->P);
    return c;
}

void nsd_destroy(NsdContext* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    for (auto e : c->evs) (void)hipEventDestroy(e);
    void* bufs[] = {c->state, c->acc, c->stagebuf, c->sfd, c->nl, c->twx, c->twy, c->sc,
                    c->mixed[0], c->mixed[2], c->stH[0], c->stH[1],
                    c->stH[2], c->stH[3], c->stR[0], c->stR[1], c->stR[2], c->stR[3], c->fmodes,
                    c->fpartner, c->fou, c->fwork, c->fbox, c->m_inv_re, c->m_seed, c->m_ic,
                    c->cl_m, c->cl_y, c->cl_src, c->ec_prev, c->ec_ymid, c->ec_part};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (c->fstream) (void)hipStreamSynchronize(c->fstream);
    for (auto e : c->fev)
        if (e) (void)hipEventDestroy(e);
    if (c->fstream) (void)hipStreamDestroy(c->fstream);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* nsd_error(const NsdContext* c) { return c ? c->err.c_str() : "null context"; }
int nsd_precision(const NsdContext* c) { return c ? c->prec : -1; }

int nsd_set_params(NsdContext* c, const NsdParams* p) {
    if (!c || !p) return ST_PARAM;
    if (!(p->Re > 0.0)) {
        c->err = "Re must be positive";
        return ST_PARAM;
    }
    if (p->scheme == NSD_SCHEME_IF && p->sfd_on) {
        c->err = "selective frequency damping requires the classic scheme";
        return ST_PARAM;
    }
    if (!(p->drag >= 0.0) || (p->hv_on && !(p->hv_coeff >= 0.0))) {
        c->err = "drag and the hyperviscosity coefficient must not be negative";
        return ST_PARAM;
    }
    if (p->errctl) {
        if (!(p->err_rtol >= 0.0) || !(p->err_atol >= 0.0) ||
            !(p->err_rtol > 0.0 || p->err_atol > 0.0)) {
            c->err = "error tolerances must be non-negative and not both zero";
            return ST_PARAM;
        }
        if (!(p->err_safety > 0.0 && p->err_safety <= 1.0)) {
            c->err = "error controller safety factor must be in (0, 1]";
            return ST_PARAM;
        }
        if (!(p->dt_min > 0.0) || !(p->dt_max >= p->dt_min)) {
            c->err = "error controller needs 0 < dt_min <= dt_max";
            return ST_PARAM;
        }
        if (p->sfd_on) {
            c->err = "the error controller does not keep the SFD state of a rejected step; "
                     "switch selective frequency damping off";
            return ST_PARAM;
        }
    }
    c->P = *p;
    Phys& f = c->phys;
    f.inv_Re = 1.0 / p->Re;
    f.Re = p->Re;
    f.drag = p->drag;
    f.scheme = p->scheme;
    f.hv_on = p->hv_on && p->hv_coeff > 0.0;
    f.hv_order = p->hv_order > 0 ? p->hv_order : 1;
    f.hv_coeff = p->hv_coeff;
    f.hv_kc = p->hv_kc;
    f.hv_invbw = p->hv_invbw;
    f.sfd_on = p->sfd_on && p->sfd_strength > 0.0;
    f.sfd_strength = p->sfd_strength;
    f.sfd_tau = p->sfd_tau;
    f.cfl = p->cfl_factor;
    f.dxmin = fmin(c->G.Lx / c->G.nx, c->G.Ly / c->G.ny);
    f.force_on = 0;
    c->gen++;
    {  // error controller: buffers on first use, fresh scalars when switched on
        const int was_on = f.errctl;
        f.errctl = p->errctl ? 1 : 0;
        f.err_rtol = p->err_rtol;
        f.err_atol = p->err_atol;
        f.err_safety = p->err_safety;
        f.dt_lo = p->errctl ? p->dt_min : 1e-8;
        f.dt_hi = p->errctl ? p->dt_max : 0.1;
        if (f.errctl) {
            const size_t C = C_of(c) * c->esz * c->nb;
            if (!c->ec_prev) CK(hipMalloc(&c->ec_prev, C));
            if (!c->ec_ymid) CK(hipMalloc(&c->ec_ymid, C));
            if (!c->ec_part)
                CK(hipMalloc((void**)&c->ec_part, 2 * sizeof(double) * c->G.nkx_keep * c->nb));
            if (!was_on) {
                hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(256), 0, c->stream, c->sc, c->nb, 4,
                                   0.0, 0.0, 0ull, f.dt_hi);
                if (lastk(c) || sync(c)) return ST_MEM;
            }
        }
    }
    // forcing: active modes in reference order (forcing.c:80-93)
    for (void** b : {(void**)&c->fmodes, (void**)&c->fpartner, (void**)&c->fou,
                     (void**)&c->fwork, &c->fbox}) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    c->fA = 0;
    if (p->force_on) {
        const Geo& G = c->G;
        const double kmin = fmax(p->force_kmin, 0.0), kmax = fmax(p->force_kmax, kmin);
        std::vector<int> modes, partner;
        int fbx = 1, fby = 0;
        for (int j = 0; j < G.ny; ++j)
            for (int i = 0; i < G.nkx; ++i) {
                const double kx = 2.0 * kPi * i / G.Lx;
                const int b = j <= G.ny / 2 ? j : j - G.ny;
                const double ky = 2.0 * kPi * b / G.Ly;
                const double km = std::sqrt(kx * kx + ky * ky);
                if (km >= kmin && km <= kmax) {
                    modes.push_back(j * G.nkx + i);
                    fbx = std::max(fbx, i + 1);
                    fby = std::max(fby, std::abs(b));
                }
            }
        const int A = (int)modes.size();
        if (A == 0) {
            c->err = "forcing shell contains no modes";
            return ST_PARAM;
        }
        partner.assign(A, -1);
        for (int r = 0; r < A; ++r) {
            const int j = modes[r] / G.nkx, i = modes[r] % G.nkx;
            if (i != 0 && 2 * i != G.nx) continue;
            if (j == 0 || 2 * j == G.ny) continue;
            const int want = (G.ny - j) * G.nkx + i;
            for (int q = 0; q < A; ++q)
                if (modes[q] == want) partner[r] = q;
        }
        c->fA = A;
        c->fbx = fbx;
        c->fby = fby;
        const size_t box = (size_t)fbx * (2 * fby + 1);
        CK(hipMalloc(&c->fmodes, sizeof(int) * A));
        CK(hipMalloc(&c->fpartner, sizeof(int) * A));
        const size_t nb = (size_t)c->nb;
        CK(hipMalloc(&c->fou, 2 * sizeof(double) * A * nb));
        CK(hipMalloc(&c->fwork, 4 * sizeof(double) * A * nb));
        CK(hipMalloc(&c->fbox, c->esz * box * nb));
        CK(hipMemcpy(c->fmodes, modes.data(), sizeof(int) * A, hipMemcpyHostToDevice));
        CK(hipMemcpy(c->fpartner, partner.data(), sizeof(int) * A, hipMemcpyHostToDevice));
        CK(hipMemset(c->fou, 0, 2 * sizeof(double) * A * nb));
        CK(hipMemset(c->fbox, 0, c->esz * box * nb));
        f.force_on = 1;
        f.fbx = fbx;
        f.fby = fby;
    }
    return ST_OK;
}

#define TYPED(fn, ...) (c->prec == NSD_FP32 ? fn<float>(__VA_ARGS__) : fn<double>(__VA_ARGS__))

int nsd_state_upload(NsdContext* c, const double* w) {
    if (!c || !w) return ST_PARAM;
    return TYPED(op_upload, c, w);
}

int nsd_state_download(NsdContext* c, double* w) {
    if (!c || !w) return ST_PARAM;
    return TYPED(op_download, c, w, c->state);
}

int nsd_sfd_upload(NsdContext* c, const double* s) {
    if (!c || !s) return ST_PARAM;
    return TYPED(op_sfd_upload, c, s);
}

int nsd_sfd_download(NsdContext* c, double* s) {
    if (!c || !s) return ST_PARAM;
    return TYPED(op_download, c, s, c->sfd);
}

int nsd_state_random_phase(NsdContext* c, unsigned long long seed, double k_peak, double energy) {
    if (!c || !(k_peak > 0.0) || !(energy >= 0.0)) return ST_PARAM;
    return TYPED(op_random_phase, c, seed, nullptr, k_peak, energy);
}

int nsd_set_time(NsdContext* c, double t, double dt) {
    if (!c) return ST_PARAM;
    // a new time also restarts the error controller (first step limited by
    // dt_max and the CFL rule alone)
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(256), 0, c->stream, c->sc, c->nb, 1 | 4, t,
                       dt, 0ull, c->phys.dt_hi);
    if (lastk(c)) return ST_MEM;
    return sync(c);
}

int nsd_get_time(NsdContext* c, double* t, double* dt, int* status) {
    if (!c) return ST_PARAM;
    DevScalars s;
    CK(hipMemcpyAsync(&s, c->sc, sizeof(s), hipMemcpyDeviceToHost, c->stream));
    if (sync(c)) return ST_MEM;
    if (t) *t = s.t;
    if (dt) *dt = s.dt;
    if (status) *status = s.status;
    return ST_OK;
}

int nsd_set_forcing_counter(NsdContext* c, unsigned long long calls) {
    if (!c) return ST_PARAM;
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(256), 0, c->stream, c->sc, c->nb, 2, 0.0, 0.0,
                       calls, 0.0);
    if (lastk(c)) return ST_MEM;
    return sync(c);
}

int nsd_get_forcing_counter(NsdContext* c, unsigned long long* calls) {
    if (!c || !calls) return ST_PARAM;
    CK(hipMemcpyAsync(calls, &c->sc->fcount, sizeof(*calls), hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

int nsd_forcing_ou(NsdContext* c, double* ou, int count, int to_device) {
    if (!c) return ST_PARAM;
    if (count != 2 * c->fA || (count > 0 && !ou)) {
        c->err = "forcing OU state: count must be twice the number of active modes";
        return ST_PARAM;
    }
    if (count == 0) return ST_OK;
    const size_t b = sizeof(double) * (size_t)count;
    if (to_device) CK(hipMemcpyAsync(c->fou, ou, b, hipMemcpyHostToDevice, c->stream));
    else CK(hipMemcpyAsync(ou, c->fou, b, hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

int nsd_forcing_modes(const NsdContext* c) { return c ? c->fA : 0; }

int nsd_errctl_stats(NsdContext* c, int member, unsigned long long* accepted,
                     unsigned long long* rejected, double* last_err, double* dt_next) {
    if (!c || member < 0 || member >= c->nb) return ST_PARAM;
    DevScalars s;
    CK(hipMemcpyAsync(&s, c->sc + member, sizeof(s), hipMemcpyDeviceToHost, c->stream));
    if (sync(c)) return ST_MEM;
    if (accepted) *accepted = s.n_acc;
    if (rejected) *rejected = s.n_rej;
    if (last_err) *last_err = s.err;
    if (dt_next) *dt_next = s.dt_err;
    return ST_OK;
}

int nsd_row_stagger(NsdContext* c, int set) {
    if (!c) return -1;
    if (set >= 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return -1;
        c->stagger = set && dispatch_breadth_first(dev, c->ncu);
        c->gen++;  // kernel arguments changed: re-capture the step graph
    }
    return c->stagger;
}

int nsd_advance(NsdContext* c, int nsteps, int adaptive) {
    if (!c || nsteps < 0) return ST_PARAM;
    if (nsteps == 0) return ST_OK;
    return TYPED(op_advance, c, nsteps, adaptive ? 1 : 0);
}

int nsd_advance_graph(NsdContext* c, int nsteps, int adaptive) {
    if (!c || nsteps < 0) return ST_PARAM;
    if (nsteps == 0) return ST_OK;
    return TYPED(op_advance_graph, c, nsteps, adaptive ? 1 : 0);
}

int nsd_state_derived(NsdContext* c, double* psi_hat, double* u_hat, double* v_hat, double* u,
                      double* v, double* omega) {
    if (!c) return ST_PARAM;
    return TYPED(op_state_derived, c, psi_hat, u_hat, v_hat, u, v, omega);
}

int nsd_state_energy(NsdContext* c, double out[2]) {
    if (!c || !out) return ST_PARAM;
    if (c->nb != 1) {
        c->err = "nsd_state_energy: batch context (use nsd_batch_energy)";
        return ST_PARAM;
    }
    return TYPED(op_state_energy, c, out);
}

int nsd_attach_closure(NsdContext* c, NsdClosure* net) {
    if (!c) return ST_PARAM;
    if (!net) {
        c->closure = nullptr;
        c->gen++;
        return ST_OK;
    }
    if (c->prec != NSD_FP32 || c->G.nx != c->G.ny || nsd_closure_size(net) != c->G.nx ||
        nsd_closure_max_batch(net) < c->nb) {
        c->err = "closure needs FP32 fields on a square grid of the network's size, and a "
                 "network batch of at least the context's members";
        return ST_PARAM;
    }
    const size_t R = R_of(c) * c->nb;
    if (!c->cl_m) CK(hipMalloc(&c->cl_m, M_of(c) * c->esz * c->nb));
    if (!c->cl_y) CK(hipMalloc(&c->cl_y, sizeof(float) * R));
    if (!c->cl_src) CK(hipMalloc(&c->cl_src, C_of(c) * c->esz * c->nb));
    c->closure = net;
    c->gen++;
    return ST_OK;
}

int nsd_subgrid_sample(NsdContext* fine, int member, NsdContext* coarse, float* x, float* t) {
    if (!fine || !coarse || !x || !t || member < 0 || member >= fine->nb) return ST_PARAM;
    NsdContext* c = coarse;
    if (c->nb != 1 || c->G.nx > fine->G.nx || c->G.ny > fine->G.ny || c->G.Lx != fine->G.Lx ||
        c->G.Ly != fine->G.Ly) {
        c->err = "subgrid sample: coarse context must be one field on a coarser grid of the same domain";
        return ST_PARAM;
    }
    if (ensure_staging(fine)) {
        c->err = fine->err;
        return ST_MEM;
    }
    if (ensure_staging(c)) return ST_MEM;
    {
        NsdContext* f = fine;
        int rc = f->prec == NSD_FP32 ? op_nl_member<float>(f, member, f->stH[1], member == 0)
                                     : op_nl_member<double>(f, member, f->stH[1], member == 0);
        if (!rc) rc = f->prec == NSD_FP32 ? op_scatter_member<float>(f, member, f->stH[0])
                                          : op_scatter_member<double>(f, member, f->stH[0]);
        if (rc) {
            c->err = f->err;
            return rc;
        }
        const size_t Hc = H_of(c);
        hipLaunchKernelGGL(k_filter_ref, dim3(grid_for(Hc)), dim3(256), 0, f->stream, f->G, c->G,
                           f->stH[0], c->stH[0]);
        hipLaunchKernelGGL(k_filter_ref, dim3(grid_for(Hc)), dim3(256), 0, f->stream, f->G, c->G,
                           f->stH[1], c->stH[2]);
        if (lastk(f) || sync(f)) {
            c->err = f->err;
            return ST_MEM;
        }
    }
    int rc = TYPED(op_gather_prime, c, c->stH[0]);
    if (!rc) rc = TYPED(op_nl_member, c, 0, c->stH[1], 1);
    if (rc) return rc;
    const size_t Hc = H_of(c), Rc = R_of(c);
    hipLaunchKernelGGL(k_sub_d, dim3(grid_for(2 * Hc)), dim3(256), 0, c->stream, 2 * Hc,
                       c->stH[2], c->stH[1], c->stH[3]);
    const int ops[1] = {OP_ID0};
    double* o0[1] = {c->stR[0]};
    double* o1[1] = {c->stR[1]};
    rc = TYPED(op_inverse, c, c->stH[0], nullptr, 1, ops, o0);
    if (!rc) rc = TYPED(op_inverse, c, c->stH[3], nullptr, 1, ops, o1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_d2f, dim3(grid_for(Rc)), dim3(256), 0, c->stream, Rc, c->stR[0], x);
    hipLaunchKernelGGL(k_d2f, dim3(grid_for(Rc)), dim3(256), 0, c->stream, Rc, c->stR[1], t);
    if (lastk(c)) return ST_MEM;
    // the inverse transforms above reused the coarse context's mixed buffers:
    // restore the invariant that they hold the derivative fields of its state
    // (the filtered fine field), so that a step taken on it next starts right
    rc = TYPED(restore_mixed, c);
    if (rc) return rc;
    return sync(c);
}

int nsd_subgrid_dataset(NsdClosure* net, NsdContext* fine, NsdContext* coarse, int snapshots,
                        int stride, int first) {
    if (!net || !fine || !coarse || snapshots < 1 || stride < 0 || first < 0) return ST_PARAM;
    for (int s = 0; s < snapshots; ++s) {
        if (stride > 0) {
            int rc = nsd_advance_graph(fine, stride, 1);
            if (rc) return rc;
        }
        for (int m = 0; m < fine->nb; ++m) {
            float *x, *t;
            if (nsd_closure_dataset_slot(net, first + s * fine->nb + m, &x, &t)) {
                coarse->err = "subgrid data set: slot out of range";
                return ST_PARAM;
            }
            int rc = nsd_subgrid_sample(fine, m, coarse, x, t);
            if (rc) return rc;
        }
    }
    return ST_OK;
}

int nsd_batch_members(const NsdContext* c) { return c ? c->nb : 0; }

int nsd_autotune(NsdContext* c, int reps, int force, int* row_variant, int* col_variant) {
    if (!c || reps < 1) return ST_PARAM;
    const int rc = TYPED(op_autotune, c, reps, force);
    if (row_variant) *row_variant = c->vrow;
    if (col_variant) *col_variant = c->vcol;
    return rc;
}

int nsd_set_variants(NsdContext* c, int row_variant, int col_variant) {
    if (!c) return ST_PARAM;
    const bool ok = c->prec == NSD_FP32
                        ? stage_variant_valid<float>(c->G.nx, row_variant, c->maxlds) &&
                              stage_variant_valid<float>(c->G.ny, col_variant, c->maxlds)
                        : stage_variant_valid<double>(c->G.nx, row_variant, c->maxlds) &&
                              stage_variant_valid<double>(c->G.ny, col_variant, c->maxlds);
    if (!ok) {
        c->err = "launch variant not available for this size";
        return ST_PARAM;
    }
    const bool pairs_before = pairs_ok(c);
    c->vrow = row_variant;
    c->vcol = col_variant;
    c->gen++;
    // the mixed fields' layout follows the variants: rewrite them from the
    // state when it changes (no length has variants of both layouts today)
    if (pairs_ok(c) != pairs_before) return TYPED(op_prime, c);
    return ST_OK;
}

int nsd_set_member_params(NsdContext* c, const double* Re, const unsigned long long* force_seed) {
    if (!c) return ST_PARAM;
    const size_t nb = (size_t)c->nb;
    if (Re) {
        std::vector<double> inv(nb);
        for (size_t m = 0; m < nb; ++m) {
            if (!(Re[m] > 0.0)) {
                c->err = "member Re must be positive";
                return ST_PARAM;
            }
            inv[m] = 1.0 / Re[m];
        }
        if (!c->m_inv_re) CK(hipMalloc(&c->m_inv_re, sizeof(double) * nb));
        CK(hipMemcpy(c->m_inv_re, inv.data(), sizeof(double) * nb, hipMemcpyHostToDevice));
    }
    if (force_seed) {
        std::vector<unsigned long long> sd(force_seed, force_seed + nb);
        for (auto& v : sd)
            if (v == 0) v = 0x123456789abcdefull;  // forcing.c:99
        if (!c->m_seed) CK(hipMalloc(&c->m_seed, sizeof(unsigned long long) * nb));
        CK(hipMemcpy(c->m_seed, sd.data(), sizeof(unsigned long long) * nb, hipMemcpyHostToDevice));
    }
    c->gen++;  // kernel arguments changed: re-capture the step graph
    return ST_OK;
}

int nsd_batch_random_phase(NsdContext* c, const unsigned long long* seeds, double k_peak,
                           double energy) {
    if (!c || !seeds || !(k_peak > 0.0) || !(energy >= 0.0)) return ST_PARAM;
    const size_t nb = (size_t)c->nb;
    if (!c->m_ic) CK(hipMalloc(&c->m_ic, sizeof(unsigned long long) * nb));
    CK(hipMemcpy(c->m_ic, seeds, sizeof(unsigned long long) * nb, hipMemcpyHostToDevice));
    return TYPED(op_random_phase, c, 0ull, c->m_ic, k_peak, energy);
}

int nsd_batch_energy(NsdContext* c, double* E, double* Z) {
    if (!c) return ST_PARAM;
    std::vector<double> out(2 * (size_t)c->nb);
    int rc = TYPED(op_state_energy, c, out.data());
    if (rc) return rc;
    for (int m = 0; m < c->nb; ++m) {
        if (E) E[m] = out[2 * m];
        if (Z) Z[m] = out[2 * m + 1];
    }
    return ST_OK;
}

int nsd_batch_scalars(NsdContext* c, double* t, double* dt, int* status) {
    if (!c) return ST_PARAM;
    std::vector<DevScalars> h(c->nb);
    if (down(c, h.data(), c->sc, sizeof(DevScalars) * c->nb) || sync(c)) return ST_MEM;
    for (int m = 0; m < c->nb; ++m) {
        if (t) t[m] = h[m].t;
        if (dt) dt[m] = h[m].dt;
        if (status) status[m] = h[m].status;
    }
    return ST_OK;
}

int nsd_member_upload(NsdContext* c, int member, const double* w) {
    if (!c || !w || member < 0 || member >= c->nb) return ST_PARAM;
    return TYPED(op_upload, c, w, member);
}

int nsd_member_download(NsdContext* c, int member, double* w) {
    if (!c || !w || member < 0 || member >= c->nb) return ST_PARAM;
    void* src = (char*)c->state + (size_t)member * C_of(c) * c->esz;
    return TYPED(op_download, c, w, src);
}

int nsd_synchronize(NsdContext* c) {
    if (!c) return ST_PARAM;
    return sync(c);
}

int nsd_profile(NsdContext* c, int on) {
    if (!c) return ST_PARAM;
    c->prof = on ? 1 : 0;
    return ST_OK;
}

int nsd_profile_read(NsdContext* c, double* row_ms, double* col_ms, double* step_ms, int* steps) {
    if (!c) return ST_PARAM;
    if (row_ms) *row_ms = c->prof_row;
    if (col_ms) *col_ms = c->prof_col;
    if (step_ms) *step_ms = c->prof_step;
    if (steps) *steps = c->prof_steps;
    return ST_OK;
}

int nsd_bandwidth_probe(NsdContext* c, size_t bytes, int reps, double* ms) {
    if (!c || bytes < 16 || reps < 1) return ST_PARAM;
    const size_t n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0, e1;
    CK(hipMalloc(&a, n * 16));
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        c->err = "bandwidth probe allocation failed";
        return ST_MEM;
    }
    CK(hipMemsetAsync(a, 0, n * 16, c->stream));
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const unsigned grid = 256 * 32;
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream, (const float4*)a, (float4*)b, n);
    CK(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream, (const float4*)a,
                           (float4*)b, n);
    CK(hipEventRecord(e1, c->stream));
    CK(hipEventSynchronize(e1));
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return lastk(c);
}

int nsd_grids(NsdContext* c, double* x, double* y, double* kx, double* ky, double* k2, int* mask) {
    if (!c) return ST_PARAM;
    if (ensure_staging(c)) return ST_MEM;
    const Geo& G = c->G;
    double* dx = c->stR[0];
    double* dy = dx + G.nx;
    double* dkx = dy + G.ny;
    double* dky = dkx + G.nkx;
    double* dk2 = c->stH[0];
    int* dmask = (int*)c->stH[1];
    CK(hipMemsetAsync(&c->sc->redmax[0], 0, sizeof(unsigned long long), c->stream));
    const unsigned g = std::max(grid_for(H_of(c)), grid_for((size_t)std::max(G.nx, G.ny)));
    hipLaunchKernelGGL(k_grids, dim3(g), dim3(256), 0, c->stream, G, dx, dy, dkx, dky, dk2, dmask,
                       &c->sc->redmax[0]);
    if (lastk(c) || down(c, x, dx, sizeof(double) * G.nx) || down(c, y, dy, sizeof(double) * G.ny) ||
        down(c, kx, dkx, sizeof(double) * G.nkx) || down(c, ky, dky, sizeof(double) * G.ny) ||
        down(c, k2, dk2, sizeof(double) * H_of(c)) || down(c, mask, dmask, sizeof(int) * H_of(c)))
        return ST_MEM;
    return sync(c);
}

int nsd_hv_weights(NsdContext* c, int enabled, double coefficient, int order, double cutoff_ratio,
                   double* weights, double* kmax_out, double* kc_out, double* invbw_out) {
    if (!c) return ST_PARAM;
    if (ensure_staging(c)) return ST_MEM;
    const Geo& G = c->G;
    CK(hipMemsetAsync(&c->sc->redmax[0], 0, sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_grids, dim3(grid_for(H_of(c))), dim3(256), 0, c->stream, G, nullptr,
                       nullptr, nullptr, nullptr, nullptr, nullptr, &c->sc->redmax[0]);
    unsigned long long kb = 0;
    if (lastk(c) || down(c, &kb, &c->sc->redmax[0], sizeof(kb)) || sync(c)) return ST_MEM;
    double kmax;
    memcpy(&kmax, &kb, sizeof(kmax));
    // hyperviscosity.c:68-79
    const double ratio = cutoff_ratio < 0.0 ? 0.0 : (cutoff_ratio > 1.0 ? 1.0 : cutoff_ratio);
    const double kc = ratio * kmax;
    const double bw = fmax(kmax - kc, 1e-12);
    const double invbw = bw > 0.0 ? 1.0 / bw : 0.0;
    if (kmax_out) *kmax_out = kmax;
    if (kc_out) *kc_out = kc;
    if (invbw_out) *invbw_out = invbw;
    if (!weights) return ST_OK;
    Phys P;
    memset(&P, 0, sizeof(P));
    P.hv_on = (enabled && coefficient > 0.0 && kmax > 0.0) ? 1 : 0;
    P.hv_coeff = coefficient;
    P.hv_order = order > 0 ? order : 1;
    P.hv_kc = kc;
    P.hv_invbw = invbw;
    hipLaunchKernelGGL(k_hv_weights, dim3(grid_for(H_of(c))), dim3(256), 0, c->stream, G, P,
                       c->stH[0]);
    if (lastk(c) || down(c, weights, c->stH[0], sizeof(double) * H_of(c))) return ST_MEM;
    return sync(c);
}

int nsd_physical_to_spectral(NsdContext* c, const double* field, double* field_hat) {
    if (!c || !field || !field_hat) return ST_PARAM;
    if (c->nb != 1) return batch_refused(c);
    if (ensure_staging(c) || up(c, c->stR[0], field, sizeof(double) * R_of(c))) return ST_MEM;
    if (TYPED(op_forward, c, c->stR[0], c->stH[0], 1.0)) return ST_MEM;
    if (down(c, field_hat, c->stH[0], 2 * sizeof(double) * H_of(c))) return ST_MEM;
    if (TYPED(restore_mixed, c)) return ST_MEM;
    return sync(c);
}

int nsd_spectral_to_physical(NsdContext* c, const double* field_hat, double* field) {
    if (!c || !field_hat || !field) return ST_PARAM;
    if (c->nb != 1) return batch_refused(c);
    if (ensure_staging(c) || up(c, c->stH[0], field_hat, 2 * sizeof(double) * H_of(c)))
        return ST_MEM;
    const int ops[1] = {OP_ID0};
    double* outs[1] = {c->stR[0]};
    if (TYPED(op_inverse, c, c->stH[0], nullptr, 1, ops, outs)) return ST_MEM;
    if (down(c, field, c->stR[0], sizeof(double) * R_of(c))) return ST_MEM;
    if (TYPED(restore_mixed, c)) return ST_MEM;
    return sync(c);
}

int nsd_dealias(NsdContext* c, double* f) {
    if (!c || !f) return ST_PARAM;
    if (ensure_staging(c) || up(c, c->stH[0], f, 2 * sizeof(double) * H_of(c))) return ST_MEM;
    hipLaunchKernelGGL(k_dealias_ref, dim3(grid_for(H_of(c))), dim3(256), 0, c->stream, c->G,
                       c->stH[0]);
    if (lastk(c) || down(c, f, c->stH[0], 2 * sizeof(double) * H_of(c))) return ST_MEM;
    return sync(c);
}

int nsd_velocity(NsdContext* c, const double* w, double* psi, double* u, double* v) {
    if (!c || !w) return ST_PARAM;
    const size_t hb = 2 * sizeof(double) * H_of(c);
    if (ensure_staging(c) || up(c, c->stH[0], w, hb)) return ST_MEM;
    hipLaunchKernelGGL(k_velocity_ref, dim3(grid_for(H_of(c))), dim3(256), 0, c->stream, c->G,
                       c->stH[0], c->stH[1], c->stH[2], c->stH[3]);
    if (lastk(c) || down(c, psi, c->stH[1], hb) || down(c, u, c->stH[2], hb) ||
        down(c, v, c->stH[3], hb))
        return ST_MEM;
    return sync(c);
}

int nsd_derivatives(NsdContext* c, const double* w, double* psi, double* uh, double* vh, double* u,
                    double* v, double* om) {
    if (!c || !w) return ST_PARAM;
    if (c->nb != 1) return batch_refused(c);
    const size_t hb = 2 * sizeof(double) * H_of(c), rb = sizeof(double) * R_of(c);
    if (ensure_staging(c) || up(c, c->stH[0], w, hb)) return ST_MEM;
    hipLaunchKernelGGL(k_velocity_ref, dim3(grid_for(H_of(c))), dim3(256), 0, c->stream, c->G,
                       c->stH[0], c->stH[1], c->stH[2], c->stH[3]);
    const int ops[3] = {OP_U0, OP_V0, OP_ID0};  // spectral_core.c:40-42
    double* outs[3] = {c->stR[0], c->stR[1], c->stR[2]};
    if (TYPED(op_inverse, c, c->stH[0], nullptr, 3, ops, outs)) return ST_MEM;
    if (down(c, psi, c->stH[1], hb) || down(c, uh, c->stH[2], hb) || down(c, vh, c->stH[3], hb) ||
        down(c, u, c->stR[0], rb) || down(c, v, c->stR[1], rb) || down(c, om, c->stR[2], rb))
        return ST_MEM;
    if (TYPED(restore_mixed, c)) return ST_MEM;
    return sync(c);
}

int nsd_nonlinear(NsdContext* c, const double* w, const double* psi, double* out) {
    if (!c || !w || !psi || !out) return ST_PARAM;
    if (c->nb != 1) return batch_refused(c);
    const size_t hb = 2 * sizeof(double) * H_of(c);
    if (ensure_staging(c) || up(c, c->stH[0], w, hb) || up(c, c->stH[1], psi, hb)) return ST_MEM;
    if (TYPED(op_nonlinear, c, c->stH[0], c->stH[1], c->stH[2])) return ST_MEM;
    if (down(c, out, c->stH[2], hb)) return ST_MEM;
    if (TYPED(restore_mixed, c)) return ST_MEM;
    return sync(c);
}

int nsd_hv_apply(NsdContext* c, double* rhs, const double* st, const double* sfd) {
    if (!c || !rhs || !st) return ST_PARAM;
    const size_t hb = 2 * sizeof(double) * H_of(c);
    if (ensure_staging(c) || up(c, c->stH[0], rhs, hb) || up(c, c->stH[1], st, hb) ||
        up(c, c->stH[2], sfd, hb))
        return ST_MEM;
    Phys P = c->phys;
    if (!sfd) P.sfd_on = 0;
    hipLaunchKernelGGL(k_hv_apply_ref, dim3(grid_for(H_of(c))), dim3(256), 0, c->stream, c->G, P,
                       c->stH[0], c->stH[1], c->stH[2]);
    if (lastk(c) || down(c, rhs, c->stH[0], hb)) return ST_MEM;
    return sync(c);
}

int nsd_hv_post_step(NsdContext* c, const double* st, double* sfd, double dt) {
    if (!c || !st || !sfd) return ST_PARAM;
    const size_t hb = 2 * sizeof(double) * H_of(c);
    if (!(c->phys.sfd_on && c->phys.sfd_tau > 0.0) || dt <= 0.0) return ST_OK;
    double alpha = dt / c->phys.sfd_tau;
    alpha = alpha > 1.0 ? 1.0 : (alpha < 0.0 ? 0.0 : alpha);
    if (ensure_staging(c) || up(c, c->stH[0], st, hb) || up(c, c->stH[1], sfd, hb)) return ST_MEM;
    hipLaunchKernelGGL(k_hv_post_ref, dim3(grid_for(H_of(c))), dim3(256), 0, c->stream, H_of(c),
                       alpha, c->stH[0], c->stH[1]);
    if (lastk(c) || down(c, sfd, c->stH[1], hb)) return ST_MEM;
    return sync(c);
}

int nsd_taylor_green(NsdContext* c, double* u, double* v, double* w, double* p) {
    if (!c) return ST_PARAM;
    const size_t rb = sizeof(double) * R_of(c);
    if (ensure_staging(c)) return ST_MEM;
    hipLaunchKernelGGL(k_taylor_green, dim3(grid_for(R_of(c))), dim3(256), 0, c->stream, c->G,
                       c->stR[0], c->stR[1], c->stR[2], c->stR[3]);
    if (lastk(c) || down(c, u, c->stR[0], rb) || down(c, v, c->stR[1], rb) ||
        down(c, w, c->stR[2], rb) || down(c, p, c->stR[3], rb))
        return ST_MEM;
    return sync(c);
}

int nsd_phys_reduce(NsdContext* c, const double* u, const double* v, const double* w,
                    double out[3]) {
    if (!c || !out) return ST_PARAM;
    const size_t rb = sizeof(double) * R_of(c);
    if (ensure_staging(c) || up(c, c->stR[0], u, rb) || up(c, c->stR[1], v, rb) ||
        up(c, c->stR[2], w, rb))
        return ST_MEM;
    CK(hipMemsetAsync(c->sc->red, 0, sizeof(c->sc->red) + sizeof(c->sc->redmax), c->stream));
    hipLaunchKernelGGL(k_phys_reduce, dim3(grid_for(R_of(c))), dim3(256), 0, c->stream, R_of(c),
                       u ? c->stR[0] : nullptr, v ? c->stR[1] : nullptr, w ? c->stR[2] : nullptr,
                       c->sc);
    DevScalars s;
    if (lastk(c) || down(c, &s, c->sc, sizeof(s)) || sync(c)) return ST_MEM;
    memcpy(&out[0], &s.redmax[0], sizeof(double));
    out[1] = s.red[0];
    out[2] = s.red[1];
    return ST_OK;
}

int nsd_spec_reduce(NsdContext* c, const double* w, double out[5]) {
    if (!c || !w || !out) return ST_PARAM;
    const size_t hb = 2 * sizeof(double) * H_of(c);
    if (ensure_staging(c) || up(c, c->stH[0], w, hb)) return ST_MEM;
    CK(hipMemsetAsync(c->sc->red, 0, sizeof(c->sc->red) + sizeof(c->sc->redmax), c->stream));
    hipLaunchKernelGGL(k_spec_reduce, dim3(grid_for(H_of(c))), dim3(256), 0, c->stream, c->G,
                       c->stH[0], c->sc);
    DevScalars s;
    if (lastk(c) || down(c, &s, c->sc, sizeof(s)) || sync(c)) return ST_MEM;
    out[0] = s.red[0];
    out[1] = s.red[1];
    memcpy(&out[2], &s.redmax[0], sizeof(double));
    out[3] = s.red[2];
    out[4] = s.red[3];
    return ST_OK;
}

int nsd_spectrum(NsdContext* c, const double* uh, const double* vh, double* ek, int* ck,
                 int max_k) {
    if (!c || !uh || !vh || !ek || !ck || max_k < 0) return ST_PARAM;
    const size_t hb = 2 * sizeof(double) * H_of(c);
    if (ensure_staging(c) || up(c, c->stH[0], uh, hb) || up(c, c->stH[1], vh, hb)) return ST_MEM;
    if (max_k == 0) return ST_OK;
    double* dek = c->stR[0];
    int* dck = (int*)c->stR[1];
    if ((size_t)max_k > R_of(c)) return ST_PARAM;
    CK(hipMemsetAsync(dek, 0, sizeof(double) * max_k, c->stream));
    CK(hipMemsetAsync(dck, 0, sizeof(int) * max_k, c->stream));
    hipLaunchKernelGGL(k_spectrum, dim3(grid_for(H_of(c))), dim3(256), 0, c->stream, c->G,
                       c->stH[0], c->stH[1], dek, dck, max_k);
    if (lastk(c) || down(c, ek, dek, sizeof(double) * max_k) ||
        down(c, ck, dck, sizeof(int) * max_k))
        return ST_MEM;
    return sync(c);
}

int nsd_forcing_generate(int H, const int* active_idx, int A, unsigned long long seed,
                         unsigned long long calls, int colored, double alpha, double beta,
                         double energy_target, double* col, double* out, char* err,
                         size_t errlen) {
    if (H <= 0 || A <= 0 || !active_idx || !out || (colored && !col)) {
        if (err && errlen) snprintf(err, errlen, "invalid forcing arguments");
        return ST_PARAM;
    }
    int* dm = nullptr;
    double *dc = nullptr, *dout = nullptr;
    const size_t hb = 2 * sizeof(double) * (size_t)H;
    hipError_t e = hipMalloc(&dm, sizeof(int) * A);
    if (e == hipSuccess) e = hipMalloc(&dout, hb);
    if (e == hipSuccess && colored) e = hipMalloc(&dc, hb);
    if (e == hipSuccess) e = hipMemcpy(dm, active_idx, sizeof(int) * A, hipMemcpyHostToDevice);
    if (e == hipSuccess && colored) e = hipMemcpy(dc, col, hb, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fill, dim3(grid_for(2 * (size_t)H)), dim3(256), 0, 0, dout,
                           2 * (size_t)H, 0.0);
        hipLaunchKernelGGL(k_force_ref, dim3(1), dim3(256), 0, 0, A, dm, seed, calls, colored,
                           alpha, beta, energy_target, dc, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, hb, hipMemcpyDeviceToHost);
    if (e == hipSuccess && colored) e = hipMemcpy(col, dc, hb, hipMemcpyDeviceToHost);
    if (dm) (void)hipFree(dm);
    if (dc) (void)hipFree(dc);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) {
        if (err && errlen) snprintf(err, errlen, "HIP: %s", hipGetErrorString(e));
        return ST_MEM;
    }
    return ST_OK;
}

}  // extern "C"
