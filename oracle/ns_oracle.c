/*
 * ns_oracle.c -- FP64 CPU oracle (test infrastructure only; see ns_oracle.h).
 *
 * A plain restatement of the reference's spectral solver.  Citations are
 * path:line in the reference repository (DiogoRibeiro7/navier-stokes-solvers).
 * The reference uses FFTW3 (absent here); the transforms below are a
 * project-written radix-2 complex FFT composed into the same unnormalised
 * 2-D r2c / c2r conventions as fftw_plan_dft_{r2c,c2r}_2d(ny, nx, ...):
 *   r2c: out[j*nkx+i] = sum_{y,x} in[y*nx+x] e^{-2 pi i (i x/nx + j y/ny)}
 *   c2r: out[y*nx+x]  = sum over the Hermitian-extended spectrum, with the
 *        imaginary parts of the kx = 0 and kx = nx/2 mixed values ignored
 *        (FFTW's c2r convention for the last dimension).
 */
#include "ns_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define OC_MIN_DT 1e-8 /* include/common_types.h:32 */
#define OC_MAX_DT 0.1  /* include/common_types.h:33 */

/* ------------------------------------------------------------------ FFT */

typedef struct {
    int n;
    double *twr, *twi; /* exp(-2 pi i k / n), k < n/2 */
    int *rev;
} Fft1;

struct OcFFT {
    Fft1 fx, fy;
};

/* exp(-2 pi i m / n) with exact quadrant reduction of the integer ratio. */
static void oc_twiddle(long m, long n, double *c, double *s) {
    m %= n;
    if (m < 0) m += n;
    long q = (4 * m) / n;
    long r = 4 * m - q * n; /* angle = (pi/2) (q + r/n) */
    double c0, s0;
    if (2 * r <= n) {
        double a = (M_PI / 2.0) * ((double)r / (double)n);
        c0 = cos(a);
        s0 = sin(a);
    } else {
        double a = (M_PI / 2.0) * ((double)(n - r) / (double)n);
        c0 = sin(a);
        s0 = cos(a);
    }
    double cc, ss;
    switch (q) {
    case 0: cc = c0; ss = s0; break;
    case 1: cc = -s0; ss = c0; break;
    case 2: cc = -c0; ss = -s0; break;
    default: cc = s0; ss = -c0; break;
    }
    *c = cc;
    *s = -ss; /* exp(-i angle) */
}

static int fft1_init(Fft1 *p, int n) {
    p->n = n;
    int half = n / 2 > 0 ? n / 2 : 1;
    p->twr = malloc(sizeof(double) * half);
    p->twi = malloc(sizeof(double) * half);
    p->rev = malloc(sizeof(int) * n);
    if (!p->twr || !p->twi || !p->rev) return -1;
    for (int k = 0; k < n / 2; ++k) oc_twiddle(k, n, &p->twr[k], &p->twi[k]);
    int bits = 0;
    while ((1 << bits) < n) ++bits;
    for (int i = 0; i < n; ++i) {
        int r = 0;
        for (int b = 0; b < bits; ++b)
            if (i & (1 << b)) r |= 1 << (bits - 1 - b);
        p->rev[i] = r;
    }
    return 0;
}

static void fft1_free(Fft1 *p) {
    free(p->twr);
    free(p->twi);
    free(p->rev);
}

/* In-place unnormalised complex DFT; sign -1 (forward) or +1 (inverse). */
static void fft1_exec(const Fft1 *p, double *re, double *im, int inverse) {
    const int n = p->n;
    for (int i = 0; i < n; ++i) {
        int j = p->rev[i];
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    const double sgn = inverse ? -1.0 : 1.0;
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int j = 0; j < half; ++j) {
                const double wr = p->twr[j * step], wi = sgn * p->twi[j * step];
                const int a = i + j, b = a + half;
                const double vr = re[b] * wr - im[b] * wi;
                const double vi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - vr;
                im[b] = im[a] - vi;
                re[a] += vr;
                im[a] += vi;
            }
        }
    }
}

static int is_pow2(int n) { return n >= 2 && (n & (n - 1)) == 0; }

/* fftw_plan_dft_r2c_2d(ny, nx) semantics (spectral_initialization.c:66-68). */
void oc_fft_r2c_2d(OcData *d, const double *in, oc_complex *out) {
    const int nx = d->nx, ny = d->ny, nkx = d->nkx;
    d->fft_count++;
#pragma omp parallel
    {
        double *re = malloc(sizeof(double) * (nx > ny ? nx : ny));
        double *im = malloc(sizeof(double) * (nx > ny ? nx : ny));
#pragma omp for schedule(static)
        for (int j = 0; j < ny; ++j) {
            for (int i = 0; i < nx; ++i) { re[i] = in[(size_t)j * nx + i]; im[i] = 0.0; }
            fft1_exec(&d->fft->fx, re, im, 0);
            for (int i = 0; i < nkx; ++i) out[(size_t)j * nkx + i] = re[i] + I * im[i];
        }
#pragma omp for schedule(static)
        for (int i = 0; i < nkx; ++i) {
            for (int j = 0; j < ny; ++j) {
                oc_complex z = out[(size_t)j * nkx + i];
                re[j] = creal(z); im[j] = cimag(z);
            }
            fft1_exec(&d->fft->fy, re, im, 0);
            for (int j = 0; j < ny; ++j) out[(size_t)j * nkx + i] = re[j] + I * im[j];
        }
        free(re);
        free(im);
    }
}

/* fftw_plan_dft_c2r_2d(ny, nx) semantics (spectral_initialization.c:70-72). */
void oc_fft_c2r_2d(OcData *d, const oc_complex *in, double *out) {
    const int nx = d->nx, ny = d->ny, nkx = d->nkx;
    d->fft_count++;
    oc_complex *tmp = malloc(sizeof(oc_complex) * (size_t)nkx * ny);
    memcpy(tmp, in, sizeof(oc_complex) * (size_t)nkx * ny);
#pragma omp parallel
    {
        double *re = malloc(sizeof(double) * (nx > ny ? nx : ny));
        double *im = malloc(sizeof(double) * (nx > ny ? nx : ny));
#pragma omp for schedule(static)
        for (int i = 0; i < nkx; ++i) {
            for (int j = 0; j < ny; ++j) {
                oc_complex z = tmp[(size_t)j * nkx + i];
                re[j] = creal(z); im[j] = cimag(z);
            }
            fft1_exec(&d->fft->fy, re, im, 1);
            for (int j = 0; j < ny; ++j) tmp[(size_t)j * nkx + i] = re[j] + I * im[j];
        }
#pragma omp for schedule(static)
        for (int j = 0; j < ny; ++j) {
            const oc_complex *row = tmp + (size_t)j * nkx;
            re[0] = creal(row[0]); im[0] = 0.0;
            for (int k = 1; k < nx / 2; ++k) {
                re[k] = creal(row[k]); im[k] = cimag(row[k]);
                re[nx - k] = creal(row[k]); im[nx - k] = -cimag(row[k]);
            }
            re[nx / 2] = creal(row[nx / 2]); im[nx / 2] = 0.0;
            fft1_exec(&d->fft->fx, re, im, 1);
            for (int x = 0; x < nx; ++x) out[(size_t)j * nx + x] = re[x];
        }
        free(re);
        free(im);
    }
    free(tmp);
}

/* --------------------------------------------------------------- memory */

/* spectral_memory.c:6-68 */
OcData *oc_allocate(int nx, int ny, double Lx, double Ly, double Re) {
    if (!is_pow2(nx) || !is_pow2(ny)) return NULL;
    OcData *d = calloc(1, sizeof(OcData));
    if (!d) return NULL;
    d->nx = nx;
    d->ny = ny;
    d->nkx = nx / 2 + 1;
    d->nky = ny;
    d->Lx = Lx;
    d->Ly = Ly;
    d->Re = Re;
    d->dt = 0.01;       /* spectral_memory.c:17 */
    d->t = 0.0;
    d->cfl_factor = 0.3; /* spectral_memory.c:19 */
    const size_t R = (size_t)nx * ny, H = (size_t)d->nkx * d->nky;
    d->u = calloc(R, sizeof(double));
    d->v = calloc(R, sizeof(double));
    d->p = calloc(R, sizeof(double));
    d->omega = calloc(R, sizeof(double));
    d->psi = calloc(R, sizeof(double));
    d->x = calloc(nx, sizeof(double));
    d->y = calloc(ny, sizeof(double));
    d->u_hat = calloc(H, sizeof(oc_complex));
    d->v_hat = calloc(H, sizeof(oc_complex));
    d->omega_hat = calloc(H, sizeof(oc_complex));
    d->psi_hat = calloc(H, sizeof(oc_complex));
    d->nonlinear_hat = calloc(H, sizeof(oc_complex));
    d->rk1_hat = calloc(H, sizeof(oc_complex));
    d->rk2_hat = calloc(H, sizeof(oc_complex));
    d->rk3_hat = calloc(H, sizeof(oc_complex));
    d->rk4_hat = calloc(H, sizeof(oc_complex));
    d->omega_tmp_hat = calloc(H, sizeof(oc_complex));
    d->sfd_hat = calloc(H, sizeof(oc_complex));
    d->kx = calloc(d->nkx, sizeof(double));
    d->ky = calloc(d->nky, sizeof(double));
    d->k2 = calloc(H, sizeof(double));
    d->dealias_mask = calloc(H, sizeof(int));
    d->hv_weights = calloc(H, sizeof(double));
    d->fft = calloc(1, sizeof(OcFFT));
    if (!d->u || !d->v || !d->p || !d->omega || !d->psi || !d->x || !d->y || !d->u_hat ||
        !d->v_hat || !d->omega_hat || !d->psi_hat || !d->nonlinear_hat || !d->rk1_hat ||
        !d->rk2_hat || !d->rk3_hat || !d->rk4_hat || !d->omega_tmp_hat || !d->sfd_hat ||
        !d->kx || !d->ky || !d->k2 || !d->dealias_mask || !d->hv_weights || !d->fft ||
        fft1_init(&d->fft->fx, nx) || fft1_init(&d->fft->fy, ny)) {
        oc_free(d);
        return NULL;
    }
    oc_hyperviscosity_default(&d->hv); /* spectral_memory.c:63 */
    return d;
}

void oc_free(OcData *d) {
    if (!d) return;
    free(d->u); free(d->v); free(d->p); free(d->omega); free(d->psi);
    free(d->x); free(d->y);
    free(d->u_hat); free(d->v_hat); free(d->omega_hat); free(d->psi_hat);
    free(d->nonlinear_hat); free(d->rk1_hat); free(d->rk2_hat); free(d->rk3_hat);
    free(d->rk4_hat); free(d->omega_tmp_hat); free(d->sfd_hat);
    free(d->kx); free(d->ky); free(d->k2); free(d->dealias_mask); free(d->hv_weights);
    if (d->fft) {
        fft1_free(&d->fft->fx);
        fft1_free(&d->fft->fy);
        free(d->fft);
    }
    free(d);
}

/* ----------------------------------------------------------------- init */

/* spectral_initialization.c:10-54 */
void oc_initialize_grids(OcData *d) {
    const int nx = d->nx, ny = d->ny;
    for (int i = 0; i < nx; ++i) d->x[i] = i * d->Lx / nx;
    for (int j = 0; j < ny; ++j) d->y[j] = j * d->Ly / ny;
    for (int i = 0; i < d->nkx; ++i) d->kx[i] = 2.0 * M_PI * i / d->Lx;
    for (int j = 0; j < d->nky; ++j)
        d->ky[j] = (j <= ny / 2) ? 2.0 * M_PI * j / d->Ly : 2.0 * M_PI * (j - ny) / d->Ly;
    for (int j = 0; j < d->nky; ++j) {
        for (int i = 0; i < d->nkx; ++i) {
            const int idx = j * d->nkx + i;
            d->k2[idx] = d->kx[i] * d->kx[i] + d->ky[j] * d->ky[j];
            const int dealias_x = (i <= (2 * d->nkx) / 3);          /* :46 (quirk Q1) */
            const int band = (j <= ny / 2) ? j : j - ny;
            const int dealias_y = (abs(band) <= (2 * ny) / 6);      /* :48 */
            d->dealias_mask[idx] = dealias_x && dealias_y;
        }
    }
    oc_hyperviscosity_update_weights(d); /* :53 */
}

/* spectral_initialization.c:72-95 */
void oc_initialize_taylor_green(OcData *d) {
    const int nx = d->nx, ny = d->ny;
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            const int idx = j * nx + i;
            const double x = d->x[i], y = d->y[j];
            d->u[idx] = sin(2 * M_PI * x / d->Lx) * cos(2 * M_PI * y / d->Ly);
            d->v[idx] = -cos(2 * M_PI * x / d->Lx) * sin(2 * M_PI * y / d->Ly);
            d->omega[idx] = 2.0 * (2 * M_PI / d->Lx) * cos(2 * M_PI * x / d->Lx) *
                            cos(2 * M_PI * y / d->Ly);
            d->p[idx] = 0.0;
        }
    }
    oc_physical_to_spectral(d, d->omega, d->omega_hat);
    oc_physical_to_spectral(d, d->u, d->u_hat);
    oc_physical_to_spectral(d, d->v, d->v_hat);
}

/* ----------------------------------------------------------- transforms */

/* spectral_transforms.c:9-22 */
void oc_physical_to_spectral(OcData *d, double *field, oc_complex *field_hat) {
    oc_fft_r2c_2d(d, field, field_hat);
    const double norm = 1.0 / (d->nx * d->ny);
    const int H = d->nkx * d->nky;
#pragma omp parallel for
    for (int k = 0; k < H; ++k) field_hat[k] *= norm;
    oc_apply_dealiasing(d, field_hat);
}

/* spectral_transforms.c:25-32 (scratch is nonlinear_hat: quirk Q5) */
void oc_spectral_to_physical(OcData *d, oc_complex *field_hat, double *field) {
    const int H = d->nkx * d->nky;
    oc_complex *scratch = d->nonlinear_hat;
    if (scratch != field_hat) memcpy(scratch, field_hat, (size_t)H * sizeof(oc_complex));
    oc_apply_dealiasing(d, scratch);
    oc_fft_c2r_2d(d, scratch, field);
}

/* spectral_transforms.c:35-43 */
void oc_apply_dealiasing(OcData *d, oc_complex *field_hat) {
    const int H = d->nkx * d->nky;
#pragma omp parallel for
    for (int k = 0; k < H; ++k)
        if (!d->dealias_mask[k]) field_hat[k] = 0.0;
}

/* ------------------------------------------------------------------ core */

/* spectral_core.c:8-34 (psi_hat = -omega_hat/k^2: quirk Q2) */
void oc_compute_velocity_from_vorticity(OcData *d) {
#pragma omp parallel for
    for (int j = 0; j < d->nky; ++j) {
        for (int i = 0; i < d->nkx; ++i) {
            const int idx = j * d->nkx + i;
            d->psi_hat[idx] = (d->k2[idx] > 1e-12) ? -d->omega_hat[idx] / d->k2[idx] : 0.0;
        }
    }
#pragma omp parallel for
    for (int j = 0; j < d->nky; ++j) {
        for (int i = 0; i < d->nkx; ++i) {
            const int idx = j * d->nkx + i;
            d->u_hat[idx] = I * d->ky[j] * d->psi_hat[idx];
            d->v_hat[idx] = -I * d->kx[i] * d->psi_hat[idx];
        }
    }
}

/* spectral_core.c:37-43 */
void oc_compute_derivatives(OcData *d) {
    oc_compute_velocity_from_vorticity(d);
    oc_spectral_to_physical(d, d->u_hat, d->u);
    oc_spectral_to_physical(d, d->v_hat, d->v);
    oc_spectral_to_physical(d, d->omega_hat, d->omega);
}

/* spectral_core.c:46-112.  With skip_discarded, the two transforms whose
 * results the reference overwrites (:51-52) are not executed. */
static void oc_nonlinear_impl(OcData *d, int skip_discarded) {
    const int nx = d->nx, ny = d->ny, H = d->nkx * d->nky;
    if (!skip_discarded) {
        oc_spectral_to_physical(d, d->psi_hat, d->psi);
        oc_spectral_to_physical(d, d->omega_hat, d->omega);
    }
    double *psi_x = d->u, *psi_y = d->v, *omega_x = d->psi, *omega_y = d->omega;
    double *jac = d->p;
    oc_complex *scratch = d->nonlinear_hat;
#pragma omp parallel for
    for (int j = 0; j < d->nky; ++j)
        for (int i = 0; i < d->nkx; ++i) {
            const int idx = j * d->nkx + i;
            scratch[idx] = I * d->kx[i] * d->psi_hat[idx];
        }
    oc_spectral_to_physical(d, scratch, psi_x);
#pragma omp parallel for
    for (int j = 0; j < d->nky; ++j)
        for (int i = 0; i < d->nkx; ++i) {
            const int idx = j * d->nkx + i;
            scratch[idx] = I * d->ky[j] * d->psi_hat[idx];
        }
    oc_spectral_to_physical(d, scratch, psi_y);
#pragma omp parallel for
    for (int j = 0; j < d->nky; ++j)
        for (int i = 0; i < d->nkx; ++i) {
            const int idx = j * d->nkx + i;
            scratch[idx] = I * d->kx[i] * d->omega_hat[idx];
        }
    oc_spectral_to_physical(d, scratch, omega_x);
#pragma omp parallel for
    for (int j = 0; j < d->nky; ++j)
        for (int i = 0; i < d->nkx; ++i) {
            const int idx = j * d->nkx + i;
            scratch[idx] = I * d->ky[j] * d->omega_hat[idx];
        }
    oc_spectral_to_physical(d, scratch, omega_y);
#pragma omp parallel for
    for (int k = 0; k < nx * ny; ++k) jac[k] = psi_x[k] * omega_y[k] - psi_y[k] * omega_x[k];
    oc_physical_to_spectral(d, jac, d->nonlinear_hat);
#pragma omp parallel for
    for (int k = 0; k < H; ++k) d->nonlinear_hat[k] = -d->nonlinear_hat[k];
}

void oc_compute_nonlinear_terms(OcData *d) { oc_nonlinear_impl(d, 0); }

/* spectral_time_integration.c:9-31 */
double oc_compute_cfl_timestep(OcData *d) {
    const int nx = d->nx, ny = d->ny;
    double max_vel = 0.0;
#pragma omp parallel for reduction(max : max_vel)
    for (int k = 0; k < nx * ny; ++k) {
        const double vm = sqrt(d->u[k] * d->u[k] + d->v[k] * d->v[k]);
        if (vm > max_vel) max_vel = vm;
    }
    const double dx_min = fmin(d->Lx / nx, d->Ly / ny);
    const double dt_cfl = d->cfl_factor * dx_min / (max_vel + 1e-12);
    const double dt_visc = 0.5 * d->Re * dx_min * dx_min;
    double dt = fmin(dt_cfl, dt_visc);
    dt = fmax(dt, OC_MIN_DT);
    dt = fmin(dt, OC_MAX_DT);
    return dt;
}

/* stage right-hand side of spectral_time_integration.c:50-53 plus
 * hyperviscosity_apply (hyperviscosity.c:104-150) */
static void oc_stage_rhs(OcData *d, oc_complex *k_out) {
    const int H = d->nkx * d->nky;
    const double inv_Re = 1.0 / d->Re;
#pragma omp parallel for
    for (int k = 0; k < H; ++k)
        k_out[k] = d->nonlinear_hat[k] - d->k2[k] * d->omega_hat[k] * inv_Re;
    oc_hyperviscosity_apply(d, k_out, d->omega_hat);
}

/* spectral_time_integration.c:34-106 (reference call sequence, 40 transforms) */
static void oc_rk4_impl(OcData *d, int minimal) {
    const int H = d->nkx * d->nky;
    const double dt = d->dt;
    oc_complex *k1 = d->rk1_hat, *k2 = d->rk2_hat, *k3 = d->rk3_hat, *k4 = d->rk4_hat;
    oc_complex *w0 = d->omega_tmp_hat;
    memcpy(w0, d->omega_hat, (size_t)H * sizeof(oc_complex));

    if (minimal) oc_compute_velocity_from_vorticity(d);
    oc_nonlinear_impl(d, minimal);
    oc_stage_rhs(d, k1);

#pragma omp parallel for
    for (int k = 0; k < H; ++k) d->omega_hat[k] = w0[k] + 0.5 * dt * k1[k];
    if (minimal) oc_compute_velocity_from_vorticity(d); else oc_compute_derivatives(d);
    oc_nonlinear_impl(d, minimal);
    oc_stage_rhs(d, k2);

#pragma omp parallel for
    for (int k = 0; k < H; ++k) d->omega_hat[k] = w0[k] + 0.5 * dt * k2[k];
    if (minimal) oc_compute_velocity_from_vorticity(d); else oc_compute_derivatives(d);
    oc_nonlinear_impl(d, minimal);
    oc_stage_rhs(d, k3);

#pragma omp parallel for
    for (int k = 0; k < H; ++k) d->omega_hat[k] = w0[k] + dt * k3[k];
    if (minimal) oc_compute_velocity_from_vorticity(d); else oc_compute_derivatives(d);
    oc_nonlinear_impl(d, minimal);
    oc_stage_rhs(d, k4);

#pragma omp parallel for
    for (int k = 0; k < H; ++k)
        d->omega_hat[k] = w0[k] + dt / 6.0 * (k1[k] + 2.0 * k2[k] + 2.0 * k3[k] + k4[k]);
    oc_hyperviscosity_post_step(d, d->omega_hat, dt);
    d->t += dt;
    oc_compute_derivatives(d);
}

void oc_advance_rk4(OcData *d) { oc_rk4_impl(d, 0); }
void oc_advance_rk4_minimal(OcData *d) { oc_rk4_impl(d, 1); }

/* ----------------------------------------------------------- diagnostics */

/* spectral_analysis.c:8-19 */
double oc_compute_energy(OcData *d) {
    const int total = d->nx * d->ny;
    double e = 0.0;
#pragma omp parallel for reduction(+ : e)
    for (int k = 0; k < total; ++k) e += 0.5 * (d->u[k] * d->u[k] + d->v[k] * d->v[k]);
    return e / total;
}

/* spectral_analysis.c:22-33 */
double oc_compute_enstrophy(OcData *d) {
    const int total = d->nx * d->ny;
    double z = 0.0;
#pragma omp parallel for reduction(+ : z)
    for (int k = 0; k < total; ++k) z += 0.5 * d->omega[k] * d->omega[k];
    return z / total;
}

/* spectral_analysis.c:36-66 (max_k = sqrt(k2[last]): quirk Q7) */
int oc_check_resolution(OcData *d, double *high_k_fraction) {
    const int H = d->nkx * d->nky;
    double high = 0.0, total = 0.0;
    const double max_k = sqrt(d->k2[H - 1]);
    for (int idx = 0; idx < H; ++idx) {
        const double m = cabs(d->omega_hat[idx]);
        const double m2 = m * m;
        total += m2;
        const double kn = (max_k > 0.0) ? sqrt(d->k2[idx]) / max_k : 0.0;
        if (kn > 0.8) high += m2;
    }
    const double frac = (total > 1e-16) ? high / total : 0.0;
    if (high_k_fraction) *high_k_fraction = frac;
    return frac > 1e-6 ? 0 : 1;
}

/* spectral_analysis.c:69-100; out = {E, Z, max|w_hat|, spectral_error, t, dt} */
void oc_analyze_convergence(OcData *d, double out[6]) {
    const int H = d->nkx * d->nky;
    double maxw = 0.0, es = 0.0;
    int high = 0;
    for (int idx = 0; idx < H; ++idx) {
        const double m = cabs(d->omega_hat[idx]);
        if (m > maxw) maxw = m;
        if (d->k2[idx] > 0.25 * d->k2[H - 1]) {
            es += m * m;
            high++;
        }
    }
    out[0] = oc_compute_energy(d);
    out[1] = oc_compute_enstrophy(d);
    out[2] = maxw;
    out[3] = high > 0 ? sqrt(es / high) : 0.0;
    out[4] = d->t;
    out[5] = d->dt;
}

/* spectral_analysis.c:103-142.  Returns max_k; energy_k[k] holds the shell
 * average for 1 <= k < first empty shell (the reference's print loop),
 * and the raw shell sum elsewhere. */
int oc_analyze_spectrum(OcData *d, double *energy_k, int *count_k, int max_bins) {
    const int nkx = d->nkx, nky = d->nky;
    int max_k = (int)fmin(nkx, nky / 2);
    if (max_k > max_bins) max_k = max_bins;
    for (int k = 0; k < max_k; ++k) { energy_k[k] = 0.0; count_k[k] = 0; }
    for (int j = 0; j < nky; ++j)
        for (int i = 0; i < nkx; ++i) {
            const int idx = j * nkx + i;
            const int bin = (int)(sqrt(d->k2[idx]) + 0.5);
            if (bin < max_k) {
                const double um = cabs(d->u_hat[idx]), vm = cabs(d->v_hat[idx]);
                energy_k[bin] += 0.5 * (um * um + vm * vm);
                count_k[bin]++;
            }
        }
    for (int k = 1; k < max_k && count_k[k] > 0; ++k) energy_k[k] /= count_k[k];
    return max_k;
}

/* -------------------------------------------------------- hyperviscosity */

static double oc_clamp(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* hyperviscosity.c:15-27 */
void oc_hyperviscosity_default(OcHyperviscosity *c) {
    if (!c) return;
    c->enabled = 0;
    c->coefficient = 0.0;
    c->order = 8;
    c->cutoff_ratio = 0.65;
    c->sfd_strength = 0.0;
    c->sfd_timescale = 1.0;
    c->sfd_enabled = 0;
    c->kmax = 0.0;
    c->cutoff_wavenumber = 0.0;
    c->inv_bandwidth = 0.0;
}

/* hyperviscosity.c:29-58 */
void oc_hyperviscosity_set(OcData *d, const OcHyperviscosity *c) {
    if (!d || !c) return;
    const size_t H = (size_t)d->nkx * d->nky;
    d->hv = *c;
    if (!d->hv.enabled) memset(d->hv_weights, 0, H * sizeof(double));
    if (!d->hv.sfd_enabled || d->hv.sfd_strength <= 0.0) {
        d->hv.sfd_enabled = 0;
        d->hv.sfd_strength = 0.0;
        memset(d->sfd_hat, 0, H * sizeof(oc_complex));
    }
    oc_hyperviscosity_update_weights(d);
    if (d->hv.sfd_enabled) memcpy(d->sfd_hat, d->omega_hat, H * sizeof(oc_complex));
}

/* hyperviscosity.c:60-102 (kmax over every stored mode: quirk Q7) */
void oc_hyperviscosity_update_weights(OcData *d) {
    const int H = d->nkx * d->nky;
    OcHyperviscosity *c = &d->hv;
    double kmax = 0.0;
    for (int idx = 0; idx < H; ++idx) kmax = fmax(kmax, sqrt(d->k2[idx]));
    c->kmax = kmax;
    const double ratio = oc_clamp(c->cutoff_ratio, 0.0, 1.0);
    c->cutoff_wavenumber = ratio * kmax;
    const double bw = fmax(kmax - c->cutoff_wavenumber, 1e-12);
    c->inv_bandwidth = (bw > 0.0) ? 1.0 / bw : 0.0;
    if (!c->enabled || c->coefficient <= 0.0 || kmax <= 0.0) {
        for (int idx = 0; idx < H; ++idx) d->hv_weights[idx] = 0.0;
        return;
    }
    const int order = (c->order > 0) ? c->order : 1;
    for (int idx = 0; idx < H; ++idx) {
        const double km = sqrt(d->k2[idx]);
        double w = 0.0;
        if (km > c->cutoff_wavenumber)
            w = c->coefficient * pow((km - c->cutoff_wavenumber) * c->inv_bandwidth, order);
        d->hv_weights[idx] = w;
    }
}

static double oc_sfd_weight(const OcData *d, int idx) {
    const OcHyperviscosity *c = &d->hv;
    const int order = (c->order > 0) ? c->order : 1;
    double ratio = 0.0;
    if (c->inv_bandwidth > 0.0) {
        const double km = sqrt(d->k2[idx]);
        if (km > c->cutoff_wavenumber) {
            double s = (km - c->cutoff_wavenumber) * c->inv_bandwidth;
            ratio = pow(oc_clamp(s, 0.0, 1.0), order);
        }
    }
    return c->sfd_strength * ratio;
}

/* hyperviscosity.c:104-150 */
void oc_hyperviscosity_apply(OcData *d, oc_complex *rhs, const oc_complex *state) {
    if (!d || !rhs || !state) return;
    const OcHyperviscosity *c = &d->hv;
    const int apply_visc = (c->enabled && c->coefficient > 0.0);
    const int apply_sfd = (c->sfd_enabled && c->sfd_strength > 0.0);
    if (!apply_visc && !apply_sfd) return;
    const int H = d->nkx * d->nky;
#pragma omp parallel for
    for (int idx = 0; idx < H; ++idx) {
        oc_complex damping = 0.0;
        if (apply_visc) damping += d->hv_weights[idx] * state[idx];
        if (apply_sfd) {
            const double sw = oc_sfd_weight(d, idx);
            if (sw > 0.0) damping += sw * (state[idx] - d->sfd_hat[idx]);
        }
        rhs[idx] -= damping;
    }
}

/* hyperviscosity.c:152-178 */
void oc_hyperviscosity_post_step(OcData *d, const oc_complex *state, double dt) {
    if (!d || !state || dt <= 0.0) return;
    const OcHyperviscosity *c = &d->hv;
    if (!c->sfd_enabled || c->sfd_strength <= 0.0 || c->sfd_timescale <= 0.0) return;
    double alpha = oc_clamp(dt / c->sfd_timescale, 0.0, 1.0);
    const int H = d->nkx * d->nky;
#pragma omp parallel for
    for (int idx = 0; idx < H; ++idx)
        d->sfd_hat[idx] = (1.0 - alpha) * d->sfd_hat[idx] + alpha * state[idx];
}

/* ---------------------------------------------------------------- forcing */

/* forcing.c:11-16: SplitMix64 (Steele, Lea & Flood 2014, Mix13 finaliser) */
uint64_t oc_splitmix64_next(uint64_t *state) {
    uint64_t z = (*state += UINT64_C(0x9E3779B97F4A7C15));
    z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
    return z ^ (z >> 31);
}

/* Counter form of the same stream: the k-th output (k >= 1) of a generator
 * seeded with `seed`, mapped to [0,1) as forcing.c:18-21 does. */
double oc_counter_uniform(uint64_t seed, uint64_t k) {
    uint64_t s = seed + (k - 1) * UINT64_C(0x9E3779B97F4A7C15);
    uint64_t b = oc_splitmix64_next(&s);
    return (double)(b >> 11) * (1.0 / 9007199254740992.0);
}

static double oc_forcing_uniform(OcForcingState *s) {
    uint64_t st = s->rng_state;
    uint64_t b = oc_splitmix64_next(&st);
    s->rng_state = st;
    return (double)(b >> 11) * (1.0 / 9007199254740992.0);
}

/* forcing.c:23-39 (Box-Muller, cosine first, sine kept as the spare) */
static double oc_forcing_gaussian(OcForcingState *s) {
    if (s->have_spare) {
        s->have_spare = 0;
        return s->spare_gaussian;
    }
    double u1 = oc_forcing_uniform(s);
    double u2 = oc_forcing_uniform(s);
    u1 = fmax(u1, DBL_MIN);
    const double radius = sqrt(-2.0 * log(u1));
    const double theta = 2.0 * M_PI * u2;
    s->spare_gaussian = radius * sin(theta);
    s->have_spare = 1;
    return radius * cos(theta);
}

/* forcing.c:41-51 */
void oc_forcing_config_default(OcForcingConfig *c) {
    if (!c) return;
    c->kmin = 0.0;
    c->kmax = 4.0;
    c->energy_target = 1.0;
    c->correlation_time = 0.5;
    c->seed = 0;
    c->type = 0;
}

/* forcing.c:53-103 */
int oc_forcing_state_init(OcForcingState *s, const OcForcingConfig *c, const double *k2, int nkx,
                          int nky) {
    if (!s || !c || !k2 || nkx <= 0 || nky <= 0) return -1;
    memset(s, 0, sizeof(*s));
    s->config = *c;
    s->nkx = nkx;
    s->nky = nky;
    s->spectral_size = nkx * nky;
    s->weights = calloc((size_t)s->spectral_size, sizeof(double));
    if (!s->weights) return -1;
    if (c->type == 1) {
        s->colored_state = calloc((size_t)s->spectral_size, sizeof(oc_complex));
        if (!s->colored_state) { oc_forcing_state_free(s); return -1; }
    }
    const double kmin = fmax(c->kmin, 0.0);
    const double kmax = fmax(c->kmax, kmin);
    int active = 0;
    for (int j = 0; j < nky; ++j)
        for (int i = 0; i < nkx; ++i) {
            const int idx = j * nkx + i;
            const double km = sqrt(k2[idx]);
            if (km >= kmin && km <= kmax) { s->weights[idx] = 1.0; active++; }
            else s->weights[idx] = 0.0;
        }
    s->active_modes = active;
    s->have_spare = 0;
    s->rng_state = c->seed ? (uint64_t)c->seed : (uint64_t)0x123456789abcdefULL; /* :99 */
    return active == 0 ? -1 : 0;
}

/* forcing.c:105-115 */
void oc_forcing_state_free(OcForcingState *s) {
    if (!s) return;
    free(s->weights);
    s->weights = NULL;
    free(s->colored_state);
    s->colored_state = NULL;
    s->spectral_size = 0;
    s->active_modes = 0;
}

/* forcing.c:117-184 */
void oc_forcing_generate(OcForcingState *s, double dt, const double *k2, oc_complex *out) {
    if (!s || !out || !k2) return;
    const int size = s->spectral_size;
    if (size == 0 || s->active_modes == 0) {
        memset(out, 0, (size_t)size * sizeof(oc_complex));
        return;
    }
    const OcForcingConfig *c = &s->config;
    if (c->energy_target <= 0.0) {
        memset(out, 0, (size_t)size * sizeof(oc_complex));
        if (s->colored_state) memset(s->colored_state, 0, (size_t)size * sizeof(oc_complex));
        return;
    }
    const double tau = fmax(c->correlation_time, 1e-6);
    const int colored = (c->type == 1);
    double alpha = 0.0, beta = 1.0;
    if (colored) {
        alpha = exp(-dt / tau);
        beta = sqrt(fmax(1.0 - alpha * alpha, 0.0));
    }
    double total = 0.0;
    for (int idx = 0; idx < size; ++idx) {
        if (s->weights[idx] <= 0.0) { out[idx] = 0.0; continue; }
        const double g1 = oc_forcing_gaussian(s);
        const double g2 = oc_forcing_gaussian(s);
        oc_complex noise = g1 + I * g2;
        oc_complex value;
        if (colored) {
            s->colored_state[idx] = alpha * s->colored_state[idx] + beta * noise;
            value = s->colored_state[idx];
        } else {
            value = noise;
        }
        out[idx] = s->weights[idx] * value;
        total += pow(s->weights[idx], 2.0) * (creal(value) * creal(value) + cimag(value) * cimag(value));
    }
    total = fmax(total, 1e-16);
    const double scale = sqrt(c->energy_target / total);
    for (int idx = 0; idx < size; ++idx) {
        out[idx] *= scale;
        if (colored) s->colored_state[idx] *= scale;
    }
}

/* --------------------------------------------- build-defined extensions */

/* Build-defined forcing application (no reference counterpart: the
 * reference step never applies forcing).  (1) The kx = 0 and kx = nx/2
 * columns are projected onto Hermitian symmetry, f(0,-ky) = conj f(0,ky);
 * (2) modes outside the dealias mask and the k = 0 mode are dropped;
 * (3) f is rescaled so that the energy injection rate of the update
 * omega_hat += sqrt(dt) f is eps = 1/2 sum_full |f_k|^2 / |k|^2. */
void oc_forcing_project_and_scale(OcData *d, oc_complex *f, double eps) {
    const int nkx = d->nkx, ny = d->ny;
    const int cols[2] = {0, d->nx / 2};
    for (int c = 0; c < 2; ++c) {
        const int i = cols[c];
        if (c == 1 && i == 0) break;
        for (int j = 1; j < ny / 2; ++j) {
            const int a = j * nkx + i, b = (ny - j) * nkx + i;
            const oc_complex avg = 0.5 * (f[a] + conj(f[b]));
            f[a] = avg;
            f[b] = conj(avg);
        }
        f[i] = creal(f[i]);
        f[(ny / 2) * nkx + i] = creal(f[(ny / 2) * nkx + i]);
    }
    double rate = 0.0;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nkx; ++i) {
            const int idx = j * nkx + i;
            if (!d->dealias_mask[idx] || d->k2[idx] <= 1e-12) { f[idx] = 0.0; continue; }
            const double w = (i == 0 || 2 * i == d->nx) ? 1.0 : 2.0;
            const double m = cabs(f[idx]);
            rate += 0.5 * w * m * m / d->k2[idx];
        }
    const double s = (rate > 0.0 && eps > 0.0) ? sqrt(eps / rate) : 0.0;
    for (int k = 0; k < nkx * ny; ++k) f[k] *= s;
}

/* Build-defined: per-mode linear operator L = -(k^2/Re + w_hv + mu). */
static double oc_linear_rate(const OcData *d, int idx) {
    return -(d->k2[idx] / d->Re + d->hv_weights[idx] + d->drag);
}

/* Nonlinear term of a stage (spectral_core.c:46-112 without the two
 * discarded transforms), psi_hat re-derived from omega_hat first. */
static void oc_stage_nonlinear(OcData *d) {
    oc_compute_velocity_from_vorticity(d);
    oc_nonlinear_impl(d, 1);
}

/* Build-defined integrating-factor RK4 (Lawson).  With v = e^{-Lt} omega,
 * classical RK4 on v gives, with E = e^{L dt}, E2 = e^{L dt/2}:
 *   a = w_n;            k1 = N(a)
 *   b = E2 (a + h/2 k1); k2 = N(b)
 *   c = E2 a + h/2 k2;   k3 = N(c)
 *   d = E a + h E2 k3;   k4 = N(d)
 *   w_{n+1} = E a + h/6 (E k1 + 2 E2 (k2 + k3) + k4)
 * then the white-in-time forcing update w += sqrt(h) f (if f != NULL). */
/* Half-plane Parseval weight of column i: 1 on kx = 0 and kx = nx/2, else 2. */
static double oc_plane_weight(const OcData *d, int idx) {
    const int i = idx % d->nkx;
    return (i == 0 || 2 * i == d->nx) ? 1.0 : 2.0;
}

/* Embedded error estimate of the build-defined adaptive-dt controller: the
 * step's second stage gives a midpoint solution y2 (order 2) beside the RK4
 * solution y4; est = y4 - y2 = O(dt^3).  norms[0] = sum w |est|^2 and
 * norms[1] = sum w |y4|^2 over the stored modes (w: oc_plane_weight), y4
 * before the forcing kick.  ymid holds y2. */
static void oc_est_norms(const OcData *d, const oc_complex *y4, const oc_complex *ymid,
                         double norms[2]) {
    const int H = d->nkx * d->nky;
    double e = 0.0, z = 0.0;
    for (int k = 0; k < H; ++k) {
        const double w = oc_plane_weight(d, k);
        const oc_complex est = y4[k] - ymid[k];
        e += w * (creal(est) * creal(est) + cimag(est) * cimag(est));
        z += w * (creal(y4[k]) * creal(y4[k]) + cimag(y4[k]) * cimag(y4[k]));
    }
    norms[0] = e;
    norms[1] = z;
}

static void oc_if_rk4_impl(OcData *d, const oc_complex *f, double *norms);

void oc_advance_if_rk4(OcData *d, const oc_complex *f) { oc_if_rk4_impl(d, f, NULL); }

/* oc_advance_if_rk4 plus the embedded estimate: y2 is the Lawson midpoint
 * solution E a + h E2 N(b), from the step's own second stage */
void oc_advance_if_rk4_est(OcData *d, const oc_complex *f, double norms[2]) {
    oc_if_rk4_impl(d, f, norms);
}

static void oc_if_rk4_impl(OcData *d, const oc_complex *f, double *norms) {
    const int H = d->nkx * d->nky;
    const double h = d->dt;
    oc_complex *a = d->omega_tmp_hat, *acc = d->rk1_hat, *w = d->omega_hat;
    oc_complex *ymid = d->rk2_hat;
    memcpy(a, w, (size_t)H * sizeof(oc_complex));

    oc_stage_nonlinear(d); /* k1 at a */
#pragma omp parallel for
    for (int k = 0; k < H; ++k) {
        const double L = oc_linear_rate(d, k);
        const double E = exp(L * h), E2 = exp(L * 0.5 * h);
        const oc_complex k1 = d->nonlinear_hat[k];
        acc[k] = E * a[k] + (h / 6.0) * E * k1;
        w[k] = E2 * (a[k] + 0.5 * h * k1);
    }
    oc_stage_nonlinear(d); /* k2 at b */
#pragma omp parallel for
    for (int k = 0; k < H; ++k) {
        const double L = oc_linear_rate(d, k);
        const double E2 = exp(L * 0.5 * h);
        const oc_complex k2 = d->nonlinear_hat[k];
        acc[k] += (h / 3.0) * E2 * k2;
        w[k] = E2 * a[k] + 0.5 * h * k2;
        if (norms) ymid[k] = exp(L * h) * a[k] + h * E2 * k2;
    }
    oc_stage_nonlinear(d); /* k3 at c */
#pragma omp parallel for
    for (int k = 0; k < H; ++k) {
        const double L = oc_linear_rate(d, k);
        const double E = exp(L * h), E2 = exp(L * 0.5 * h);
        const oc_complex k3 = d->nonlinear_hat[k];
        acc[k] += (h / 3.0) * E2 * k3;
        w[k] = E * a[k] + h * E2 * k3;
    }
    oc_stage_nonlinear(d); /* k4 at d */
    const double sq = sqrt(h);
#pragma omp parallel for
    for (int k = 0; k < H; ++k) w[k] = acc[k] + (h / 6.0) * d->nonlinear_hat[k];
    if (norms) oc_est_norms(d, w, ymid, norms);
    if (f)
#pragma omp parallel for
        for (int k = 0; k < H; ++k) w[k] += sq * f[k];
    d->t += h;
    oc_compute_derivatives(d);
}

/* Build-defined classical RK4 with drag and forcing: the reference stage
 * algebra (spectral_time_integration.c:44-100) with k_s = N - (k^2/Re + mu) w
 * - HV(w), SFD post-step (:102), then w += sqrt(dt) f.  With mu = 0 and
 * f = NULL this is oc_advance_rk4_minimal. */
static void oc_classic_ext_impl(OcData *d, const oc_complex *f, double *norms);

void oc_advance_classic_ext(OcData *d, const oc_complex *f) { oc_classic_ext_impl(d, f, NULL); }

/* oc_advance_classic_ext plus the embedded estimate: y2 = w0 + dt k2, the
 * explicit midpoint solution from the step's own second stage */
void oc_advance_classic_ext_est(OcData *d, const oc_complex *f, double norms[2]) {
    oc_classic_ext_impl(d, f, norms);
}

static void oc_classic_ext_impl(OcData *d, const oc_complex *f, double *norms) {
    const int H = d->nkx * d->nky;
    const double dt = d->dt;
    oc_complex *ks[4] = {d->rk1_hat, d->rk2_hat, d->rk3_hat, d->rk4_hat};
    oc_complex *w0 = d->omega_tmp_hat, *w = d->omega_hat;
    const double cs[4] = {0.5, 0.5, 1.0, 0.0};
    memcpy(w0, w, (size_t)H * sizeof(oc_complex));
    for (int s = 0; s < 4; ++s) {
        oc_stage_nonlinear(d);
        oc_stage_rhs(d, ks[s]);
#pragma omp parallel for
        for (int k = 0; k < H; ++k) ks[s][k] -= d->drag * w[k];
        if (s < 3)
#pragma omp parallel for
            for (int k = 0; k < H; ++k) w[k] = w0[k] + cs[s] * dt * ks[s][k];
    }
    const double sq = sqrt(dt);
#pragma omp parallel for
    for (int k = 0; k < H; ++k) {
        w[k] = w0[k] + dt / 6.0 * (ks[0][k] + 2.0 * ks[1][k] + 2.0 * ks[2][k] + ks[3][k]);
    }
    if (norms) {
        /* the midpoint solution overwrites k1, which is no longer needed */
        oc_complex *ymid = ks[0];
#pragma omp parallel for
        for (int k = 0; k < H; ++k) ymid[k] = w0[k] + dt * ks[1][k];
        oc_est_norms(d, w, ymid, norms);
    }
    oc_hyperviscosity_post_step(d, w, dt);
    if (f)
        for (int k = 0; k < H; ++k) w[k] += sq * f[k];
    d->t += dt;
    oc_compute_derivatives(d);
}

/* Build-defined random-phase initial vorticity: E(k) ~ k^4 exp(-(k/k_peak)^2),
 * |w_hat_k|^2 ~ k E(k) / pi on every kept mode, phase 2 pi U where U is the
 * (idx+1)-th output of the counter SplitMix64 stream seeded with `seed`
 * (idx = j*nkx + i in the reference layout).  The kx = 0 column is made
 * Hermitian from its ky > 0 half; the field is scaled so that
 * 1/2 <u^2 + v^2> = energy; then the canonical compute_derivatives runs. */
void oc_initialize_random_phase(OcData *d, uint64_t seed, double k_peak, double energy) {
    const int nkx = d->nkx, ny = d->ny, H = nkx * ny;
    oc_complex *w = d->omega_hat;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nkx; ++i) {
            const int idx = j * nkx + i;
            const double k2 = d->k2[idx];
            if (!d->dealias_mask[idx] || k2 <= 1e-12) { w[idx] = 0.0; continue; }
            const double km = sqrt(k2);
            const double r = km / k_peak;
            const double Ek = km * km * km * km * exp(-r * r);
            const double amp = sqrt(km * Ek / M_PI);
            const double ph = 2.0 * M_PI * oc_counter_uniform(seed, (uint64_t)idx + 1);
            w[idx] = amp * (cos(ph) + I * sin(ph));
        }
    for (int j = ny / 2 + 1; j < ny; ++j) w[j * nkx] = conj(w[(ny - j) * nkx]);
    w[(ny / 2) * nkx] = creal(w[(ny / 2) * nkx]);
    double e = 0.0;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nkx; ++i) {
            const int idx = j * nkx + i;
            if (d->k2[idx] <= 1e-12) continue;
            const double wt = (i == 0 || 2 * i == d->nx) ? 1.0 : 2.0;
            const double m = cabs(w[idx]);
            e += 0.5 * wt * m * m / d->k2[idx];
        }
    const double s = e > 0.0 ? sqrt(energy / e) : 0.0;
    for (int k = 0; k < H; ++k) w[k] *= s;
    oc_compute_derivatives(d);
}
