/*
 * spectral_api.c -- the reference's spectral solver API (ns_spectral_solver.h)
 * on MI355X.  Each entry point moves the host arrays it reads to the GPU,
 * runs the HIP kernels through the thin layer of ns_device.h, and writes its
 * results back into the host arrays, so reference call sequences keep their
 * meaning.  The additions at the end keep the state on the device across
 * many steps (the fast path).
 */
#include "ns_spectral_solver.h"

#include <stdint.h>

#include "ns_device.h"
#include "spectral_internal.h"

static struct ns_fft_plan_s { int unused; } ns_plan_token;

static size_t H_of(const NSSpectralData *d) { return (size_t)d->nkx * (size_t)d->nky; }
static size_t R_of(const NSSpectralData *d) { return (size_t)d->nx * (size_t)d->ny; }

static SolverStatus set_status(NSSpectralData *d, int rc) {
    if (!d || !d->device) return SOLVER_INVALID_PARAM;
    d->device->status = (SolverStatus)rc;
    if (rc != 0 && d->device->ctx)
        snprintf(d->device->err, sizeof(d->device->err), "%s", nsd_error(d->device->ctx));
    else if (rc == 0)
        d->device->err[0] = '\0';
    return (SolverStatus)rc;
}

static int ready(const NSSpectralData *d) { return d && d->device && d->device->ctx; }

static void build_params(const NSSpectralData *d, NsdParams *p) {
    memset(p, 0, sizeof(*p));
    const NSSpectralHyperviscosityConfig *h = &d->hyperviscosity;
    p->Re = d->Re;
    p->drag = d->device->drag;
    p->cfl_factor = d->cfl_factor;
    p->scheme = d->device->scheme == NS_SCHEME_IF_RK4 ? NSD_SCHEME_IF : NSD_SCHEME_CLASSIC;
    p->hv_on = (h->enabled && h->coefficient > 0.0) ? 1 : 0;
    p->hv_order = h->order > 0 ? h->order : 1;
    p->hv_coeff = h->coefficient;
    p->hv_kc = h->cutoff_wavenumber;
    p->hv_invbw = h->inv_bandwidth;
    p->sfd_on = (h->sfd_enabled && h->sfd_strength > 0.0) ? 1 : 0;
    p->sfd_strength = h->sfd_strength;
    p->sfd_tau = h->sfd_timescale;
    if (d->device->force_on) {
        const NSForcingConfig *f = &d->device->force_cfg;
        p->force_on = 1;
        p->force_colored = f->type == NS_FORCING_COLORED_NOISE;
        p->force_kmin = f->kmin;
        p->force_kmax = f->kmax;
        p->force_energy_target = f->energy_target;
        p->force_tau = f->correlation_time;
        p->force_rate = d->device->force_rate;
        p->force_seed = (unsigned long long)f->seed;
    }
    if (d->device->errctl) {
        p->errctl = 1;
        p->err_rtol = d->device->err_rtol;
        p->err_atol = d->device->err_atol;
        p->err_safety = d->device->err_safety;
        p->dt_min = d->device->dt_min;
        p->dt_max = d->device->dt_max;
    }
}

static int sync_params(NSSpectralData *d) {
    NsdParams p;
    build_params(d, &p);
    if (d->device->applied_valid && memcmp(&p, &d->device->applied, sizeof(p)) == 0) return 0;
    int rc = nsd_set_params(d->device->ctx, &p);
    if (rc == 0) {
        d->device->applied = p;
        d->device->applied_valid = 1;
    }
    return rc;
}

/* ------------------------------------------------------------ memory */

NSSpectralData *ns_spectral_allocate(int nx, int ny, double Lx, double Ly, double Re) {
    NSSpectralData *d = (NSSpectralData *)calloc(1, sizeof(NSSpectralData));
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
    d->device = (struct NSSpectralDevice *)calloc(1, sizeof(struct NSSpectralDevice));
    if (!d->device || nx <= 0 || ny <= 0) {
        ns_spectral_free(d);
        return NULL;
    }
    d->device->ctx = nsd_create(nx, ny, Lx, Ly, NSD_FP64, d->device->err, sizeof(d->device->err));
    if (!d->device->ctx) {
        fprintf(stderr, "ns_spectral_allocate: %s\n", d->device->err);
        ns_spectral_free(d);
        return NULL;
    }
    const size_t R = R_of(d), H = H_of(d);
    d->u = (double *)calloc(R, sizeof(double));
    d->v = (double *)calloc(R, sizeof(double));
    d->p = (double *)calloc(R, sizeof(double));
    d->omega = (double *)calloc(R, sizeof(double));
    d->psi = (double *)calloc(R, sizeof(double));
    d->x = (double *)calloc((size_t)nx, sizeof(double));
    d->y = (double *)calloc((size_t)ny, sizeof(double));
    d->u_hat = (fftw_complex *)calloc(H, sizeof(fftw_complex));
    d->v_hat = (fftw_complex *)calloc(H, sizeof(fftw_complex));
    d->omega_hat = (fftw_complex *)calloc(H, sizeof(fftw_complex));
    d->psi_hat = (fftw_complex *)calloc(H, sizeof(fftw_complex));
    d->nonlinear_hat = (fftw_complex *)calloc(H, sizeof(fftw_complex));
    d->rk1_hat = (fftw_complex *)calloc(H, sizeof(fftw_complex));
    d->rk2_hat = (fftw_complex *)calloc(H, sizeof(fftw_complex));
    d->rk3_hat = (fftw_complex *)calloc(H, sizeof(fftw_complex));
    d->rk4_hat = (fftw_complex *)calloc(H, sizeof(fftw_complex));
    d->omega_tmp_hat = (fftw_complex *)calloc(H, sizeof(fftw_complex));
    d->sfd_hat = (fftw_complex *)calloc(H, sizeof(fftw_complex));
    d->kx = (double *)calloc((size_t)d->nkx, sizeof(double));
    d->ky = (double *)calloc((size_t)d->nky, sizeof(double));
    d->k2 = (double *)calloc(H, sizeof(double));
    d->dealias_mask = (int *)calloc(H, sizeof(int));
    d->hyperviscosity_weights = (double *)calloc(H, sizeof(double));
    if (!d->u || !d->v || !d->p || !d->omega || !d->psi || !d->x || !d->y || !d->u_hat ||
        !d->v_hat || !d->omega_hat || !d->psi_hat || !d->nonlinear_hat || !d->rk1_hat ||
        !d->rk2_hat || !d->rk3_hat || !d->rk4_hat || !d->omega_tmp_hat || !d->sfd_hat || !d->kx ||
        !d->ky || !d->k2 || !d->dealias_mask || !d->hyperviscosity_weights) {
        ns_spectral_free(d);
        return NULL;
    }
    ns_spectral_hyperviscosity_default(&d->hyperviscosity); /* spectral_memory.c:63 */
    d->device->precision = NS_PRECISION_FP64;
    d->device->scheme = NS_SCHEME_RK4;
    return d;
}

void ns_spectral_free(NSSpectralData *d) {
    if (!d) return;
    if (d->device) {
        if (d->device->ctx) nsd_destroy(d->device->ctx);
        free(d->device);
    }
    free(d->u); free(d->v); free(d->p); free(d->omega); free(d->psi);
    free(d->x); free(d->y);
    free(d->u_hat); free(d->v_hat); free(d->omega_hat); free(d->psi_hat);
    free(d->nonlinear_hat); free(d->rk1_hat); free(d->rk2_hat); free(d->rk3_hat);
    free(d->rk4_hat); free(d->omega_tmp_hat); free(d->sfd_hat);
    free(d->kx); free(d->ky); free(d->k2); free(d->dealias_mask);
    free(d->hyperviscosity_weights);
    free(d);
}

/* ------------------------------------------------------------ init */

void ns_spectral_initialize_grids(NSSpectralData *d) {
    if (!ready(d)) return;
    int rc = nsd_grids(d->device->ctx, d->x, d->y, d->kx, d->ky, d->k2, d->dealias_mask);
    set_status(d, rc);
    if (rc == 0) ns_spectral_hyperviscosity_update_weights(d); /* spectral_initialization.c:53 */
}

void ns_spectral_initialize_fft_plans(NSSpectralData *d) {
    if (!d) return;
    d->forward_plan = &ns_plan_token;
    d->backward_plan = &ns_plan_token;
}

void ns_spectral_initialize_taylor_green(NSSpectralData *d) {
    if (!ready(d)) return;
    int rc = nsd_taylor_green(d->device->ctx, d->u, d->v, d->omega, d->p);
    if (rc == 0) rc = nsd_physical_to_spectral(d->device->ctx, d->omega, (double *)d->omega_hat);
    if (rc == 0) rc = nsd_physical_to_spectral(d->device->ctx, d->u, (double *)d->u_hat);
    if (rc == 0) rc = nsd_physical_to_spectral(d->device->ctx, d->v, (double *)d->v_hat);
    set_status(d, rc);
}

/* ------------------------------------------------------------ transforms */

void ns_spectral_physical_to_spectral(NSSpectralData *d, double *field, fftw_complex *field_hat) {
    if (!ready(d) || !field || !field_hat) return;
    set_status(d, nsd_physical_to_spectral(d->device->ctx, field, (double *)field_hat));
}

void ns_spectral_spectral_to_physical(NSSpectralData *d, fftw_complex *field_hat, double *field) {
    if (!ready(d) || !field || !field_hat) return;
    set_status(d, nsd_spectral_to_physical(d->device->ctx, (const double *)field_hat, field));
}

void ns_spectral_apply_dealiasing(NSSpectralData *d, fftw_complex *field_hat) {
    if (!ready(d) || !field_hat) return;
    set_status(d, nsd_dealias(d->device->ctx, (double *)field_hat));
}

/* ------------------------------------------------------------ core */

void ns_spectral_compute_velocity_from_vorticity(NSSpectralData *d) {
    if (!ready(d)) return;
    set_status(d, nsd_velocity(d->device->ctx, (const double *)d->omega_hat,
                               (double *)d->psi_hat, (double *)d->u_hat, (double *)d->v_hat));
}

void ns_spectral_compute_derivatives(NSSpectralData *d) {
    if (!ready(d)) return;
    set_status(d, nsd_derivatives(d->device->ctx, (const double *)d->omega_hat,
                                  (double *)d->psi_hat, (double *)d->u_hat, (double *)d->v_hat,
                                  d->u, d->v, d->omega));
}

void ns_spectral_compute_nonlinear_terms(NSSpectralData *d) {
    if (!ready(d)) return;
    set_status(d, nsd_nonlinear(d->device->ctx, (const double *)d->omega_hat,
                                (const double *)d->psi_hat, (double *)d->nonlinear_hat));
}

/* ------------------------------------------------------------ time integration */

void ns_spectral_advance_rk4(NSSpectralData *d) {
    if (!ready(d)) return;
    int rc = ns_spectral_sync_device(d);
    if (rc == 0) rc = ns_spectral_advance_n(d, 1, 0);
    if (rc == 0) rc = ns_spectral_sync_host(d);
    set_status(d, rc);
}

double ns_spectral_compute_cfl_timestep(NSSpectralData *d) {
    if (!ready(d)) return MIN_DT;
    double r[3] = {0, 0, 0};
    int rc = nsd_phys_reduce(d->device->ctx, d->u, d->v, NULL, r);
    set_status(d, rc);
    /* spectral_time_integration.c:22-30 */
    const double max_vel = sqrt(r[0]);
    const double dx_min = fmin(d->Lx / d->nx, d->Ly / d->ny);
    const double dt_cfl = d->cfl_factor * dx_min / (max_vel + 1e-12);
    const double dt_visc = 0.5 * d->Re * dx_min * dx_min;
    double dt = fmin(dt_cfl, dt_visc);
    dt = fmax(dt, MIN_DT);
    dt = fmin(dt, MAX_DT);
    return dt;
}

/* ------------------------------------------------------------ diagnostics */

double ns_spectral_compute_energy(NSSpectralData *d) {
    if (!ready(d)) return 0.0;
    double r[3] = {0, 0, 0};
    set_status(d, nsd_phys_reduce(d->device->ctx, d->u, d->v, NULL, r));
    return r[1] / (double)R_of(d);
}

double ns_spectral_compute_enstrophy(NSSpectralData *d) {
    if (!ready(d)) return 0.0;
    double r[3] = {0, 0, 0};
    set_status(d, nsd_phys_reduce(d->device->ctx, NULL, NULL, d->omega, r));
    return r[2] / (double)R_of(d);
}

int ns_spectral_check_resolution(NSSpectralData *d) {
    if (!ready(d)) return 0;
    double s[5] = {0, 0, 0, 0, 0};
    if (set_status(d, nsd_spec_reduce(d->device->ctx, (const double *)d->omega_hat, s))) return 0;
    /* spectral_analysis.c:58-65 */
    const double frac = (s[0] > 1e-16) ? s[1] / s[0] : 0.0;
    if (frac > 1e-6) {
        printf("Warning: %.2e%% of energy in high wavenumbers\n", frac * 100);
        return 0;
    }
    return 1;
}

/* spectral_analysis.c:71-96: E and Z of the physical arrays, then over the
 * stored modes max|omega_hat| and the RMS of |omega_hat| where
 * k2 > 0.25 k2[last] (k2[last] is the last stored mode, quirk Q7) */
int ns_spectral_convergence_metrics(NSSpectralData *d, double out[6]) {
    if (!ready(d) || !out) return SOLVER_INVALID_PARAM;
    const double energy = ns_spectral_compute_energy(d);
    if (ns_spectral_last_status(d) != SOLVER_SUCCESS) return ns_spectral_last_status(d);
    const double enstrophy = ns_spectral_compute_enstrophy(d);
    if (ns_spectral_last_status(d) != SOLVER_SUCCESS) return ns_spectral_last_status(d);
    double s[5] = {0, 0, 0, 0, 0};
    int rc = nsd_spec_reduce(d->device->ctx, (const double *)d->omega_hat, s);
    if (rc) return set_status(d, rc);
    out[0] = energy;
    out[1] = enstrophy;
    out[2] = s[2];
    out[3] = (s[4] > 0.0) ? sqrt(s[3] / s[4]) : 0.0; /* spectral_analysis.c:96 */
    out[4] = d->t;
    out[5] = d->dt;
    return set_status(d, 0);
}

void ns_spectral_analyze_convergence(NSSpectralData *d, int step) {
    if (!ready(d) || step % 10 != 0) return;
    double m[6];
    if (ns_spectral_convergence_metrics(d, m)) return;
    /* spectral_analysis.c:98-99 */
    printf("Step %d: t=%.4f, E=%.6e, \xce\xa9=%.6e, max|\xcf\x89\xcc\x82|=%.2e, err=%.2e, dt=%.2e\n",
           step, m[4], m[0], m[1], m[2], m[3], m[5]);
    fflush(stdout);
}

void ns_spectral_analyze_spectrum(NSSpectralData *d) {
    if (!ready(d)) return;
    int max_k = (int)fmin(d->nkx, d->nky / 2);
    if (max_k <= 0) return;
    double *ek = (double *)calloc((size_t)max_k, sizeof(double));
    int *ck = (int *)calloc((size_t)max_k, sizeof(int));
    if (!ek || !ck) {
        free(ek);
        free(ck);
        set_status(d, SOLVER_MEM_ERROR);
        return;
    }
    int rc = nsd_spectrum(d->device->ctx, (const double *)d->u_hat, (const double *)d->v_hat, ek,
                          ck, max_k);
    set_status(d, rc);
    if (rc == 0) {
        /* spectral_analysis.c:125-139 */
        printf("\nEnergy Spectrum (t=%.4f):\n", d->t);
        printf("k\tE(k)\t\tSlope\n");
        for (int k = 1; k < max_k && ck[k] > 0; k++) {
            ek[k] /= ck[k];
            double slope = 0.0;
            if (k > 1 && ek[k - 1] > 1e-16 && ek[k] > 1e-16)
                slope = log(ek[k] / ek[k - 1]) / log((double)k / (k - 1));
            if (k <= 10 || k % 5 == 0) printf("%d\t%.2e\t%.2f\n", k, ek[k], slope);
        }
    }
    free(ek);
    free(ck);
}

/* ------------------------------------------------------------ hyperviscosity */

void ns_spectral_hyperviscosity_default(NSSpectralHyperviscosityConfig *c) {
    if (!c) return;
    /* hyperviscosity.c:20-26 */
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

void ns_spectral_hyperviscosity_set(NSSpectralData *d, const NSSpectralHyperviscosityConfig *c) {
    if (!d || !c) return;
    /* hyperviscosity.c:34-57 */
    d->hyperviscosity = *c;
    if (!d->hyperviscosity.enabled)
        memset(d->hyperviscosity_weights, 0, H_of(d) * sizeof(double));
    if (!d->hyperviscosity.sfd_enabled || d->hyperviscosity.sfd_strength <= 0.0) {
        d->hyperviscosity.sfd_enabled = 0;
        d->hyperviscosity.sfd_strength = 0.0;
        memset(d->sfd_hat, 0, H_of(d) * sizeof(fftw_complex));
    }
    ns_spectral_hyperviscosity_update_weights(d);
    if (d->hyperviscosity.sfd_enabled && d->sfd_hat && d->omega_hat)
        memcpy(d->sfd_hat, d->omega_hat, H_of(d) * sizeof(fftw_complex));
}

void ns_spectral_hyperviscosity_update_weights(NSSpectralData *d) {
    if (!ready(d) || !d->hyperviscosity_weights || !d->k2) return;
    NSSpectralHyperviscosityConfig *c = &d->hyperviscosity;
    double kmax = 0.0, kc = 0.0, invbw = 0.0;
    int rc = nsd_hv_weights(d->device->ctx, c->enabled, c->coefficient, c->order, c->cutoff_ratio,
                            d->hyperviscosity_weights, &kmax, &kc, &invbw);
    set_status(d, rc);
    if (rc == 0) {
        c->kmax = kmax;
        c->cutoff_wavenumber = kc;
        c->inv_bandwidth = invbw;
    }
}

void ns_spectral_hyperviscosity_apply(NSSpectralData *d, fftw_complex *rhs,
                                      const fftw_complex *state) {
    if (!ready(d) || !rhs || !state) return;
    const NSSpectralHyperviscosityConfig *c = &d->hyperviscosity;
    const int apply_visc = (c->enabled && c->coefficient > 0.0);
    const int apply_sfd = (c->sfd_enabled && c->sfd_strength > 0.0);
    if (!apply_visc && !apply_sfd) return; /* hyperviscosity.c:120-122 */
    int rc = sync_params(d);
    if (rc == 0)
        rc = nsd_hv_apply(d->device->ctx, (double *)rhs, (const double *)state,
                          apply_sfd ? (const double *)d->sfd_hat : NULL);
    set_status(d, rc);
}

void ns_spectral_hyperviscosity_post_step(NSSpectralData *d, const fftw_complex *state,
                                          double dt) {
    if (!ready(d) || !state || dt <= 0.0) return;
    const NSSpectralHyperviscosityConfig *c = &d->hyperviscosity;
    if (!c->sfd_enabled || c->sfd_strength <= 0.0 || c->sfd_timescale <= 0.0) return;
    int rc = sync_params(d);
    if (rc == 0)
        rc = nsd_hv_post_step(d->device->ctx, (const double *)state, (double *)d->sfd_hat, dt);
    set_status(d, rc);
}

/* ------------------------------------------------------------ output */

void ns_spectral_output_solution(NSSpectralData *d, const char *filename) {
    if (!ready(d) || !filename) return;
    FILE *fp = fopen(filename, "w");
    if (!fp) return;
    /* spectral_output.c:10-24 */
    fprintf(fp, "# Spectral Navier-Stokes Solution\n");
    fprintf(fp, "# x y u v omega psi\n");
    set_status(d, nsd_spectral_to_physical(d->device->ctx, (const double *)d->psi_hat, d->psi));
    for (int j = 0; j < d->ny; j++) {
        for (int i = 0; i < d->nx; i++) {
            const size_t idx = (size_t)j * d->nx + i;
            fprintf(fp, "%.8f %.8f %.8e %.8e %.8e %.8e\n", d->x[i], d->y[j], d->u[idx], d->v[idx],
                    d->omega[idx], d->psi[idx]);
        }
        fprintf(fp, "\n");
    }
    fclose(fp);
}

/* ------------------------------------------------------------ additions */

SolverStatus ns_spectral_last_status(const NSSpectralData *d) {
    if (!d || !d->device) return SOLVER_INVALID_PARAM;
    return d->device->status;
}

const char *ns_spectral_last_error(const NSSpectralData *d) {
    if (!d || !d->device) return "invalid instance";
    return d->device->err;
}

int ns_spectral_set_precision(NSSpectralData *d, NSPrecision precision) {
    if (!ready(d)) return SOLVER_INVALID_PARAM;
    if (precision != NS_PRECISION_FP64 && precision != NS_PRECISION_FP32)
        return set_status(d, SOLVER_INVALID_PARAM);
    if ((int)precision == d->device->precision) return set_status(d, 0);
    NsdContext *c = nsd_create(d->nx, d->ny, d->Lx, d->Ly,
                               precision == NS_PRECISION_FP32 ? NSD_FP32 : NSD_FP64,
                               d->device->err, sizeof(d->device->err));
    if (!c) {
        d->device->status = SOLVER_MEM_ERROR;
        return SOLVER_MEM_ERROR;
    }
    nsd_destroy(d->device->ctx);
    d->device->ctx = c;
    d->device->precision = precision;
    d->device->applied_valid = 0;
    return set_status(d, 0);
}

int ns_spectral_set_scheme(NSSpectralData *d, NSTimeScheme scheme) {
    if (!ready(d)) return SOLVER_INVALID_PARAM;
    if (scheme != NS_SCHEME_RK4 && scheme != NS_SCHEME_IF_RK4)
        return set_status(d, SOLVER_INVALID_PARAM);
    d->device->scheme = scheme;
    return set_status(d, 0);
}

int ns_spectral_set_drag(NSSpectralData *d, double mu) {
    if (!ready(d)) return SOLVER_INVALID_PARAM;
    if (!(mu >= 0.0)) return set_status(d, SOLVER_INVALID_PARAM);
    d->device->drag = mu;
    return set_status(d, 0);
}

int ns_spectral_set_forcing(NSSpectralData *d, const NSForcingConfig *cfg, double rate) {
    if (!ready(d)) return SOLVER_INVALID_PARAM;
    if (!cfg) {
        d->device->force_on = 0;
        return set_status(d, 0);
    }
    if (!(rate >= 0.0) || !(cfg->kmax >= cfg->kmin)) return set_status(d, SOLVER_INVALID_PARAM);
    d->device->force_cfg = *cfg;
    d->device->force_rate = rate;
    d->device->force_on = 1;
    return set_status(d, sync_params(d));
}

int ns_spectral_initialize_random_phase(NSSpectralData *d, unsigned long long seed, double k_peak,
                                        double energy) {
    if (!ready(d)) return SOLVER_INVALID_PARAM;
    int rc = sync_params(d);
    if (rc == 0) rc = nsd_state_random_phase(d->device->ctx, seed, k_peak, energy);
    if (rc == 0) rc = nsd_state_download(d->device->ctx, (double *)d->omega_hat);
    if (rc == 0)
        rc = nsd_state_derived(d->device->ctx, (double *)d->psi_hat, (double *)d->u_hat,
                               (double *)d->v_hat, d->u, d->v, d->omega);
    return set_status(d, rc);
}

int ns_spectral_sync_device(NSSpectralData *d) {
    if (!ready(d)) return SOLVER_INVALID_PARAM;
    int rc = sync_params(d);
    if (rc == 0) rc = nsd_state_upload(d->device->ctx, (const double *)d->omega_hat);
    if (rc == 0 && d->hyperviscosity.sfd_enabled)
        rc = nsd_sfd_upload(d->device->ctx, (const double *)d->sfd_hat);
    if (rc == 0) rc = nsd_set_time(d->device->ctx, d->t, d->dt);
    return set_status(d, rc);
}

int ns_spectral_advance_n(NSSpectralData *d, int nsteps, int adaptive) {
    if (!ready(d)) return SOLVER_INVALID_PARAM;
    if (nsteps < 0) return set_status(d, SOLVER_INVALID_PARAM);
    int rc = sync_params(d);
    if (rc == 0) rc = nsd_advance_graph(d->device->ctx, nsteps, adaptive);
    if (rc == 0) {
        int st = 0;
        rc = nsd_get_time(d->device->ctx, &d->t, &d->dt, &st);
        if (rc == 0 && st != 0) rc = st;
    }
    return set_status(d, rc);
}

int ns_spectral_sync_host(NSSpectralData *d) {
    if (!ready(d)) return SOLVER_INVALID_PARAM;
    int st = 0;
    int rc = nsd_get_time(d->device->ctx, &d->t, &d->dt, &st);
    if (rc == 0) rc = nsd_state_download(d->device->ctx, (double *)d->omega_hat);
    if (rc == 0)
        rc = nsd_state_derived(d->device->ctx, (double *)d->psi_hat, (double *)d->u_hat,
                               (double *)d->v_hat, d->u, d->v, d->omega);
    if (rc == 0 && d->hyperviscosity.sfd_enabled)
        rc = nsd_sfd_download(d->device->ctx, (double *)d->sfd_hat);
    if (rc == 0 && st != 0) rc = st;
    return set_status(d, rc);
}

int ns_spectral_set_error_control(NSSpectralData *d, const AdaptiveTimeParams *params,
                                  double rtol, double atol) {
    if (!ready(d)) return SOLVER_INVALID_PARAM;
    if (!params || !params->adapt_enabled) {
        d->device->errctl = 0;
        return set_status(d, sync_params(d));
    }
    if (!(rtol >= 0.0) || !(atol >= 0.0) || !(rtol > 0.0 || atol > 0.0) ||
        !(params->safety_factor > 0.0 && params->safety_factor <= 1.0) ||
        !(params->dt_min > 0.0) || !(params->dt_max >= params->dt_min) ||
        !(params->cfl_factor > 0.0)) {
        snprintf(d->device->err, sizeof(d->device->err),
                 "error control needs rtol, atol >= 0 (not both 0), safety in (0, 1], "
                 "0 < dt_min <= dt_max and cfl_factor > 0");
        d->device->status = SOLVER_INVALID_PARAM;
        return SOLVER_INVALID_PARAM;
    }
    d->cfl_factor = params->cfl_factor;
    d->device->errctl = 1;
    d->device->err_rtol = rtol;
    d->device->err_atol = atol;
    d->device->err_safety = params->safety_factor;
    d->device->dt_min = params->dt_min;
    d->device->dt_max = params->dt_max;
    int rc = sync_params(d);
    if (rc) d->device->errctl = 0;
    return set_status(d, rc);
}

int ns_spectral_error_control_stats(NSSpectralData *d, unsigned long long *accepted,
                                    unsigned long long *rejected, double *last_err,
                                    double *dt_next) {
    if (!ready(d)) return SOLVER_INVALID_PARAM;
    return set_status(d, nsd_errctl_stats(d->device->ctx, 0, accepted, rejected, last_err, dt_next));
}

int ns_spectral_device_diagnostics(NSSpectralData *d, double *energy, double *enstrophy) {
    if (!ready(d)) return SOLVER_INVALID_PARAM;
    double out[2] = {0, 0};
    int rc = nsd_state_energy(d->device->ctx, out);
    if (energy) *energy = out[0];
    if (enstrophy) *enstrophy = out[1];
    return set_status(d, rc);
}
