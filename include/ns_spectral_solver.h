/*
 * ns_spectral_solver.h -- 2-D periodic pseudospectral vorticity solver on
 * MI355X (gfx950).
 *
 * Drop-in replacement for the reference's src/spectral/ns_spectral_solver.h.
 * The struct keeps every reference field at the same position and meaning
 * (ns_spectral_solver.h:8-68); callers keep reading and writing the host
 * arrays (data->u, data->omega_hat, data->dt, ...) between calls.  Each
 * reference entry point (ns_spectral_solver.h:70-113) uploads the host
 * arrays it reads, runs on the GPU, and writes its results back into the
 * host arrays, so call sequences written for the reference behave the same.
 *
 * Differences a caller can observe:
 *   - nx and ny must be powers of two in [16, 8192] (allocate returns NULL
 *     otherwise) and a HIP device must be present (likewise);
 *   - ns_spectral_spectral_to_physical and ns_spectral_compute_nonlinear_terms
 *     do not leave intermediate values in nonlinear_hat, u, v, psi, omega, p
 *     as the reference's scratch reuse does (spectral_transforms.c:27-28,
 *     spectral_core.c:54-59); nonlinear_hat holds the result of
 *     compute_nonlinear_terms as in the reference;
 *   - ns_spectral_advance_rk4 derives psi_hat from omega_hat at every stage;
 *     the reference reads whatever psi_hat holds at stage 1
 *     (spectral_core.c:51,66,76), which agrees whenever compute_derivatives
 *     (or a previous step) last wrote it;
 *   - modes outside the 2/3 mask are not part of the device state: the
 *     reference keeps them at zero in every sequence its own
 *     initialisers start from, and so does this solver.
 *
 * fftw_complex is defined here without fftw3.h, with the layout of the
 * reference's (double _Complex: real and imaginary parts interleaved).
 */
#ifndef NS_SPECTRAL_SOLVER_H
#define NS_SPECTRAL_SOLVER_H

#include "common_types.h"
#include "forcing.h"
#include "ns_config.h"
#include "time_integration.h"

#ifdef __cplusplus
typedef double fftw_complex[2];
extern "C" {
#else
#include <complex.h>
typedef double _Complex fftw_complex;
#endif

typedef struct ns_fft_plan_s *fftw_plan; /* opaque; non-NULL once plans are "made" */

struct NSSpectralDevice;

/* ns_spectral_solver.h:8-19 */
typedef struct {
    int enabled;
    double coefficient;
    int order;
    double cutoff_ratio;
    double sfd_strength;
    double sfd_timescale;
    int sfd_enabled;
    double kmax;
    double cutoff_wavenumber;
    double inv_bandwidth;
} NSSpectralHyperviscosityConfig;

/* ns_spectral_solver.h:22-68 (fields in the reference order) */
typedef struct {
    int nx, ny;
    int nkx, nky;
    double Lx, Ly;
    double Re;
    double dt;
    double t;
    double cfl_factor;

    double *NS_RESTRICT u;
    double *NS_RESTRICT v;
    double *NS_RESTRICT p;
    double *NS_RESTRICT omega;
    double *NS_RESTRICT psi;
    double *NS_RESTRICT x;
    double *NS_RESTRICT y;

    fftw_complex *NS_RESTRICT u_hat;
    fftw_complex *NS_RESTRICT v_hat;
    fftw_complex *NS_RESTRICT omega_hat;
    fftw_complex *NS_RESTRICT psi_hat;
    fftw_complex *NS_RESTRICT nonlinear_hat;
    fftw_complex *NS_RESTRICT rk1_hat;
    fftw_complex *NS_RESTRICT rk2_hat;
    fftw_complex *NS_RESTRICT rk3_hat;
    fftw_complex *NS_RESTRICT rk4_hat;
    fftw_complex *NS_RESTRICT omega_tmp_hat;

    double *NS_RESTRICT kx;
    double *NS_RESTRICT ky;
    double *NS_RESTRICT k2;

    fftw_plan forward_plan, backward_plan;

    int *NS_RESTRICT dealias_mask;

    NSSpectralHyperviscosityConfig hyperviscosity;
    double *NS_RESTRICT hyperviscosity_weights;
    fftw_complex *NS_RESTRICT sfd_hat;

    /* appended: device state (no reference counterpart) */
    struct NSSpectralDevice *device;
} NSSpectralData;

/* ---- reference entry points (ns_spectral_solver.h:70-113) */

/* spectral_memory.c:6-68 */
NSSpectralData *ns_spectral_allocate(int nx, int ny, double Lx, double Ly, double Re);
/* spectral_memory.c:71-104 (NULL-safe; no global cleanup, instances are independent) */
void ns_spectral_free(NSSpectralData *data);

/* spectral_initialization.c:10-54 */
void ns_spectral_initialize_grids(NSSpectralData *data);
/* spectral_initialization.c:57-69 (no planning is needed; marks plans ready) */
void ns_spectral_initialize_fft_plans(NSSpectralData *data);
/* spectral_initialization.c:72-95 */
void ns_spectral_initialize_taylor_green(NSSpectralData *data);

/* spectral_transforms.c:9-22: r2c, x 1/(nx ny), dealias */
void ns_spectral_physical_to_spectral(NSSpectralData *data, double *field, fftw_complex *field_hat);
/* spectral_transforms.c:25-32: dealiased copy, unnormalised c2r */
void ns_spectral_spectral_to_physical(NSSpectralData *data, fftw_complex *field_hat, double *field);
/* spectral_transforms.c:35-43 */
void ns_spectral_apply_dealiasing(NSSpectralData *data, fftw_complex *field_hat);

/* spectral_core.c:37-43 */
void ns_spectral_compute_derivatives(NSSpectralData *data);
/* spectral_core.c:8-34 */
void ns_spectral_compute_velocity_from_vorticity(NSSpectralData *data);
/* spectral_core.c:46-112 */
void ns_spectral_compute_nonlinear_terms(NSSpectralData *data);

/* spectral_time_integration.c:34-106 (one step with data->dt) */
void ns_spectral_advance_rk4(NSSpectralData *data);
/* spectral_time_integration.c:9-31 */
double ns_spectral_compute_cfl_timestep(NSSpectralData *data);

/* spectral_analysis.c:8-19, 22-33, 69-100, 103-142, 36-66 */
double ns_spectral_compute_energy(NSSpectralData *data);
double ns_spectral_compute_enstrophy(NSSpectralData *data);
void ns_spectral_analyze_convergence(NSSpectralData *data, int step);
void ns_spectral_analyze_spectrum(NSSpectralData *data);
int ns_spectral_check_resolution(NSSpectralData *data);

/* hyperviscosity.c:15-27, 29-58, 60-102, 104-150, 152-178 */
void ns_spectral_hyperviscosity_default(NSSpectralHyperviscosityConfig *config);
void ns_spectral_hyperviscosity_set(NSSpectralData *data,
                                    const NSSpectralHyperviscosityConfig *config);
void ns_spectral_hyperviscosity_update_weights(NSSpectralData *data);
void ns_spectral_hyperviscosity_apply(NSSpectralData *data, fftw_complex *rhs_hat,
                                      const fftw_complex *state_hat);
void ns_spectral_hyperviscosity_post_step(NSSpectralData *data, const fftw_complex *state_hat,
                                          double dt);

/* spectral_output.c:4-27 */
void ns_spectral_output_solution(NSSpectralData *data, const char *filename);

/* ---- additions (no reference counterpart) */

typedef enum { NS_PRECISION_FP64 = 0, NS_PRECISION_FP32 = 1 } NSPrecision;
typedef enum { NS_SCHEME_RK4 = 0, NS_SCHEME_IF_RK4 = 1 } NSTimeScheme;

/* status of the last call on this instance (the reference's spectral calls
 * are void and report nothing; SOLVER_DIVERGED after a non-finite field) */
SolverStatus ns_spectral_last_status(const NSSpectralData *data);
const char *ns_spectral_last_error(const NSSpectralData *data);

/* arithmetic precision of the device state and transforms (default FP64);
 * changing it discards the device state (sync_device uploads it again) */
int ns_spectral_set_precision(NSSpectralData *data, NSPrecision precision);
/* RK4 with explicit linear terms (reference) or integrating-factor RK4 for
 * stiff hyperviscosity (Lawson); default NS_SCHEME_RK4 */
int ns_spectral_set_scheme(NSSpectralData *data, NSTimeScheme scheme);
/* linear drag: d omega/dt += -mu omega */
int ns_spectral_set_drag(NSSpectralData *data, double mu);
/* white/coloured forcing of forcing.h applied each step as
 * omega_hat += sqrt(dt) f, f made Hermitian on the kx = 0 column and scaled
 * to the energy injection rate `injection_rate`; cfg == NULL disables */
int ns_spectral_set_forcing(NSSpectralData *data, const NSForcingConfig *cfg,
                            double injection_rate);
/* random-phase vorticity with E(k) ~ k^4 exp(-(k/k_peak)^2), 1/2<u^2+v^2> = energy,
 * phases from the counter SplitMix64 stream of `seed`; ends with
 * compute_derivatives like the canonical init */
int ns_spectral_initialize_random_phase(NSSpectralData *data, unsigned long long seed,
                                        double k_peak, double energy);

/* device-resident fast path: upload omega_hat, t, dt (and sfd_hat) once,
 * advance on the device, download when needed */
int ns_spectral_sync_device(NSSpectralData *data);
/* nsteps RK4 steps; adaptive != 0 recomputes dt = cfl_timestep before every
 * step (as main_spectral.c:46-48 does), else every step uses data->dt */
int ns_spectral_advance_n(NSSpectralData *data, int nsteps, int adaptive);
/* omega_hat, psi_hat, u_hat, v_hat, u, v, omega, sfd_hat, t, dt <- device */
int ns_spectral_sync_host(NSSpectralData *data);
/* energy and enstrophy of the device state (Parseval, FP64 sums) */
int ns_spectral_device_diagnostics(NSSpectralData *data, double *energy, double *enstrophy);
/* Adaptive-dt error controller for ns_spectral_advance_n(..., adaptive = 1),
 * with the parameters the reference declares and never uses
 * (AdaptiveTimeParams, include/time_integration.h:16-22): every step takes
 * dt = clamp(min(cfl rule with params->cfl_factor, dt_err), dt_min, dt_max)
 * and computes, in FP64, err = |y4 - y2| / (atol + rtol |y4|) from the RK4
 * solution y4 and the midpoint solution y2 its own second stage gives
 * (Lawson midpoint with the integrating factor); err > 1 rejects the step
 * (state and t return to the step before, unless dt is at dt_min) and the
 * next steps may use dt_err = dt clamp(safety_factor err^(-1/3), 0.2, 5).
 * adapt_enabled = 0 (or params == NULL) switches it off.  Not with
 * selective frequency damping. */
int ns_spectral_set_error_control(NSSpectralData *data, const AdaptiveTimeParams *params,
                                  double rtol, double atol);
/* judged steps accepted and rejected since sync_device, the last error norm
 * and the step it allows next */
int ns_spectral_error_control_stats(NSSpectralData *data, unsigned long long *accepted,
                                    unsigned long long *rejected, double *last_err,
                                    double *dt_next);
/* the numbers ns_spectral_analyze_convergence prints (spectral_analysis.c:69-100),
 * from the host arrays: out = {E, Z, max|omega_hat|, spectral error (RMS of
 * |omega_hat| over the modes with k^2 > k2[last] / 4), t, dt} */
int ns_spectral_convergence_metrics(NSSpectralData *data, double out[6]);

/* apply ns_config keys (fluid.Re, time.dt, spectral.*, forcing.*; see the
 * key table in ns_config.c) to an allocated instance */
int ns_spectral_configure(NSSpectralData *data, const NSConfig *config);
/* allocate + grids + plans from grid.nx, grid.ny, domain.Lx, domain.Ly,
 * fluid.Re, then ns_spectral_configure */
NSSpectralData *ns_spectral_create_from_config(const NSConfig *config);

#ifdef __cplusplus
}
#endif

#endif /* NS_SPECTRAL_SOLVER_H */
