/*
 * ns_oracle.h -- FP64 CPU oracle for the Fourier pseudospectral vorticity step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library (libnsspectral.so)
 * never links, loads or calls it.
 *
 * It restates, in plain C11 with a project-written radix-2 FFT, the
 * algorithm of the reference's src/spectral (all .c files) and src/common/forcing.c.
 * Every function cites the reference file:line it follows.  Extensions
 * that have no reference counterpart (integrating-factor RK4, linear drag,
 * forcing application, random-phase initial condition) are marked
 * "build-defined".
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   FFT r2c/c2r      numpy.fft.rfft2/irfft2 (pocketfft) + brute-force DFT at 8x8
 *   mask, grids      closed-form kept-mode counts; test_solvers.c:450-471,513-545
 *   velocity         single-mode closed forms (sign quirk psi_hat = -omega_hat/k^2)
 *   nonlinear term   two-mode analytic Jacobian evaluated pointwise on the grid
 *   classic RK4      linear decay omega_n = R(z)^n omega_0, R = 1+z+z^2/2+z^3/6+z^4/24
 *   IF-RK4           linear decay exp(L t) exactly; Lawson stage algebra vs classic
 *                    RK4 on the transformed variable at small N
 *   CFL              Taylor-Green closed form (umax = 1)
 *   hyperviscosity   closed-form weights; test_solvers.c:616-664
 *   forcing draws    bit-exact against the reference forcing.c compiled into
 *                    oracle/_ref/ and against committed golden vectors
 *   energy/enstrophy Taylor-Green closed forms E = 0.25 e^{-4 nu t}, Z = 0.5 e^{-4 nu t}
 *   spectrum         shell sums vs numpy binning of the same modes
 *   RK4, generic J   the whole step on a J != 0 field: one-step error against 64
 *                    finer steps falls as dt^5 (ratio 32 per halving; a stage
 *                    built from the wrong field gives 8); inviscid, energy and
 *                    enstrophy change by the truncation error only
 *
 * No compiled reference step exists to compare the whole step with (the
 * reference's spectral path needs FFTW, absent here); the properties above and
 * the test that the minimal step equals the reference's own 40-transform call
 * sequence pin it instead.
 */
#ifndef NS_ORACLE_H
#define NS_ORACLE_H

#include <complex.h>
#include <stdint.h>

typedef double _Complex oc_complex;

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
} OcHyperviscosity;

typedef struct OcFFT OcFFT;

typedef struct {
    int nx, ny, nkx, nky;
    double Lx, Ly, Re, dt, t, cfl_factor;
    double *u, *v, *p, *omega, *psi, *x, *y;
    oc_complex *u_hat, *v_hat, *omega_hat, *psi_hat, *nonlinear_hat;
    oc_complex *rk1_hat, *rk2_hat, *rk3_hat, *rk4_hat, *omega_tmp_hat;
    double *kx, *ky, *k2;
    int *dealias_mask;
    OcHyperviscosity hv;
    double *hv_weights;
    oc_complex *sfd_hat;
    OcFFT *fft;
    /* build-defined extension parameters */
    double drag;          /* linear drag mu: d omega/dt += -mu omega            */
    long long fft_count;  /* number of 2-D transforms executed (for timing) */
} OcData;

/* forcing (src/common/forcing.h:7-32) */
typedef struct {
    double kmin, kmax, energy_target, correlation_time;
    unsigned long seed;
    int type; /* 0 white, 1 colored */
} OcForcingConfig;

typedef struct {
    OcForcingConfig config;
    int nkx, nky, spectral_size, active_modes;
    double *weights;
    oc_complex *colored_state;
    unsigned int have_spare;
    double spare_gaussian;
    unsigned long long rng_state;
} OcForcingState;

/* memory / init */
OcData *oc_allocate(int nx, int ny, double Lx, double Ly, double Re);
void oc_free(OcData *d);
void oc_initialize_grids(OcData *d);
void oc_initialize_taylor_green(OcData *d);

/* transforms */
void oc_physical_to_spectral(OcData *d, double *field, oc_complex *field_hat);
void oc_spectral_to_physical(OcData *d, oc_complex *field_hat, double *field);
void oc_apply_dealiasing(OcData *d, oc_complex *field_hat);
void oc_fft_r2c_2d(OcData *d, const double *in, oc_complex *out);  /* unnormalised */
void oc_fft_c2r_2d(OcData *d, const oc_complex *in, double *out);  /* unnormalised */

/* core */
void oc_compute_velocity_from_vorticity(OcData *d);
void oc_compute_derivatives(OcData *d);
void oc_compute_nonlinear_terms(OcData *d);
void oc_advance_rk4(OcData *d);                 /* reference sequence, 40 transforms */
void oc_advance_rk4_minimal(OcData *d);         /* same arithmetic, 20 transforms + 3 */
double oc_compute_cfl_timestep(OcData *d);

/* diagnostics */
double oc_compute_energy(OcData *d);
double oc_compute_enstrophy(OcData *d);
int oc_check_resolution(OcData *d, double *high_k_fraction);
void oc_analyze_convergence(OcData *d, double out[6]);
int oc_analyze_spectrum(OcData *d, double *energy_k, int *count_k, int max_bins);

/* hyperviscosity */
void oc_hyperviscosity_default(OcHyperviscosity *c);
void oc_hyperviscosity_set(OcData *d, const OcHyperviscosity *c);
void oc_hyperviscosity_update_weights(OcData *d);
void oc_hyperviscosity_apply(OcData *d, oc_complex *rhs, const oc_complex *state);
void oc_hyperviscosity_post_step(OcData *d, const oc_complex *state, double dt);

/* forcing */
uint64_t oc_splitmix64_next(uint64_t *state);
double oc_counter_uniform(uint64_t seed, uint64_t k);
void oc_forcing_config_default(OcForcingConfig *c);
int oc_forcing_state_init(OcForcingState *s, const OcForcingConfig *c, const double *k2, int nkx, int nky);
void oc_forcing_state_free(OcForcingState *s);
void oc_forcing_generate(OcForcingState *s, double dt, const double *k2, oc_complex *out);

/* build-defined extensions */
void oc_forcing_project_and_scale(OcData *d, oc_complex *f, double injection_rate);
void oc_advance_if_rk4(OcData *d, const oc_complex *forcing_scaled);
void oc_advance_classic_ext(OcData *d, const oc_complex *forcing_scaled);
void oc_initialize_random_phase(OcData *d, uint64_t seed, double k_peak, double energy);
/* the same two steps returning the embedded error estimate of the
 * adaptive-dt controller: norms = {sum w |y4 - y2|^2, sum w |y4|^2} */
void oc_advance_if_rk4_est(OcData *d, const oc_complex *forcing_scaled, double norms[2]);
void oc_advance_classic_ext_est(OcData *d, const oc_complex *forcing_scaled, double norms[2]);

#endif
