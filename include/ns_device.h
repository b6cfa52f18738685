/*
 * ns_device.h -- thin C-ABI between the C host library and the HIP kernels.
 *
 * Every function takes plain pointers and sizes.  Spectral arrays passed
 * here are interleaved complex FP64 in the reference layout [ny][nx/2+1]
 * (fftw_complex == double _Complex, ns_spectral_solver.h:5-6); physical
 * arrays are FP64 [ny][nx].  All functions return 0 on success and a
 * non-zero SolverStatus value on failure (nsd_error() holds the message);
 * none of them aborts the process.  One context owns one HIP stream; calls
 * that return host data synchronise that stream.
 *
 * These entry points are the device half of the reference's spectral API
 * (src/spectral/ns_spectral_solver.h:70-113); the mapping of each public
 * ns_spectral_* function onto them is in the host library
 * (navier-stokes-solvers_amd/csrc/host/spectral_api.c).
 */
#ifndef NS_DEVICE_H
#define NS_DEVICE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct NsdContext NsdContext;

enum { NSD_FP64 = 0, NSD_FP32 = 1 };
enum { NSD_SCHEME_CLASSIC = 0, NSD_SCHEME_IF = 1 };

typedef struct {
    double Re;
    double drag;            /* linear drag mu (build-defined extension)            */
    double cfl_factor;      /* ns_spectral_solver.h:29                              */
    int scheme;             /* NSD_SCHEME_*                                          */
    /* hyperviscosity.c:60-102 parameterisation, after update_weights            */
    int hv_on;
    int hv_order;
    double hv_coeff, hv_kc, hv_invbw;
    int sfd_on;
    double sfd_strength, sfd_tau;
    /* forcing.h:12-19 plus the build-defined injection rate                     */
    int force_on;
    int force_colored;
    double force_kmin, force_kmax;
    double force_energy_target;
    double force_tau;
    double force_rate;      /* energy injection rate of omega_hat += sqrt(dt) f   */
    unsigned long long force_seed;
    /* adaptive-dt error controller (build-defined; AdaptiveTimeParams of
     * include/time_integration.h:16-22): with errctl on, adaptive steps take
     * dt = clamp(min(CFL dt, dt_err), dt_min, dt_max), where dt_err follows
     * an embedded estimate; a step whose error exceeds the tolerance
     * err_atol + err_rtol |omega_hat| is rejected and taken again (see
     * nsd_errctl_stats).  Not with selective frequency damping. */
    int errctl;
    double err_rtol, err_atol;
    double err_safety;      /* AdaptiveTimeParams.safety_factor, in (0, 1]       */
    double dt_min, dt_max;  /* AdaptiveTimeParams.dt_min, dt_max                 */
} NsdParams;

NsdContext *nsd_create(int nx, int ny, double Lx, double Ly, int precision, char *err,
                       size_t errlen);
void nsd_destroy(NsdContext *c);
const char *nsd_error(const NsdContext *c);
int nsd_set_params(NsdContext *c, const NsdParams *p);
int nsd_precision(const NsdContext *c);

/* ---- device-resident state (fast path) */
int nsd_state_upload(NsdContext *c, const double *omega_hat);
int nsd_state_download(NsdContext *c, double *omega_hat);
int nsd_sfd_upload(NsdContext *c, const double *sfd_hat);
int nsd_sfd_download(NsdContext *c, double *sfd_hat);
/* random-phase initial vorticity scaled to `energy` (the same field for the
 * same seed on every call) */
int nsd_state_random_phase(NsdContext *c, unsigned long long seed, double k_peak, double energy);
int nsd_set_time(NsdContext *c, double t, double dt);
int nsd_get_time(NsdContext *c, double *t, double *dt, int *status);
int nsd_set_forcing_counter(NsdContext *c, unsigned long long calls);
int nsd_get_forcing_counter(NsdContext *c, unsigned long long *calls);
/* coloured-forcing (OU) state of member 0: 2 * nsd_forcing_modes() doubles,
 * interleaved complex in the active-mode order of forcing.c:80-93;
 * to_device != 0 uploads, else downloads */
int nsd_forcing_ou(NsdContext *c, double *ou, int count, int to_device);
int nsd_forcing_modes(const NsdContext *c);
/* nsteps steps (adaptive != 0: CFL dt, and the error controller if on;
 * with the controller each call ends with the last step judged, so a
 * rejected step leaves the state and t of the step before it) */
int nsd_advance(NsdContext *c, int nsteps, int adaptive);
int nsd_advance_graph(NsdContext *c, int nsteps, int adaptive);
/* error controller of member `member`: judged steps accepted and rejected,
 * the last judged error norm (1 = the tolerance) and the step it allows next */
int nsd_errctl_stats(NsdContext *c, int member, unsigned long long *accepted,
                     unsigned long long *rejected, double *last_err, double *dt_next);
int nsd_state_derived(NsdContext *c, double *psi_hat, double *u_hat, double *v_hat, double *u,
                      double *v, double *omega);
/* out = {E, Z} from the spectral state by Parseval, FP64 accumulation in a
 * fixed order (the same bits on every call for the same state) */
int nsd_state_energy(NsdContext *c, double out[2]);
int nsd_synchronize(NsdContext *c);

/* timing of the fast path: average duration (ms) of each kernel family over
 * the last nsd_advance call made with profiling on */
int nsd_profile(NsdContext *c, int on);
int nsd_profile_read(NsdContext *c, double *row_ms, double *col_ms, double *step_ms,
                     int *steps);

/* copy `bytes` between two fresh device buffers `reps` times with a plain
 * streaming kernel (k_copy); *ms = average duration of one copy.  Used to
 * measure attainable bandwidth and to calibrate HBM byte counters. */
int nsd_bandwidth_probe(NsdContext *c, size_t bytes, int reps, double *ms);

/* ---- batches (ensembles).  No reference counterpart: the reference solves
 * one field per process.  A batch context holds `members` independent fields
 * of one grid size; every time-step kernel runs all of them in one launch
 * (members along the grid's y dimension), each with its own t, dt, CFL
 * reduction, 1/Re and forcing stream.  nsd_advance / nsd_advance_graph
 * advance every member; nsd_state_upload/download address member 0;
 * the compat operations below refuse batch contexts. */
NsdContext *nsd_create_batch(int nx, int ny, double Lx, double Ly, int precision, int members,
                             char *err, size_t errlen);
int nsd_batch_members(const NsdContext *c);
/* per-member Re (NULL: keep the NsdParams value) and forcing seeds (NULL:
 * keep NsdParams.force_seed; 0 maps to 0x123456789abcdef as forcing.c:99) */
int nsd_set_member_params(NsdContext *c, const double *Re, const unsigned long long *force_seed);
/* random-phase initial field of every member from its own seed */
int nsd_batch_random_phase(NsdContext *c, const unsigned long long *seeds, double k_peak,
                           double energy);
/* E[m], Z[m] (Parseval, FP64 sums); t[m], dt[m], status[m]; either may be NULL */
int nsd_batch_energy(NsdContext *c, double *E, double *Z);
int nsd_batch_scalars(NsdContext *c, double *t, double *dt, int *status);
int nsd_member_upload(NsdContext *c, int member, const double *omega_hat);
int nsd_member_download(NsdContext *c, int member, double *omega_hat);

/* Launch variants of the two time-step kernels (workgroup size and whether
 * inverse transforms run in lock-step pairs; lengths 64..1024 have several).
 * nsd_autotune times every variant on this context's shape (reps launches
 * each) and keeps the fastest; results are cached per (length, precision,
 * batch members) for the process and, if the environment names a file in
 * NSD_TUNE_FILE (lines "row|col L precision members variant"),
 * across processes (force != 0 re-times).  It zeroes the state: call it
 * before setting the initial field.  Every variant gives identical results. */
int nsd_autotune(NsdContext *c, int reps, int force, int *row_variant, int *col_variant);
/* Staggered start of the 4096-point FP32 row stage: the second resident
 * workgroup of each compute unit starts about 5 us late, so that the two
 * workgroups of a CU reach their gathers at different times.  It is on when a
 * probe launch at context creation shows that the device gives every CU one
 * workgroup before any gets a second (the dispatch order the delay assumes).
 * set = 0 turns it off, 1 turns it on if that check passed, -1 only queries;
 * returns the state (1 on, 0 off), -1 on error.  Results do not depend on it. */
int nsd_row_stagger(NsdContext *c, int set);
int nsd_set_variants(NsdContext *c, int row_variant, int col_variant);

/* One captured hipGraph that advances several batch contexts (different grid
 * sizes, same precision) by one step per launch; each context's kernels run
 * on its own stream inside the graph, so the sizes overlap on the device.
 * The graph is re-captured when a context's parameters change. */
typedef struct NsdMulti NsdMulti;
NsdMulti *nsd_multi_create(NsdContext *const *contexts, int n);
void nsd_multi_destroy(NsdMulti *m);
const char *nsd_multi_error(const NsdMulti *m);
int nsd_multi_advance(NsdMulti *m, int nsteps, int adaptive);
int nsd_multi_synchronize(NsdMulti *m);

/* ---- learned subgrid closure (ns_closure.h); no reference counterpart.
 * nsd_attach_closure: every step of this context (FP32 fields on a square
 * grid of the network's size; a batch context needs a network batch of at
 * least its members) first evaluates the network on each member's vorticity
 * and adds its tendency to the nonlinear term of all four stages; NULL
 * detaches.  nsd_subgrid_sample: from member `member` of a fine context and
 * a coarse single-field context of the same domain, the coarse vorticity
 * (the fine modes the coarse 2/3 mask keeps) and the subgrid tendency
 * filter(N_fine) - N_coarse(filtered vorticity) as physical FP32 fields
 * (device pointers x, t); the nonlinear fields of all members are computed
 * when member == 0, so call members in order.  nsd_subgrid_dataset: advance
 * the fine context `stride` steps, sample every member into the network's
 * data set from slot `first` on, `snapshots` times. */
struct NsdClosure;
int nsd_attach_closure(NsdContext *c, struct NsdClosure *net);
int nsd_subgrid_sample(NsdContext *fine, int member, NsdContext *coarse, float *x, float *t);
int nsd_subgrid_dataset(struct NsdClosure *net, NsdContext *fine, NsdContext *coarse,
                        int snapshots, int stride, int first);

/* ---- compat operations on host arrays (reference layouts, FP64) */
int nsd_grids(NsdContext *c, double *x, double *y, double *kx, double *ky, double *k2, int *mask);
int nsd_hv_weights(NsdContext *c, int enabled, double coefficient, int order, double cutoff_ratio,
                   double *weights, double *kmax, double *kc, double *invbw);
int nsd_physical_to_spectral(NsdContext *c, const double *field, double *field_hat);
int nsd_spectral_to_physical(NsdContext *c, const double *field_hat, double *field);
int nsd_dealias(NsdContext *c, double *field_hat);
int nsd_velocity(NsdContext *c, const double *omega_hat, double *psi_hat, double *u_hat,
                 double *v_hat);
int nsd_derivatives(NsdContext *c, const double *omega_hat, double *psi_hat, double *u_hat,
                    double *v_hat, double *u, double *v, double *omega);
int nsd_nonlinear(NsdContext *c, const double *omega_hat, const double *psi_hat,
                  double *nonlinear_hat);
int nsd_hv_apply(NsdContext *c, double *rhs, const double *state, const double *sfd);
int nsd_hv_post_step(NsdContext *c, const double *state, double *sfd, double dt);
int nsd_taylor_green(NsdContext *c, double *u, double *v, double *omega, double *p);
/* out = {max(u^2+v^2), sum 1/2(u^2+v^2), sum 1/2 omega^2} over R points */
int nsd_phys_reduce(NsdContext *c, const double *u, const double *v, const double *omega,
                    double out[3]);
/* out = {sum|w|^2, sum|w|^2 over sqrt(k2)/max_k > 0.8, max|w|,
 *        sum|w|^2 over k2 > 0.25 k2[last], count of those modes} */
int nsd_spec_reduce(NsdContext *c, const double *omega_hat, double out[5]);
int nsd_spectrum(NsdContext *c, const double *u_hat, const double *v_hat, double *energy_k,
                 int *count_k, int max_k);
/* forcing.c:117-184 semantics for one generate call on `calls` previous
 * calls: active-mode list in reference order, out (H complex) overwritten,
 * colored (H complex, may be NULL for white noise) updated in place.  It
 * needs no solver context (NSForcingState carries none), so it runs on the
 * null stream of the current device. */
int nsd_forcing_generate(int spectral_size, const int *active_idx, int nactive,
                         unsigned long long seed_eff, unsigned long long calls, int colored_on,
                         double alpha, double beta, double energy_target, double *colored,
                         double *out, char *err, size_t errlen);

/* ---- allocation-free forward transform.  No counterpart in the reference;
 * for callers that cannot allocate (code under HIP graph capture, serving
 * loops that forbid hipMalloc).  The same transform as
 * nsd_physical_to_spectral -- r2c, x 1/(nx ny), 2/3 mask
 * (spectral_transforms.c:9-22) -- on device arrays: `field` FP64 [ny][nx],
 * `field_hat` FP64 interleaved complex [ny][nx/2+1], queued on the HIP
 * stream `stream` (NULL: the null stream) of the current device, which must
 * be synchronised before reading `field_hat`.  Needs no context and calls no
 * allocator: its intermediate field and twiddles live in one 6 MiB device
 * array per GPU, which bounds nx and ny to powers of two in [16, 1024]
 * (larger sizes return SOLVER_INVALID_PARAM with a message in `err`).  Calls
 * that overlap on one GPU share that array and race. */
int nsd_physical_to_spectral_noalloc(int nx, int ny, const double *field, double *field_hat,
                                     void *stream, char *err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif
