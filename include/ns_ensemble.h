/*
 * ns_ensemble.h -- ensembles of independent 2-D periodic vorticity solves of
 * different grid sizes, seeds and viscosities, advanced together on one GPU.
 *
 * No reference counterpart: the reference advances one NSSpectralData per
 * process (src/spectral/main_spectral.c:26-100).  Each member here follows the
 * same step as ns_spectral_advance_n with the integrating-factor scheme (or
 * the reference's classic RK4), adaptive dt at its own CFL limit, optional
 * del^8 hyperviscosity, linear drag and band forcing from its own seed.
 * Members of equal size share batched kernel launches; all sizes are advanced
 * by one captured hipGraph per step.
 */
#ifndef NS_ENSEMBLE_H
#define NS_ENSEMBLE_H

#include "ns_spectral_solver.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double Lx, Ly;              /* domain, default 2 pi x 2 pi                        */
    NSPrecision precision;      /* default NS_PRECISION_FP32                          */
    NSTimeScheme scheme;        /* default NS_SCHEME_IF_RK4                           */
    double cfl_factor;          /* default 0.5                                        */
    double drag;                /* linear drag mu, default 0                          */
    double nu4;                 /* del^8 coefficient: damping nu4 |k|^8, 0 = off      */
    int forcing_enabled;        /* default 0                                          */
    NSForcingConfig forcing;    /* shell, energy_target, type; the seed is per member */
    double injection_rate;      /* energy injection rate of the forcing               */
    double k_peak;              /* random-phase initial field: E(k) ~ k^4 e^-(k/k_peak)^2 */
    double energy;              /* ... normalised to this energy                       */
    int autotune;               /* choose kernel launch variants per size, default 1  */
} NSEnsembleConfig;

typedef struct NSEnsemble NSEnsemble;

void ns_ensemble_config_default(NSEnsembleConfig *cfg);
/* members fields; member m has grid n[m] x n[m] (power of two in [16, 8192]),
 * Reynolds number Re[m] and seed seeds[m] (initial field and forcing).
 * Returns NULL on failure (ns_ensemble_create_error() says why). */
NSEnsemble *ns_ensemble_create(int members, const int *n, const double *Re,
                               const unsigned long long *seeds, const NSEnsembleConfig *cfg);
const char *ns_ensemble_create_error(void);
void ns_ensemble_free(NSEnsemble *e);
int ns_ensemble_members(const NSEnsemble *e);
/* nsteps adaptive steps of every member; returns once queued */
int ns_ensemble_advance(NSEnsemble *e, int nsteps);
int ns_ensemble_synchronize(NSEnsemble *e);
/* per member: simulation time, energy, enstrophy (Parseval, FP64 sums) and
 * SolverStatus; any output may be NULL */
int ns_ensemble_diagnostics(NSEnsemble *e, double *t, double *energy, double *enstrophy,
                            int *status);
/* omega_hat of one member in the reference layout [n][n/2+1], interleaved FP64 */
int ns_ensemble_member_state(NSEnsemble *e, int member, double *omega_hat);
const char *ns_ensemble_last_error(const NSEnsemble *e);

#ifdef __cplusplus
}
#endif

#endif /* NS_ENSEMBLE_H */
