/*
 * forcing.h -- band-limited stochastic forcing generator.
 * Same API and semantics as the reference's src/common/forcing.h:7-46:
 * SplitMix64 draws, Box-Muller pairs, white or Ornstein-Uhlenbeck coloured
 * noise on the shell kmin <= |k| <= kmax, rescaled so that
 * sum |f_k|^2 = energy_target.  ns_forcing_generate runs on the GPU through
 * the counter form of the SplitMix64 stream.  Two fields are appended after
 * the reference's: the active-mode index list and the generate-call counter.
 */
#ifndef NS_COMMON_FORCING_H
#define NS_COMMON_FORCING_H

#include "common_types.h"
#ifndef __cplusplus
#include <complex.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    NS_FORCING_WHITE_NOISE = 0,
    NS_FORCING_COLORED_NOISE = 1
} NSForcingType;

typedef struct {
    double kmin;
    double kmax;
    double energy_target;
    double correlation_time;
    unsigned long seed;
    NSForcingType type;
} NSForcingConfig;

typedef struct {
    NSForcingConfig config;
    int nkx;
    int nky;
    int spectral_size;
    int active_modes;
    double *weights;
#ifdef __cplusplus
    double *colored_state;
#else
    double complex *colored_state;
#endif
    unsigned int have_spare;
    double spare_gaussian;
    unsigned long long rng_state;
    /* appended (no reference counterpart) */
    int *active_index;            /* reference index j*nkx+i of each active mode */
    unsigned long long calls;     /* completed ns_forcing_generate calls          */
} NSForcingState;

void ns_forcing_config_default(NSForcingConfig *config);                     /* forcing.h:34 */
int ns_forcing_state_init(NSForcingState *state, const NSForcingConfig *config,
                          const double *k2, int nkx, int nky);                /* forcing.h:36-39 */
void ns_forcing_state_free(NSForcingState *state);                           /* forcing.h:41 */
#ifdef __cplusplus
void ns_forcing_generate(NSForcingState *state, double dt, const double *k2, double *forcing_out);
#else
void ns_forcing_generate(NSForcingState *state, double dt, const double *k2,
                         double complex *forcing_out);                        /* forcing.h:43-46 */
#endif

#ifdef __cplusplus
}
#endif

#endif /* NS_COMMON_FORCING_H */
