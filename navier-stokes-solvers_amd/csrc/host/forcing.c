/*
 * forcing.c -- band-limited stochastic forcing (include/forcing.h).
 *
 * Semantics of the reference's src/common/forcing.c:41-184.  The draws run
 * on the GPU (nsd_forcing_generate) through the counter form of the
 * SplitMix64 stream: the k-th output of a generator seeded with s is
 * mix(s + k * 0x9E3779B97F4A7C15), so call c (0-based) with A active modes
 * gives mode rank a the uniforms 2(cA+a)+1 and 2(cA+a)+2 -- exactly the pair
 * the reference's sequential Box-Muller consumes for that mode.  The host
 * keeps rng_state advanced as the reference's would be.
 */
#include "forcing.h"

#include <stdint.h>

#include "ns_device.h"

#define SPLITMIX_GAMMA 0x9E3779B97F4A7C15ULL

void ns_forcing_config_default(NSForcingConfig *c) {
    if (!c) return;
    /* forcing.c:45-50 */
    c->kmin = 0.0;
    c->kmax = 4.0;
    c->energy_target = 1.0;
    c->correlation_time = 0.5;
    c->seed = 0;
    c->type = NS_FORCING_WHITE_NOISE;
}

static unsigned long long effective_seed(const NSForcingConfig *c) {
    return c->seed ? (unsigned long long)c->seed : 0x123456789abcdefULL; /* forcing.c:99 */
}

int ns_forcing_state_init(NSForcingState *s, const NSForcingConfig *c, const double *k2, int nkx,
                          int nky) {
    if (!s || !c || !k2 || nkx <= 0 || nky <= 0) return -1;
    memset(s, 0, sizeof(*s));
    s->config = *c;
    s->nkx = nkx;
    s->nky = nky;
    s->spectral_size = nkx * nky;
    s->weights = (double *)calloc((size_t)s->spectral_size, sizeof(double));
    s->active_index = (int *)calloc((size_t)s->spectral_size, sizeof(int));
    if (!s->weights || !s->active_index) {
        ns_forcing_state_free(s);
        return -1;
    }
    if (c->type == NS_FORCING_COLORED_NOISE) {
        s->colored_state = calloc((size_t)s->spectral_size, 2 * sizeof(double));
        if (!s->colored_state) {
            ns_forcing_state_free(s);
            return -1;
        }
    }
    /* shell selection, forcing.c:75-93 */
    const double kmin = fmax(c->kmin, 0.0);
    const double kmax = fmax(c->kmax, kmin);
    int active = 0;
    for (int idx = 0; idx < s->spectral_size; ++idx) {
        const double km = sqrt(k2[idx]);
        const int on = (km >= kmin && km <= kmax);
        s->weights[idx] = on ? 1.0 : 0.0;
        if (on) s->active_index[active++] = idx;
    }
    s->active_modes = active;
    s->have_spare = 0;
    s->rng_state = effective_seed(c);
    s->calls = 0;
    return active == 0 ? -1 : 0;
}

void ns_forcing_state_free(NSForcingState *s) {
    if (!s) return;
    free(s->weights);
    s->weights = NULL;
    free(s->colored_state);
    s->colored_state = NULL;
    free(s->active_index);
    s->active_index = NULL;
    s->spectral_size = 0;
    s->active_modes = 0;
}

void ns_forcing_generate(NSForcingState *s, double dt, const double *k2,
                         double complex *forcing_out) {
    if (!s || !forcing_out || !k2) return;
    const int size = s->spectral_size;
    const NSForcingConfig *c = &s->config;
    if (size == 0 || s->active_modes == 0) {
        memset(forcing_out, 0, (size_t)size * sizeof(double complex));
        return;
    }
    if (c->energy_target <= 0.0) { /* forcing.c:128-134 */
        memset(forcing_out, 0, (size_t)size * sizeof(double complex));
        if (s->colored_state) memset(s->colored_state, 0, (size_t)size * sizeof(double complex));
        return;
    }
    const int colored = (c->type == NS_FORCING_COLORED_NOISE);
    double alpha = 0.0, beta = 1.0;
    if (colored) { /* forcing.c:140-143 */
        const double tau = fmax(c->correlation_time, 1e-6);
        alpha = exp(-dt / tau);
        beta = sqrt(fmax(1.0 - alpha * alpha, 0.0));
    }
    char err[128];
    const int rc = nsd_forcing_generate(size, s->active_index, s->active_modes, effective_seed(c),
                                        s->calls, colored, alpha, beta, c->energy_target,
                                        (double *)s->colored_state, (double *)forcing_out, err,
                                        sizeof(err));
    if (rc != 0) {
        fprintf(stderr, "ns_forcing_generate: %s\n", err);
        memset(forcing_out, 0, (size_t)size * sizeof(double complex));
        return;
    }
    s->calls++;
    s->rng_state += 2ULL * (unsigned long long)s->active_modes * SPLITMIX_GAMMA;
}
