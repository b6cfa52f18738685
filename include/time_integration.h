/*
 * time_integration.h -- time-step helpers.
 * Replaces the reference's include/time_integration.h:7-27.  The three
 * helpers are declared there but defined nowhere in the reference; their
 * meaning is taken from docs/api_reference.md:178-193 and they are defined
 * in navier-stokes-solvers_amd/csrc/host/time_integration.c.
 */
#ifndef TIME_INTEGRATION_H
#define TIME_INTEGRATION_H

#include "common_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    TIME_EULER_EXPLICIT,
    TIME_EULER_IMPLICIT,
    TIME_CRANK_NICOLSON,
    TIME_RK4,
    TIME_ADAMS_BASHFORTH
} TimeScheme;

typedef struct {
    double cfl_factor;
    double dt_min;
    double dt_max;
    double safety_factor;
    int adapt_enabled;
} AdaptiveTimeParams;

/* advective limit cfl / (max_u/dx + max_v/dy)                 (api_reference.md:181-182) */
double compute_cfl_condition(double max_u, double max_v, double dx, double dy, double cfl_factor);
/* viscous limit 0.5 Re dx dy / (dx + dy)                       (api_reference.md:184-185) */
double compute_viscous_timestep(double dx, double dy, double Re);
/* safety * min(advective, viscous), clamped to [dt_min, dt_max] (api_reference.md:187-193) */
double adaptive_timestep(double max_vel, double dx, double dy, double Re, AdaptiveTimeParams *params);

#ifdef __cplusplus
}
#endif

#endif /* TIME_INTEGRATION_H */
