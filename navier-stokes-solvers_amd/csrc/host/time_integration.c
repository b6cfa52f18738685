/*
 * time_integration.c -- the three helpers of include/time_integration.h.
 * The reference declares them (include/time_integration.h:24-26) but never
 * defines them; docs/api_reference.md:178-193 gives their meaning.
 */
#include "time_integration.h"

double compute_cfl_condition(double max_u, double max_v, double dx, double dy, double cfl_factor) {
    const double rate = fabs(max_u) / dx + fabs(max_v) / dy;
    if (!(rate > 0.0)) return MAX_DT;
    return cfl_factor / rate;
}

double compute_viscous_timestep(double dx, double dy, double Re) {
    return 0.5 * Re * dx * dy / (dx + dy);
}

double adaptive_timestep(double max_vel, double dx, double dy, double Re, AdaptiveTimeParams *p) {
    const double cfl = p ? p->cfl_factor : 0.5;
    const double lo = p ? p->dt_min : MIN_DT, hi = p ? p->dt_max : MAX_DT;
    const double safety = (p && p->safety_factor > 0.0) ? p->safety_factor : 1.0;
    const double h = fmin(dx, dy);
    const double adv = (max_vel > 0.0) ? cfl * h / max_vel : hi;
    double dt = safety * fmin(adv, compute_viscous_timestep(dx, dy, Re));
    if (p && !p->adapt_enabled) return fmin(fmax(hi, lo), hi);
    if (dt < lo) dt = lo;
    if (dt > hi) dt = hi;
    return dt;
}
