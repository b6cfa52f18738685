/*
 * spectral_internal.h -- the device half of NSSpectralData (the appended
 * `device` member of ns_spectral_solver.h), shared by the host files.
 */
#ifndef NS_SPECTRAL_INTERNAL_H
#define NS_SPECTRAL_INTERNAL_H

#include "ns_device.h"
#include "ns_spectral_solver.h"
#include "time_integration.h"

struct NSSpectralDevice {
    NsdContext *ctx;
    int precision;
    int scheme;
    double drag;
    int force_on;
    NSForcingConfig force_cfg;
    double force_rate;
    /* adaptive-dt error controller (ns_spectral_set_error_control) */
    int errctl;
    double err_rtol, err_atol, err_safety, dt_min, dt_max;
    SolverStatus status;
    char err[256];
    NsdParams applied;
    int applied_valid;
};

#endif /* NS_SPECTRAL_INTERNAL_H */
