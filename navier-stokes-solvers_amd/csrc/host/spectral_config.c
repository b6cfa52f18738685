/*
 * spectral_config.c -- ns_config keys -> solver settings (the reference has
 * the keys, config/parameters.ini:1-26, but no driver reads them).
 */
#include "ns_spectral_solver.h"

int ns_spectral_configure(NSSpectralData *d, const NSConfig *cfg) {
    if (!d || !cfg) return SOLVER_INVALID_PARAM;
    d->Re = ns_config_get_double(cfg, "fluid.Re", d->Re);
    d->dt = ns_config_get_double(cfg, "time.dt", d->dt);
    d->cfl_factor = ns_config_get_double(cfg, "time.cfl_factor", d->cfl_factor);

    NSSpectralHyperviscosityConfig hv = d->hyperviscosity;
    const int has_coeff = ns_config_has_key(cfg, "spectral.hyperviscosity.coefficient");
    hv.coefficient = ns_config_get_double(cfg, "spectral.hyperviscosity.coefficient", hv.coefficient);
    hv.order = ns_config_get_int(cfg, "spectral.hyperviscosity.order", hv.order);
    hv.cutoff_ratio = ns_config_get_double(cfg, "spectral.hyperviscosity.cutoff_ratio", hv.cutoff_ratio);
    hv.enabled = ns_config_get_bool(cfg, "spectral.hyperviscosity.enabled",
                                    has_coeff ? hv.coefficient > 0.0 : hv.enabled);
    hv.sfd_strength = ns_config_get_double(cfg, "spectral.hyperviscosity.sfd_strength", hv.sfd_strength);
    hv.sfd_timescale = ns_config_get_double(cfg, "spectral.hyperviscosity.sfd_timescale", hv.sfd_timescale);
    hv.sfd_enabled = ns_config_get_bool(cfg, "spectral.hyperviscosity.sfd_enabled", hv.sfd_strength > 0.0);
    ns_spectral_hyperviscosity_set(d, &hv);

    const char *prec = ns_config_get_string(cfg, "spectral.precision", "fp64");
    int rc = ns_spectral_set_precision(d, (prec[0] == 'f' || prec[0] == 'F') && strstr(prec, "32")
                                              ? NS_PRECISION_FP32 : NS_PRECISION_FP64);
    if (rc) return rc;
    const char *scheme = ns_config_get_string(cfg, "spectral.scheme", "rk4");
    rc = ns_spectral_set_scheme(d, (strstr(scheme, "if") || strstr(scheme, "IF")) ? NS_SCHEME_IF_RK4
                                                                                   : NS_SCHEME_RK4);
    if (rc) return rc;
    rc = ns_spectral_set_drag(d, ns_config_get_double(cfg, "fluid.drag", 0.0));
    if (rc) return rc;

    const char *ftype = ns_config_get_string(cfg, "forcing.type", "none");
    if (strcmp(ftype, "white") == 0 || strcmp(ftype, "colored") == 0) {
        NSForcingConfig f;
        ns_forcing_config_default(&f);
        f.type = strcmp(ftype, "colored") == 0 ? NS_FORCING_COLORED_NOISE : NS_FORCING_WHITE_NOISE;
        f.kmin = ns_config_get_double(cfg, "forcing.kmin", f.kmin);
        f.kmax = ns_config_get_double(cfg, "forcing.kmax", f.kmax);
        f.energy_target = ns_config_get_double(cfg, "forcing.energy_target", f.energy_target);
        f.correlation_time = ns_config_get_double(cfg, "forcing.correlation_time", f.correlation_time);
        f.seed = (unsigned long)ns_config_get_double(cfg, "forcing.seed", 0.0);
        const double rate = ns_config_get_double(cfg, "forcing.injection_rate", f.energy_target);
        rc = ns_spectral_set_forcing(d, &f, rate);
    } else {
        rc = ns_spectral_set_forcing(d, NULL, 0.0);
    }
    return rc;
}

NSSpectralData *ns_spectral_create_from_config(const NSConfig *cfg) {
    if (!cfg) return NULL;
    const int nx = ns_config_get_int(cfg, "grid.nx", 128), ny = ns_config_get_int(cfg, "grid.ny", nx);
    const double Lx = ns_config_get_double(cfg, "domain.Lx", 2.0 * M_PI);
    const double Ly = ns_config_get_double(cfg, "domain.Ly", Lx);
    const double Re = ns_config_get_double(cfg, "fluid.Re", 1000.0);
    NSSpectralData *d = ns_spectral_allocate(nx, ny, Lx, Ly, Re);
    if (!d) return NULL;
    ns_spectral_initialize_grids(d);
    ns_spectral_initialize_fft_plans(d);
    if (ns_spectral_configure(d, cfg) != SOLVER_SUCCESS) {
        ns_spectral_free(d);
        return NULL;
    }
    return d;
}
