/*
 * ns_spectral_run.c -- config-driven driver of the MI355X spectral solver.
 *
 * The reference's src/spectral/main_spectral.c:26-100 hard-codes its run
 * (128^2, Re 1000, t_end 5, Taylor-Green, output every 50 steps) and ignores
 * the ns_config keys.  This driver reads them (INI or JSON files, later files
 * override earlier ones, then KEY=VALUE arguments):
 *
 *   solver.type                  spectral (finite_difference is refused)
 *   grid.nx, grid.ny             power-of-two grid, default 128
 *   domain.Lx, domain.Ly         default 2 pi
 *   fluid.Re, fluid.drag         default 1000, 0
 *   time.dt, time.t_end          fixed dt (when time.adaptive = 0), end time
 *   time.cfl_factor              default 0.3 (spectral_memory.c:19)
 *   time.adaptive                1: dt from the CFL limit before every step
 *   time.max_steps               stop after this many steps (0: no limit)
 *   spectral.precision           fp64 | fp32
 *   spectral.scheme              rk4 | if_rk4
 *   spectral.hyperviscosity.*    enabled, coefficient, order, cutoff_ratio,
 *                                sfd_strength, sfd_timescale, sfd_enabled
 *   forcing.*                    type (none|white|colored), kmin, kmax,
 *                                energy_target, correlation_time, seed,
 *                                injection_rate
 *   init.type                    taylor_green | random_phase
 *   init.seed, init.k_peak, init.energy   random-phase initial field
 *   output.interval              steps between outputs (0: final only)
 *   output.format                dat | vtk | vtu | none (comma list allowed)
 *   output.prefix                file name prefix, default spectral_solution
 *   output.diagnostics_interval  steps between diagnostics lines, default 10
 *   output.resolution_check      1 (default): stop like the reference when
 *                                ns_spectral_check_resolution fails
 *   checkpoint.interval, checkpoint.file, checkpoint.restart
 *
 * Between outputs the state stays on the GPU (ns_spectral_advance_n); the
 * host arrays are refreshed only when something is written.  Like the
 * reference, the run stops early when the resolution check fails or dt falls
 * below MIN_DT (main_spectral.c:53-55,67-70).  Exit status 0 on success.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "checkpoint.h"
#include "ns_spectral_solver.h"
#include "vtk_writer.h"

static double now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int has_format(const char *list, const char *f) {
    const size_t n = strlen(f);
    for (const char *p = list; (p = strstr(p, f)) != NULL; p += n)
        if ((p == list || p[-1] == ',') && (p[n] == '\0' || p[n] == ',')) return 1;
    return 0;
}

static int write_outputs(NSSpectralData *d, const char *prefix, const char *formats, int step,
                         int final) {
    char name[512];
    int rc = ns_spectral_sync_host(d);
    if (rc) return rc;
    if (has_format(formats, "dat")) {
        if (final) snprintf(name, sizeof(name), "%s_final.dat", prefix);
        else snprintf(name, sizeof(name), "%s_step_%04d.dat", prefix, step);
        ns_spectral_output_solution(d, name); /* spectral_output.c:4-27 */
        rc |= ns_spectral_last_status(d);
    }
    if (has_format(formats, "vtk")) {
        snprintf(name, sizeof(name), "%s_%06d.vtk", prefix, step);
        rc |= ns_vtk_write_spectral_legacy(name, d, d->t) != 0;
    }
    if (has_format(formats, "vtu") || has_format(formats, "vtr")) {
        snprintf(name, sizeof(name), "%s_%06d.vtr", prefix, step);
        rc |= ns_vtk_write_spectral_xml(name, d, d->t) != 0;
        char pvd[512], base[512];
        snprintf(pvd, sizeof(pvd), "%s.pvd", prefix);
        const char *slash = strrchr(name, '/');
        snprintf(base, sizeof(base), "%s", slash ? slash + 1 : name);
        rc |= ns_vtk_update_pvd(pvd, base, d->t) != 0;
    }
    return rc;
}

int main(int argc, char **argv) {
    NSConfig cfg;
    ns_config_init(&cfg);
    const char *files[64];
    size_t nfiles = 0;
    for (int i = 1; i < argc; ++i)
        if (!strchr(argv[i], '=') && nfiles < 64) files[nfiles++] = argv[i];
    char err[512] = "";
    if (nfiles && ns_config_load_files(&cfg, files, nfiles, err, sizeof(err)) != 0) {
        fprintf(stderr, "ns_spectral_run: %s\n", err);
        return 2;
    }
    for (int i = 1; i < argc; ++i) {
        char *eq = strchr(argv[i], '=');
        if (!eq) continue;
        char key[256];
        snprintf(key, sizeof(key), "%.*s", (int)(eq - argv[i]), argv[i]);
        ns_config_set(&cfg, key, eq + 1);
    }
    const char *solver = ns_config_get_string(&cfg, "solver.type", "spectral");
    if (strcmp(solver, "spectral") != 0) {
        fprintf(stderr, "ns_spectral_run: solver.type=%s: only the spectral solver runs here\n",
                solver);
        ns_config_free(&cfg);
        return 2;
    }
    NSSpectralData *d = ns_spectral_create_from_config(&cfg);
    if (!d) {
        fprintf(stderr, "ns_spectral_run: could not create the solver (grid size or device)\n");
        ns_config_free(&cfg);
        return 1;
    }
    const double t_end = ns_config_get_double(&cfg, "time.t_end", 5.0);
    const int adaptive = ns_config_get_bool(&cfg, "time.adaptive", 1);
    const int max_steps = ns_config_get_int(&cfg, "time.max_steps", 0);
    const int out_every = ns_config_get_int(&cfg, "output.interval", 50);
    const int diag_every = ns_config_get_int(&cfg, "output.diagnostics_interval", 10);
    const char *formats = ns_config_get_string(&cfg, "output.format", "dat");
    const char *prefix = ns_config_get_string(&cfg, "output.prefix", "spectral_solution");
    const int ck_every = ns_config_get_int(&cfg, "checkpoint.interval", 0);
    const char *ck_file = ns_config_get_string(&cfg, "checkpoint.file", "spectral_checkpoint.h5");
    const char *restart = ns_config_get_string(&cfg, "checkpoint.restart", "");
    const int res_check = ns_config_get_bool(&cfg, "output.resolution_check", 1);

    const char *init = ns_config_get_string(&cfg, "init.type", "taylor_green");
    int rc = 0;
    if (restart[0]) {
        if (ns_spectral_advance_n(d, 0, adaptive) == 0 &&
            ns_checkpoint_read_spectral(restart, d) == 0) {
            printf("Restarted from %s at t=%.6f\n", restart, d->t);
        } else {
            fprintf(stderr, "ns_spectral_run: restart from %s failed: %s\n", restart,
                    ns_checkpoint_error());
            rc = 1;
        }
    } else if (strcmp(init, "random_phase") == 0) {
        rc = ns_spectral_initialize_random_phase(
            d, (unsigned long long)ns_config_get_double(&cfg, "init.seed", 42.0),
            ns_config_get_double(&cfg, "init.k_peak", 10.0),
            ns_config_get_double(&cfg, "init.energy", 1.0));
    } else {
        ns_spectral_initialize_taylor_green(d);
        ns_spectral_compute_derivatives(d); /* canonical init: psi_hat, u, v valid */
        rc = ns_spectral_last_status(d);
    }
    if (rc == 0) rc = ns_spectral_sync_device(d);

    printf("Spectral Navier-Stokes Solver (MI355X)\n");
    printf("Resolution: %dx%d modes, domain [0,%.2f] x [0,%.2f], Re %.1f, %s stages\n", d->nx,
           d->ny, d->Lx, d->Ly, d->Re,
           strstr(ns_config_get_string(&cfg, "spectral.precision", "fp64"), "32") ? "FP32"
                                                                                  : "FP64");
    double E = 0.0, Z = 0.0;
    if (rc == 0) rc = ns_spectral_device_diagnostics(d, &E, &Z);
    printf("Initial: energy %.6e, enstrophy %.6e\n", E, Z);

    int step = 0, stop = 0;
    const double t0 = now();
    while (rc == 0 && !stop && d->t < t_end && (max_steps <= 0 || step < max_steps)) {
        /* advance on the device up to the next diagnostics/output/checkpoint step */
        int chunk = diag_every > 0 ? diag_every - step % diag_every : 100;
        if (out_every > 0 && out_every - step % out_every < chunk) chunk = out_every - step % out_every;
        if (ck_every > 0 && ck_every - step % ck_every < chunk) chunk = ck_every - step % ck_every;
        if (max_steps > 0 && max_steps - step < chunk) chunk = max_steps - step;
        /* the end time is checked between chunks; near it, go step by step */
        if (d->t + chunk * d->dt > t_end && d->dt > 0.0) chunk = 1;
        rc = ns_spectral_advance_n(d, chunk, adaptive);
        step += chunk;
        if (rc) break;
        if (diag_every > 0 && step % diag_every == 0) {
            ns_spectral_device_diagnostics(d, &E, &Z);
            printf("Step %d: t=%.6f, dt=%.3e, E=%.6e, Z=%.6e\n", step, d->t, d->dt, E, Z);
        }
        if (step % 100 == 0 || (out_every > 0 && step % out_every == 0)) {
            rc = ns_spectral_sync_host(d);
            if (rc == 0 && res_check && !ns_spectral_check_resolution(d)) {
                printf("Resolution check failed at step %d\n", step);
                stop = 1;
            }
        }
        if (rc == 0 && out_every > 0 && step % out_every == 0)
            rc = write_outputs(d, prefix, formats, step, 0);
        if (rc == 0 && ck_every > 0 && step % ck_every == 0) {
            rc = ns_spectral_sync_host(d);
            if (rc == 0 && ns_checkpoint_write_spectral(ck_file, d) != 0) {
                fprintf(stderr, "ns_spectral_run: checkpoint failed: %s\n", ns_checkpoint_error());
                rc = 1;
            }
        }
        if (adaptive && d->dt < MIN_DT) {
            printf("Time step too small!\n");
            stop = 1;
        }
    }
    const double wall = now() - t0;
    if (rc == 0) rc = write_outputs(d, prefix, formats, step, 1);
    if (rc == 0) rc = ns_spectral_device_diagnostics(d, &E, &Z);
    printf("\n=== SIMULATION COMPLETED ===\n");
    printf("Total steps: %d\n", step);
    printf("Final time: %.6f\n", d->t);
    printf("Final energy: %.6e\n", E);
    printf("Final enstrophy: %.6e\n", Z);
    printf("Wall time: %.3f s, steps/second: %.1f, Gpoint-steps/s: %.3f\n", wall,
           wall > 0 ? step / wall : 0.0, wall > 0 ? (double)d->nx * d->ny * step / wall / 1e9 : 0.0);
    if (rc) fprintf(stderr, "ns_spectral_run: %s\n", ns_spectral_last_error(d));
    ns_spectral_free(d);
    ns_config_free(&cfg);
    return rc ? 1 : 0;
}
