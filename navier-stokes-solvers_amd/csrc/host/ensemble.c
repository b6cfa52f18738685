/*
 * ensemble.c -- ensembles of independent solves over batch device contexts.
 *
 * Members are grouped by grid size; each group is one batch context whose
 * time-step kernels run all of its members per launch, and one NsdMulti graph
 * advances every group by a step.  Per member: Reynolds number, seed of the
 * random-phase initial field and of the forcing stream.  Shared: scheme,
 * CFL factor, drag, del^8 hyperviscosity coefficient, forcing shell.
 */
#include "ns_ensemble.h"

#include "ns_device.h"

typedef struct {
    int n;
    int count;
    int *member;   /* ensemble index of each group slot */
    NsdContext *ctx;
} Group;

struct NSEnsemble {
    int members;
    int ngroups;
    Group *groups;
    int *group_of, *slot_of;
    NsdMulti *multi;
    char err[256];
};

static char g_create_err[512];

const char *ns_ensemble_create_error(void) { return g_create_err; }

void ns_ensemble_config_default(NSEnsembleConfig *cfg) {
    if (!cfg) return;
    memset(cfg, 0, sizeof(*cfg));
    cfg->Lx = cfg->Ly = 2.0 * M_PI;
    cfg->precision = NS_PRECISION_FP32;
    cfg->scheme = NS_SCHEME_IF_RK4;
    cfg->cfl_factor = 0.5;
    ns_forcing_config_default(&cfg->forcing);
    cfg->k_peak = 10.0;
    cfg->energy = 1.0;
    cfg->autotune = 1;
}

static void params_for(const NSEnsembleConfig *cfg, int n, NsdParams *p) {
    memset(p, 0, sizeof(*p));
    p->Re = 1.0; /* replaced per member */
    p->drag = cfg->drag;
    p->cfl_factor = cfg->cfl_factor;
    p->scheme = cfg->scheme == NS_SCHEME_IF_RK4 ? NSD_SCHEME_IF : NSD_SCHEME_CLASSIC;
    if (cfg->nu4 > 0.0) {
        /* hyperviscosity.c:60-102 with cutoff 0 and order 8: kmax over the
         * stored modes, coefficient nu4 kmax^8, so the damping is nu4 |k|^8 */
        const double kx = 2.0 * M_PI / cfg->Lx * (n / 2), ky = 2.0 * M_PI / cfg->Ly * (n / 2);
        const double kmax = sqrt(kx * kx + ky * ky);
        p->hv_on = 1;
        p->hv_order = 8;
        p->hv_coeff = cfg->nu4 * pow(kmax, 8);
        p->hv_kc = 0.0;
        p->hv_invbw = 1.0 / kmax;
    }
    if (cfg->forcing_enabled) {
        const NSForcingConfig *f = &cfg->forcing;
        p->force_on = 1;
        p->force_colored = f->type == NS_FORCING_COLORED_NOISE;
        p->force_kmin = f->kmin;
        p->force_kmax = f->kmax;
        p->force_energy_target = f->energy_target;
        p->force_tau = f->correlation_time;
        p->force_rate = cfg->injection_rate;
    }
}

void ns_ensemble_free(NSEnsemble *e) {
    if (!e) return;
    if (e->multi) nsd_multi_destroy(e->multi);
    for (int g = 0; g < e->ngroups; ++g) {
        if (e->groups[g].ctx) nsd_destroy(e->groups[g].ctx);
        free(e->groups[g].member);
    }
    free(e->groups);
    free(e->group_of);
    free(e->slot_of);
    free(e);
}

static NSEnsemble *fail(NSEnsemble *e, const char *what, const char *detail) {
    snprintf(g_create_err, sizeof(g_create_err), "%s%s%s", what, detail ? ": " : "",
             detail ? detail : "");
    ns_ensemble_free(e);
    return NULL;
}

NSEnsemble *ns_ensemble_create(int members, const int *n, const double *Re,
                               const unsigned long long *seeds, const NSEnsembleConfig *cfg_in) {
    g_create_err[0] = '\0';
    if (members < 1 || !n || !Re || !seeds) return fail(NULL, "invalid arguments", NULL);
    NSEnsembleConfig cfg;
    if (cfg_in) cfg = *cfg_in;
    else ns_ensemble_config_default(&cfg);
    NSEnsemble *e = (NSEnsemble *)calloc(1, sizeof(NSEnsemble));
    if (!e) return fail(NULL, "out of memory", NULL);
    e->members = members;
    e->group_of = (int *)calloc((size_t)members, sizeof(int));
    e->slot_of = (int *)calloc((size_t)members, sizeof(int));
    e->groups = (Group *)calloc((size_t)members, sizeof(Group));
    if (!e->group_of || !e->slot_of || !e->groups) return fail(e, "out of memory", NULL);
    /* group by size, in order of first appearance */
    for (int m = 0; m < members; ++m) {
        int g = 0;
        while (g < e->ngroups && e->groups[g].n != n[m]) ++g;
        if (g == e->ngroups) {
            e->groups[g].n = n[m];
            e->groups[g].member = (int *)calloc((size_t)members, sizeof(int));
            if (!e->groups[g].member) return fail(e, "out of memory", NULL);
            e->ngroups++;
        }
        e->group_of[m] = g;
        e->slot_of[m] = e->groups[g].count;
        e->groups[g].member[e->groups[g].count++] = m;
    }
    const int prec = cfg.precision == NS_PRECISION_FP32 ? NSD_FP32 : NSD_FP64;
    NsdContext **cs = (NsdContext **)calloc((size_t)e->ngroups, sizeof(NsdContext *));
    double *re = (double *)calloc((size_t)members, sizeof(double));
    unsigned long long *sd = (unsigned long long *)calloc((size_t)members, sizeof(*sd));
    if (!cs || !re || !sd) {
        free(cs);
        free(re);
        free(sd);
        return fail(e, "out of memory", NULL);
    }
    const char *why = NULL;
    for (int g = 0; g < e->ngroups && !why; ++g) {
        Group *G = &e->groups[g];
        G->ctx = nsd_create_batch(G->n, G->n, cfg.Lx, cfg.Ly, prec, G->count, e->err,
                                  sizeof(e->err));
        if (!G->ctx) {
            why = e->err;
            break;
        }
        NsdParams p;
        params_for(&cfg, G->n, &p);
        for (int s = 0; s < G->count; ++s) {
            re[s] = Re[G->member[s]];
            sd[s] = seeds[G->member[s]];
        }
        if (nsd_set_params(G->ctx, &p) ||
            (cfg.autotune && nsd_autotune(G->ctx, 3, 0, NULL, NULL)) ||
            nsd_set_member_params(G->ctx, re, sd) ||
            nsd_batch_random_phase(G->ctx, sd, cfg.k_peak, cfg.energy) ||
            nsd_set_time(G->ctx, 0.0, 0.0))
            why = nsd_error(G->ctx);
        cs[g] = G->ctx;
    }
    if (!why) {
        e->multi = nsd_multi_create(cs, e->ngroups);
        if (!e->multi) why = "could not create the ensemble graph";
    }
    free(cs);
    free(re);
    free(sd);
    if (why) {
        char msg[256];
        snprintf(msg, sizeof(msg), "%s", why);
        return fail(e, "ns_ensemble_create", msg);
    }
    return e;
}

int ns_ensemble_members(const NSEnsemble *e) { return e ? e->members : 0; }

const char *ns_ensemble_last_error(const NSEnsemble *e) { return e ? e->err : "null ensemble"; }

int ns_ensemble_advance(NSEnsemble *e, int nsteps) {
    if (!e || nsteps < 0) return SOLVER_INVALID_PARAM;
    int rc = nsd_multi_advance(e->multi, nsteps, 1);
    if (rc) snprintf(e->err, sizeof(e->err), "%s", nsd_multi_error(e->multi));
    return rc;
}

int ns_ensemble_synchronize(NSEnsemble *e) {
    if (!e) return SOLVER_INVALID_PARAM;
    int rc = nsd_multi_synchronize(e->multi);
    if (rc) snprintf(e->err, sizeof(e->err), "%s", nsd_multi_error(e->multi));
    return rc;
}

int ns_ensemble_diagnostics(NSEnsemble *e, double *t, double *energy, double *enstrophy,
                            int *status) {
    if (!e) return SOLVER_INVALID_PARAM;
    int rc = ns_ensemble_synchronize(e);
    for (int g = 0; g < e->ngroups && rc == 0; ++g) {
        Group *G = &e->groups[g];
        double *buf = (double *)calloc(3 * (size_t)G->count, sizeof(double));
        int *st = (int *)calloc((size_t)G->count, sizeof(int));
        if (!buf || !st) rc = SOLVER_MEM_ERROR;
        if (rc == 0) rc = nsd_batch_energy(G->ctx, buf, buf + G->count);
        if (rc == 0) rc = nsd_batch_scalars(G->ctx, buf + 2 * G->count, NULL, st);
        if (rc) snprintf(e->err, sizeof(e->err), "%s", nsd_error(G->ctx));
        for (int s = 0; s < G->count && rc == 0; ++s) {
            const int m = G->member[s];
            if (energy) energy[m] = buf[s];
            if (enstrophy) enstrophy[m] = buf[G->count + s];
            if (t) t[m] = buf[2 * G->count + s];
            if (status) status[m] = st[s];
        }
        free(buf);
        free(st);
    }
    return rc;
}

int ns_ensemble_member_state(NSEnsemble *e, int member, double *omega_hat) {
    if (!e || member < 0 || member >= e->members || !omega_hat) return SOLVER_INVALID_PARAM;
    int rc = ns_ensemble_synchronize(e);
    const Group *G = &e->groups[e->group_of[member]];
    if (rc == 0) rc = nsd_member_download(G->ctx, e->slot_of[member], omega_hat);
    if (rc) snprintf(e->err, sizeof(e->err), "%s", nsd_error(G->ctx));
    return rc;
}
