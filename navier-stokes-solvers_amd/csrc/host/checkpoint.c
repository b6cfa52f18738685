/*
 * checkpoint.c -- HDF5 checkpoint/restart of the spectral state in the file
 * layout of the reference's src/common/checkpoint.c:266-360, plus what a
 * forced run needs to resume bit for bit (see include/checkpoint.h).
 *
 * The HDF5 C library is opened with dlopen on first use, so the solver
 * library has no link-time dependency on it.  Only the documented public API
 * of HDF5 1.10+ is used (hid_t is 64-bit; the predefined type and property
 * class identifiers are read from their exported *_g variables after
 * H5open, which is what the HDF5 headers' macros do).
 */
#include "checkpoint.h"

#include <dlfcn.h>
#include <stdint.h>

#include "spectral_internal.h"

typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;

static struct {
    int tried, ok;
    herr_t (*open)(void);
    herr_t (*eset_auto)(hid_t, void *, void *);
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    hid_t (*fopen)(const char *, unsigned, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*gcreate)(hid_t, const char *, hid_t, hid_t, hid_t);
    hid_t (*gopen)(hid_t, const char *, hid_t);
    herr_t (*gclose)(hid_t);
    hid_t (*screate)(int);
    hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
    herr_t (*sclose)(hid_t);
    int (*sget_ndims)(hid_t);
    int (*sget_dims)(hid_t, hsize_t *, hsize_t *);
    hid_t (*acreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t);
    hid_t (*aopen)(hid_t, const char *, hid_t);
    int (*aexists)(hid_t, const char *);
    herr_t (*awrite)(hid_t, hid_t, const void *);
    herr_t (*aread)(hid_t, hid_t, void *);
    herr_t (*aclose)(hid_t);
    hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*dopen)(hid_t, const char *, hid_t);
    int (*lexists)(hid_t, const char *, hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
    hid_t (*dget_space)(hid_t);
    herr_t (*dclose)(hid_t);
    hid_t (*tcopy)(hid_t);
    herr_t (*tset_size)(hid_t, size_t);
    herr_t (*tclose)(hid_t);
    hid_t (*pcreate)(hid_t);
    herr_t (*pset_chunk)(hid_t, int, const hsize_t *);
    herr_t (*pclose)(hid_t);
    hid_t T_DOUBLE, T_INT, T_ULLONG, T_C_S1, P_DATASET_CREATE;
} h5;

static char g_ck_err[256];

const char *ns_checkpoint_error(void) { return g_ck_err; }

static int ck_fail(const char *msg) {
    snprintf(g_ck_err, sizeof(g_ck_err), "%s", msg);
    return -1;
}

#define H5F_ACC_RDONLY_ 0x0000u
#define H5F_ACC_TRUNC_ 0x0002u
#define H5P_DEFAULT_ ((hid_t)0)
#define H5S_ALL_ ((hid_t)0)
#define H5S_SCALAR_ 0

static int h5_load(void) {
    if (h5.tried) return h5.ok;
    h5.tried = 1;
    const char *names[] = {"libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"};
    void *lib = NULL;
    for (size_t i = 0; i < sizeof(names) / sizeof(names[0]) && !lib; ++i)
        lib = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
    if (!lib) return ck_fail("HDF5 library (libhdf5.so) not found"), 0;
#define SYM(field, name)                                       \
    do {                                                       \
        *(void **)(&h5.field) = dlsym(lib, name);              \
        if (!h5.field) return ck_fail("HDF5 symbol missing: " name), 0; \
    } while (0)
    SYM(open, "H5open");
    SYM(eset_auto, "H5Eset_auto2");
    SYM(fcreate, "H5Fcreate");
    SYM(fopen, "H5Fopen");
    SYM(fclose, "H5Fclose");
    SYM(gcreate, "H5Gcreate2");
    SYM(gopen, "H5Gopen2");
    SYM(gclose, "H5Gclose");
    SYM(screate, "H5Screate");
    SYM(screate_simple, "H5Screate_simple");
    SYM(sclose, "H5Sclose");
    SYM(sget_ndims, "H5Sget_simple_extent_ndims");
    SYM(sget_dims, "H5Sget_simple_extent_dims");
    SYM(acreate, "H5Acreate2");
    SYM(aopen, "H5Aopen");
    SYM(aexists, "H5Aexists");
    SYM(awrite, "H5Awrite");
    SYM(aread, "H5Aread");
    SYM(aclose, "H5Aclose");
    SYM(dcreate, "H5Dcreate2");
    SYM(dopen, "H5Dopen2");
    SYM(lexists, "H5Lexists");
    SYM(dwrite, "H5Dwrite");
    SYM(dread, "H5Dread");
    SYM(dget_space, "H5Dget_space");
    SYM(dclose, "H5Dclose");
    SYM(tcopy, "H5Tcopy");
    SYM(tset_size, "H5Tset_size");
    SYM(tclose, "H5Tclose");
    SYM(pcreate, "H5Pcreate");
    SYM(pset_chunk, "H5Pset_chunk");
    SYM(pclose, "H5Pclose");
#undef SYM
    if (h5.open() < 0) return ck_fail("H5open failed"), 0;
    const char *gnames[] = {"H5T_NATIVE_DOUBLE_g", "H5T_NATIVE_INT_g", "H5T_NATIVE_ULLONG_g",
                            "H5T_C_S1_g", "H5P_CLS_DATASET_CREATE_ID_g"};
    hid_t *gdst[] = {&h5.T_DOUBLE, &h5.T_INT, &h5.T_ULLONG, &h5.T_C_S1, &h5.P_DATASET_CREATE};
    for (int i = 0; i < 5; ++i) {
        const hid_t *g = (const hid_t *)dlsym(lib, gnames[i]);
        if (!g) return ck_fail("HDF5 identifier missing"), 0;
        *gdst[i] = *g;
    }
    h5.eset_auto(0, NULL, NULL); /* H5E_DEFAULT: no error-stack printing */
    h5.ok = 1;
    return 1;
}

/* ---- writing (checkpoint.c:15-27, :29-58, :60-75, :104-127 of the reference) */

static int put_attr(hid_t loc, const char *name, hid_t type, const void *value) {
    hid_t sp = h5.screate(H5S_SCALAR_);
    if (sp < 0) return -1;
    hid_t a = h5.acreate(loc, name, type, sp, H5P_DEFAULT_, H5P_DEFAULT_);
    int rc = -1;
    if (a >= 0) {
        rc = h5.awrite(a, type, value) < 0 ? -1 : 0;
        h5.aclose(a);
    }
    h5.sclose(sp);
    return rc;
}

static int put_string_attr(hid_t loc, const char *name, const char *value) {
    hid_t t = h5.tcopy(h5.T_C_S1);
    if (t < 0) return -1;
    int rc = h5.tset_size(t, strlen(value) + 1) < 0 ? -1 : put_attr(loc, name, t, value);
    h5.tclose(t);
    return rc;
}

/* chunked like the reference's datasets, with chunks of whole rows (up to 64)
 * rather than 8 x 8 tiles, and without compression */
static int put_dataset(hid_t parent, const char *name, int rank, const hsize_t *dims,
                       const double *data) {
    hid_t sp = h5.screate_simple(rank, dims, NULL);
    if (sp < 0) return -1;
    hid_t pl = h5.pcreate(h5.P_DATASET_CREATE);
    hsize_t chunk[3];
    for (int i = 0; i < rank; ++i) chunk[i] = dims[i];
    chunk[0] = dims[0] < 64 ? dims[0] : 64;
    int rc = -1;
    if (pl >= 0 && h5.pset_chunk(pl, rank, chunk) >= 0) {
        hid_t ds = h5.dcreate(parent, name, h5.T_DOUBLE, sp, H5P_DEFAULT_, pl, H5P_DEFAULT_);
        if (ds >= 0) {
            rc = h5.dwrite(ds, h5.T_DOUBLE, H5S_ALL_, H5S_ALL_, H5P_DEFAULT_, data) < 0 ? -1 : 0;
            h5.dclose(ds);
        }
    }
    if (pl >= 0) h5.pclose(pl);
    h5.sclose(sp);
    return rc;
}

static int put_real(hid_t g, const char *name, const double *f, int ny, int nx) {
    const hsize_t dims[2] = {(hsize_t)ny, (hsize_t)nx};
    return put_dataset(g, name, 2, dims, f);
}

/* fftw_complex is interleaved (re, im), which is the [ny][nkx][2] layout */
static int put_complex(hid_t g, const char *name, const fftw_complex *f, int nky, int nkx) {
    const hsize_t dims[3] = {(hsize_t)nky, (hsize_t)nkx, 2};
    return put_dataset(g, name, 3, dims, (const double *)f);
}

int ns_checkpoint_write_spectral(const char *filename, const NSSpectralData *s) {
    g_ck_err[0] = '\0';
    if (!filename || !s) return ck_fail("null argument");
    if (!h5_load()) return -1;
    hid_t f = h5.fcreate(filename, H5F_ACC_TRUNC_, H5P_DEFAULT_, H5P_DEFAULT_);
    if (f < 0) return ck_fail("cannot create the checkpoint file");
    int rc = 0;
    /* checkpoint.c:270-277: attribute order solver, time, dt, Re, nx, ny, Lx, Ly */
    rc |= put_string_attr(f, "solver", "spectral");
    rc |= put_attr(f, "time", h5.T_DOUBLE, &s->t);
    rc |= put_attr(f, "dt", h5.T_DOUBLE, &s->dt);
    rc |= put_attr(f, "Re", h5.T_DOUBLE, &s->Re);
    rc |= put_attr(f, "nx", h5.T_INT, &s->nx);
    rc |= put_attr(f, "ny", h5.T_INT, &s->ny);
    rc |= put_attr(f, "Lx", h5.T_DOUBLE, &s->Lx);
    rc |= put_attr(f, "Ly", h5.T_DOUBLE, &s->Ly);
    /* additions for an exact restart */
    rc |= put_attr(f, "cfl_factor", h5.T_DOUBLE, &s->cfl_factor);
    unsigned long long calls = 0;
    double *ou = NULL;
    int nou = 0;
    if (s->device && s->device->ctx) {
        rc |= nsd_get_forcing_counter(s->device->ctx, &calls);
        if (s->device->force_on && s->device->force_cfg.type == NS_FORCING_COLORED_NOISE) {
            nou = 2 * nsd_forcing_modes(s->device->ctx);
            ou = (double *)calloc((size_t)(nou > 0 ? nou : 1), sizeof(double));
            if (!ou || nsd_forcing_ou(s->device->ctx, ou, nou, 0)) rc |= -1;
        }
    }
    rc |= put_attr(f, "forcing_calls", h5.T_ULLONG, &calls);
    hid_t g = h5.gcreate(f, "/fields", H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_);
    if (g < 0) {
        free(ou);
        h5.fclose(f);
        return ck_fail("cannot create /fields");
    }
    /* checkpoint.c:287-295 */
    rc |= put_real(g, "u", s->u, s->ny, s->nx);
    rc |= put_real(g, "v", s->v, s->ny, s->nx);
    rc |= put_real(g, "omega", s->omega, s->ny, s->nx);
    rc |= put_real(g, "psi", s->psi, s->ny, s->nx);
    rc |= put_complex(g, "u_hat", s->u_hat, s->nky, s->nkx);
    rc |= put_complex(g, "v_hat", s->v_hat, s->nky, s->nkx);
    rc |= put_complex(g, "omega_hat", s->omega_hat, s->nky, s->nkx);
    rc |= put_complex(g, "psi_hat", s->psi_hat, s->nky, s->nkx);
    if (s->hyperviscosity.sfd_enabled && s->sfd_hat)
        rc |= put_complex(g, "sfd_hat", s->sfd_hat, s->nky, s->nkx);
    if (ou && nou > 0) {
        const hsize_t dims[2] = {(hsize_t)(nou / 2), 2};
        rc |= put_dataset(g, "forcing_ou", 2, dims, ou);
    }
    free(ou);
    h5.gclose(g);
    h5.fclose(f);
    return rc ? ck_fail("HDF5 write failed") : 0;
}

/* ---- reading (checkpoint.c:304-360 of the reference) */

static int get_attr(hid_t loc, const char *name, hid_t type, void *value) {
    if (h5.aexists(loc, name) <= 0) return 1; /* absent */
    hid_t a = h5.aopen(loc, name, H5P_DEFAULT_);
    if (a < 0) return -1;
    int rc = h5.aread(a, type, value) < 0 ? -1 : 0;
    h5.aclose(a);
    return rc;
}

static int get_dataset(hid_t g, const char *name, int rank, const hsize_t *dims, double *out) {
    if (h5.lexists(g, name, H5P_DEFAULT_) <= 0) return 1; /* absent */
    hid_t ds = h5.dopen(g, name, H5P_DEFAULT_);
    if (ds < 0) return -1;
    int rc = -1;
    hid_t sp = h5.dget_space(ds);
    if (sp >= 0) {
        hsize_t d[3] = {0, 0, 0};
        if (h5.sget_ndims(sp) == rank && h5.sget_dims(sp, d, NULL) == rank) {
            int same = 1;
            for (int i = 0; i < rank; ++i) same &= d[i] == dims[i];
            if (same) rc = h5.dread(ds, h5.T_DOUBLE, H5S_ALL_, H5S_ALL_, H5P_DEFAULT_, out) < 0 ? -1 : 0;
        }
        h5.sclose(sp);
    }
    h5.dclose(ds);
    return rc;
}

int ns_checkpoint_read_spectral(const char *filename, NSSpectralData *s) {
    g_ck_err[0] = '\0';
    if (!filename || !s) return ck_fail("null argument");
    if (!h5_load()) return -1;
    hid_t f = h5.fopen(filename, H5F_ACC_RDONLY_, H5P_DEFAULT_);
    if (f < 0) return ck_fail("cannot open the checkpoint file");
    int nx = 0, ny = 0;
    double Lx = 0.0, Ly = 0.0, Re = 0.0, t = 0.0, dt = 0.0, cfl = s->cfl_factor;
    unsigned long long calls = 0;
    int rc = 0;
    rc |= get_attr(f, "nx", h5.T_INT, &nx) < 0;
    rc |= get_attr(f, "ny", h5.T_INT, &ny) < 0;
    rc |= get_attr(f, "Lx", h5.T_DOUBLE, &Lx) < 0;
    rc |= get_attr(f, "Ly", h5.T_DOUBLE, &Ly) < 0;
    rc |= get_attr(f, "Re", h5.T_DOUBLE, &Re) < 0;
    rc |= get_attr(f, "time", h5.T_DOUBLE, &t) < 0;
    rc |= get_attr(f, "dt", h5.T_DOUBLE, &dt) < 0;
    rc |= get_attr(f, "cfl_factor", h5.T_DOUBLE, &cfl) < 0;
    const int has_calls = get_attr(f, "forcing_calls", h5.T_ULLONG, &calls) == 0;
    if (rc || nx != s->nx || ny != s->ny) {
        h5.fclose(f);
        return ck_fail("checkpoint grid does not match the solver (nx, ny)");
    }
    hid_t g = h5.gopen(f, "/fields", H5P_DEFAULT_);
    if (g < 0) {
        h5.fclose(f);
        return ck_fail("checkpoint has no /fields group");
    }
    const hsize_t dr[2] = {(hsize_t)ny, (hsize_t)nx};
    const hsize_t dc[3] = {(hsize_t)s->nky, (hsize_t)s->nkx, 2};
    int miss = 0, r;
    /* the reference's datasets are required, the additions optional */
    const char *rn[4] = {"u", "v", "omega", "psi"};
    double *rf[4] = {s->u, s->v, s->omega, s->psi};
    for (int i = 0; i < 4; ++i)
        if ((r = get_dataset(g, rn[i], 2, dr, rf[i])) != 0) miss |= 1, rc |= r < 0;
    const char *cn[3] = {"u_hat", "v_hat", "omega_hat"};
    fftw_complex *cf[3] = {s->u_hat, s->v_hat, s->omega_hat};
    for (int i = 0; i < 3; ++i)
        if ((r = get_dataset(g, cn[i], 3, dc, (double *)cf[i])) != 0) miss |= 1, rc |= r < 0;
    rc |= get_dataset(g, "psi_hat", 3, dc, (double *)s->psi_hat) < 0;
    if (s->sfd_hat) rc |= get_dataset(g, "sfd_hat", 3, dc, (double *)s->sfd_hat) < 0;
    if (s->device && s->device->ctx && rc == 0 && !miss) {
        const int nou = 2 * nsd_forcing_modes(s->device->ctx);
        if (nou > 0) {
            double *ou = (double *)calloc((size_t)nou, sizeof(double));
            const hsize_t d2[2] = {(hsize_t)(nou / 2), 2};
            if (!ou) rc |= 1;
            else if (get_dataset(g, "forcing_ou", 2, d2, ou) == 0)
                rc |= nsd_forcing_ou(s->device->ctx, ou, nou, 1) != 0;
            free(ou);
        }
        if (has_calls) rc |= nsd_set_forcing_counter(s->device->ctx, calls) != 0;
    }
    h5.gclose(g);
    h5.fclose(f);
    if (rc || miss) return ck_fail(miss ? "checkpoint misses a field" : "HDF5 read failed");
    s->Lx = Lx;
    s->Ly = Ly;
    s->Re = Re;
    s->t = t;
    s->dt = dt;
    s->cfl_factor = cfl;
    return 0;
}
