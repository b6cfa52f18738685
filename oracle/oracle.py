"""ctypes binding of the FP64 CPU oracle (oracle/libnsoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the thing
measured or shipped.  The C sources restate the reference's src/spectral
and src/common/forcing.c; see ns_oracle.h for the citation map.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnsoracle.so")
REF_FORCING = os.path.join(_HERE, "_ref", "libref_forcing.so")
REF_CONFIG = os.path.join(_HERE, "_ref", "libref_config.so")

c_double_p = C.POINTER(C.c_double)
c_int_p = C.POINTER(C.c_int)


class OcHyperviscosity(C.Structure):
    _fields_ = [("enabled", C.c_int), ("coefficient", C.c_double), ("order", C.c_int),
                ("cutoff_ratio", C.c_double), ("sfd_strength", C.c_double),
                ("sfd_timescale", C.c_double), ("sfd_enabled", C.c_int), ("kmax", C.c_double),
                ("cutoff_wavenumber", C.c_double), ("inv_bandwidth", C.c_double)]


class OcData(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nkx", C.c_int), ("nky", C.c_int),
                ("Lx", C.c_double), ("Ly", C.c_double), ("Re", C.c_double), ("dt", C.c_double),
                ("t", C.c_double), ("cfl_factor", C.c_double)] + \
               [(n, c_double_p) for n in ("u", "v", "p", "omega", "psi", "x", "y")] + \
               [(n, c_double_p) for n in ("u_hat", "v_hat", "omega_hat", "psi_hat",
                                          "nonlinear_hat", "rk1_hat", "rk2_hat", "rk3_hat",
                                          "rk4_hat", "omega_tmp_hat")] + \
               [("kx", c_double_p), ("ky", c_double_p), ("k2", c_double_p),
                ("dealias_mask", c_int_p), ("hv", OcHyperviscosity), ("hv_weights", c_double_p),
                ("sfd_hat", c_double_p), ("fft", C.c_void_p), ("drag", C.c_double),
                ("fft_count", C.c_longlong)]


class OcForcingConfig(C.Structure):
    _fields_ = [("kmin", C.c_double), ("kmax", C.c_double), ("energy_target", C.c_double),
                ("correlation_time", C.c_double), ("seed", C.c_ulong), ("type", C.c_int)]


class OcForcingState(C.Structure):
    _fields_ = [("config", OcForcingConfig), ("nkx", C.c_int), ("nky", C.c_int),
                ("spectral_size", C.c_int), ("active_modes", C.c_int),
                ("weights", c_double_p), ("colored_state", c_double_p),
                ("have_spare", C.c_uint), ("spare_gaussian", C.c_double),
                ("rng_state", C.c_ulonglong)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(OcData)
        L.oc_allocate.restype = P
        L.oc_allocate.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
        for name in ("oc_free", "oc_initialize_grids", "oc_initialize_taylor_green",
                     "oc_compute_velocity_from_vorticity", "oc_compute_derivatives",
                     "oc_compute_nonlinear_terms", "oc_advance_rk4", "oc_advance_rk4_minimal",
                     "oc_hyperviscosity_update_weights"):
            getattr(L, name).argtypes = [P]
            getattr(L, name).restype = None
        for name in ("oc_compute_cfl_timestep", "oc_compute_energy", "oc_compute_enstrophy"):
            getattr(L, name).argtypes = [P]
            getattr(L, name).restype = C.c_double
        L.oc_physical_to_spectral.argtypes = [P, c_double_p, c_double_p]
        L.oc_spectral_to_physical.argtypes = [P, c_double_p, c_double_p]
        L.oc_apply_dealiasing.argtypes = [P, c_double_p]
        L.oc_fft_r2c_2d.argtypes = [P, c_double_p, c_double_p]
        L.oc_fft_c2r_2d.argtypes = [P, c_double_p, c_double_p]
        L.oc_check_resolution.argtypes = [P, c_double_p]
        L.oc_check_resolution.restype = C.c_int
        L.oc_analyze_convergence.argtypes = [P, c_double_p]
        L.oc_analyze_spectrum.argtypes = [P, c_double_p, c_int_p, C.c_int]
        L.oc_analyze_spectrum.restype = C.c_int
        L.oc_hyperviscosity_default.argtypes = [C.POINTER(OcHyperviscosity)]
        L.oc_hyperviscosity_set.argtypes = [P, C.POINTER(OcHyperviscosity)]
        L.oc_hyperviscosity_apply.argtypes = [P, c_double_p, c_double_p]
        L.oc_hyperviscosity_post_step.argtypes = [P, c_double_p, C.c_double]
        L.oc_splitmix64_next.argtypes = [C.POINTER(C.c_uint64)]
        L.oc_splitmix64_next.restype = C.c_uint64
        L.oc_counter_uniform.argtypes = [C.c_uint64, C.c_uint64]
        L.oc_counter_uniform.restype = C.c_double
        L.oc_forcing_config_default.argtypes = [C.POINTER(OcForcingConfig)]
        L.oc_forcing_state_init.argtypes = [C.POINTER(OcForcingState), C.POINTER(OcForcingConfig),
                                            c_double_p, C.c_int, C.c_int]
        L.oc_forcing_state_init.restype = C.c_int
        L.oc_forcing_state_free.argtypes = [C.POINTER(OcForcingState)]
        L.oc_forcing_generate.argtypes = [C.POINTER(OcForcingState), C.c_double, c_double_p,
                                          c_double_p]
        L.oc_forcing_project_and_scale.argtypes = [P, c_double_p, C.c_double]
        L.oc_advance_if_rk4.argtypes = [P, c_double_p]
        L.oc_advance_classic_ext.argtypes = [P, c_double_p]
        L.oc_initialize_random_phase.argtypes = [P, C.c_uint64, C.c_double, C.c_double]
        L.oc_advance_if_rk4_est.argtypes = [P, c_double_p, c_double_p]
        L.oc_advance_classic_ext_est.argtypes = [P, c_double_p, c_double_p]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(c_double_p)


class Oracle:
    """One oracle solver instance (FP64, reference layouts [ny][nx] and [ny][nx/2+1])."""

    def __init__(self, nx, ny, Lx=2 * np.pi, Ly=2 * np.pi, Re=100.0):
        self.L = lib()
        self.p = self.L.oc_allocate(nx, ny, Lx, Ly, Re)
        if not self.p:
            raise ValueError("oracle allocation failed (sizes must be powers of two)")
        self.d = self.p.contents
        self.nx, self.ny, self.nkx = nx, ny, nx // 2 + 1

    def close(self):
        if self.p:
            self.L.oc_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- array views (share memory with the C struct)
    def _real(self, name, shape):
        return np.ctypeslib.as_array(getattr(self.d, name), shape=shape)

    def _cplx(self, name):
        a = np.ctypeslib.as_array(getattr(self.d, name), shape=(self.ny, self.nkx * 2))
        return a.view(np.complex128)

    def phys(self, name):
        return self._real(name, (self.ny, self.nx))

    def spec(self, name):
        return self._cplx(name)

    @property
    def kx(self):
        return self._real("kx", (self.nkx,))

    @property
    def ky(self):
        return self._real("ky", (self.ny,))

    @property
    def k2(self):
        return self._real("k2", (self.ny, self.nkx))

    @property
    def mask(self):
        return np.ctypeslib.as_array(self.d.dealias_mask, shape=(self.ny, self.nkx))

    @property
    def hv_weights(self):
        return self._real("hv_weights", (self.ny, self.nkx))

    # -- pass-through API
    def init_grids(self):
        self.L.oc_initialize_grids(self.p)

    def init_taylor_green(self):
        self.L.oc_initialize_taylor_green(self.p)

    def init_random_phase(self, seed, k_peak=10.0, energy=1.0):
        self.L.oc_initialize_random_phase(self.p, seed, k_peak, energy)

    def compute_derivatives(self):
        self.L.oc_compute_derivatives(self.p)

    def velocity_from_vorticity(self):
        self.L.oc_compute_velocity_from_vorticity(self.p)

    def nonlinear(self):
        self.L.oc_compute_nonlinear_terms(self.p)

    def advance_rk4(self, minimal=False):
        (self.L.oc_advance_rk4_minimal if minimal else self.L.oc_advance_rk4)(self.p)

    def advance_if(self, forcing=None):
        self.L.oc_advance_if_rk4(self.p, None if forcing is None else _dp(forcing))

    def advance_classic_ext(self, forcing=None):
        self.L.oc_advance_classic_ext(self.p, None if forcing is None else _dp(forcing))

    def advance_est(self, integrating_factor, forcing=None):
        """One build-defined step (IF-RK4 or classic with drag) returning the
        embedded estimate's norms {sum w |y4 - y2|^2, sum w |y4|^2}."""
        norms = np.zeros(2)
        fn = self.L.oc_advance_if_rk4_est if integrating_factor else \
            self.L.oc_advance_classic_ext_est
        fn(self.p, None if forcing is None else _dp(forcing), _dp(norms))
        return norms

    def cfl(self):
        return self.L.oc_compute_cfl_timestep(self.p)

    def energy(self):
        return self.L.oc_compute_energy(self.p)

    def enstrophy(self):
        return self.L.oc_compute_enstrophy(self.p)

    def check_resolution(self):
        f = C.c_double(0.0)
        ok = self.L.oc_check_resolution(self.p, C.byref(f))
        return ok, f.value

    def analyze_convergence(self):
        out = np.zeros(6)
        self.L.oc_analyze_convergence(self.p, _dp(out))
        return out

    def analyze_spectrum(self, max_bins=100000):
        e = np.zeros(max_bins)
        c = np.zeros(max_bins, dtype=np.int32)
        m = self.L.oc_analyze_spectrum(self.p, _dp(e), c.ctypes.data_as(c_int_p), max_bins)
        return e[:m].copy(), c[:m].copy()

    def p2s(self, field):
        field = np.ascontiguousarray(field, dtype=np.float64)
        out = np.zeros((self.ny, self.nkx), dtype=np.complex128)
        self.L.oc_physical_to_spectral(self.p, _dp(field), _dp(out))
        return out

    def s2p(self, field_hat):
        fh = np.ascontiguousarray(field_hat, dtype=np.complex128).copy()
        out = np.zeros((self.ny, self.nx))
        self.L.oc_spectral_to_physical(self.p, _dp(fh), _dp(out))
        return out

    def r2c(self, field):
        field = np.ascontiguousarray(field, dtype=np.float64)
        out = np.zeros((self.ny, self.nkx), dtype=np.complex128)
        self.L.oc_fft_r2c_2d(self.p, _dp(field), _dp(out))
        return out

    def c2r(self, field_hat):
        fh = np.ascontiguousarray(field_hat, dtype=np.complex128)
        out = np.zeros((self.ny, self.nx))
        self.L.oc_fft_c2r_2d(self.p, _dp(fh), _dp(out))
        return out

    def dealias(self, field_hat):
        fh = np.ascontiguousarray(field_hat, dtype=np.complex128).copy()
        self.L.oc_apply_dealiasing(self.p, _dp(fh))
        return fh

    def hv_set(self, **kw):
        c = OcHyperviscosity()
        self.L.oc_hyperviscosity_default(C.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        self.L.oc_hyperviscosity_set(self.p, C.byref(c))

    def hv_apply(self, rhs, state):
        r = np.ascontiguousarray(rhs, dtype=np.complex128).copy()
        s = np.ascontiguousarray(state, dtype=np.complex128)
        self.L.oc_hyperviscosity_apply(self.p, _dp(r), _dp(s))
        return r

    def hv_post_step(self, state, dt):
        s = np.ascontiguousarray(state, dtype=np.complex128)
        self.L.oc_hyperviscosity_post_step(self.p, _dp(s), dt)

    def project_forcing(self, f, eps):
        f = np.ascontiguousarray(f, dtype=np.complex128).copy()
        self.L.oc_forcing_project_and_scale(self.p, _dp(f), eps)
        return f


class OracleForcing:
    """Sequential reference-semantics forcing generator (forcing.c:53-184)."""

    def __init__(self, k2, nkx, nky, kmin, kmax, energy_target=1.0, seed=0, colored=False,
                 correlation_time=0.5, library=None):
        self.L = library if library is not None else lib()
        self.cfg = OcForcingConfig(kmin, kmax, energy_target, correlation_time, seed,
                                   1 if colored else 0)
        self.state = OcForcingState()
        self.k2 = np.ascontiguousarray(k2, dtype=np.float64).ravel()
        self.n = nkx * nky
        self.shape = (nky, nkx)
        fn = getattr(self.L, "oc_forcing_state_init", None) or self.L.ns_forcing_state_init
        self.rc = fn(C.byref(self.state), C.byref(self.cfg), _dp(self.k2), nkx, nky)

    def generate(self, dt):
        out = np.zeros(self.n, dtype=np.complex128)
        fn = getattr(self.L, "oc_forcing_generate", None) or self.L.ns_forcing_generate
        fn(C.byref(self.state), C.c_double(dt), _dp(self.k2), _dp(out))
        return out.reshape(self.shape)

    def close(self):
        fn = getattr(self.L, "oc_forcing_state_free", None) or self.L.ns_forcing_state_free
        fn(C.byref(self.state))


def ref_forcing_lib():
    """The reference's own forcing.c compiled into oracle/_ref/ (None if absent)."""
    if not os.path.exists(REF_FORCING):
        return None
    L = C.CDLL(REF_FORCING)
    L.ns_forcing_state_init.argtypes = [C.POINTER(OcForcingState), C.POINTER(OcForcingConfig),
                                        c_double_p, C.c_int, C.c_int]
    L.ns_forcing_state_init.restype = C.c_int
    L.ns_forcing_state_free.argtypes = [C.POINTER(OcForcingState)]
    L.ns_forcing_generate.argtypes = [C.POINTER(OcForcingState), C.c_double, c_double_p,
                                      c_double_p]
    return L


FAC_MIN, FAC_MAX = 0.2, 5.0


class OracleErrorControl:
    """Build-defined adaptive-dt error controller, restated step for step from
    the device's sequence.  The reference declares its parameters
    (AdaptiveTimeParams, include/time_integration.h:16-22: cfl_factor,
    dt_min, dt_max, safety_factor, adapt_enabled) and never uses them.

    Every attempt first judges the previous candidate step (if one is
    pending): err = |y4 - y2| / (atol + rtol |y4|) in the half-plane
    Parseval norm, fac = clamp(safety err^(-1/3), 0.2, 5) (fac = 5 for
    err = 0).  err <= 1 (or dt at dt_min) accepts; otherwise t goes back by
    dt and this attempt restores the previous state instead of stepping (a
    rejected candidate's forcing draw is discarded).  Either way the next
    step may use dt_err = dt fac.  A stepping attempt takes
    dt = clamp(min(CFL dt, dt_err), dt_min, dt_max), draws the forcing (if
    any) for that dt, keeps the state and steps.  finish() judges the last
    candidate, so that the state left behind is always an accepted one."""

    def __init__(self, o, rtol, atol=0.0, safety=0.9, dt_min=1e-8, dt_max=0.1,
                 integrating_factor=True, forcing=None, rate=0.0):
        self.o, self.rtol, self.atol, self.safety = o, rtol, atol, safety
        self.dt_min, self.dt_max = dt_min, dt_max
        self.integrating_factor = integrating_factor
        self.forcing, self.rate = forcing, rate
        self.pending = self.redo = False
        self.dt_err = dt_max
        self.accepted = self.rejected = 0
        self.err = 0.0
        self.norms = None
        self.prev = None
        self.log = []

    def _judge(self):
        if not self.pending:
            self.redo = False
            return
        num = float(np.sqrt(self.norms[0]))
        den = self.atol + self.rtol * float(np.sqrt(self.norms[1]))
        err = num / den if den > 0 else (float("inf") if num > 0 else 0.0)
        fac = FAC_MAX if err == 0 else min(FAC_MAX, max(FAC_MIN, self.safety * err ** (-1.0 / 3.0)))
        dt = self.o.d.dt
        self.err = err
        self.pending = False
        if err <= 1.0 or dt <= self.dt_min:
            self.accepted += 1
            self.redo = False
        else:
            self.rejected += 1
            self.o.d.t -= dt
            self.redo = True
        self.dt_err = dt * fac
        self.log.append((dt, err, not self.redo))

    def attempt(self):
        self._judge()
        o = self.o
        if self.redo:
            o.spec("omega_hat")[:] = self.prev
            o.compute_derivatives()
            return
        o.d.dt = min(max(min(o.cfl(), self.dt_err), self.dt_min), self.dt_max)
        self.prev = o.spec("omega_hat").copy()
        f = None
        if self.forcing is not None:
            f = o.project_forcing(self.forcing.generate(o.d.dt), self.rate)
        self.norms = o.advance_est(self.integrating_factor, f)
        self.pending = True

    def finish(self):
        self._judge()
        if self.redo:
            self.o.spec("omega_hat")[:] = self.prev
            self.o.compute_derivatives()
        self.redo = False
