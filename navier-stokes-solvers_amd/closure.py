"""Learned subgrid closure workflow for coarse runs (host orchestration of the
C-ABI in include/ns_closure.h and include/ns_device.h; every field operation
runs on the GPU).

  generate()  fine forced runs (a batch of seeds) sampled every `stride` steps:
              coarse vorticity and subgrid tendency pairs in the network's data set
  Net.train() Adam on the bf16 MFMA network
  evaluate()  a fine run against plain and closure-augmented coarse runs from
              the same filtered initial field: shell energy spectra and
              enstrophy (1/2 <omega^2>, spectral_analysis.c:22-33, by Parseval)
              against the filtered fine run, wall time per simulated time; the
              closure is accepted only if its enstrophy stays within
              `max_drift` (10 %) of the filtered fine run, else the fine solve
              (the judge, filtered to the coarse grid) is the result: the
              fallback

Physics of every run: forced 2-D turbulence as the 4096^2 throughput
workload (Re 1e4, drag 0.1, del^8 nu4 = 1e-16 through the integrating
factor, white forcing on 4 <= |k| <= 6 at injection rate 1e-3), FP32 stages.
"""
import ctypes as C
import time

import numpy as np

from . import NSD_FP32, NSD_FP64, NsdParams, c_double_p, load

PHYSICS = dict(Re=1e4, drag=0.1, nu4=1e-16, kmin=4.0, kmax=6.0, rate=1e-3, cfl=0.5, k_peak=10.0)
TWO_PI = 2 * np.pi


def params(n, phys=PHYSICS, force_seed=0):
    p = NsdParams()
    p.force_seed = int(force_seed)
    p.Re, p.drag, p.cfl_factor, p.scheme = phys["Re"], phys["drag"], phys["cfl"], 1
    kmax = np.sqrt(2.0) * n / 2.0  # hyperviscosity.c:68-73, L = 2 pi
    p.hv_on, p.hv_order, p.hv_kc, p.hv_invbw = 1, 8, 0.0, 1.0 / kmax
    p.hv_coeff = phys["nu4"] * kmax ** 8
    if phys.get("rate", 0.0) > 0.0:
        p.force_on, p.force_kmin, p.force_kmax = 1, phys["kmin"], phys["kmax"]
        p.force_energy_target, p.force_rate = 1.0, phys["rate"]
    return p


def _check(L, rc, ctx=None, what=""):
    if rc != 0:
        msg = L.nsd_error(ctx).decode() if ctx else ""
        raise RuntimeError(f"{what} failed ({rc}) {msg}")


def context(n, precision=NSD_FP32, members=1, phys=PHYSICS, seeds=None, force_seed=0):
    L = load()
    err = C.create_string_buffer(256)
    h = L.nsd_create_batch(n, n, TWO_PI, TWO_PI, precision, members, err, 256)
    if not h:
        raise RuntimeError(err.value.decode())
    p = params(n, phys, force_seed)
    _check(L, L.nsd_set_params(h, C.byref(p)), h, "nsd_set_params")
    # launch shapes timed on this batch (lengths 64..1024; the tuner clears
    # the state, so it runs before the initial field), as ensembles do
    _check(L, L.nsd_autotune(h, 3, 0, None, None), h, "nsd_autotune")
    if seeds is not None:
        sd = (C.c_ulonglong * members)(*[int(s) for s in seeds])
        _check(L, L.nsd_set_member_params(h, None, sd), h, "member params")
        _check(L, L.nsd_batch_random_phase(h, sd, phys["k_peak"], 1.0), h, "random phase")
    return h


class Net:
    """The closure network (ns_closure.h)."""

    def __init__(self, n, max_batch=16, seed=1):
        self.L = load()
        err = C.create_string_buffer(256)
        self.h = self.L.nsd_closure_create(n, max_batch, seed, err, 256)
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.n, self.count = n, 0

    def close(self):
        if self.h:
            self.L.nsd_closure_destroy(self.h)
            self.h = None

    def reserve(self, count):
        _check(self.L, self.L.nsd_closure_dataset_reserve(self.h, count), None, "reserve")
        self.count = count

    def train(self, epochs, batch=16, lr=2e-3, seed=3, count=None):
        l0, l1, lv = C.c_double(), C.c_double(), C.c_double()
        rc = self.L.nsd_closure_train(self.h, count or self.count, epochs, batch, lr, seed,
                                      C.byref(l0), C.byref(l1), C.byref(lv))
        if rc:
            raise RuntimeError(self.L.nsd_closure_error(self.h).decode())
        return l0.value, l1.value, lv.value

    def params(self):
        p = np.zeros(self.L.nsd_closure_param_count(self.h), np.float32)
        self.L.nsd_closure_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float)), 0)
        return p


def generate(net, nf, seeds, snapshots, stride, spinup=0, phys=PHYSICS):
    """Sample `snapshots` times every `stride` fine steps, all seeds at once
    (one batch context); returns the number of pairs written."""
    L = load()
    fine = context(nf, NSD_FP32, len(seeds), phys, seeds)
    coarse = context(net.n, NSD_FP64, 1, phys)
    try:
        if spinup:
            _check(L, L.nsd_advance_graph(fine, spinup, 1), fine, "spin-up")
        rc = L.nsd_subgrid_dataset(net.h, fine, coarse, snapshots, stride, 0)
        _check(L, rc, coarse, "subgrid data set")
    finally:
        L.nsd_destroy(fine)
        L.nsd_destroy(coarse)
    return snapshots * len(seeds)


def filter_to(w_fine, nc):
    """Fine reference-layout spectrum -> the coarse grid's kept modes."""
    nf = w_fine.shape[0]
    ky = np.where(np.arange(nc) <= nc // 2, np.arange(nc), np.arange(nc) - nc)
    rows = np.where(ky >= 0, ky, ky + nf)
    out = w_fine[rows, : nc // 2 + 1].copy()
    kx = np.arange(nc // 2 + 1)
    keep = (kx[None, :] <= (2 * (nc // 2 + 1)) // 3) & (np.abs(ky)[:, None] <= (2 * nc) // 6)
    return np.where(keep, out, 0)


def spectrum(w, kmax):
    """Shell energy spectrum E(k) = sum over |k| in [k - 1/2, k + 1/2) of
    1/2 |w|^2 / k^2 (half-plane weights), k = 1..kmax, from a reference-layout
    vorticity spectrum."""
    n = w.shape[0]
    kx = np.arange(w.shape[1])
    ky = np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)
    k2 = kx[None, :] ** 2 + ky[:, None] ** 2
    wt = np.where((kx == 0) | (2 * kx == n), 1.0, 2.0)[None, :]
    e = np.where(k2 > 0, 0.5 * wt * np.abs(w) ** 2 / np.maximum(k2, 1), 0.0)
    b = np.rint(np.sqrt(k2)).astype(int)
    E = np.bincount(b.ravel(), weights=e.ravel(), minlength=kmax + 1)
    return E[1:kmax + 1]


def _state(L, h, n, member=0):
    out = np.zeros((n, n // 2 + 1), np.complex128)
    _check(L, L.nsd_member_download(h, member, out.ctypes.data_as(c_double_p)), h, "download")
    return out


def _time(L, h):
    t, dt, st = C.c_double(), C.c_double(), C.c_int()
    L.nsd_get_time(h, C.byref(t), C.byref(dt), C.byref(st))
    return t.value, dt.value, st.value


def _enstrophy(w):
    n = w.shape[0]
    kx = np.arange(w.shape[1])
    wt = np.where((kx == 0) | (2 * kx == n), 1.0, 2.0)[None, :]
    return float(0.5 * np.sum(wt * np.abs(w) ** 2))


def evaluate(net, nf, seeds, t_end, phys=PHYSICS, max_drift=0.10):
    """Ensembles of fine runs, plain coarse runs and closure coarse runs (one
    member per seed, batched) from the filtered fine initial fields, to t_end
    with fixed steps (coarse dt = fine dt x nf / nc).  Errors of the ensemble
    mean shell spectrum and enstrophy against the filtered fine ensemble, and
    wall times per simulated time.  res["fields"] holds the result's member
    spectra on the coarse grid: the closure runs if accepted, else the fine
    runs filtered to the coarse grid (res["result"] names which)."""
    L = load()
    nc, M = net.n, len(seeds)
    fine = context(nf, NSD_FP32, M, phys, seeds)
    # fixed step from the initial CFL limits (the flows decay, so it stays stable)
    _check(L, L.nsd_advance(fine, 1, 1), fine, "first step")
    dts = np.zeros(M)
    L.nsd_batch_scalars(fine, None, dts.ctypes.data_as(c_double_p), None)
    dt_f = 0.8 * float(dts.min())
    ratio = nf // nc
    steps_c = int(np.ceil(t_end / (dt_f * ratio)))
    steps_f = steps_c * ratio
    w0 = [_state(L, fine, nf, m) for m in range(M)]
    L.nsd_destroy(fine)
    res = {"nf": nf, "nc": nc, "members": M, "t_end": steps_f * dt_f, "dt_fine": dt_f,
           "steps_fine": steps_f, "steps_coarse": steps_c}
    runs = {}
    sd = (C.c_ulonglong * M)(*[int(x) for x in seeds])
    for name, n, steps, dt in (("fine", nf, steps_f, dt_f), ("plain", nc, steps_c, dt_f * ratio),
                               ("closure", nc, steps_c, dt_f * ratio)):
        h = context(n, NSD_FP32, M, phys)
        _check(L, L.nsd_set_member_params(h, None, sd), h, "forcing seeds")

        def reset():
            for m in range(M):
                w = w0[m] if n == nf else filter_to(w0[m], nc)
                _check(L, L.nsd_member_upload(h, m, np.ascontiguousarray(w).ctypes.data_as(c_double_p)),
                       h, "upload")
            _check(L, L.nsd_set_forcing_counter(h, 1), h, "forcing counter")
            _check(L, L.nsd_set_time(h, 0.0, dt), h, "time")

        reset()
        if name == "closure":
            _check(L, L.nsd_attach_closure(h, net.h), h, "attach")
        _check(L, L.nsd_advance_graph(h, 1, 0), h, "warm-up step")  # graph capture, untimed
        reset()
        _check(L, L.nsd_synchronize(h), h, "sync")
        t0 = time.perf_counter()
        _check(L, L.nsd_advance_graph(h, steps, 0), h, name)
        _check(L, L.nsd_synchronize(h), h, "sync")
        wall = time.perf_counter() - t0
        st = np.zeros(M, np.int32)
        L.nsd_batch_scalars(h, None, None, st.ctypes.data_as(C.POINTER(C.c_int)))
        ws = [_state(L, h, n, m) for m in range(M)]
        runs[name] = ([filter_to(w, nc) for w in ws] if n == nf else ws, wall, int(st.max()))
        L.nsd_destroy(h)
    kmax = int((2 * nc) // 6)
    Ef = np.mean([spectrum(w, kmax) for w in runs["fine"][0]], axis=0)
    Zf = np.mean([_enstrophy(w) for w in runs["fine"][0]])
    res["wall_fine_s"] = runs["fine"][1]
    for name in ("plain", "closure"):
        ws, wall, st = runs[name]
        E = np.mean([spectrum(w, kmax) for w in ws], axis=0)
        rel = np.abs(E - Ef) / np.maximum(Ef, 1e-300)
        res[name] = {"spectrum_mean_rel_err": float(np.mean(rel)),
                     "spectrum_max_rel_err": float(np.max(rel)),
                     # shell of the largest error, and the fine spectrum there
                     # relative to its peak
                     "spectrum_max_shell": int(np.argmax(rel)) + 1,
                     "spectrum_max_shell_level": float(Ef[int(np.argmax(rel))] / np.max(Ef)),
                     "enstrophy_rel_err": abs(np.mean([_enstrophy(w) for w in ws]) / Zf - 1.0),
                     "wall_s": wall, "speedup_vs_fine": runs["fine"][1] / wall, "status": st}
    res["closure_accepted"] = res["closure"]["enstrophy_rel_err"] <= max_drift and \
        res["closure"]["status"] == 0
    # The fine solve is the judge and the fallback: a rejected closure run is
    # replaced by the fine runs (filtered to the coarse grid, the grid of the
    # result either way).
    res["fine_filtered"] = runs["fine"][0]
    if res["closure_accepted"]:
        res["result"] = "closure"
        res["fields"] = runs["closure"][0]
        res["result_speedup_vs_fine"] = res["closure"]["speedup_vs_fine"]
    else:
        res["result"] = "fine (fallback)"
        res["fields"] = runs["fine"][0]
        res["result_speedup_vs_fine"] = 1.0
    return res
