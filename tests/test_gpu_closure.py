"""GPU closure network (bf16 MFMA, FP32 accumulation) against the FP64 CPU
oracle (oracle/closure_oracle.py).  bf16 keeps 8 significant bits: inputs,
weights and every stored activation are rounded to it, so the tolerances are
set from that precision (relative L2 2e-2 for outputs and gradients;
a wrong tap, channel, sign or lane layout gives errors of order 1)."""
import ctypes as C

import numpy as np
import pytest

import closure_oracle as CO
from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu] {request.node.name}", flush=True)


def fp(a):
    return np.ascontiguousarray(a, dtype=np.float32).ctypes.data_as(C.POINTER(C.c_float))


def make(ns, n, batch, seed=3):
    L = ns.load()
    err = C.create_string_buffer(256)
    h = L.nsd_closure_create(n, batch, seed, err, 256)
    assert h, err.value
    return L, h


# n = 16, 32: per-tile kernel; n = 64, 128: the LDS-tiled forward kernels; at
# 128 x 16 the persistent one walks two tiles per workgroup on a 256-CU device,
# the second patch loaded while the first is computed
@pytest.mark.parametrize("n,batch", [(16, 1), (32, 3), (64, 2), (128, 2), (128, 16)])
def test_forward_matches_oracle(ns, n, batch):
    L, h = make(ns, n, batch)
    rng = np.random.default_rng(n)
    P = L.nsd_closure_param_count(h)
    assert P == CO.param_count()
    p = (rng.standard_normal(P) * 0.3).astype(np.float32)
    assert L.nsd_closure_params(h, fp(p), 1) == 0
    x = rng.standard_normal((batch, n, n)).astype(np.float32)
    y = np.zeros((batch, n, n), np.float32)
    assert L.nsd_closure_forward(h, fp(x), batch, y.ctypes.data_as(C.POINTER(C.c_float))) == 0
    y_ref, _ = CO.forward(p.astype(np.float64), x.astype(np.float64))
    err = rel_l2(y, y_ref)
    print(f"n={n} batch={batch}: forward rel L2 {err:.2e}", flush=True)
    assert err <= 2e-2
    L.nsd_closure_destroy(h)


# n = 16, 32: per-block weight-gradient kernel; n = 64, 128: the LDS-tiled one
# (with 16 and 64 tiles per member) for the 32-channel-input layers
@pytest.mark.parametrize("n,batch", [(16, 2), (32, 2), (64, 1), (128, 2)])
def test_gradient_matches_oracle(ns, n, batch):
    L, h = make(ns, n, batch, seed=11)
    rng = np.random.default_rng(100 + n)
    P = L.nsd_closure_param_count(h)
    p = (rng.standard_normal(P) * 0.3).astype(np.float32)
    assert L.nsd_closure_params(h, fp(p), 1) == 0
    x = rng.standard_normal((batch, n, n)).astype(np.float32)
    t = rng.standard_normal((batch, n, n)).astype(np.float32)
    g = np.zeros(P, np.float32)
    loss = C.c_double()
    assert L.nsd_closure_gradient(h, fp(x), fp(t), batch, C.byref(loss),
                                  g.ctypes.data_as(C.POINTER(C.c_float))) == 0
    loss_ref, g_ref = CO.loss_and_grad(p.astype(np.float64), x.astype(np.float64),
                                       t.astype(np.float64))
    print(f"n={n}: loss {loss.value:.6e} vs {loss_ref:.6e}", flush=True)
    assert loss.value == pytest.approx(loss_ref, rel=3e-2)
    off = 0
    for l, (ci, co) in enumerate(zip(CO.CIN, CO.COUT)):
        nw = co * 9 * ci
        ew = rel_l2(g[off:off + nw], g_ref[off:off + nw])
        eb = rel_l2(g[off + nw:off + nw + co], g_ref[off + nw:off + nw + co])
        print(f"layer {l}: weight grad rel L2 {ew:.2e}, bias grad {eb:.2e}", flush=True)
        assert ew <= 2e-2 and eb <= 2e-2
        off += nw + co
    L.nsd_closure_destroy(h)


# ------------------------------------------------------------ subgrid data

def ctx_of(ns, n, fp32, Re=1e3):
    s = ns.Spectral(n, n, 2 * np.pi, 2 * np.pi, Re)
    s.init(taylor_green=False)
    s.set_precision(fp32)
    return s


def test_subgrid_sample_matches_oracle(ns, oc):
    """x = filtered vorticity, t = filter(N_fine) - N_coarse(x), in physical
    space on the coarse grid, against the oracle's nonlinear term and
    transforms (FP64 contexts; outputs are FP32)."""
    nf, nc = 64, 32
    f = ctx_of(ns, nf, False)
    f.random_phase(21, k_peak=6.0)
    f.sync_device()
    c = ctx_of(ns, nc, False)
    c.sync_device()
    L = ns.load()
    x = np.zeros((nc, nc), np.float32)
    t = np.zeros((nc, nc), np.float32)
    # device outputs: a network's data-set slot
    err = C.create_string_buffer(256)
    net = L.nsd_closure_create(nc, 1, 1, err, 256)
    assert net and L.nsd_closure_dataset_reserve(net, 1) == 0
    px, pt = C.c_void_p(), C.c_void_p()
    assert L.nsd_closure_dataset_slot(net, 0, C.byref(px), C.byref(pt)) == 0
    assert L.nsd_subgrid_sample(f.ctx, 0, c.ctx, px, pt) == 0, c.L.nsd_error(c.ctx)
    assert L.nsd_closure_dataset_read(net, 0, x.ctypes.data_as(C.POINTER(C.c_float)),
                                      t.ctypes.data_as(C.POINTER(C.c_float))) == 0
    # oracle
    of = oc.Oracle(nf, nf, 2 * np.pi, 2 * np.pi, 1e3)
    of.init_grids()
    of.init_random_phase(21, 6.0, 1.0)
    wf = of.spec("omega_hat").copy()
    of.velocity_from_vorticity()
    of.nonlinear()
    Nf = of.spec("nonlinear_hat").copy()
    o = oc.Oracle(nc, nc, 2 * np.pi, 2 * np.pi, 1e3)
    o.init_grids()
    mask = o.mask.astype(bool)
    kyc = np.where(np.arange(nc) <= nc // 2, np.arange(nc), np.arange(nc) - nc)

    def filt(a):
        out = np.zeros((nc, nc // 2 + 1), complex)
        rows = np.where(kyc >= 0, kyc, kyc + nf)
        out[:, :] = a[rows, : nc // 2 + 1]
        return np.where(mask, out, 0)

    wbar = filt(wf)
    o.spec("omega_hat")[:] = wbar
    o.velocity_from_vorticity()
    o.nonlinear()
    S = filt(Nf) - o.spec("nonlinear_hat")
    x_ref = o.c2r(wbar)
    t_ref = o.c2r(S)
    ex, et = rel_l2(x, x_ref), rel_l2(t, t_ref)
    print(f"sample vs oracle: vorticity {ex:.1e}, subgrid tendency {et:.1e}", flush=True)
    assert ex <= 1e-6 and et <= 1e-6
    L.nsd_closure_destroy(net)


def oracle_sample(oc, nf, nc, seed, k_peak):
    """x = filtered vorticity and t = filter(N_fine) - N_coarse(x) in physical
    space on the coarse grid, from the oracle's random-phase field, nonlinear
    term and transforms."""
    of = oc.Oracle(nf, nf, 2 * np.pi, 2 * np.pi, 1e3)
    of.init_grids()
    of.init_random_phase(seed, k_peak, 1.0)
    wf = of.spec("omega_hat").copy()
    of.velocity_from_vorticity()
    of.nonlinear()
    Nf = of.spec("nonlinear_hat").copy()
    o = oc.Oracle(nc, nc, 2 * np.pi, 2 * np.pi, 1e3)
    o.init_grids()
    mask = o.mask.astype(bool)
    kyc = np.where(np.arange(nc) <= nc // 2, np.arange(nc), np.arange(nc) - nc)

    def filt(a):
        rows = np.where(kyc >= 0, kyc, kyc + nf)
        return np.where(mask, a[rows, : nc // 2 + 1], 0)

    wbar = filt(wf)
    o.spec("omega_hat")[:] = wbar
    o.velocity_from_vorticity()
    o.nonlinear()
    return o.c2r(wbar), o.c2r(filt(Nf) - o.spec("nonlinear_hat"))


def test_subgrid_sample_from_a_4096_fine_grid_vs_oracle(ns, oc):
    """Samples from a 4096^2 fine grid against the oracle: FP64, and FP32,
    where the row stage writes J in 2 x 2 blocks and the forward column
    transform reads them so.  Outputs are FP32: the FP64 sample is held to
    1e-6; the FP32 one keeps the vorticity to 1e-6 and the subgrid tendency, a
    difference of two nonlinear terms each good to about 5e-7, to 1e-4 (a J
    read in the wrong layout scrambles it at order 1)."""
    # k_peak near the coarse cutoff (85), so that the subgrid tendency is of
    # the order of the nonlinear term itself
    nf, nc, seed, kp = 4096, 256, 8, 48.0
    x_ref, t_ref = oracle_sample(oc, nf, nc, seed, kp)
    L = ns.load()
    err = C.create_string_buffer(256)
    net = L.nsd_closure_create(nc, 1, 1, err, 256)
    assert net and L.nsd_closure_dataset_reserve(net, 2) == 0
    errs = []
    for i, fp32 in enumerate((False, True)):
        f = ctx_of(ns, nf, fp32)
        f.random_phase(seed, k_peak=kp)
        f.sync_device()
        c = ctx_of(ns, nc, True)
        c.sync_device()
        px, pt = C.c_void_p(), C.c_void_p()
        assert L.nsd_closure_dataset_slot(net, i, C.byref(px), C.byref(pt)) == 0
        assert L.nsd_subgrid_sample(f.ctx, 0, c.ctx, px, pt) == 0, c.L.nsd_error(c.ctx)
        x = np.zeros((nc, nc), np.float32)
        t = np.zeros((nc, nc), np.float32)
        assert L.nsd_closure_dataset_read(net, i, x.ctypes.data_as(C.POINTER(C.c_float)),
                                          t.ctypes.data_as(C.POINTER(C.c_float))) == 0
        errs.append((rel_l2(x, x_ref), rel_l2(t, t_ref)))
        print(f"4096^2 {'FP32' if fp32 else 'FP64'} fine grid vs oracle: vorticity "
              f"{errs[-1][0]:.1e}, subgrid tendency {errs[-1][1]:.1e}", flush=True)
        f.close()
        c.close()
    L.nsd_closure_destroy(net)
    assert errs[0][0] <= 1e-6 and errs[0][1] <= 1e-6
    assert errs[1][0] <= 1e-6 and errs[1][1] <= 1e-4


def test_closure_source_enters_the_step(ns):
    """A network whose only non-zero parameter is the last bias b outputs the
    constant tendency b * out_scale; without drag the mean vorticity (the
    k = 0 mode, which the nonlinear term never changes) then grows as
    b * out_scale * t."""
    n = 64
    s = ctx_of(ns, n, True, Re=500.0)
    s.set_scheme(True)
    s.random_phase(5, k_peak=6.0)
    s.d.dt = 1e-3
    s.sync_device()
    L = ns.load()
    err = C.create_string_buffer(256)
    net = L.nsd_closure_create(n, 1, 1, err, 256)
    P = L.nsd_closure_param_count(net)
    p = np.zeros(P, np.float32)
    p[-1] = 0.25
    assert L.nsd_closure_params(net, fp(p), 1) == 0
    assert L.nsd_closure_set_scales(net, 1.0, 2.0) == 0
    assert L.nsd_attach_closure(s.ctx, net) == 0
    assert s.advance_n(40, False) == 0
    assert s.sync_host() == 0
    w00 = s.spec("omega_hat")[0, 0]
    print(f"mean vorticity {w00.real:.7f} at t = {s.d.t:.4f} (expected {0.5 * s.d.t:.7f})", flush=True)
    assert abs(w00.real - 0.5 * s.d.t) <= 1e-5 and abs(w00.imag) <= 1e-6
    assert L.nsd_attach_closure(s.ctx, None) == 0
    s.close()
    L.nsd_closure_destroy(net)


def test_dataset_and_training_reduce_the_loss(ns):
    nf, nc, members = 256, 64, 4
    L = ns.load()
    err = C.create_string_buffer(256)
    fine = L.nsd_create_batch(nf, nf, 2 * np.pi, 2 * np.pi, ns.NSD_FP32, members, err, 256)
    assert fine, err.value
    p = ns.NsdParams()
    p.Re, p.drag, p.cfl_factor, p.scheme = 1e4, 0.1, 0.5, 1
    kmax = np.sqrt(2.0) * nf / 2
    p.hv_on, p.hv_order, p.hv_coeff, p.hv_kc, p.hv_invbw = 1, 8, 1e-16 * kmax ** 8, 0.0, 1 / kmax
    p.force_on, p.force_kmin, p.force_kmax, p.force_energy_target, p.force_rate = 1, 4.0, 6.0, 1.0, 1e-3
    assert L.nsd_set_params(fine, C.byref(p)) == 0
    seeds = (C.c_ulonglong * members)(*range(1, members + 1))
    assert L.nsd_set_member_params(fine, None, seeds) == 0
    assert L.nsd_batch_random_phase(fine, seeds, 10.0, 1.0) == 0
    coarse = L.nsd_create(nc, nc, 2 * np.pi, 2 * np.pi, ns.NSD_FP64, err, 256)
    assert coarse, err.value
    net = L.nsd_closure_create(nc, 8, 7, err, 256)
    snaps = 16
    assert L.nsd_closure_dataset_reserve(net, snaps * members) == 0
    assert L.nsd_subgrid_dataset(net, fine, coarse, snaps, 10, 0) == 0, L.nsd_error(coarse)
    l0, l1, lv = C.c_double(), C.c_double(), C.c_double()
    assert L.nsd_closure_train(net, snaps * members, 30, 8, 2e-3, 3, C.byref(l0), C.byref(l1),
                               C.byref(lv)) == 0, L.nsd_closure_error(net)
    print(f"training loss {l0.value:.3f} -> {l1.value:.3f}, validation {lv.value:.3f}", flush=True)
    assert l1.value < 0.7 * l0.value and lv.value < 1.0
    for h in (fine, coarse):
        L.nsd_destroy(h)
    L.nsd_closure_destroy(net)


def test_trained_closure_beats_the_plain_coarse_solve(ns):
    """A posteriori: coarse runs with the trained network against the plain
    coarse solve, both against the filtered fine ensemble (1024^2 -> 128^2,
    8 members): the closure must at least halve the error of the ensemble
    mean spectrum, keep enstrophy within the 10 % acceptance bound, and be
    accepted (else the workflow falls back to the plain solve)."""
    import importlib
    cl = importlib.import_module(ns.__name__ + ".closure")
    net = cl.Net(128, max_batch=16, seed=1)
    seeds = list(range(1, 33))
    net.reserve(len(seeds) * 32)
    cl.generate(net, 1024, seeds, 32, 10)
    l0, l1, lv = net.train(8, batch=16, lr=2e-3)
    r = cl.evaluate(net, 1024, list(range(2001, 2009)), 0.3)
    print(f"loss {l0:.3f} -> {l1:.3f} (validation {lv:.3f}); plain: spectrum "
          f"{r['plain']['spectrum_mean_rel_err']:.3f}, enstrophy {r['plain']['enstrophy_rel_err']:.3f}; "
          f"closure: spectrum {r['closure']['spectrum_mean_rel_err']:.3f}, enstrophy "
          f"{r['closure']['enstrophy_rel_err']:.3f}, {r['closure']['speedup_vs_fine']:.1f}x the fine runs",
          flush=True)
    assert lv < 1.0
    assert r["closure"]["spectrum_mean_rel_err"] < 0.5 * r["plain"]["spectrum_mean_rel_err"]
    assert r["closure_accepted"] and r["result"] == "closure"
    net.close()


def test_rejected_closure_falls_back_to_the_fine_solve(ns):
    """A network that drives the enstrophy away from the fine solve (its only
    non-zero parameter is a large output bias, which raises the mean
    vorticity at 400 per unit time) is rejected by the 10 % enstrophy-drift
    rule, and the result is the fine solve itself (filtered to the coarse
    grid), not the closure run and not the plain coarse solve."""
    import importlib
    cl = importlib.import_module(ns.__name__ + ".closure")
    net = cl.Net(64, max_batch=4, seed=1)
    P = net.L.nsd_closure_param_count(net.h)
    p = np.zeros(P, np.float32)
    p[-1] = 1.0
    assert net.L.nsd_closure_params(net.h, fp(p), 1) == 0
    assert net.L.nsd_closure_set_scales(net.h, 1.0, 400.0) == 0
    r = cl.evaluate(net, 256, [5, 6, 7, 8], 0.05)
    print(f"closure enstrophy drift {r['closure']['enstrophy_rel_err']:.3f} -> {r['result']}",
          flush=True)
    assert r["closure"]["enstrophy_rel_err"] > 0.10
    assert not r["closure_accepted"] and r["result"] == "fine (fallback)"
    assert r["result_speedup_vs_fine"] == 1.0
    assert len(r["fields"]) == 4
    for got, fine in zip(r["fields"], r["fine_filtered"]):
        assert got.shape == (64, 33) and np.array_equal(got, fine)
    net.close()
