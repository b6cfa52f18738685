"""Pins the oracle's adaptive-dt error controller (build-defined; the
reference declares AdaptiveTimeParams in include/time_integration.h:16-22
and uses it nowhere).

The embedded estimate is est = y4 - y2, the RK4 solution minus the midpoint
solution formed from the step's own second stage.  On linear decay
(Taylor-Green is one shell, so J = 0) with the classic scheme and rate
lambda = -(k^2/Re + mu), z = lambda dt:
    y4 = R(z) y,  R = 1 + z + z^2/2 + z^3/6 + z^4/24
    y2 = (1 + z + z^2/2) y        (explicit midpoint)
    est = (z^3/6 + z^4/24) y
and with the integrating factor both solutions are e^{lambda dt} y, est = 0
(up to the round-off of J ~ 1e-16 that the transforms leave).
The controller's decisions then follow from the scalar closed form alone.
"""
import numpy as np
import pytest

from conftest import rel_l2

TWO_PI = 2 * np.pi


def tg_oracle(oc, n=16, Re=100.0, mu=0.0, scale=1.0):
    o = oc.Oracle(n, n, TWO_PI, TWO_PI, Re)
    o.init_grids()
    o.d.drag = mu
    o.init_taylor_green()
    o.spec("omega_hat")[:] *= scale
    o.compute_derivatives()
    return o


def R4(z):
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


def test_classic_estimate_linear_decay_closed_form(oc):
    Re, mu, dt = 100.0, 3.0, 0.05
    o = tg_oracle(oc, Re=Re, mu=mu)
    y = o.spec("omega_hat").copy()
    o.d.dt = dt
    norms = o.advance_est(False)
    z = -(2.0 / Re + mu) * dt
    wt = np.where(np.arange(o.nkx) == 0, 1.0, 2.0)[None, :]
    y2 = np.sum(wt * np.abs(y) ** 2)
    assert norms[0] == pytest.approx((z ** 3 / 6 + z ** 4 / 24) ** 2 * y2, rel=1e-12)
    assert norms[1] == pytest.approx(R4(z) ** 2 * y2, rel=1e-12)
    assert rel_l2(o.spec("omega_hat"), R4(z) * y) < 1e-14


def test_integrating_factor_estimate_vanishes_on_linear_decay(oc):
    o = tg_oracle(oc, Re=100.0, mu=3.0)
    o.d.dt = 0.05
    norms = o.advance_est(True)
    assert norms[1] > 0.0 and norms[0] <= 1e-30 * norms[1]


def test_integrating_factor_estimate_is_the_lawson_midpoint(oc):
    """Nonlinear field: est equals y4 minus E (w0 + h k2) of the classical
    scheme written on v = e^{-L t} omega (k2 its second stage)."""
    n = 16
    o = oc.Oracle(n, n, Re=20.0)
    o.init_grids()
    o.hv_set(enabled=1, coefficient=40.0, order=4, cutoff_ratio=0.0)
    o.d.drag = 0.3
    o.init_random_phase(11, k_peak=2.0)
    w0 = o.spec("omega_hat").copy()
    L = -(o.k2 / 20.0 + o.hv_weights + 0.3)
    h = 0.02
    helper = oc.Oracle(n, n, Re=20.0)
    helper.init_grids()

    def N(w):
        helper.spec("omega_hat")[:] = w
        helper.velocity_from_vorticity()
        helper.nonlinear()
        return helper.spec("nonlinear_hat").copy()

    def f(t, v):
        return np.exp(-L * t) * N(np.exp(L * t) * v)

    k1 = f(0, w0)
    k2 = f(h / 2, w0 + h / 2 * k1)
    k3 = f(h / 2, w0 + h / 2 * k2)
    k4 = f(h, w0 + h * k3)
    y4 = np.exp(L * h) * (w0 + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4))
    est = y4 - np.exp(L * h) * (w0 + h * k2)
    wt = np.where(np.arange(o.nkx) == 0, 1.0, 2.0)[None, :]
    o.d.dt = h
    norms = o.advance_est(True)
    assert norms[0] == pytest.approx(np.sum(wt * np.abs(est) ** 2), rel=1e-10)
    assert norms[1] == pytest.approx(np.sum(wt * np.abs(y4) ** 2), rel=1e-12)
    assert norms[0] > 1e-12 * norms[1]


def scalar_controller(lam, y0, attempts, rtol, safety, dt_min, dt_max, dt_cap):
    """The controller on y' = lam y with the classic scheme, from the closed
    forms alone; dt_cap is the step limit the CFL rule gives (constant here)."""
    pending, redo, dt_err, t, y, prev, dt = False, False, dt_max, 0.0, y0, y0, None
    est = sol = None
    log = []

    def judge():
        nonlocal pending, redo, dt_err, t
        if not pending:
            redo = False
            return
        err = abs(est) / (rtol * abs(sol))
        fac = 5.0 if err == 0 else min(5.0, max(0.2, safety * err ** (-1.0 / 3.0)))
        pending = False
        redo = not (err <= 1.0 or dt <= dt_min)
        if redo:
            t -= dt
        dt_err = dt * fac
        log.append((dt, err, not redo))

    for _ in range(attempts):
        judge()
        if redo:
            y = prev
            continue
        dt = min(max(min(dt_cap, dt_err), dt_min), dt_max)
        prev = y
        z = lam * dt
        est, sol = (z ** 3 / 6 + z ** 4 / 24) * y, R4(z) * y
        y, t, pending = sol, t + dt, True
    judge()
    if redo:
        y = prev
    return log, t, y


def test_controller_rejects_and_recovers_on_linear_decay(oc):
    """Classic scheme on Taylor-Green with strong drag: the first attempts at
    dt_max are rejected (t rolls back, the state is restored), dt shrinks by
    at most 5x per rejection, then grows again; the sequence of (dt, err,
    accepted) and the final state match the scalar closed form."""
    Re, mu, rtol, safety = 100.0, 20.0, 1e-4, 0.9
    o = tg_oracle(oc, Re=Re, mu=mu, scale=1e-9)  # tiny velocity: the CFL rule gives MAX_DT
    y0 = o.spec("omega_hat").copy()
    assert o.cfl() == 0.1
    ctl = oc.OracleErrorControl(o, rtol, 0.0, safety, 1e-8, 0.1, integrating_factor=False)
    attempts = 40
    for _ in range(attempts):
        ctl.attempt()
    ctl.finish()
    lam = -(2.0 / Re + mu)
    log, t, yfac = scalar_controller(lam, 1.0, attempts, rtol, safety, 1e-8, 0.1, 0.1)
    assert len(ctl.log) == len(log)
    for (dt_o, err_o, acc_o), (dt_s, err_s, acc_s) in zip(ctl.log, log):
        assert acc_o == acc_s
        assert dt_o == pytest.approx(dt_s, rel=1e-12)
        assert err_o == pytest.approx(err_s, rel=1e-9)
    assert ctl.rejected >= 2 and ctl.accepted + ctl.rejected == len(log)
    assert o.d.t == pytest.approx(t, rel=1e-12)
    assert rel_l2(o.spec("omega_hat"), yfac * y0) < 1e-12


def test_controller_never_limits_an_exact_step(oc):
    """Integrating factor on linear decay: every estimate is zero to
    round-off, every step is accepted, and dt is the CFL rule's (clamped to
    dt_max), the controller's growth factor being at its cap of 5."""
    o = tg_oracle(oc, Re=100.0, mu=5.0)
    ctl = oc.OracleErrorControl(o, 1e-10, 0.0, 0.9, 1e-8, 0.05, integrating_factor=True)
    dts = []
    for _ in range(12):
        expect = min(o.cfl(), 0.05)
        ctl.attempt()
        dts.append((o.d.dt, expect))
    ctl.finish()
    assert ctl.rejected == 0 and ctl.accepted == 12
    assert all(e <= 1e-5 for _, e, _ in ctl.log)
    for got, want in dts:
        assert got == want
