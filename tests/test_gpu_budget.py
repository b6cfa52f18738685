"""Energy and enstrophy budgets of the GPU state (the forced integrating-factor
workload of the benchmark, at 256^2 and 512^2).

With the half-spectrum weights w (1 on the kx = 0 column, 2 elsewhere) the
energy and enstrophy of a state are E = 1/2 sum w |W|^2 / k^2 and
Z = 1/2 sum w |W|^2 (Parseval), and the equation the step integrates,
dW/dt = N(W) - (nu k^2 + mu + h(k)) W with the hyperviscous rate
h(k) = nu4 |k|^8, gives

  dE/dt = -2 nu Z - 2 mu E - sum w h |W|^2 / k^2
  dZ/dt = -2 nu P - 2 mu Z - sum w h |W|^2,     P = 1/2 sum w k^2 |W|^2,

because the dealiased Jacobian conserves both (up to the single extra x
column the reference keeps, quirk Q1, whose modes hold nothing here).  The
reference pins only monotone decay (tests/test_conservation.c:110-186); these
tests close the budgets.

Forcing off: E(t_N) - E(t_0) against minus the composite Simpson integral of
the dissipation over N fixed steps (quadrature error O(dt^4), relative ~1e-10
here; RK4 truncation is smaller still).  Forcing on: each step adds
sqrt(dt) f after the deterministic update (f drawn by the same counter
SplitMix64 stream as the oracle's generator, projected to injection rate
eps = 1/2 sum w |f|^2 / k^2), so W_det = W_{n+1} - sqrt(dt) f must satisfy
the deterministic budget over the step (trapezoid, relative error
~dt^2 / (12 tau^2) at dt = 5e-5) and the kick must inject eps dt plus the
cross term.  Tolerances are stated per test; a dropped term (drag,
hyperviscosity, viscosity, the kick's sqrt(dt)) misses them by orders of
magnitude.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi
RE, DRAG, NU4, RATE = 1e4, 0.1, 1e-16, 1e-3


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu] {request.node.name}", flush=True)


def wavenumbers(n):
    kx = np.arange(n // 2 + 1, dtype=float)
    ky = np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n).astype(float)
    k2 = kx[None, :] ** 2 + ky[:, None] ** 2
    w = np.where(kx == 0, 1.0, 2.0)[None, :] * np.ones_like(k2)
    return k2, w


def budget_terms(W, k2, w):
    """E, Z, dE/dt and dZ/dt of a reference-layout spectrum (FP64)."""
    a2 = np.abs(W) ** 2
    inv_k2 = np.where(k2 > 0, 1.0 / np.maximum(k2, 1e-300), 0.0)
    h = NU4 * k2 ** 4
    rate = k2 / RE + DRAG + h
    E = 0.5 * np.sum(w * a2 * inv_k2)
    Z = 0.5 * np.sum(w * a2)
    dE = -np.sum(w * rate * a2 * inv_k2)
    dZ = -np.sum(w * rate * a2)
    return E, Z, dE, dZ


def make(ns, n, fp32, forced, seed=7):
    s = ns.Spectral(n, n, TWO_PI, TWO_PI, RE)
    s.init(taylor_green=False)
    s.hv_set(enabled=1, coefficient=NU4 * (np.sqrt(2) * n / 2) ** 8, order=8, cutoff_ratio=0.0)
    s.set_precision(fp32)
    s.set_scheme(True)
    s.set_drag(DRAG)
    if forced:
        s.set_forcing(4.0, 6.0, RATE, seed=42)
    s.random_phase(seed, k_peak=10.0, energy=1.0)
    s.sync_device()
    return s


def download(ns, s):
    out = np.zeros((s.ny, s.nkx), np.complex128)
    s.check(s.L.nsd_state_download(s.ctx, ns._dp(out)))
    return out


@pytest.mark.parametrize("n,fp32,tol", [(256, False, 1e-8), (512, False, 1e-8),
                                        (256, True, 2e-4)])
def test_unforced_budget_closes(ns, n, fp32, tol):
    steps, dt = 40, 2e-4
    s = make(ns, n, fp32, forced=False)
    k2, w = wavenumbers(n)
    s.d.dt = dt
    s.check(s.L.nsd_set_time(s.ctx, 0.0, dt))
    hist = []
    for i in range(steps + 1):
        W = download(ns, s)
        hist.append(budget_terms(W, k2, w))
        if i == 0:
            # device Parseval sums agree with the host ones
            e_dev, z_dev = s.device_diagnostics()
            assert e_dev == pytest.approx(hist[0][0], rel=1e-6 if fp32 else 1e-13)
            assert z_dev == pytest.approx(hist[0][1], rel=1e-6 if fp32 else 1e-13)
        if i < steps:
            assert s.advance_n(1, False) == 0
    E, Z, dE, dZ = (np.array(c) for c in zip(*hist))
    simpson = np.ones(steps + 1)
    simpson[1:-1:2], simpson[2:-1:2] = 4.0, 2.0
    intE, intZ = dt / 3.0 * np.sum(simpson * dE), dt / 3.0 * np.sum(simpson * dZ)
    rE = abs((E[-1] - E[0]) - intE) / abs(intE)
    rZ = abs((Z[-1] - Z[0]) - intZ) / abs(intZ)
    print(f"{n}^2 {'FP32' if fp32 else 'FP64'}: E {E[0]:.6f} -> {E[-1]:.6f}, "
          f"Z {Z[0]:.4f} -> {Z[-1]:.4f}; budget residual / dissipation: "
          f"energy {rE:.2e}, enstrophy {rZ:.2e}", flush=True)
    assert rE <= tol and rZ <= tol
    s.close()


def test_forced_budget_closes_per_step(ns, oc):
    """Forcing on, FP64, 256^2: with the oracle's draw f, every step's state
    minus the kick sqrt(dt) f must close the deterministic budget over the
    step (trapezoid rule); a kick of the wrong size or sign leaves a residual
    of the order of the kick's energy, sqrt(dt) sum w Re(conj(W) f) / k^2
    + eps dt, which is larger than the step's dissipation here."""
    n, steps, dt = 256, 20, 5e-5
    s = make(ns, n, False, forced=True)
    k2, w = wavenumbers(n)
    s.check(s.L.nsd_set_time(s.ctx, 0.0, dt))
    o = oc.Oracle(n, n, TWO_PI, TWO_PI, RE)
    o.init_grids()
    gen = oc.OracleForcing(o.k2, n // 2 + 1, n, 4.0, 6.0, energy_target=1.0, seed=42)
    inv_k2 = np.where(k2 > 0, 1.0 / np.maximum(k2, 1e-300), 0.0)
    W = download(ns, s)
    worst_E = worst_Z = 0.0
    total_kick = total_eps = 0.0
    for _ in range(steps):
        E0, Z0, dE0, dZ0 = budget_terms(W, k2, w)
        assert s.advance_n(1, False) == 0
        W1 = download(ns, s)
        f = o.project_forcing(gen.generate(dt), RATE)
        eps = 0.5 * np.sum(w * np.abs(f) ** 2 * inv_k2)
        assert eps == pytest.approx(RATE, rel=1e-12)
        Wd = W1 - np.sqrt(dt) * f
        E1, Z1, dE1, dZ1 = budget_terms(Wd, k2, w)
        rE = abs((E1 - E0) - 0.5 * dt * (dE0 + dE1)) / abs(dt * dE0)
        rZ = abs((Z1 - Z0) - 0.5 * dt * (dZ0 + dZ1)) / abs(dt * dZ0)
        worst_E, worst_Z = max(worst_E, rE), max(worst_Z, rZ)
        total_kick += budget_terms(W1, k2, w)[0] - E1
        total_eps += eps * dt
        W = W1
    print(f"forced 256^2 FP64, {steps} steps of dt {dt}: deterministic budget residual "
          f"per step / dissipation: energy {worst_E:.2e}, enstrophy {worst_Z:.2e}; "
          f"kick energy {total_kick:.3e} of which eps dt {total_eps:.3e}", flush=True)
    assert worst_E <= 1e-6 and worst_Z <= 1e-6
    gen.close()
    o.close()
    s.close()
