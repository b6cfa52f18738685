"""The adaptive-dt error controller on the GPU against the oracle's
restatement (oracle.OracleErrorControl), through the C-ABI.

The controller is build-defined (the reference declares AdaptiveTimeParams,
include/time_integration.h:16-22, and uses it nowhere).  Each step judges
the previous candidate from the embedded estimate |y4 - y2| (FP64 sums of
the stage kernels' workgroup partials, in a fixed order), rejects it when
the error exceeds atol + rtol |y4| (t rolls back, the next step restores the
state kept before the candidate), and sets the next step limit from
safety err^(-1/3).  With rtol = 1e-6 the first two candidates of these runs
are rejected and the rest accepted, the error then settling at safety^3.

FP64: the same decisions, t and dt to 1e-12, omega_hat to 1e-11 relative L2.
FP32 stages: the estimate |y4 - y2| ~ 7e-7 |y| at this tolerance is the
difference of two FP32 solutions, so every judged error carries ~1e-4
relative round-off, which the controller feeds back into dt (dt_err =
dt safety err^(-1/3)).  The decisions are the same (no error lies near 1),
the error settles at the same value, and the dt sequence drifts from the
FP64 one as a random walk of those perturbations: t within 1e-2 after the 58
judged steps, energy and enstrophy within 1e-3 of the oracle's (whose t
differs by that drift).
"""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu] {request.node.name}", flush=True)


CASES = {
    # integrating factor, forced (the benchmark's physics at 64^2)
    "if_forced": dict(integ=True, forced=True, Re=1e4),
    # classic RK4 with drag, decaying
    "classic": dict(integ=False, forced=False, Re=200.0),
}


def hv(n):
    return dict(enabled=1, coefficient=1e-16 * (np.sqrt(2) * n / 2) ** 8, order=8,
                cutoff_ratio=0.0)


def oracle_run(oc, case, n, attempts, rtol):
    c = CASES[case]
    o = oc.Oracle(n, n, TWO_PI, TWO_PI, c["Re"])
    o.init_grids()
    if c["integ"]:
        o.hv_set(**hv(n))
    o.d.drag = 0.1
    o.d.cfl_factor = 0.5
    o.init_random_phase(42, 10.0, 1.0)
    gen = oc.OracleForcing(o.k2, n // 2 + 1, n, 4.0, 6.0, energy_target=1.0,
                           seed=42) if c["forced"] else None
    ctl = oc.OracleErrorControl(o, rtol, 0.0, 0.9, 1e-8, 0.1, integrating_factor=c["integ"],
                                forcing=gen, rate=1e-3)
    for _ in range(attempts):
        ctl.attempt()
    ctl.finish()
    return o, ctl


def gpu_run(ns, case, n, attempts, rtol, fp32=False, graph=True, calls=1, control=True):
    c = CASES[case]
    s = ns.Spectral(n, n, TWO_PI, TWO_PI, c["Re"])
    s.init(taylor_green=False)
    if c["integ"]:
        s.hv_set(**hv(n))
    s.set_precision(fp32)
    s.set_scheme(c["integ"])
    s.set_drag(0.1)
    if c["forced"]:
        s.set_forcing(4.0, 6.0, 1e-3, seed=42)
    s.d.cfl_factor = 0.5
    if control:
        s.set_error_control(rtol, 0.0, 0.9, 1e-8, 0.1)
    s.random_phase(42, k_peak=10.0, energy=1.0)
    s.sync_device()
    per = attempts // calls
    for _ in range(calls):
        if graph:
            assert s.advance_n(per, True) == 0
        else:
            assert s.L.nsd_advance(s.ctx, per, 1) == 0
    assert s.sync_host() == 0
    return s


@pytest.mark.parametrize("case", ["if_forced", "classic"])
def test_error_controller_fp64_matches_oracle(ns, oc, case):
    n, attempts, rtol = 64, 60, 1e-6
    o, ctl = oracle_run(oc, case, n, attempts, rtol)
    s = gpu_run(ns, case, n, attempts, rtol)
    acc, rej, err, dt_next = s.error_control_stats()
    w_err = rel_l2(s.spec("omega_hat"), o.spec("omega_hat"))
    print(f"{case}: GPU {acc} accepted / {rej} rejected, oracle {ctl.accepted} / {ctl.rejected}; "
          f"t {s.d.t:.12f} vs {o.d.t:.12f}; last err {err:.4f} vs {ctl.err:.4f}; "
          f"omega_hat rel L2 {w_err:.2e}", flush=True)
    assert ctl.rejected >= 2 and ctl.accepted >= 50
    assert (acc, rej) == (ctl.accepted, ctl.rejected)
    assert s.d.t == pytest.approx(o.d.t, rel=1e-12)
    assert s.d.dt == pytest.approx(o.d.dt, rel=1e-12)
    assert err == pytest.approx(ctl.err, rel=1e-9)
    assert dt_next == pytest.approx(ctl.dt_err, rel=1e-9)
    assert w_err <= 1e-11
    s.close()


def test_error_controller_fp64_4096_matches_oracle(ns, oc):
    """The 4096^2 stage kernels with the controller (their wave-split
    transforms, the embedded estimate, the judge and the restoring step)
    against the oracle.  At rtol 1e-8 the first candidate's error is about
    1.6: it is rejected, the step restored and retried with a smaller dt, and
    the next candidates are accepted.  The estimate |y4 - y2| is then ~1e-8
    of |y|, so the judged errors agree to the FP64 round-off of that
    difference (1e-6 relative), far from the threshold at 1."""
    n, attempts, rtol = 4096, 3, 1e-8
    o, ctl = oracle_run(oc, "if_forced", n, attempts, rtol)
    s = gpu_run(ns, "if_forced", n, attempts, rtol)
    acc, rej, err, dt_next = s.error_control_stats()
    w_err = rel_l2(s.spec("omega_hat"), o.spec("omega_hat"))
    print(f"4096^2: GPU {acc} accepted / {rej} rejected, oracle {ctl.accepted} / {ctl.rejected}; "
          f"t {s.d.t:.12e} vs {o.d.t:.12e}; last err {err:.6f} vs {ctl.err:.6f}; "
          f"omega_hat rel L2 {w_err:.2e}", flush=True)
    assert ctl.rejected >= 1 and ctl.accepted >= 1
    assert (acc, rej) == (ctl.accepted, ctl.rejected)
    assert s.d.t == pytest.approx(o.d.t, rel=1e-12)
    assert s.d.dt == pytest.approx(o.d.dt, rel=1e-12)
    assert err == pytest.approx(ctl.err, rel=1e-6)
    assert dt_next == pytest.approx(ctl.dt_err, rel=1e-6)
    assert w_err <= 1e-11
    s.close()


def test_error_controller_fp32_4096_loose_tolerance_equals_plain_steps(ns):
    """The FP32 4096^2 stage kernels with the controller (pair layout,
    wave-split transforms, the embedded estimate and its reductions).  At
    rtol 1e-2 the FP32 estimate (~1e-6 of |y|) accepts every candidate and
    the error's step limit stays above the CFL one, so the controlled run
    takes the plain run's steps: same dt, same t, and the same field to FP32
    round-off."""
    n, steps = 4096, 3
    ctl = gpu_run(ns, "if_forced", n, steps, 1e-2, fp32=True)
    acc, rej, err, dt_next = ctl.error_control_stats()
    plain = gpu_run(ns, "if_forced", n, steps, 1e-2, fp32=True, control=False)
    w_err = rel_l2(ctl.spec("omega_hat"), plain.spec("omega_hat"))
    print(f"FP32 4096^2: {acc} accepted / {rej} rejected, last err {err:.3e}, dt_next "
          f"{dt_next:.3e}; t {ctl.d.t:.12e} vs {plain.d.t:.12e}; omega_hat rel L2 {w_err:.2e}",
          flush=True)
    assert rej == 0 and acc in (steps - 1, steps)  # the last may still await its judgment
    assert err < 0.1
    assert ctl.d.t == pytest.approx(plain.d.t, rel=1e-12)
    assert ctl.d.dt == pytest.approx(plain.d.dt, rel=1e-12)
    assert w_err <= 1e-6
    ctl.close()
    plain.close()


def test_error_controller_fp32_stages(ns, oc):
    n, attempts, rtol = 64, 60, 1e-6
    o, ctl = oracle_run(oc, "if_forced", n, attempts, rtol)
    s = gpu_run(ns, "if_forced", n, attempts, rtol, fp32=True)
    acc, rej, err, _ = s.error_control_stats()
    w_err = rel_l2(s.spec("omega_hat"), o.spec("omega_hat"))
    print(f"FP32: {acc} / {rej} (oracle {ctl.accepted} / {ctl.rejected}); t {s.d.t:.9f} vs "
          f"{o.d.t:.9f}; err {err:.4f} vs {ctl.err:.4f}; omega_hat rel L2 {w_err:.2e}", flush=True)
    assert (acc, rej) == (ctl.accepted, ctl.rejected)
    assert err == pytest.approx(ctl.err, rel=1e-2)
    assert s.d.t == pytest.approx(o.d.t, rel=1e-2)
    e, z = s.device_diagnostics()
    assert e == pytest.approx(o.energy(), rel=1e-3)
    assert z == pytest.approx(o.enstrophy(), rel=1e-3)
    s.close()


def test_error_controller_graph_replay_equals_eager(ns):
    """The judge, the restoring steps and the finalize pass are the same in a
    replayed step graph as in eager launches: identical bits."""
    runs = [gpu_run(ns, "if_forced", 64, 40, 1e-6, graph=g) for g in (True, False)]
    a, b = runs
    assert a.error_control_stats() == b.error_control_stats()
    assert a.d.t == b.d.t
    assert np.array_equal(a.spec("omega_hat"), b.spec("omega_hat"))
    for s in runs:
        s.close()


def test_error_controller_parameters_are_validated(ns):
    s = ns.Spectral(64, 64)
    s.init(taylor_green=False)
    P = ns.AdaptiveTimeParams
    L = s.L
    bad = [(P(0.5, 1e-8, 0.1, 0.9, 1), 0.0, 0.0),      # no tolerance
           (P(0.5, 1e-8, 0.1, 1.5, 1), 1e-6, 0.0),     # safety > 1
           (P(0.5, 0.1, 1e-3, 0.9, 1), 1e-6, 0.0)]     # dt_min > dt_max
    for p, rtol, atol in bad:
        assert L.ns_spectral_set_error_control(s.p, C.byref(p), rtol, atol) == \
            ns.SOLVER_INVALID_PARAM
    # selective frequency damping cannot be restored after a rejection
    s.hv_set(enabled=1, coefficient=0.05, order=4, cutoff_ratio=0.6, sfd_enabled=1,
             sfd_strength=0.3, sfd_timescale=0.4)
    rc = L.ns_spectral_set_error_control(s.p, C.byref(P(0.5, 1e-8, 0.1, 0.9, 1)), 1e-6, 0.0)
    assert rc == ns.SOLVER_INVALID_PARAM and "selective frequency damping" in s.error()
    # adapt_enabled = 0 switches it off
    assert L.ns_spectral_set_error_control(s.p, C.byref(P(0.5, 1e-8, 0.1, 0.9, 0)), 1e-6,
                                           0.0) == 0
    s.close()
