"""Checkpoint and restart on the GPU: a run interrupted by
ns_checkpoint_write_spectral and resumed in a fresh solver through
ns_checkpoint_read_spectral continues exactly as the uninterrupted run (bit
for bit), for forced integrating-factor runs in FP32 and FP64 with white and
coloured forcing (generator counter and OU state in the file) and for the
reference's classic scheme with selective frequency damping (sfd_hat)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu] {request.node.name}", flush=True)


def configured(ns, n, fp32, colored, classic):
    s = ns.Spectral(n, n, TWO_PI, TWO_PI, 2e3)
    s.init(taylor_green=False)
    if classic:
        s.hv_set(enabled=1, coefficient=0.05, order=4, cutoff_ratio=0.6, sfd_strength=0.3,
                 sfd_timescale=0.4, sfd_enabled=1)
    else:
        kmax = np.sqrt(2.0) * n / 2.0
        s.hv_set(enabled=1, coefficient=1e-16 * kmax ** 8, order=8, cutoff_ratio=0.0)
        s.set_scheme(True)
        s.set_drag(0.1)
        s.set_forcing(4.0, 6.0, 1e-3, seed=17, colored=colored, correlation_time=0.2)
    s.set_precision(fp32)
    s.d.cfl_factor = 0.5
    return s


@pytest.mark.parametrize("fp32,colored,classic", [(True, False, False), (False, True, False),
                                                  (True, True, False), (False, False, True)])
def test_restart_is_bit_exact(ns, tmp_path, fp32, colored, classic):
    n, first, second = 64, 20, 30
    fn = str(tmp_path / "ck.h5").encode()
    a = configured(ns, n, fp32, colored, classic)
    a.random_phase(5, k_peak=8.0)
    a.sync_device()
    assert a.advance_n(first, True) == 0
    assert a.sync_host() == 0
    rc = a.L.ns_checkpoint_write_spectral(fn, a.p)
    assert rc == 0, a.L.ns_checkpoint_error()
    assert a.advance_n(second, True) == 0
    assert a.sync_host() == 0
    w_full, t_full = a.spec("omega_hat").copy(), a.d.t
    a.close()

    b = configured(ns, n, fp32, colored, classic)
    assert b.L.ns_spectral_advance_n(b.p, 0, 1) == 0  # parameters onto the device
    rc = b.L.ns_checkpoint_read_spectral(fn, b.p)
    assert rc == 0, b.L.ns_checkpoint_error()
    b.sync_device()
    assert b.advance_n(second, True) == 0
    assert b.sync_host() == 0
    w_res, t_res = b.spec("omega_hat").copy(), b.d.t
    b.close()
    print(f"resume after {first} + {second} steps: t {t_res:.9f} vs {t_full:.9f}, "
          f"max |diff| {np.abs(w_res - w_full).max():.1e}", flush=True)
    assert t_res == t_full
    assert np.array_equal(w_res, w_full)
