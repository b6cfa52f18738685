"""Pins the CPU oracle's restatement of the reference's spectral solver with
closed forms, invariants and the reference's own assertions (test_solvers.c,
test_conservation.c), so that a dropped term, wrong sign or index fails."""
import numpy as np
import pytest

from conftest import quadratic_invariants, rel_l2, rk4_one_step_errors

TWO_PI = 2 * np.pi


def kept_count(n):
    nkx = n // 2 + 1
    return ((2 * nkx) // 3 + 1) * (2 * ((2 * n) // 6) + 1)


def test_mask_counts_closed_form(oc):
    # spectral_initialization.c:46-49; counts quoted for 64^2, 1024^2, 4096^2
    assert kept_count(64) == 989 and kept_count(1024) == 234269 and kept_count(4096) == 3733277
    for n in (16, 64, 128):
        o = oc.Oracle(n, n)
        o.init_grids()
        assert o.mask.sum() == kept_count(n)
        # quirk Q1: x keeps i <= (2 nkx)/3, one column more than N/3 at 64
        assert o.mask[0, (2 * (n // 2 + 1)) // 3] == 1
        assert o.mask[0, (2 * (n // 2 + 1)) // 3 + 1] == 0


def test_grid_initialization(oc):
    # test_solvers.c:450-471
    o = oc.Oracle(32, 32)
    o.init_grids()
    assert abs(o.kx[0]) < 1e-10 and abs(o.kx[1] - 1.0) < 1e-10
    assert abs(o.k2[0, 0]) < 1e-10 and o.mask[0, 0] == 1
    assert o.ky[17] == pytest.approx(-15.0)


def test_dealiasing_zeroes_every_masked_mode(oc):
    # test_solvers.c:513-545
    o = oc.Oracle(32, 32)
    o.init_grids()
    f = o.dealias(np.ones((32, 17), complex))
    assert np.all(np.abs(f[o.mask == 0]) < 1e-15)
    assert np.all(f[o.mask == 1] == 1)


def test_velocity_single_mode_closed_form(oc):
    # spectral_core.c:8-34 with psi_hat = -omega_hat / k^2 (quirk Q2)
    n = 32
    o = oc.Oracle(n, n)
    o.init_grids()
    w = o.spec("omega_hat")
    w[:] = 0
    j0, i0 = n - 3, 2  # ky = -3, kx = 2
    w[j0, i0] = 0.7 - 0.2j
    o.velocity_from_vorticity()
    k2 = 13.0
    psi = -(0.7 - 0.2j) / k2
    assert o.spec("psi_hat")[j0, i0] == pytest.approx(psi)
    assert o.spec("u_hat")[j0, i0] == pytest.approx(1j * (-3.0) * psi)
    assert o.spec("v_hat")[j0, i0] == pytest.approx(-1j * 2.0 * psi)
    assert np.count_nonzero(o.spec("psi_hat")) == 1


def test_uniform_vorticity_gives_zero_velocity(oc):
    # test_solvers.c:547-577
    o = oc.Oracle(32, 32)
    o.init_grids()
    w = o.spec("omega_hat")
    w[:] = 0
    w[0, 0] = 1.0
    o.velocity_from_vorticity()
    assert np.abs(o.s2p(o.spec("u_hat"))).max() < 1e-10
    assert np.abs(o.s2p(o.spec("v_hat"))).max() < 1e-10


def analytic_two_mode_jacobian(n, a, k1, b, k2):
    """omega = a cos(k1.x) + b sin(k2.x); psi_ref = -sum omega_k/|k|^2 (quirk Q2);
    returns omega and J = psi_x omega_y - psi_y omega_x evaluated pointwise."""
    x = TWO_PI * np.arange(n) / n
    X, Y = np.meshgrid(x, x)  # [y][x]
    p1 = k1[0] * X + k1[1] * Y
    p2 = k2[0] * X + k2[1] * Y
    n1 = k1[0] ** 2 + k1[1] ** 2
    n2 = k2[0] ** 2 + k2[1] ** 2
    om = a * np.cos(p1) + b * np.sin(p2)
    om_x = -a * k1[0] * np.sin(p1) + b * k2[0] * np.cos(p2)
    om_y = -a * k1[1] * np.sin(p1) + b * k2[1] * np.cos(p2)
    # psi = -(a/n1) cos(p1) - (b/n2) sin(p2)
    ps_x = (a / n1) * k1[0] * np.sin(p1) - (b / n2) * k2[0] * np.cos(p2)
    ps_y = (a / n1) * k1[1] * np.sin(p1) - (b / n2) * k2[1] * np.cos(p2)
    return om, ps_x * om_y - ps_y * om_x


@pytest.mark.parametrize("k1,k2", [((1, 2), (3, -1)), ((2, 0), (0, 3)), ((4, 1), (1, -5))])
def test_nonlinear_two_mode_closed_form(oc, k1, k2):
    # spectral_core.c:46-112: nonlinear_hat = -dealias(FFT(J)/N^2)
    n = 32
    o = oc.Oracle(n, n)
    o.init_grids()
    om, J = analytic_two_mode_jacobian(n, 0.8, k1, -1.3, k2)
    o.spec("omega_hat")[:] = o.p2s(om)
    o.velocity_from_vorticity()
    o.nonlinear()
    expect = -np.fft.rfft2(J) / n ** 2 * o.mask
    got = o.spec("nonlinear_hat")
    assert np.abs(got - expect).max() < 1e-13
    assert np.abs(expect).max() > 1e-2  # the Jacobian is not trivially zero


def test_nonlinear_single_mode_vanishes(oc):
    n = 32
    o = oc.Oracle(n, n)
    o.init_grids()
    o.init_taylor_green()
    o.compute_derivatives()
    o.nonlinear()
    assert np.abs(o.spec("nonlinear_hat")).max() < 1e-15


def test_taylor_green_rk4_closed_form(oc):
    # config 1: 64^2, nu = 0.01, dt = 1e-3, 100 steps; linear decay
    # omega_n = R(z)^n omega_0, R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24, z = -k^2 dt / Re
    n = 64
    o = oc.Oracle(n, n, Re=100.0)
    o.init_grids()
    o.init_taylor_green()
    o.compute_derivatives()
    w0 = o.spec("omega_hat").copy()
    E0, Z0 = o.energy(), o.enstrophy()
    assert E0 == pytest.approx(0.25, rel=1e-14) and Z0 == pytest.approx(0.5, rel=1e-13)
    o.d.dt = 1e-3
    for _ in range(100):
        o.advance_rk4()
    z = -2.0 * 1e-3 / 100.0
    R = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    assert rel_l2(o.spec("omega_hat"), R ** 100 * w0) < 1e-14
    t = o.d.t
    assert t == pytest.approx(0.1, rel=1e-12)
    x = TWO_PI * np.arange(n) / n
    exact = 2 * np.cos(x)[None, :] * np.cos(x)[:, None] * np.exp(-2 * 0.01 * t)
    assert np.abs(o.phys("omega") - exact).max() < 1e-13
    assert o.energy() == pytest.approx(0.25 * np.exp(-4 * 0.01 * t), rel=1e-12)
    assert o.enstrophy() == pytest.approx(0.5 * np.exp(-4 * 0.01 * t), rel=1e-12)


def test_minimal_step_equals_reference_sequence(oc):
    # the 40-transform reference sequence and the 20-transform step do the same arithmetic
    a, b = oc.Oracle(32, 32, Re=50.0), oc.Oracle(32, 32, Re=50.0)
    for o in (a, b):
        o.init_grids()
        o.init_random_phase(7, k_peak=3.0)
        o.d.dt = 2e-3
    for _ in range(5):
        a.advance_rk4()
        b.advance_rk4(minimal=True)
    assert np.array_equal(a.spec("omega_hat"), b.spec("omega_hat"))
    # 3 transforms of the initial compute_derivatives, then per step 40 (36 c2r + 4 r2c)
    # for the reference sequence and 4 x 5 + 3 for the minimal one
    assert a.d.fft_count == 3 + 5 * 40 and b.d.fft_count == 3 + 5 * 23


def test_cfl_taylor_green(oc):
    # spectral_time_integration.c:9-31 with umax = 1 for TG after compute_derivatives
    n = 64
    o = oc.Oracle(n, n, Re=100.0)
    o.init_grids()
    o.init_taylor_green()
    o.compute_derivatives()
    dx = TWO_PI / n
    assert o.cfl() == pytest.approx(min(0.3 * dx / (1 + 1e-12), 0.5 * 100 * dx * dx), rel=1e-12)
    o.d.Re = 1e-6  # viscous limit below MIN_DT -> clamp
    assert o.cfl() == pytest.approx(1e-8)


def test_hyperviscosity_closed_form_and_reference_assertions(oc):
    n = 32
    o = oc.Oracle(n, n)
    o.init_grids()
    # test_solvers.c:616-664
    o.hv_set(enabled=1, coefficient=0.05, order=4, cutoff_ratio=0.6, sfd_enabled=1,
             sfd_strength=0.3, sfd_timescale=0.4)
    kmag = np.sqrt(o.k2)
    kmax = np.sqrt(2) * n / 2
    kc = 0.6 * kmax
    expect = np.where(kmag > kc, 0.05 * ((kmag - kc) / (kmax - kc)) ** 4, 0.0)
    assert np.abs(o.hv_weights - expect).max() < 1e-15
    assert o.d.hv.kmax == pytest.approx(kmax)
    assert o.hv_weights[0, 0] == 0.0 and o.hv_weights[-1, -1] > 0.0
    o.spec("omega_hat")[:] = 1.0
    o.spec("sfd_hat")[:] = 0.25
    rhs = o.hv_apply(np.zeros((n, 17), complex), np.ones((n, 17), complex))
    sfdw = np.where(kmag > kc, 0.3 * np.clip((kmag - kc) / (kmax - kc), 0, 1) ** 4, 0.0)
    assert np.abs(rhs - (-(expect * 1.0 + sfdw * (1.0 - 0.25)))).max() < 1e-14
    assert abs(rhs[0, 0]) < 1e-8 and abs(rhs[-1, -1]) > 1e-3
    o.hv_post_step(np.ones((n, 17), complex), 0.1)
    assert o.spec("sfd_hat")[-1, -1] == pytest.approx(0.75 * 0.25 + 0.25)


def test_if_rk4_linear_decay_is_exponential(oc):
    # integrating factor: with J = 0 the step is exact, omega_n = e^{L t} omega_0
    n = 32
    o = oc.Oracle(n, n, Re=100.0)
    o.init_grids()
    o.hv_set(enabled=1, coefficient=3.0, order=8, cutoff_ratio=0.0)
    o.d.drag = 0.1
    o.init_taylor_green()
    o.compute_derivatives()
    w0 = o.spec("omega_hat").copy()
    o.d.dt = 0.01
    for _ in range(50):
        o.advance_if()
    L = -(o.k2 / 100.0 + o.hv_weights + 0.1)
    assert rel_l2(o.spec("omega_hat"), np.exp(L * 0.5) * w0) < 1e-14


def test_if_rk4_equals_rk4_on_transformed_variable(oc):
    # Lawson stage algebra vs classical RK4 written here on v = e^{-L t} omega
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
    expect = np.exp(L * h) * (w0 + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4))
    o.d.dt = h
    o.advance_if()
    assert rel_l2(o.spec("omega_hat"), expect) < 1e-13


def test_classic_ext_reduces_to_reference_step(oc):
    a, b = oc.Oracle(32, 32, Re=80.0), oc.Oracle(32, 32, Re=80.0)
    for o in (a, b):
        o.init_grids()
        o.hv_set(enabled=1, coefficient=0.02, order=4, cutoff_ratio=0.5)
        o.init_random_phase(3, k_peak=4.0)
        o.d.dt = 1e-3
    for _ in range(3):
        a.advance_rk4()
        b.advance_classic_ext()
    assert rel_l2(b.spec("omega_hat"), a.spec("omega_hat")) < 1e-14


def test_drag_linear_decay_classic(oc):
    n = 32
    o = oc.Oracle(n, n, Re=100.0)
    o.init_grids()
    o.d.drag = 0.5
    o.init_taylor_green()
    o.compute_derivatives()
    w0 = o.spec("omega_hat").copy()
    o.d.dt = 0.01
    for _ in range(10):
        o.advance_classic_ext()
    z = -(2.0 / 100.0 + 0.5) * 0.01
    R = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    assert rel_l2(o.spec("omega_hat"), R ** 10 * w0) < 1e-14


def test_spectrum_matches_numpy_binning(oc):
    n = 64
    o = oc.Oracle(n, n, Re=100.0)
    o.init_grids()
    o.init_random_phase(5, k_peak=6.0)
    ek, ck = o.analyze_spectrum()
    bins = (np.sqrt(o.k2) + 0.5).astype(int)
    e = 0.5 * (np.abs(o.spec("u_hat")) ** 2 + np.abs(o.spec("v_hat")) ** 2)
    max_k = min(n // 2 + 1, n // 2)
    for k in range(1, max_k):
        sel = bins == k
        assert ck[k] == sel.sum()
        assert ek[k] == pytest.approx(e[sel].sum() / sel.sum(), rel=1e-12)


def test_resolution_check_flags_high_k_energy(oc):
    # spectral_analysis.c:36-66; failure injected on the CPU
    o = oc.Oracle(32, 32)
    o.init_grids()
    o.init_taylor_green()
    ok, frac = o.check_resolution()
    assert ok == 1 and frac < 1e-20
    o.spec("omega_hat")[16, 16] = 1.0
    ok, frac = o.check_resolution()
    assert ok == 0 and frac > 1e-6


def test_random_phase_initial_condition(oc):
    n = 64
    o = oc.Oracle(n, n, Re=1e4)
    o.init_grids()
    o.init_random_phase(42, k_peak=10.0, energy=1.0)
    w = o.spec("omega_hat")
    assert o.energy() == pytest.approx(1.0, rel=1e-12)  # physical-space 1/2<u^2+v^2>
    col = w[:, 0]
    assert np.allclose(col[1:][::-1], np.conj(col[1:]), atol=1e-15)
    assert np.all(w[o.mask == 0] == 0)
    # shell amplitude |w_k|^2 proportional to k E(k)/pi with E ~ k^4 exp(-(k/10)^2)
    km = np.sqrt(o.k2)
    sel = (o.mask == 1) & (km > 0)
    ratio = np.abs(w[sel]) ** 2 / (km[sel] ** 5 * np.exp(-(km[sel] / 10) ** 2))
    assert ratio.std() / ratio.mean() < 1e-12
    # omega from the counter stream: phase of mode idx uses uniform idx+1
    j, i = 3, 5
    ph = 2 * np.pi * oc.lib().oc_counter_uniform(42, j * (n // 2 + 1) + i + 1)
    assert np.angle(w[j, i]) == pytest.approx(np.angle(np.exp(1j * ph)), abs=1e-12)


def test_forcing_projection_and_injection_rate(oc):
    n = 32
    o = oc.Oracle(n, n)
    o.init_grids()
    f = oc.OracleForcing(o.k2, 17, n, 2.0, 5.0, energy_target=1.0, seed=9)
    raw = f.generate(0.01)
    proj = o.project_forcing(raw, 1e-3)
    col = proj[:, 0]
    assert np.allclose(col[1:][::-1], np.conj(col[1:]), atol=1e-16)
    w = np.where(np.arange(17)[None, :] == 0, 1.0, 2.0)
    k2 = np.where(o.k2 > 0, o.k2, 1.0)
    rate = 0.5 * np.sum(w * np.abs(proj) ** 2 / k2)
    assert rate == pytest.approx(1e-3, rel=1e-13)
    # interior modes are the raw draws up to the common scale factor
    sel = (np.abs(raw) > 0) & (np.arange(17)[None, :] > 0)
    s = proj[sel] / raw[sel]
    assert np.allclose(s, s[0], rtol=1e-13)


def test_config2_fixture_reproduces_and_decays(oc):
    """The 1024^2 decaying-turbulence fixture (tests/golden/make_config2.py):
    unit initial energy, monotone decay of energy and enstrophy without
    forcing, shell counts of the spectrum, and the first 100 oracle steps
    recomputed here."""
    import os
    from conftest import GOLDEN
    G = np.load(os.path.join(GOLDEN, "config2_golden.npz"))
    N, Re, seed, steps, every, klow = G["meta"]
    N, seed, every = int(N), int(seed), int(every)
    h = G["history"]
    assert h[0, 3] == pytest.approx(1.0, rel=1e-12)
    assert np.all(np.diff(h[:, 3]) < 0) and np.all(np.diff(h[:, 4]) < 0)
    kx = np.arange(N // 2 + 1)
    ky = np.where(np.arange(N) <= N // 2, np.arange(N), np.arange(N) - N)
    bins = (np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2) + 0.5).astype(int)
    counts = np.bincount(bins.ravel(), minlength=len(G["counts"]))[: len(G["counts"])]
    assert np.array_equal(counts, G["counts"])
    o = oc.Oracle(N, N, 2 * np.pi, 2 * np.pi, float(Re))
    o.init_grids()
    kmax = np.sqrt(2.0) * N / 2.0
    o.hv_set(enabled=1, coefficient=1e-16 * kmax ** 8, order=8, cutoff_ratio=0.0)
    o.d.cfl_factor = 0.5
    o.init_random_phase(seed, 10.0, 1.0)
    for _ in range(every):
        o.d.dt = o.cfl()
        o.advance_if()
    assert o.d.t == pytest.approx(h[1, 1], rel=1e-13)
    assert o.energy() == pytest.approx(h[1, 3], rel=1e-12)
    assert o.enstrophy() == pytest.approx(h[1, 4], rel=1e-12)


def test_analyze_convergence_closed_forms(oc):
    """spectral_analysis.c:69-100: max|w_hat| over the stored modes and the RMS
    of |w_hat| over the modes with k^2 > k2[last]/4, where k2[last] is the last
    stored mode (quirk Q7: (N/2)^2 + 1, not the corner 2 (N/2)^2)."""
    n = 16
    o = oc.Oracle(n, n)
    o.init_grids()
    w = o.spec("omega_hat")
    # (a) constant modulus c on every stored mode, arbitrary phases: max = err = c
    c = 0.37
    phase = np.exp(1j * np.linspace(0.0, 5.0, w.size)).reshape(w.shape)
    w[:] = c * phase
    out = o.analyze_convergence()
    assert out[2] == pytest.approx(c, rel=1e-15)
    assert out[3] == pytest.approx(c, rel=1e-15)
    # (b) two modes straddling the threshold k2[last]/4 = (64 + 1)/4 = 16.25:
    # (kx, ky) = (4, 0), k^2 = 16, amplitude 3 (max, not counted) and
    # (4, 1), k^2 = 17, amplitude 2 (counted).  Stored modes: 9 x 16 = 144;
    # lattice points with kx >= 0 and k^2 <= 16: 9 + 7 + 7 + 5 + 1 = 29.
    w[:] = 0
    w[0, 4] = 3.0
    w[1, 4] = 2.0j
    out = o.analyze_convergence()
    assert out[2] == pytest.approx(3.0, rel=1e-15)
    assert out[3] == pytest.approx(np.sqrt(4.0 / (144 - 29)), rel=1e-14)
    # a mode below the threshold alone gives zero error
    w[1, 4] = 0
    assert o.analyze_convergence()[3] == 0.0
    # E, Z come from the physical arrays; t and dt are passed through
    o.init_taylor_green()
    o.compute_derivatives()
    o.d.t, o.d.dt = 0.25, 1e-3
    out = o.analyze_convergence()
    assert out[0] == pytest.approx(0.25, rel=1e-13) and out[1] == pytest.approx(0.5, rel=1e-13)
    assert out[4] == 0.25 and out[5] == 1e-3


def test_if_rk4_energy_enstrophy_budget_closes(oc):
    """The oracle's integrating-factor step conserves what the dealiased
    Jacobian conserves: E(T) - E(0) equals minus the Simpson integral of
    2 nu Z + 2 mu E + sum w nu4 k^8 |W|^2 / k^2 (and likewise for Z), to the
    quadrature's O(dt^4) (about 1e-12 here).  Dropping the drag, the
    hyperviscous or the viscous term misses by more than 1e-4.  At 256^2 the
    modes of the extra x column the mask keeps (quirk Q1, k ~ 86) hold
    e^-74 of the energy; at 128^2 (k ~ 43) their aliased transfer already
    shows at 1e-10 of the budget."""
    n, steps, dt, Re, mu, nu4 = 256, 20, 2e-4, 1e4, 0.1, 1e-16
    o = oc.Oracle(n, n, TWO_PI, TWO_PI, Re)
    o.init_grids()
    o.hv_set(enabled=1, coefficient=nu4 * (np.sqrt(2) * n / 2) ** 8, order=8, cutoff_ratio=0.0)
    o.d.drag = mu
    o.init_random_phase(7, 10.0, 1.0)
    kx = np.arange(n // 2 + 1, dtype=float)
    ky = np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n).astype(float)
    k2 = kx[None, :] ** 2 + ky[:, None] ** 2
    wt = np.where(kx == 0, 1.0, 2.0)[None, :]
    inv_k2 = np.where(k2 > 0, 1.0 / np.maximum(k2, 1e-300), 0.0)
    rate = k2 / Re + mu + nu4 * k2 ** 4

    def terms():
        a2 = np.abs(o.spec("omega_hat")) ** 2
        return (0.5 * np.sum(wt * a2 * inv_k2), 0.5 * np.sum(wt * a2),
                -np.sum(wt * rate * a2 * inv_k2), -np.sum(wt * rate * a2))

    hist = [terms()]
    for _ in range(steps):
        o.d.dt = dt
        o.advance_if()
        hist.append(terms())
    E, Z, dE, dZ = (np.array(c) for c in zip(*hist))
    simpson = np.ones(steps + 1)
    simpson[1:-1:2], simpson[2:-1:2] = 4.0, 2.0
    for X, dX in ((E, dE), (Z, dZ)):
        integral = dt / 3.0 * np.sum(simpson * dX)
        assert abs((X[-1] - X[0]) - integral) <= 1e-9 * abs(integral)


def test_classic_rk4_generic_field_is_fourth_order_and_conservative(oc):
    """The classic RK4 step on a generic field (J != 0) as one composite
    (spectral_time_integration.c:34-106), pinned by what the mathematics
    fixes rather than by a second copy of the formula:
    - its one-step error against the same system stepped 64 times finer falls
      as dt^5, so each halving of dt divides it by 32. A step whose stages take
      the wrong field (stage 3 built from k1 instead of k2) is second order and
      gives 8, planted below in a composition of the oracle's own nonlinear
      term; weights other than (1, 2, 2, 1)/6 drop the order the same way;
    - inviscid, the dealiased Jacobian conserves energy and enstrophy, so one
      step changes them by the step's truncation error only. A dropped or
      transposed product in J changes them at first order in dt."""
    n = 32
    o = oc.Oracle(n, n, Re=1e300)  # 1/Re = 1e-300: the inviscid system
    o.init_grids()
    o.init_random_phase(5, k_peak=3.0, energy=1.0)
    w0 = o.spec("omega_hat").copy()
    k2 = o.k2.reshape(w0.shape)

    def oracle_steps(w, dt, m):
        o.spec("omega_hat")[:] = w
        o.compute_derivatives()
        for _ in range(m):
            o.d.dt = dt / m
            o.advance_rk4()
        return o.spec("omega_hat").copy()

    dts = (0.02, 0.01, 0.005)
    errs = rk4_one_step_errors(oracle_steps, w0, dts)
    ratios = errs[:-1] / errs[1:]
    assert np.all((ratios > 28.0) & (ratios < 36.0)), ratios
    assert errs[-1] < 1e-8
    E0, Z0 = quadratic_invariants(w0, k2)
    for dt, bound in ((0.01, 1e-8), (0.005, 3e-10)):
        E1, Z1 = quadratic_invariants(oracle_steps(w0, dt, 1), k2)
        assert abs(E1 / E0 - 1) < bound and abs(Z1 / Z0 - 1) < 10 * bound

    # the same measure on a planted mistake: stage 3 from k1 instead of k2
    def f(w):
        o.spec("omega_hat")[:] = w
        o.compute_derivatives()
        o.nonlinear()
        return o.spec("nonlinear_hat").copy()

    def planted_steps(w, dt, m):
        if m > 1:
            return oracle_steps(w, dt, m)
        k1 = f(w)
        k2_ = f(w + dt / 2 * k1)
        k3 = f(w + dt / 2 * k1)
        k4 = f(w + dt * k3)
        return w + dt / 6 * (k1 + 2 * k2_ + 2 * k3 + k4)

    bad = rk4_one_step_errors(planted_steps, w0, dts)
    assert np.all(bad[:-1] / bad[1:] < 10.0)
