"""Pins the oracle's 2-D transforms (the FFTW boundary of spectral_transforms.c:9-32,
spectral_initialization.c:66-72) against independent implementations."""
import numpy as np
import pytest


def dft2_brute(f):
    ny, nx = f.shape
    y = np.arange(ny)[:, None]
    x = np.arange(nx)[None, :]
    out = np.zeros((ny, nx // 2 + 1), complex)
    for j in range(ny):
        for i in range(nx // 2 + 1):
            out[j, i] = np.sum(f * np.exp(-2j * np.pi * (i * x / nx + j * y / ny)))
    return out


@pytest.mark.parametrize("ny,nx", [(8, 8), (16, 8), (8, 32), (64, 64), (32, 128)])
def test_r2c_matches_numpy_and_brute_force(oc, ny, nx):
    o = oc.Oracle(nx, ny)
    f = np.random.default_rng(nx * 7 + ny).standard_normal((ny, nx))
    X = o.r2c(f)
    assert np.abs(X - np.fft.rfft2(f)).max() < 1e-12 * nx * ny
    if nx * ny <= 256:
        assert np.abs(X - dft2_brute(f)).max() < 1e-11


@pytest.mark.parametrize("n", [16, 64, 256])
def test_c2r_matches_numpy(oc, n):
    o = oc.Oracle(n, n)
    f = np.random.default_rng(n).standard_normal((n, n))
    X = np.fft.rfft2(f)
    assert np.abs(o.c2r(X) - f * n * n).max() < 1e-10 * n * n


def test_c2r_ignores_imaginary_dc_and_nyquist(oc):
    # FFTW's c2r convention: Im of the kx = 0 and kx = nx/2 mixed values is ignored
    n = 16
    o = oc.Oracle(n, n)
    X = np.zeros((n, n // 2 + 1), complex)
    X[0, 0] = 1 + 5j
    X[0, n // 2] = 2 - 3j
    out = o.c2r(X)
    x = np.arange(n)
    expect = 1.0 + 2.0 * (-1.0) ** x
    assert np.allclose(out, np.broadcast_to(expect, (n, n)), atol=1e-13)


def test_roundtrip_sin_cos(oc):
    # test_solvers.c:473-511: r2c -> c2r of sin(x) cos(y) at 32^2 within 1e-10
    n = 32
    o = oc.Oracle(n, n)
    o.init_grids()
    x = o._real("x", (n,))
    y = o._real("y", (n,))
    f = np.sin(x)[None, :] * np.cos(y)[:, None]
    back = o.s2p(o.p2s(f))
    assert np.abs(back - f).max() < 1e-10
