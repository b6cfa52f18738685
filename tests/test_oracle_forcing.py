"""Pins the oracle's forcing restatement (forcing.c:11-184) bit-for-bit against
the reference's own forcing.c (golden vectors written by tests/golden/make_golden.py
from oracle/_ref/libref_forcing.so, and the live build when it is present)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

G = np.load(os.path.join(GOLDEN, "forcing_golden.npz"))


def k2_grid(n):
    nkx = n // 2 + 1
    kx = np.arange(nkx, dtype=float)
    j = np.arange(n)
    ky = np.where(j <= n // 2, j, j - n).astype(float)
    return kx[None, :] ** 2 + ky[:, None] ** 2, nkx


@pytest.mark.parametrize("n", [16, 32, 64])
@pytest.mark.parametrize("colored", [0, 1])
def test_oracle_forcing_bit_exact_vs_reference_golden(oc, n, colored):
    k2, nkx = k2_grid(n)
    f = oc.OracleForcing(k2, nkx, n, 4.0, 6.0, energy_target=1.5, seed=42,
                         colored=bool(colored), correlation_time=0.2)
    assert f.rc == 0
    got = np.stack([f.generate(0.05) for _ in range(3)])
    ref = G[f"n{n}_c{colored}"]
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    assert f.state.rng_state == int(G[f"n{n}_c{colored}_rng_state"][0])
    # 37 active half-plane modes on the shell [4, 6]
    assert f.state.active_modes == 37


def test_oracle_forcing_normalisation_4x4(oc):
    # test_solvers.c:239-295: sum |f|^2 = energy_target for white and coloured noise
    k2 = np.array([[i * i + j * j for i in range(4)] for j in range(4)], dtype=float)
    for colored, dt, tau in ((0, 0.1, 0.5), (1, 0.05, 0.2)):
        f = oc.OracleForcing(k2, 4, 4, 0.5, 2.5, energy_target=1.5, seed=42,
                             colored=bool(colored), correlation_time=tau)
        got = np.stack([f.generate(dt) for _ in range(2)])
        assert np.array_equal(got.view(np.uint64), G[f"t4_c{colored}"].view(np.uint64))
        assert np.sum(np.abs(got[0]) ** 2) == pytest.approx(1.5, abs=1e-6)


def test_counter_form_equals_sequential_stream(oc):
    # uniform #k of the counter form == k-th output of the sequential generator
    import ctypes as C
    L = oc.lib()
    st = C.c_uint64(42)
    for k in range(1, 200):
        b = L.oc_splitmix64_next(C.byref(st))
        assert L.oc_counter_uniform(42, k) == (b >> 11) * (1.0 / 9007199254740992.0)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(GOLDEN), "..", "oracle",
                                                    "_ref", "libref_forcing.so")),
                    reason="reference build oracle/_ref/libref_forcing.so not present")
def test_oracle_forcing_vs_live_reference_build(oc):
    ref = oc.ref_forcing_lib()
    k2, nkx = k2_grid(48 if False else 32)
    for seed in (0, 1, 12345):
        a = oc.OracleForcing(k2, nkx, 32, 1.0, 9.5, energy_target=0.7, seed=seed, colored=True,
                             correlation_time=0.1)
        b = oc.OracleForcing(k2, nkx, 32, 1.0, 9.5, energy_target=0.7, seed=seed, colored=True,
                             correlation_time=0.1, library=ref)
        for _ in range(4):
            assert np.array_equal(a.generate(0.03).view(np.uint64),
                                  b.generate(0.03).view(np.uint64))
