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


@pytest.mark.parametrize("n,batch", [(16, 1), (32, 3), (64, 2)])
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


@pytest.mark.parametrize("n,batch", [(16, 2), (32, 2)])
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
