"""CPU oracle of the subgrid-closure network -- test infrastructure only.

FP64 numpy restatement of the 3x3 periodic convolutional network of
include/ns_closure.h (1 -> 32 -> 32 -> 32 -> 32 -> 1 channels: four hidden
layers with GELU in its tanh form, then a linear output layer) and of its
mean-squared-error gradient.
There is no reference counterpart (the reference has no closure); the
gradient is pinned by central finite differences and the convolution by a
direct sum in tests/test_oracle_closure.py.

Parameter layout (flat): per layer the weights [cout][tap][cin], tap = 3 ky + kx
for the offset (ky - 1, kx - 1), then the biases [cout].
"""
import numpy as np

HIDDEN = 32
CIN = [1, HIDDEN, HIDDEN, HIDDEN, HIDDEN]
COUT = [HIDDEN, HIDDEN, HIDDEN, HIDDEN, 1]
LAYERS = len(CIN)
C_GELU = np.sqrt(2.0 / np.pi)


def param_count():
    return sum(co * 9 * ci + co for ci, co in zip(CIN, COUT))


def unpack(params):
    out, o = [], 0
    for ci, co in zip(CIN, COUT):
        w = params[o:o + co * 9 * ci].reshape(co, 9, ci)
        o += co * 9 * ci
        b = params[o:o + co]
        o += co
        out.append((w, b))
    return out


def shifted(x, tap, s):
    """x[b, y + s dy, x + s dx, c] for tap's offset (dy, dx), periodic."""
    dy, dx = (tap // 3 - 1) * s, (tap % 3 - 1) * s
    return np.roll(x, shift=(-dy, -dx), axis=(1, 2))


def conv(x, w, b):
    z = np.zeros(x.shape[:3] + (w.shape[0],))
    for tap in range(9):
        z += shifted(x, tap, 1) @ w[:, tap, :].T
    return z + b


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def gelu_d(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * (1 + t) + 0.5 * z * (1 - t * t) * C_GELU * (1 + 3 * 0.044715 * z * z)


def forward(params, x):
    """x: [batch, n, n] -> y: [batch, n, n]; also the per-layer cache."""
    a = np.asarray(x, dtype=np.float64)[..., None]
    cache = []
    for l, (w, b) in enumerate(unpack(np.asarray(params, dtype=np.float64))):
        z = conv(a, w, b)
        cache.append((a, z))
        a = gelu(z) if l < LAYERS - 1 else z
    return a[..., 0], cache


def loss_and_grad(params, x, t):
    params = np.asarray(params, dtype=np.float64)
    y, cache = forward(params, x)
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(d * d))
    layers = unpack(params)
    grads = []
    dz = (2.0 * d / d.size)[..., None]
    for l in range(LAYERS - 1, -1, -1):
        a_in, _ = cache[l]
        w, _ = layers[l]
        dw = np.zeros_like(w)
        for tap in range(9):
            dw[:, tap, :] = np.einsum("bijo,bijc->oc", dz, shifted(a_in, tap, 1))
        grads.append((dw, dz.sum(axis=(0, 1, 2))))
        if l > 0:
            dx = np.zeros(a_in.shape)
            for tap in range(9):
                dx += shifted(dz, tap, -1) @ w[:, tap, :]
            dz = dx * gelu_d(cache[l - 1][1])
    flat = []
    for dw, db in reversed(grads):
        flat += [dw.ravel(), db.ravel()]
    return loss, np.concatenate(flat)
