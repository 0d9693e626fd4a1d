"""CPU oracle of the learned propagator's network (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py as the checker;
the product (mavi_hip / libmavi_hip.so) never imports it.

The network is the project's own (the reference has no learned path), so
there is no reference file to follow; the oracle restates the textbook MLP
forward/backward:

    Z_l = H_{l-1} W_l^T + b_l,  H_l = GELU(Z_l)  (l = 0, 1, 2;  H_{-1} = X)
    Y   = H_2 W_3^T + b_3,      L = 1/(2M) sum (Y - T)^2
    dY = (Y - T)/M,  dW_l = dZ_l^T H_{l-1},  db_l = sum_m dZ_l,
    dZ_{l-1} = (dZ_l W_l) * GELU'(Z_{l-1})

with GELU in its sigmoid form, g(z) = z sigma(1.702 z).
With ``bf16=True`` it rounds to BF16 exactly where
the device kernels store BF16 (weights, hidden activations, dY and dZ), so
the device gradients are compared against the same rounded quantities.

Parity pinned by: (1) central finite differences of the FP64 loss
(bf16=False) against the analytic gradient, (2) torch.autograd in float64 on
the CPU (tests/test_oracle_mlp.py).
"""
from __future__ import annotations

import math

import numpy as np

N_IN, N_HID, N_OUT = 160, 128, 2
SHAPES = [(N_HID, N_IN), (N_HID, N_HID), (N_HID, N_HID), (N_OUT, N_HID)]  # W_l^T: out x in


def offsets():
    """Flat offsets of (W_l^T, b_l) in the parameter vector."""
    out, o = [], 0
    for (nout, nin) in SHAPES:
        out.append((o, o + nout * nin))
        o += nout * nin + nout
    return out, o


OFFS, N_PARAMS = offsets()


def unpack(p):
    ws, bs = [], []
    for (ow, ob), (nout, nin) in zip(OFFS, SHAPES):
        ws.append(np.asarray(p[ow:ow + nout * nin], dtype=np.float64).reshape(nout, nin))
        bs.append(np.asarray(p[ob:ob + nout], dtype=np.float64))
    return ws, bs


def pack(ws, bs):
    p = np.zeros(N_PARAMS)
    for (ow, ob), w, b in zip(OFFS, ws, bs):
        p[ow:ow + w.size] = w.ravel()
        p[ob:ob + b.size] = b
    return p


def f32_to_bf16_bits(x) -> np.ndarray:
    """Round float32 to BF16 (nearest even), as bit patterns (uint16)."""
    u = np.ascontiguousarray(np.asarray(x, dtype=np.float32)).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def bf16_bits_to_f64(b) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)


def rbf16(x) -> np.ndarray:
    """Round to BF16 through float32 and back to float64."""
    return bf16_bits_to_f64(f32_to_bf16_bits(x))


GS = 1.702


def _sigmoid(x):
    from scipy.special import expit

    return expit(x)


def gelu(z):
    return z * _sigmoid(GS * z)


def gelu_grad(z):
    s = _sigmoid(GS * z)
    return s + GS * z * s * (1.0 - s)


def forward(params, x, bf16=True):
    """Returns (Y, cache).  x: float64 features (already BF16 values when
    bf16=True), params: flat FP32/FP64 master parameters."""
    ws, bs = unpack(params)
    if bf16:
        ws = [rbf16(w) for w in ws]
        bs = [np.asarray(b, dtype=np.float32).astype(np.float64) for b in bs]
    h = np.asarray(x, dtype=np.float64)
    Hs, Zs = [h], []
    for l in range(3):
        z = h @ ws[l].T + bs[l]
        a = gelu(z)
        h = rbf16(a) if bf16 else a
        Zs.append(z)
        Hs.append(h)
    y = h @ ws[3].T + bs[3]
    return y, (ws, Hs, Zs)


def loss_grad(params, x, t, bf16=True):
    """(loss, flat gradient) of L = 1/(2M) sum (Y - T)^2."""
    y, (ws, Hs, Zs) = forward(params, x, bf16)
    m = len(x)
    d = y - t
    loss = float(np.sum(d * d) / (2.0 * m))
    dY = d / m
    if bf16:
        dY = rbf16(dY)
    gws, gbs = [None] * 4, [None] * 4
    gws[3] = dY.T @ Hs[3]
    gbs[3] = dY.sum(axis=0)
    dZ = (dY @ ws[3]) * gelu_grad(Zs[2])
    if bf16:
        dZ = rbf16(dZ)
    for l in (2, 1, 0):
        gws[l] = dZ.T @ Hs[l]
        gbs[l] = dZ.sum(axis=0)
        if l == 0:
            break
        dZ = (dZ @ ws[l]) * gelu_grad(Zs[l - 1])
        if bf16:
            dZ = rbf16(dZ)
    return loss, pack(gws, gbs)


def init_params(seed: int) -> np.ndarray:
    """Glorot-uniform weights, zero biases (numpy RNG; any init works for the
    checks, the device has its own)."""
    rng = np.random.default_rng(seed)
    ws, bs = [], []
    for (nout, nin) in SHAPES:
        a = math.sqrt(6.0 / (nin + nout))
        ws.append(rng.uniform(-a, a, size=(nout, nin)))
        bs.append(np.zeros(nout))
    return pack(ws, bs).astype(np.float32)
