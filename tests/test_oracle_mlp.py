"""Pins the MLP oracle (oracle/mlp_oracle.py) before it checks the device:
BF16 rounding against torch's float32 -> bfloat16 conversion, the analytic
gradient against central finite differences and against torch.autograd in
float64."""
import numpy as np
import pytest

import mlp_oracle as mo


def test_bf16_rounding_known_values_and_torch():
    import torch

    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-9, -2.5, 3.14159265, 1e-30, 65504.0, 0.0, -0.0],
                 dtype=np.float32)
    bits = mo.f32_to_bf16_bits(x)
    assert bits[0] == 0x3F80
    assert bits[1] == 0x3F80          # halfway: ties to even
    assert bits[2] == 0x3F81          # 1 + 3/512 = 1 + 0.75 ulp rounds up to 1 + 1/128
    assert bits[3] == 0xC020
    rng = np.random.default_rng(0)
    y = (rng.standard_normal(20000) * np.exp(rng.uniform(-20, 20, 20000))).astype(np.float32)
    want = torch.from_numpy(y).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(mo.f32_to_bf16_bits(y), want)
    assert np.array_equal(mo.bf16_bits_to_f64(want), torch.from_numpy(want.view(np.int16)).view(
        torch.bfloat16).to(torch.float64).numpy())


def small_problem(seed=0, m=6):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed).astype(np.float64)
    p += rng.normal(scale=0.05, size=p.size)  # non-zero biases
    x = rng.normal(size=(m, mo.N_IN))
    t = rng.normal(size=(m, mo.N_OUT))
    return p, x, t


def test_gradient_matches_finite_differences():
    p, x, t = small_problem()
    _, g = mo.loss_grad(p, x, t, bf16=False)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(mo.N_PARAMS, 60, replace=False),
                          [o[1] for o in mo.OFFS],                 # first bias of every layer
                          [mo.OFFS[3][0], mo.N_PARAMS - 1]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_grad(pp, x, t, bf16=False)[0] - mo.loss_grad(pm, x, t, bf16=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (int(i), fd, g[i])


def test_gradient_matches_torch_autograd():
    import torch

    p, x, t = small_problem(seed=3, m=9)
    loss, g = mo.loss_grad(p, x, t, bf16=False)
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    h = torch.tensor(x, dtype=torch.float64)
    for l, (nout, nin) in enumerate(mo.SHAPES):
        ow, ob = mo.OFFS[l]
        W = pt[ow:ow + nout * nin].view(nout, nin)
        b = pt[ob:ob + nout]
        z = h @ W.T + b
        h = z * torch.sigmoid(1.702 * z) if l < 3 else z
    L = ((h - torch.tensor(t)) ** 2).sum() / (2 * len(x))
    L.backward()
    assert abs(L.item() - loss) <= 1e-12 * abs(loss)
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-9, atol=1e-13)


def test_bf16_mode_close_to_exact():
    """The BF16-rounded oracle stays within BF16 precision of the exact one."""
    p, x, t = small_problem(seed=5, m=32)
    xb = mo.rbf16(x)
    l0, g0 = mo.loss_grad(p, xb, t, bf16=False)
    l1, g1 = mo.loss_grad(p, xb, t, bf16=True)
    assert abs(l1 - l0) <= 0.02 * abs(l0)
    assert np.linalg.norm(g1 - g0) <= 0.03 * np.linalg.norm(g0)
