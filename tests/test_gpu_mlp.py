"""The learned propagator's network on the matrix cores against the CPU MLP
oracle (oracle/mlp_oracle.py, pinned by finite differences and torch
autograd in tests/test_oracle_mlp.py).

Tolerances: the device accumulates BF16 products in FP32 and rounds hidden
activations, dY and dZ to BF16 exactly where the oracle does; the remaining
differences are FP32 vs FP64 accumulation and the rare BF16 rounding flip
they cause, so
  GEMM       |C - C_ref| <= 2^-20 * sum_k |A_ik B_kj| + 1e-30
  forward    max|Y - Y_ref| <= 1e-2 * max|Y_ref|
  gradient   ||g_l - g_l,ref|| <= 5e-3 * ||g_l,ref|| for every weight/bias block
"""
import sys

import numpy as np
import pytest

import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rand_bits(rng, shape, scale=1.0):
    return mo.f32_to_bf16_bits(rng.normal(scale=scale, size=shape).astype(np.float32))


@pytest.mark.parametrize("m,n,k", [(64, 64, 32), (37, 50, 70), (130, 16, 257), (16, 128, 512)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_tiles(gpu_ctx, m, n, k, ta, tb):
    from mavi_hip import learned

    say(f"[gemm] m={m} n={n} k={k} ta={ta} tb={tb}")
    rng = np.random.default_rng(m * 1000 + n + k + 7 * ta + 13 * tb)
    a = rand_bits(rng, (k, m) if ta else (m, k))
    b = rand_bits(rng, (n, k) if tb else (k, n))
    A = mo.bf16_bits_to_f64(a)
    B = mo.bf16_bits_to_f64(b)
    A = A.T if ta else A
    B = B.T if tb else B
    ref = A @ B
    bound = np.abs(A) @ np.abs(B) * 2.0**-20 + 1e-30
    c = learned.gemm_check(gpu_ctx, a, b, m, n, k, ta, tb)
    assert np.all(np.abs(c - ref) <= bound), float(np.max(np.abs(c - ref) / bound))


def problem(seed, m):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed)
    p = (p + rng.normal(scale=0.05, size=p.size)).astype(np.float32)
    x = rand_bits(rng, (m, mo.N_IN))
    t = rng.normal(size=(m, mo.N_OUT)).astype(np.float32)
    return p, x, t


def test_forward_fused_and_training_path(gpu_ctx):
    from mavi_hip import learned

    say("[mlp forward]")
    p, x, _ = problem(1, 1000)
    net = learned.MLP(ctx=gpu_ctx)
    net.params = p
    assert np.array_equal(net.params, p)
    y_fused = net.forward(x, fused=True)
    y_gemm = net.forward(x, fused=False)
    y_ref, _ = mo.forward(p, mo.bf16_bits_to_f64(x), bf16=True)
    scale = np.abs(y_ref).max()
    assert np.abs(y_gemm - y_ref).max() <= 1e-2 * scale
    assert np.abs(y_fused - y_ref).max() <= 1e-2 * scale
    assert np.abs(y_fused - y_gemm).max() <= 1e-2 * scale
    # a row's output does not depend on the batch it is in (tiles are independent)
    assert np.array_equal(net.forward(x[:77], fused=True), y_fused[:77])


@pytest.mark.parametrize("m", [777, 4096])
def test_gradients_vs_oracle(gpu_ctx, m):
    from mavi_hip import learned

    say(f"[mlp grad] m={m}")
    p, x, t = problem(2 + m, m)
    net = learned.MLP(ctx=gpu_ctx)
    net.params = p
    loss, g = net.loss_grad(x, t)
    loss_ref, g_ref = mo.loss_grad(p.astype(np.float64), mo.bf16_bits_to_f64(x), t.astype(np.float64), bf16=True)
    assert abs(loss - loss_ref) <= 1e-4 * abs(loss_ref)
    for (ow, ob), (nout, nin) in zip(mo.OFFS, mo.SHAPES):
        for lo, hi in ((ow, ow + nout * nin), (ob, ob + nout)):
            d = np.linalg.norm(g[lo:hi] - g_ref[lo:hi])
            assert d <= 5e-3 * np.linalg.norm(g_ref[lo:hi]) + 1e-12, (lo, hi, d, np.linalg.norm(g_ref[lo:hi]))


def test_training_reduces_loss_and_is_deterministic(gpu_ctx):
    """Adam on a smooth synthetic regression held in the device data set."""
    from mavi_hip import learned

    say("[mlp train]")
    rng = np.random.default_rng(11)
    n = 20000
    xf = rng.uniform(-1, 1, size=(n, mo.N_IN)).astype(np.float32)
    t = np.stack([np.sin(2 * xf[:, 0]) + 0.5 * xf[:, 1] * xf[:, 2], np.cos(xf[:, 3]) - xf[:, 4] ** 2],
                 axis=1).astype(np.float32)
    x = mo.f32_to_bf16_bits(xf)
    nets = []
    for _ in range(2):
        net = learned.MLP(ctx=gpu_ctx, seed=5)
        net.dataset_append(x, t)
        assert net.dataset_size() == n
        xs, ts = net.dataset_get(100, 3)
        assert np.array_equal(xs, x[100:103]) and np.array_equal(ts, t[100:103])
        losses = net.train(steps=400, batch=1024, lr=2e-3, seed=9)
        nets.append((net, losses))
    l0 = nets[0][1]
    say(f"[mlp train] loss {l0[:5].mean():.4f} -> {l0[-20:].mean():.5f}")
    assert l0[-20:].mean() < 0.1 * l0[:5].mean()
    assert np.array_equal(nets[0][1], nets[1][1])
    assert np.array_equal(nets[0][0].params, nets[1][0].params)
    # the trained network generalises to fresh rows
    xv = rng.uniform(-1, 1, size=(2000, mo.N_IN)).astype(np.float32)
    tv = np.stack([np.sin(2 * xv[:, 0]) + 0.5 * xv[:, 1] * xv[:, 2], np.cos(xv[:, 3]) - xv[:, 4] ** 2], axis=1)
    yv = nets[0][0].forward(mo.f32_to_bf16_bits(xv))
    mse = float(np.mean((yv - tv) ** 2))
    base = float(np.mean((tv - tv.mean(axis=0)) ** 2))
    assert mse < 0.2 * base, (mse, base)
