"""mavi_mlp_forward_noalloc: the fused inference kernel with the library's
fixed per-GPU staging array instead of per-call allocations.  Its output must
equal mavi_mlp_forward (fused) bit for bit, on two calls in a row (the second
smaller, so stale rows of the first call would show), at the row limit, and
sizes beyond the limit fail with a clear error."""
import sys

import numpy as np
import pytest

import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def features(m, seed):
    rng = np.random.default_rng(seed)
    return mo.f32_to_bf16_bits(rng.standard_normal((m, 160)).astype(np.float32))


def test_noalloc_forward_equals_forward(gpu_ctx):
    import mavi_hip as mv
    from mavi_hip import learned

    net = learned.MLP(ctx=gpu_ctx, seed=7)
    for m, seed in ((5000, 1), (1234, 2), (learned.NOALLOC_ROWS, 3), (1, 4)):
        say(f"[no-alloc] {m} rows")
        x = features(m, seed)
        y_ref = net.forward(x, fused=True)
        y = net.forward_noalloc(x)
        assert y.shape == (m, 2)
        assert np.array_equal(y, y_ref), (m, float(np.abs(y - y_ref).max()))
        assert np.all(np.isfinite(y))
    say("[no-alloc] beyond the limit")
    with pytest.raises(mv.MaviError, match="NOALLOC_ROWS"):
        net.forward_noalloc(features(learned.NOALLOC_ROWS + 1, 5))
    # still usable after the refused call
    x = features(300, 6)
    assert np.array_equal(net.forward_noalloc(x), net.forward(x, fused=True))
