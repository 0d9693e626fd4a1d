"""Objects of the C ABI may be destroyed in any order (a garbage collector,
Python's at interpreter exit or Julia's finalizers, picks the order): a
context outlives its systems, batches and networks, and a system or network
outlives the learned propagators bound to it, until the last of them is
destroyed."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def build(ctx):
    from mavi_hip import learned, rings
    from mavi_hip import workloads as wl

    st, sp, cfg, ic = wl.rings_mixture(nrings=60, seed=4, chunk=1.25)
    s = rings.RingsSystem(state=st, space_cfg=sp, dynamic_cfg=cfg, int_cfg=ic, seed=4, ctx=ctx)
    net = learned.MLP(ctx=ctx, seed=1)
    lp = learned.Learned(s.dev, net, k=10, r_feat=1.25)
    return s, net, lp


@pytest.mark.parametrize("order", [("ctx", "sys", "net", "lp"), ("sys", "net", "ctx", "lp"),
                                   ("lp", "ctx", "net", "sys"), ("net", "lp", "sys", "ctx")])
def test_destroy_in_any_order(order):
    from mavi_hip import _lib

    say(f"[lifetimes] destroy order {order}")
    ctx = _lib.Context(0)  # a context of its own (not the shared default)
    s, net, lp = build(ctx)
    lp.run(1)  # use all of them once
    objs = {"ctx": ctx, "sys": s.dev, "net": net, "lp": lp}
    for name in order:
        objs[name].__del__()  # calls the C destroy and clears the handle
        assert objs[name].handle is None
    # the library keeps working on a fresh context afterwards
    ctx2 = _lib.Context(0)
    s2, net2, lp2 = build(ctx2)
    assert lp2.run(1) >= 0
    s2.download()
    assert np.all(np.isfinite(s2.state.pos))
