"""The allocation-free pair stage (centre_dev_pair_entropy_noalloc) against
the allocating entry point (centre_dev_pair_entropy) on the same inputs:
S1, S2 and MI within 1e-10 relative (MI 1e-10 absolute), on calls in a row
whose second input is smaller, so the static scratch is reused with the
previous call's contents left in it."""
import numpy as np
import pytest

from centre_amd import hip

pytestmark = pytest.mark.gpu

REL = 1e-10


def rel_err(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


def system(seed, D, F, nb, stride):
    rng = np.random.default_rng(seed)
    codes = np.zeros((D, stride), np.uint8)
    c = rng.integers(0, nb, size=(D, F), dtype=np.uint8)
    c[1::2] = (c[0::2][: len(c[1::2])] + rng.integers(0, 3, size=F)) % nb
    codes[:, :F] = c
    types = np.array([1, 2, 3] * (D // 3 + 1), np.uint8)[:D]
    ext = np.zeros((D, 4))
    ext[:, 0] = rng.uniform(0.5, 1.0, D)
    ext[:, 1] = ext[:, 0] + rng.uniform(0.2, 1.5, D)
    return codes, types, ext


def run_both(ctx, D, F, nb, S, ests, stride, pairs):
    codes, types, ext = system(D * 7 + nb, D, F, nb, stride)
    plan = hip.PairPlan(ctx, D, nb, pairs)
    P = plan.records
    nest = len(ests)
    dc, dt, dx = ctx.to_device(codes), ctx.to_device(types), ctx.to_device(ext)
    out = []
    for noalloc in (False, True):
        d_s1 = ctx.alloc((D, nest, S), np.float64)
        d_s2 = ctx.alloc((P, nest, S), np.float64)
        d_mi = ctx.alloc((P, S), np.float64)
        f = ctx.dev_pair_entropy_noalloc if noalloc else ctx.dev_pair_entropy
        f(plan, dc.ptr, stride, F, dt.ptr, dx.ptr, S, ests, True, False, d_s1.ptr, d_s2.ptr,
          d_mi.ptr)
        out.append((d_s1.numpy(), d_s2.numpy(), d_mi.numpy()))
    plan.close()
    (a1, a2, am), (b1, b2, bm) = out
    assert rel_err(b1, a1) < REL and rel_err(b2, a2) < REL
    assert np.max(np.abs(bm - am)) < 1e-10
    assert np.all(np.isfinite(b2))


# (nb, D, F, S, ests, code row stride, explicit pairs): all-pairs plans on the
# eight-wave tile, a misaligned stride (step-aligned copy in the scratch), an
# explicit list (canonical outputs in the scratch), and a small-nb plan on the
# small-bin tile
CASES = [
    (50, 64, 20000, 4, (0, 1), 20000, False),
    (50, 23, 9001, 3, (1,), 9003, False),
    (30, 40, 6000, 5, (0,), 6000, True),
    (5, 30, 12000, 2, (0, 1), 12000, False),
]


@pytest.mark.parametrize("first,second", [(0, 1), (2, 3)])
def test_noalloc_matches_allocating_entry(first, second):
    with hip.Context(0) as ctx:
        for k in (first, second):
            nb, D, F, S, ests, stride, explicit = CASES[k]
            pairs = None
            if explicit:
                rng = np.random.default_rng(k)
                pairs = rng.integers(0, D, size=(5 * D, 2))
                pairs = pairs[pairs[:, 0] != pairs[:, 1]].astype(np.uint32)
            print(f"noalloc case {k}: nb={nb} D={D} F={F} S={S} ests={ests} stride={stride} "
                  f"explicit={explicit}", flush=True)
            run_both(ctx, D, F, nb, S, ests, stride, pairs)


def test_noalloc_rejects_calls_beyond_the_scratch():
    """The marginals of 12,000 DOFs x 2 estimators x 255 steps alone (73 MB)
    exceed the 64 MiB array: the call returns -3 before any launch; so does
    an estimator the allocation-free path does not take."""
    D, F, nb, S = 12000, 255, 10, 255
    with hip.Context(0) as ctx:
        plan = hip.PairPlan(ctx, D, nb, np.array([[0, 1]], np.uint32))
        dc = ctx.alloc((D, F), np.uint8)
        dt = ctx.to_device(np.ones(D, np.uint8))
        dx = ctx.to_device(np.tile([1.0, 2.0, 0.0, 0.0], (D, 1)))
        d_s2 = ctx.alloc((1, 2, S), np.float64)
        with pytest.raises(hip.CentreError) as e:
            ctx.dev_pair_entropy_noalloc(plan, dc.ptr, F, F, dt.ptr, dx.ptr, S, (0, 1), True,
                                         False, 0, d_s2.ptr, 0)
        assert e.value.code == -3
        with pytest.raises(hip.CentreError) as e:
            ctx.dev_pair_entropy_noalloc(plan, dc.ptr, F, F, dt.ptr, dx.ptr, S, (2,), True,
                                         False, 0, d_s2.ptr, 0)
        assert e.value.code == -3
        plan.close()
