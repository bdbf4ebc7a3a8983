"""Batched pair stage (one persistent launch over many systems) against the
single-system path: the same S1, S2 and MI for every system, whatever the
tile order of the shared work queue and the tile shape (DOFs per tile side),
including the autotuned ones; and against the oracle."""
import numpy as np
import pytest

from centre_amd import hip

pytestmark = pytest.mark.gpu


def make_system(ctx, seed, nbd, nan, ndh, F, nb):
    bat = hip.synth_bat(seed, nbd, nan, ndh, F)
    D = nbd + nan + ndh
    d_bat = ctx.to_device(bat)
    d_codes = ctx.alloc((D, F), np.uint8)
    d_ext = ctx.alloc((D, 4), np.float64)
    types = np.array([1] * nbd + [2] * nan + [3] * ndh, np.uint8)
    col = 0
    for t, n in ((1, nbd), (2, nan), (3, ndh)):
        if n:
            ctx.dev_discretize_hist(d_bat.ptr + 8 * col, D, n, F, t, [nb], nb, True,
                                    d_codes.ptr + col * F, d_ext.ptr + 32 * col)
        col += n
    ctx.sync()
    d_bat.free()
    return dict(D=D, F=F, codes=d_codes, ext=d_ext, types=ctx.to_device(types))


@pytest.mark.parametrize("nb,ests", [(50, (0,)), (30, (1, 0)), (5, (0, 1))])
def test_batch_matches_single_system_path(nb, ests):
    ctx = hip.Context(0)
    S = 3
    rng = np.random.default_rng(nb)
    sizes = [(int(rng.integers(3, 14)), int(rng.integers(3, 14)), int(rng.integers(3, 14)),
              int(rng.integers(1000, 3000)) * 4 + 2) for _ in range(6)]
    systems = [make_system(ctx, 0xCE27E + 200 + k, *sz, nb) for k, sz in enumerate(sizes)]
    nest = len(ests)
    want = []
    for s in systems:
        D, F = s["D"], s["F"]
        P = D * (D - 1) // 2
        plan = hip.PairPlan(ctx, D, nb)
        d1, d2, dm = (ctx.alloc((D, nest, S), np.float64), ctx.alloc((P, nest, S), np.float64),
                      ctx.alloc((P, S), np.float64))
        ctx.dev_pair_entropy(plan, s["codes"].ptr, F, F, s["types"].ptr, s["ext"].ptr, S, list(ests),
                             True, False, d1.ptr, d2.ptr, dm.ptr)
        plan.close()
        want.append((d1.numpy(), d2.numpy(), dm.numpy()))
    R = int(hip.lib().centre_pair_dofs_per_tile(nb))
    shapes = [R] if nb <= 8 else [R, R - 1, 2]
    for order, shape in [(1, 0), (2, 0), (3, 0), (0, 0)] + [(1, r) for r in shapes]:
        print(f"batch nb={nb} order={order} tile_dofs={shape} sizes={sizes}", flush=True)
        outs, descs = [], []
        for s in systems:
            D, F = s["D"], s["F"]
            P = D * (D - 1) // 2
            o = (ctx.alloc((D, nest, S), np.float64), ctx.alloc((P, nest, S), np.float64),
                 ctx.alloc((P, S), np.float64))
            outs.append(o)
            descs.append(hip.BatchSystem(s["codes"].ptr, F, D, F, s["types"].ptr, s["ext"].ptr,
                                         o[0].ptr, o[1].ptr, o[2].ptr))
        sys_ms, tot, used, rused = ctx.batch_pair_entropy(descs, nb, S, ests, order=order,
                                                          tile_dofs=shape, shape_used=True)
        assert used in (1, 2, 3) and (order == 0 or used == order)
        assert rused == shape if shape else rused in (R, R - 1)
        assert tot > 0 and np.all(sys_ms > 0) and np.all(sys_ms <= tot * 1.05 + 0.05)
        for (a1, a2, am), o in zip(want, outs):
            # S1 comes from the same sequential per-DOF sums: identical.  The
            # pair sums are exact integer sums of c log c held in a fixed
            # point of 2^-K, K set by the largest frame count of the launch;
            # a batch takes K from its largest system, so a smaller system's
            # table entries are rounded at a coarser K than in its own call:
            # at most 2,500 x 2^-K apart in T, ~1e-13 of S2 for these sizes.
            np.testing.assert_array_equal(o[0].numpy(), a1)
            np.testing.assert_allclose(o[1].numpy(), a2, rtol=1e-12, atol=1e-12)
            np.testing.assert_allclose(o[2].numpy(), am, rtol=1e-12, atol=1e-12)
    ctx.close()


@pytest.mark.parametrize("fmt", [0, 1], ids=["fp4", "int8"])
@pytest.mark.parametrize("nb", [50, 6])
def test_batch_vs_oracle_and_bad_shape(nb, fmt):
    """Partial last blocks on the persistent eight-wave tile (D not a
    multiple of the tile's DOFs; the small-bin tile at nb 6) against the
    oracle directly, with FP4 and with int8 one-hot operands; tile shapes
    outside 1..R are refused."""
    import oracle as O
    ctx = hip.Context(0)
    ctx.set_mfma_format(fmt)
    S = 2
    R = int(hip.lib().centre_pair_dofs_per_tile(nb))
    systems = [make_system(ctx, 0xCE27E + 300 + k, 4, 3, 4 + k, 4000 + 1000 * k, nb) for k in range(3)]
    outs, descs = [], []
    for s in systems:
        D, F = s["D"], s["F"]
        P = D * (D - 1) // 2
        o = (ctx.alloc((D, 1, S), np.float64), ctx.alloc((P, 1, S), np.float64), ctx.alloc((P, S), np.float64))
        outs.append(o)
        descs.append(hip.BatchSystem(s["codes"].ptr, F, D, F, s["types"].ptr, s["ext"].ptr,
                                     o[0].ptr, o[1].ptr, o[2].ptr))
    print(f"batch vs oracle nb={nb} fmt={fmt}", flush=True)
    ctx.batch_pair_entropy(descs, nb, S, [0], order=1, tile_dofs=R)
    for s, o in zip(systems, outs):
        D = s["D"]
        codes, types, ext = s["codes"].numpy(), s["types"].numpy(), s["ext"].numpy()
        iu, ju = np.triu_indices(D, 1)
        want = O.pair_entropies(codes, types, ext, np.stack([iu, ju], 1), S, [nb])
        np.testing.assert_allclose(o[1].numpy()[:, 0, :], want[:, 0, 0, :], rtol=1e-10, atol=1e-12)
        s1 = O.dof_entropies(codes, types, ext, S, [nb])
        mi = s1[iu, 0, 0, :] + s1[ju, 0, 0, :] - want[:, 0, 0, :]
        np.testing.assert_allclose(o[2].numpy(), mi, rtol=1e-9, atol=1e-11)
    for bad in (R + 1, -1):
        with pytest.raises(RuntimeError):
            ctx.batch_pair_entropy(descs, nb, S, [0], order=1, tile_dofs=bad)
    ctx.close()
    ctx.close()
