"""Benchmark: full pairwise MIE pair stage on one MI355X.

Workload (the headline configuration): 600 DOF (200 bonds, 200 angles,
200 torsions) x 100,000 float64 frames from the seeded synthetic generator,
histogram discretisation with 50 bins, 10 convergence steps, ML estimator
with the Jacobian, all 179,700 DOF pairs.  One step = one pass of the pair
stage from device-resident uint8 codes to per-pair S2 (10 steps each) and MI
in HBM: marginal counts and S1, the one-hot MFMA joint counts (FP4 e2m1
operands, exact f32 counts) with the fused FP64 entropy epilogue, and MI.  Discretisation and MIST run once
outside the timed region and are reported as separate stage times.

Prints ONE JSON line (rank 0).  --gpus N > 1 is refused: the library runs
on a single GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np


METRIC = "pair entropies/sec @100k frames, 50 bins; screened vs plain at <=0.1% S error"

# Dense MFMA peaks of the MI355X (TFLOP/s, or TOP/s for int8).
PEAK = {"fp4": 10100.0, "int8": 5000.0}


def mfma_form(n_max):
    """Operand format of the pair tiles for a launch whose largest joint count
    is n_max (steps x frames a step): FP4 while f32 accumulation is exact
    (n_max <= 2^24), else int8 (the library's default rule)."""
    return "fp4" if n_max <= (1 << 24) else "int8"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-bond", type=int, default=200)
    ap.add_argument("--n-angle", type=int, default=200)
    ap.add_argument("--n-dih", type=int, default=200)
    ap.add_argument("--frames", type=int, default=100000)
    ap.add_argument("--nbins", type=int, default=50)
    ap.add_argument("--nsteps", type=int, default=10)
    ap.add_argument("--cpu-pairs", type=int, default=0,
                    help="pairs in the CPU-baseline sample (0 = all pairs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-screen", action="store_true",
                    help="leave out the screened-vs-plain measurement of the default line")
    ap.add_argument("--workload", default="pairs",
                    choices=["pairs", "screen", "batch", "vmkde", "host"],
                    help="pairs: the headline full pairwise pair stage (default); "
                         "screen: learned pair screen vs the plain path at 6,000 DOF; "
                         "batch: 64 uneven systems in one persistent launch; "
                         "vmkde: von Mises KDE discretisation of 3,000 torsions x 50k "
                         "frames, then A-MIE / A-MIST over index neighbours; "
                         "host: the C++ host end to end with file I/O on the headline "
                         "workload (BAT file in, bin/entcontri/CSV files out)")
    ap.add_argument("--batch-systems", type=int, default=64, help="systems (batch)")
    ap.add_argument("--batch-dof", type=int, nargs=2, default=[200, 2000],
                    help="DOF range of the batch systems (default 200 2000)")
    ap.add_argument("--screen-train", type=int, default=20, help="training systems (screen)")
    ap.add_argument("--screen-dof", type=int, default=6000, help="evaluation DOFs (screen)")
    ap.add_argument("--screen-sub", type=int, default=8000,
                    help="frames of the screen's subsample features (screen)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after timing, copy 256 MiB device to device once (the "
                         "known-size kernel the HBM counters are scaled by)")
    return ap.parse_args()


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(codes, types, ext, iu, ju, S, nb, n_pairs):
    """The oracle's per-pair loop (binData2D + entropy2D restated in C,
    OpenMP dynamic over pairs) on an evenly strided sample of the same pairs
    and codes: -O3 and -Ofast builds (-Ofast is the reference's release flag),
    on every host CPU this process may use up to the per-GPU share
    OMP_NUM_THREADS, and on one thread; median of 5 timed runs after one
    warm-up each."""
    import oracle as O
    visible = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or visible
    threads = max(1, min(visible, share))
    P = len(iu)
    k = min(P, n_pairs)
    pick = np.linspace(0, P - 1, k).astype(np.int64)
    pairs = np.stack([iu[pick], ju[pick]], 1)
    k1 = min(k, max(64, k // 40))
    pairs1 = pairs[np.linspace(0, k - 1, k1).astype(np.int64)]
    res = {}
    for v in ("O3", "Ofast"):
        def run(pp, nt):
            t0 = time.perf_counter()
            O.pair_entropies(codes, types, ext, pp, S, [nb], nthreads=nt, variant=v)
            return time.perf_counter() - t0
        run(pairs, threads)
        tn = statistics.median(run(pairs, threads) for _ in range(5))
        run(pairs1, 1)
        t1 = statistics.median(run(pairs1, 1) for _ in range(5))
        res[v] = {"pairs_per_s": k / tn, "one_thread_pairs_per_s": k1 / t1}
    best = max(res, key=lambda v: res[v]["pairs_per_s"])
    return {"value": res[best]["pairs_per_s"], "unit": "pair entropies/s", "cores": threads,
            "kind": "port",
            "sample": f"{k} of the {P} pairs (evenly strided; one thread: {k1}), same codes, "
                      f"binData2D + entropy2D restated in C, OpenMP dynamic over pairs, "
                      f"-{best} build, median of 5 runs after 1 warm-up",
            "cpu_model": cpu_model(), "cpus_visible": visible, "threads_used": threads,
            "build": best, "by_build": res,
            "one_thread_value": res[best]["one_thread_pairs_per_s"]}


PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                       "pair_tile8_kernel_600dof_100kframes_50bins.md")


def traffic_from_profile(workload_is_headline: bool):
    """HBM bytes per pair-kernel launch measured by the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs of this benchmark (scaled by the copy-kernel
    calibration); None when the profile summary is absent or the workload
    differs from the profiled one."""
    if not workload_is_headline or not os.path.exists(PROFILE):
        return None
    import re
    with open(PROFILE) as f:
        m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read(), re.M)
    return float(m.group(1)) if m else None


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the library runs on one GPU only (the pair stage "
              "shards by pair tiles across GPUs by design, but no multi-GPU "
              "path is built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from centre_amd import hip

    if args.workload == "screen":
        return screen_workload(args, hip)
    if args.workload == "batch":
        return batch_workload(args, hip)
    if args.workload == "vmkde":
        return vmkde_workload(args, hip)
    if args.workload == "host":
        return host_workload(args, hip)
    nb, S, F = args.nbins, args.nsteps, args.frames
    nbd, nan, ndh = args.n_bond, args.n_angle, args.n_dih
    D = nbd + nan + ndh
    P = D * (D - 1) // 2
    seed = 0xCE27E + 1

    t0 = time.perf_counter()
    bat = hip.synth_bat(seed, nbd, nan, ndh, F)
    t_gen = time.perf_counter() - t0

    ctx = hip.Context(0)
    d_bat = ctx.to_device(bat)
    d_codes = ctx.alloc((D, F), np.uint8)
    d_ext = ctx.alloc((D, 4), np.float64)
    types = np.array([1] * nbd + [2] * nan + [3] * ndh, np.uint8)
    d_types = ctx.to_device(types)
    ctx.device_sync()

    # ---- discretisation (once; reported separately) ----------------------
    t0 = time.perf_counter()
    col = 0
    for t, n in ((1, nbd), (2, nan), (3, ndh)):
        if n:
            ctx.dev_discretize_hist(d_bat.ptr + 8 * col, D, n, F, t, [nb],
                                    nb, True, d_codes.ptr + col * F,
                                    d_ext.ptr + 32 * col)
        col += n
    ctx.sync()
    t_disc = time.perf_counter() - t0

    plan = hip.PairPlan(ctx, D, nb)
    d_s1 = ctx.alloc((D, 1, S), np.float64)
    d_s2 = ctx.alloc((P, 1, S), np.float64)
    d_mi = ctx.alloc((P, S), np.float64)

    def step():
        ctx.dev_pair_entropy(plan, d_codes.ptr, F, F, d_types.ptr, d_ext.ptr, S,
                             [0], True, False, d_s1.ptr, d_s2.ptr, d_mi.ptr)

    for _ in range(args.warmup):
        step()
    ctx.device_sync()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        kms.append(ctx.last_pair_kernel_ms())
    ctx.device_sync()
    t_run = time.perf_counter() - t0
    ms_step = 1000.0 * t_run / args.steps
    value = P / (t_run / args.steps)

    # ---- MIST on the last step's MI (once; reported separately) ------------
    iu, ju = np.triu_indices(D, 1)
    d_src = ctx.to_device(iu.astype(np.uint32))
    d_dst = ctx.to_device(ju.astype(np.uint32))
    mi_host = d_mi.numpy()
    w_last = ctx.to_device(np.ascontiguousarray(mi_host[:, S - 1]))
    ctx.device_sync()
    t0 = time.perf_counter()
    tree = ctx.dev_mst(d_src.ptr, d_dst.ptr, w_last.ptr, P, D)
    t_mst = time.perf_counter() - t0

    if args.calib_copy:
        nbytes = 256 << 20
        src, dst = ctx.alloc(nbytes, np.uint8), ctx.alloc(nbytes, np.uint8)
        ctx.copy_d2d(dst.ptr, src.ptr, nbytes)
        src.free()
        dst.free()

    kernel_ms = statistics.mean(kms)
    ops = 2.0 * nb * nb * (S * (F // S)) * P
    achieved = ops / (kernel_ms / 1000.0) / 1e12
    form = mfma_form(S * (F // S))
    peak = PEAK[form]

    # ---- the same pass with int8 one-hot operands and int32 counts (after
    # the timed region): reported beside the default FP4 form's figures, with
    # a check that both give bit-identical S2 and MI ------------------------
    int8_form = None
    if form == "fp4":
        d_s2b = ctx.alloc((P, 1, S), np.float64)
        d_mib = ctx.alloc((P, S), np.float64)
        ctx.set_mfma_format(1)
        try:
            def step8():
                ctx.dev_pair_entropy(plan, d_codes.ptr, F, F, d_types.ptr, d_ext.ptr, S,
                                     [0], True, False, d_s1.ptr, d_s2b.ptr, d_mib.ptr)
            step8()
            ctx.device_sync()
            kms8 = []
            t8 = time.perf_counter()
            for _ in range(args.steps):
                step8()
                kms8.append(ctx.last_pair_kernel_ms())
            ctx.device_sync()
            t8 = time.perf_counter() - t8
        finally:
            ctx.set_mfma_format(0)
        k8 = statistics.mean(kms8)
        same = (np.array_equal(d_s2b.numpy().view(np.uint64), d_s2.numpy().view(np.uint64)) and
                np.array_equal(d_mib.numpy().view(np.uint64), d_mi.numpy().view(np.uint64)))
        int8_form = {"value": P / (t8 / args.steps), "ms_per_step": 1000.0 * t8 / args.steps,
                     "kernel_ms": k8, "peak": PEAK["int8"], "unit": "TOP/s",
                     "achieved": ops / (k8 / 1000.0) / 1e12,
                     "frac": ops / (k8 / 1000.0) / 1e12 / PEAK["int8"],
                     "s2_mi_bit_identical_to_fp4": bool(same)}
        d_s2b.free()
        d_mib.free()

    # ---- CPU baseline: the oracle's per-pair loop on a bounded sample -------
    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle as O
        codes = d_codes.numpy()
        ext = d_ext.numpy()
        cpu = cpu_baseline(codes, types, ext, iu, ju, S, nb,
                           args.cpu_pairs if args.cpu_pairs > 0 else 60000)
        cpu["gpu_speedup"] = value / cpu["value"]
        # spot check of the timed GPU output against the oracle on 64 pairs
        chk = np.linspace(0, P - 1, 64).astype(np.int64)
        want = O.pair_entropies(codes, types, ext, np.stack([iu[chk], ju[chk]], 1), S, [nb],
                                nthreads=cpu["threads_used"])[:, 0, 0, :]
        got = d_s2.numpy()[chk, 0, :]
        parity = {"pairs_checked": int(len(chk)),
                  "max_rel_err_S2": float(np.max(np.abs(got - want) /
                                                 np.maximum(np.abs(want), 1.0)))}

    # ---- learned screen vs the plain pass at 6,000 DOF (the metric's second
    # clause; outside the timed region) ---------------------------------------
    scr = None
    if not args.no_screen:
        scr = screen_measure(ctx, hip, args.screen_train, args.screen_dof, args.screen_sub)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair entropies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": form,
        "dtype_note": ("one-hot MFMA operands are exactly 0 or 1 (e2m1 nibbles in the fp4 form); "
                       "joint counts are exact integers (f32 below 2^24, int32 in the int8 form); "
                       "entropy sums are FP64 and fixed-point integer; int8_form checks that S2 and "
                       "MI are bit-identical between the two forms"),
        "data": "synthetic (seeded Philox BAT generator, latent-coupled)",
        "config": {"workload": f"full pairwise MIE pair stage, {D} DOF "
                               f"({nbd}B/{nan}A/{ndh}T) x {F} frames, {nb} bins, "
                               f"{S} steps, ML + Jacobian",
                   "n_dof": D, "n_frame": F, "nbins": nb, "nsteps": S, "n_pair": P},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": traffic_from_profile((D, F, nb, S) == (600, 100000, 50, 10)),
                     "kernel": "pair_tile8_kernel", "kernel_ms": kernel_ms,
                     "ops_per_launch": ops},
        "int8_form": int8_form,
        "cpu_baseline": cpu,
        "screened_vs_plain": scr,
        "parity_sample": parity,
        "stages_s": {"generate_host": t_gen, "discretize": t_disc,
                     "mst": t_mst, "mst_edges": int(len(tree))},
    }
    print(json.dumps(out), flush=True)
    plan.close()
    ctx.close()


# ---- learned pair screen (cfg: 6,000 DOF x 100k frames, 50 bins) ---------------
def _system(ctx, hip, seed, D, F, nb):
    nbd = D // 3
    nan = D // 3
    ndh = D - nbd - nan
    bat = hip.synth_bat(seed, nbd, nan, ndh, F)
    d_bat = ctx.to_device(bat)
    del bat
    d_codes = ctx.alloc((D, F), np.uint8)
    d_ext = ctx.alloc((D, 4), np.float64)
    types = np.array([1] * nbd + [2] * nan + [3] * ndh, np.uint8)
    col = 0
    for t, n in ((1, nbd), (2, nan), (3, ndh)):
        ctx.dev_discretize_hist(d_bat.ptr + 8 * col, D, n, F, t, [nb], nb, True,
                                d_codes.ptr + col * F, d_ext.ptr + 32 * col)
        col += n
    ctx.sync()
    d_bat.free()
    return d_codes, d_ext, ctx.to_device(types)


def _pass(ctx, hip, d_codes, d_types, d_ext, D, F, nb, S, tiles=None, f0=0, nf=None):
    """Pair stage over all pairs (or the given tiles) of frames f0..f0+nf of
    rows of F frames: returns MI of the last step in triangle order (pairs
    outside the tiles: NaN, set before the timed call), S1 and seconds."""
    P = D * (D - 1) // 2
    nf = F - f0 if nf is None else nf
    d_s1 = ctx.alloc((D, 1, S), np.float64)
    d_s2 = ctx.alloc((P, 1, S), np.float64)
    d_mi = ctx.to_device(np.full((P, S), np.nan)) if tiles is not None else \
        ctx.alloc((P, S), np.float64)
    ctx.device_sync()
    t0 = time.perf_counter()
    plan = hip.PairPlan(ctx, D, nb, tiles=tiles) if tiles is not None else hip.PairPlan(ctx, D, nb)
    ctx.dev_pair_entropy(plan, d_codes.ptr + f0, F, nf, d_types.ptr, d_ext.ptr, S, [0], True,
                         False, d_s1.ptr, d_s2.ptr, d_mi.ptr)
    ctx.device_sync()
    dt = time.perf_counter() - t0
    plan.close()
    mi = d_mi.numpy()[:, S - 1].copy()
    s1 = d_s1.numpy()[:, 0, S - 1].copy()
    for a in (d_s1, d_s2, d_mi):
        a.free()
    return mi, s1, dt


def _tile_pairs(tiles, D, R):
    """Pair indices (i < j) inside the given DOF tiles."""
    a, b = np.divmod(np.arange(R * R), R)
    i = (tiles[:, 0:1] * R + a[None, :]).ravel()
    j = (tiles[:, 1:2] * R + b[None, :]).ravel()
    ok = (i < j) & (j < D)
    return i[ok], j[ok]


def _tri(i, j, D):
    i = i.astype(np.int64)
    return i * D - i * (i + 1) // 2 + (j.astype(np.int64) - i - 1)


def _mist(ctx, D, i, j, w, s1sum):
    """MIST total and edge set over edges (i, j) of weight w (NaN: absent)."""
    tree = ctx.mst(D, i.astype(np.uint32), j.astype(np.uint32), w)
    return s1sum - float(w[tree].sum()), set(zip(i[tree].tolist(), j[tree].tolist()))


def screen_measure(ctx, hip, n_train, D, n_sub=8000):
    """Learned pair screen against the plain path on a D-DOF system (100k
    frames, 50 bins, 10 steps, ML, MIST): the network is trained by the
    library's kernels on the plain-path MI of n_train seeded 1,000-DOF
    systems (a fifth of them, at most 4, held back to choose the cutoff);
    returns the plain and screened timings, MIST totals and tree agreement."""
    from centre_amd import screen
    nb, S, F = 50, 10, 100000
    topk, err_budget = 2, 5e-4
    R = int(hip.lib().centre_pair_dofs_per_tile(nb))
    t_start = time.perf_counter()
    # training data: plain-path MI of seeded 1,000-DOF systems
    feats, targets, held = [], [], []
    n_hold = min(4, max(1, n_train // 5))
    for k in range(n_train):
        d_codes, d_ext, d_types = _system(ctx, hip, 0xCE27E + 100 + k, 1000, F, nb)
        d_feat = screen.features(ctx, d_codes, F, d_types, d_ext, nb, n_sub)
        mi, s1, _ = _pass(ctx, hip, d_codes, d_types, d_ext, 1000, F, nb, S)
        if k < n_train - n_hold:
            feats.append(d_feat.numpy())
            targets.append(screen.target(mi))
        else:
            held.append((d_feat.numpy(), mi, s1))
        for a in (d_codes, d_ext, d_types, d_feat):
            a.free()
    ftr, ttr = np.concatenate(feats), np.concatenate(targets)
    net = screen.Screen(ctx, screen.init_params(ftr, 1))
    d_f, d_t = ctx.to_device(ftr), ctx.to_device(ttr)
    t0 = time.perf_counter()
    loss = net.train(d_f, d_t, len(ftr), epochs=4, batch=4096, lr=2e-3, seed=3)
    t_train = time.perf_counter() - t0
    d_f.free()
    d_t.free()
    # cutoff: the largest one whose MIST error on the held-back training
    # systems stays within half the 0.1% budget
    cut_grid = [1e-3, 5e-3, 8e-3, 1e-2, 1.1e-2, 1.2e-2, 1.3e-2, 1.4e-2, 1.5e-2, 1.6e-2, 1.7e-2,
                1.8e-2, 2e-2]
    iu, ju = np.triu_indices(1000, 1)
    chosen = cut_grid[0]
    tune = []
    for cut in cut_grid:
        worst = 0.0
        frac = []
        for fh, mih, s1h in held:
            d_fh = ctx.to_device(fh)
            d_ph = net.predict(d_fh, len(fh))
            tl = ctx.screen_tiles(d_ph.ptr, 1000, nb, cut, topk)
            d_fh.free()
            d_ph.free()
            i, j = _tile_pairs(tl, 1000, R)
            full, _ = _mist(ctx, 1000, iu, ju, mih, s1h.sum())
            scr, _ = _mist(ctx, 1000, i, j, mih[_tri(i, j, 1000)], s1h.sum())
            worst = max(worst, abs(scr - full) / abs(full))
            frac.append(len(i) / len(mih))
        tune.append({"cutoff": cut, "held_max_rel_err": worst, "pairs_computed": float(np.mean(frac))})
        if worst <= err_budget:
            chosen = cut
    # evaluation system
    P = D * (D - 1) // 2
    d_codes, d_ext, d_types = _system(ctx, hip, 0xCE27E + 4, D, F, nb)
    iu, ju = np.triu_indices(D, 1)
    mi, s1, t_pass = _pass(ctx, hip, d_codes, d_types, d_ext, D, F, nb, S)
    t0 = time.perf_counter()
    full_mist, full_tree = _mist(ctx, D, iu, ju, mi, s1.sum())
    t_plain = t_pass + time.perf_counter() - t0

    trees = {}

    def screened(cut):
        ctx.device_sync()
        t0 = time.perf_counter()
        d_feat = screen.features(ctx, d_codes, F, d_types, d_ext, nb, n_sub)
        d_pred = net.predict(d_feat, P)
        tiles = ctx.screen_tiles(d_pred.ptr, D, nb, cut, topk)
        t_sel = time.perf_counter() - t0
        d_feat.free()
        d_pred.free()
        mi_s, _, t_p = _pass(ctx, hip, d_codes, d_types, d_ext, D, F, nb, S, tiles=tiles)
        # the MIST over the pairs computed: skipped pairs are NaN, which the
        # MST leaves out, so the edge list is the same triangle as the plain one
        t1 = time.perf_counter()
        mst, tree = _mist(ctx, D, iu, ju, mi_s, s1.sum())
        t_m = time.perf_counter() - t1
        trees[cut] = tree
        tot = t_sel + t_p + t_m
        return {"cutoff": cut, "tiles": int(len(tiles)),
                "pairs_computed": float(np.count_nonzero(~np.isnan(mi_s))) / P,
                "seconds": tot, "select_s": t_sel, "pass_s": t_p, "mst_s": t_m,
                "speedup": t_plain / tot, "mist": mst,
                "mist_rel_err": abs(mst - full_mist) / abs(full_mist),
                "mst_edges_agree": len(tree & full_tree) / max(1, len(full_tree))}

    at_spec = screened(1e-3)
    at_tuned = screened(chosen)
    # How reproducible the plain tree is: the plain trees of the two halves of
    # the frames, compared with each other and with the full tree.  Edge
    # agreement below this level is the data's sampling noise (near-equal MI
    # among exchangeable DOFs), not the screen's.
    half = F // 2
    mi_a, s1_a, _ = _pass(ctx, hip, d_codes, d_types, d_ext, D, F, nb, S // 2, nf=half)
    mi_b, s1_b, _ = _pass(ctx, hip, d_codes, d_types, d_ext, D, F, nb, S // 2, f0=half, nf=half)
    _, tree_a = _mist(ctx, D, iu, ju, mi_a, s1_a.sum())
    _, tree_b = _mist(ctx, D, iu, ju, mi_b, s1_b.sum())
    del mi_a, mi_b
    n_tree = max(1, len(full_tree))
    types = d_types.numpy()
    tree_s = trees[chosen]

    def by_class(tree):
        # plain-tree edges with a torsion end, and those between bonds/angles
        tor = [e for e in full_tree if types[e[0]] == 3 or types[e[1]] == 3]
        ba = [e for e in full_tree if types[e[0]] != 3 and types[e[1]] != 3]
        return {"torsion_edges": len(tor),
                "torsion_agree": sum(e in tree for e in tor) / max(1, len(tor)),
                "bond_angle_edges": len(ba),
                "bond_angle_agree": sum(e in tree for e in ba) / max(1, len(ba))}

    reproducibility = {
        "halves_agree": len(tree_a & tree_b) / n_tree,
        "first_half_vs_full": len(tree_a & full_tree) / n_tree,
        "second_half_vs_full": len(tree_b & full_tree) / n_tree,
        "screened_by_class": by_class(tree_s),
        "first_half_by_class": by_class(tree_a),
    }
    for a in (d_codes, d_ext, d_types):
        a.free()
    return {
        "speedup": at_tuned["speedup"],
        "mist_rel_err": at_tuned["mist_rel_err"],
        "mst_edges_agree": at_tuned["mst_edges_agree"],
        "config": f"learned pair screen, {D} DOF x {F} frames, {nb} bins, {S} steps, ML, "
                  f"MIST; trained on {n_train - n_hold} + {n_hold} held-back 1,000-DOF systems",
        "n_sub": n_sub, "topk_partners": topk,
        "plain": {"seconds": t_plain, "pass_s": t_pass, "mist": full_mist, "pairs": P},
        "plain_tree_reproducibility": reproducibility,
        "screen_at_1e-3": at_spec,
        "screen_tuned": at_tuned,
        "cutoff_tuning": tune,
        "training": {"pairs": int(len(ftr)), "epochs": 4, "seconds": t_train,
                     "loss_by_epoch": loss.tolist()},
        "wall_s": time.perf_counter() - t_start,
    }


def screen_workload(args, hip):
    """Learned pair screen against the plain path (one JSON line)."""
    ctx = hip.Context(0)
    r = screen_measure(ctx, hip, args.screen_train, args.screen_dof, args.screen_sub)
    out = {
        "metric": "screened MIST speedup over the plain GPU path at <=0.1% S error",
        "value": r["speedup"], "unit": "x", "higher_is_better": True,
        "n_gpus": 1, "data": "synthetic (seeded Philox BAT generator, latent-coupled)",
        "config": {"workload": r["config"], "n_sub": r["n_sub"],
                   "topk_partners": r["topk_partners"]},
    }
    out.update({k: v for k, v in r.items() if k not in ("speedup", "config", "n_sub",
                                                          "topk_partners")})
    print(json.dumps(out), flush=True)
    ctx.close()


# ---- batching (64 systems, D ~ U[200, 2000], F ~ U[20k, 200k]) --------------------
def batch_workload(args, hip):
    """Many uneven systems: one batched call against one call per system."""
    nb, S = 50, 10
    ctx = hip.Context(0)
    rng = np.random.default_rng(0xCE27E + 3)
    n_sys = args.batch_systems
    dlo, dhi = args.batch_dof
    dims = [(int(rng.integers(dlo, dhi + 1)), int(rng.integers(20000, 200001))) for _ in range(n_sys)]
    t0 = time.perf_counter()
    systems = []
    for k, (D, F) in enumerate(dims):
        d_codes, d_ext, d_types = _system(ctx, hip, 0xCE27E + 3000 + k, D, F, nb)
        P = D * (D - 1) // 2
        systems.append(dict(D=D, F=F, P=P, codes=d_codes, ext=d_ext, types=d_types,
                            s1=ctx.alloc((D, 1, S), np.float64), s2=ctx.alloc((P, 1, S), np.float64),
                            mi=ctx.alloc((P, S), np.float64)))
    t_gen = time.perf_counter() - t0
    n_pairs = sum(s["P"] for s in systems)
    mm_ops = float(sum(2.0 * nb * nb * s["F"] * s["P"] for s in systems))
    form = mfma_form(S * max(F // S for _, F in dims))
    # one launch per system (the plain path, system after system)
    ctx.device_sync()
    t0 = time.perf_counter()
    seq_kernel_ms = 0.0
    for s in systems:
        plan = hip.PairPlan(ctx, s["D"], nb)
        ctx.dev_pair_entropy(plan, s["codes"].ptr, s["F"], s["F"], s["types"].ptr, s["ext"].ptr, S,
                             [0], True, False, s["s1"].ptr, s["s2"].ptr, s["mi"].ptr)
        seq_kernel_ms += ctx.last_pair_kernel_ms()
        plan.close()
    ctx.device_sync()
    t_seq = time.perf_counter() - t0
    ref = [systems[k]["mi"].numpy()[:, S - 1].copy() for k in (0, n_sys - 1)]
    # one persistent launch for all (tile order autotuned)
    descs = [hip.BatchSystem(s["codes"].ptr, s["F"], s["D"], s["F"], s["types"].ptr, s["ext"].ptr,
                             s["s1"].ptr, s["s2"].ptr, s["mi"].ptr) for s in systems]
    ctx.device_sync()
    t0 = time.perf_counter()
    sys_ms, kern_ms, order, shape = ctx.batch_pair_entropy(descs, nb, S, [0], order=0, shape_used=True)
    ctx.device_sync()
    t_batch = time.perf_counter() - t0
    # equal to the fixed-point rounding of the batch's coarser scale (<= 1e-12)
    same = all(np.allclose(systems[k]["mi"].numpy()[:, S - 1], r, rtol=1e-12, atol=1e-12)
               for k, r in zip((0, n_sys - 1), ref))
    per_choice = {}
    R = int(hip.lib().centre_pair_dofs_per_tile(nb))
    for r in (R, R - 1):
        for o in (1, 2, 3):
            _, ms, _ = ctx.batch_pair_entropy(descs, nb, S, [0], order=o, tile_dofs=r)
            per_choice[f"order{o}_dofs{r}"] = ms
    # MIST of every system on its last-step MI
    t0 = time.perf_counter()
    mist = []
    for s in systems:
        iu, ju = np.triu_indices(s["D"], 1)
        w = s["mi"].numpy()[:, S - 1]
        tree = ctx.mst(s["D"], iu.astype(np.uint32), ju.astype(np.uint32), w)
        mist.append(float(s["s1"].numpy()[:, 0, S - 1].sum() - w[tree].sum()))
    t_mist = time.perf_counter() - t0
    out = {
        "metric": f"batched pair entropies/sec, {n_sys} uneven systems in one call",
        "value": n_pairs / t_batch, "unit": "pair entropies/s", "higher_is_better": True,
        "n_gpus": 1, "data": "synthetic (seeded Philox BAT generator, latent-coupled)",
        "config": {"workload": f"{n_sys} systems, D ~ U[{dlo}, {dhi}], F ~ U[20k, 200k], {nb} bins, "
                               f"{S} steps, ML, MIE + MIST", "pairs": n_pairs,
                   "frames_total": int(sum(F for _, F in dims))},
        "batch": {"seconds": t_batch, "kernel_ms": kern_ms, "order_used": order,
                  "tile_dofs_used": shape, "kernel_ms_by_choice": per_choice,
                  "system_latency_ms": {"min": float(sys_ms.min()), "median": float(np.median(sys_ms)),
                                        "max": float(sys_ms.max())}},
        "one_launch_per_system": {"seconds": t_seq, "pairs_per_s": n_pairs / t_seq,
                                  "kernel_ms_sum": seq_kernel_ms},
        "speedup_vs_per_system": t_seq / t_batch,
        "kernel_speedup_vs_per_system": seq_kernel_ms / kern_ms,
        # MFMA roofline of the batch kernel: 2 nb^2 F ops per pair, every system
        "dtype": form,
        "roofline": {"bound": "mfma", "ops": mm_ops,
                     "achieved": mm_ops / (kern_ms * 1e-3) / 1e12, "peak": PEAK[form],
                     "unit": "TFLOP/s", "frac": mm_ops / (kern_ms * 1e-3) / 1e12 / PEAK[form]},
        "results_match_per_system_path": bool(same),
        "mist_all_systems_s": t_mist,
        "generate_and_discretise_s": t_gen,
    }
    print(json.dumps(out), flush=True)
    ctx.close()




def vmkde_workload(args, hip):
    """von Mises KDE discretisation (kappa 10, 64-interval grid, at most 5
    states) of 3,000 torsions x 50,000 frames, then A-MIE and A-MIST over the
    pairs within 30 DOF by index (one JSON line)."""
    D, F, G, kappa, nmax, S, cut = 3000, 50000, 64, 10.0, 5, 10, 30
    ctx = hip.Context(0)
    t0 = time.perf_counter()
    bat = hip.synth_bat(0xCE27E + 7, 0, 0, D, F)
    t_gen = time.perf_counter() - t0
    d_bat = ctx.to_device(bat)
    d_codes = ctx.alloc((D, F), np.uint8)
    d_min = ctx.alloc((D, 9), np.float64)
    d_nf = ctx.alloc(D, np.uint8)
    d_grid = ctx.alloc((D, 2, G + 1), np.float64)
    types = np.full(D, 3, np.uint8)
    d_types = ctx.to_device(types)
    d_ext = ctx.to_device(np.zeros((D, 4)))
    ii, jj = [], []
    for i in range(D):
        j = np.arange(i + 1, min(D, i + cut + 1))
        ii.append(np.full(len(j), i))
        jj.append(j)
    ii, jj = np.concatenate(ii).astype(np.uint32), np.concatenate(jj).astype(np.uint32)
    pairs = np.stack([ii, jj], 1)
    P = len(pairs)
    d_s1 = ctx.alloc((D, 1, S), np.float64)
    d_s2 = ctx.alloc((P, 1, S), np.float64)
    d_mi = ctx.alloc((P, S), np.float64)
    d_src, d_dst = ctx.to_device(ii), ctx.to_device(jj)
    d_w = ctx.alloc(P, np.float64)
    t0 = time.perf_counter()
    plan = hip.PairPlan(ctx, D, nmax, pairs)
    t_plan = time.perf_counter() - t0

    def disc():
        ctx.dev_discretize_vmkde(d_bat.ptr, D, D, F, nmax, kappa, d_codes.ptr, d_min.ptr,
                                 d_nf.ptr, grid_intervals=G, d_grid=d_grid.ptr)

    def pair():
        ctx.dev_pair_entropy(plan, d_codes.ptr, F, F, d_types.ptr, d_ext.ptr, S, [0], True, True,
                             d_s1.ptr, d_s2.ptr, d_mi.ptr)

    def mist():
        # MI of the last step as the edge weights, tree on the device
        ctx.copy_2d(d_w.ptr, 8, d_mi.ptr + 8 * (S - 1), 8 * S, 8, P)
        return ctx.dev_mst(d_src.ptr, d_dst.ptr, d_w.ptr, P, D)

    for _ in range(max(1, args.warmup)):
        disc()
        pair()
        mist()
    t_disc, t_pair, t_mst = [], [], []
    for _ in range(args.steps):
        ctx.device_sync()
        a = time.perf_counter()
        disc()
        ctx.device_sync()
        b = time.perf_counter()
        pair()
        ctx.device_sync()
        c = time.perf_counter()
        tree = mist()
        ctx.device_sync()
        d = time.perf_counter()
        t_disc.append(b - a)
        t_pair.append(c - b)
        t_mst.append(d - c)
    kernel_ms = ctx.last_pair_kernel_ms()
    plan.close()
    codes = d_codes.numpy()
    nf = d_nf.numpy()
    mi = d_mi.numpy()[:, S - 1]
    s1 = d_s1.numpy()[:, 0, S - 1]
    td, tp, tm = statistics.median(t_disc), statistics.median(t_pair), statistics.median(t_mst)
    out = {
        "metric": "vmKDE discretisation + A-MIE/A-MIST, 3,000 torsions x 50k frames",
        "value": P / tp, "unit": "neighbour pair entropies/s", "higher_is_better": True,
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": (td + tp + tm) * 1e3, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (KDE, entropies); fp4 one-hot counts",
        "data": "synthetic (seeded torsion mixtures, latent-coupled)",
        "config": {"workload": "3,000 torsions x 50,000 frames, vmKDE kappa 10, 64-interval grid, "
                               "nmaxconf 5, neighbours within 30 DOF, 10 steps, ML",
                   "pairs": P},
        "vmkde": {"ms": td * 1e3, "dof_per_s": D / td,
                  "grid_evals_per_s": D * (G + 1) * F / td,
                  "states_found": np.bincount(nf, minlength=nmax + 1).tolist()},
        "pair_stage": {"ms": tp * 1e3, "kernel_ms": kernel_ms, "plan_setup_ms": t_plan * 1e3},
        "mist_ms": tm * 1e3,
        "entropy": {"S1_sum": float(s1.sum()), "MIE": float(s1.sum() - mi.sum()),
                    "MIST": float(s1.sum() - mi[tree].sum())},
        "generate_s": t_gen,
    }
    if not args.no_cpu:
        import oracle as O
        sample = 6
        grid = d_grid.numpy()
        t_vm = 0.0
        grid_err, same_given_grid, frames_diff = 0.0, True, 0
        for d in range(sample):
            x = bat[:, d].astype(np.float32)
            t0 = time.perf_counter()
            wg = O.vmkde_grid(x, kappa, G)
            wc, _, _ = O.vmkde_states(x, wg, nmax, 5, 1e-4, 1000, G)
            t_vm += time.perf_counter() - t0
            g = grid[d].reshape(-1)
            grid_err = max(grid_err, float(np.max(np.abs(g[:G + 1] - wg[:G + 1])) /
                                           np.max(np.abs(wg[:G + 1]))))
            gc, _, _ = O.vmkde_states(x, g, nmax, 5, 1e-4, 1000, G)
            same_given_grid &= bool(np.array_equal(gc, codes[d]))
            frames_diff += int(np.count_nonzero(wc != codes[d]))
        t_vm /= sample
        ps = 3000
        t0 = time.perf_counter()
        w2 = O.pair_entropies(codes, types, np.zeros((D, 4)), pairs[:ps], S, [nmax], [0],
                              True, True, nthreads=1)
        t_pr = (time.perf_counter() - t0) / ps
        got = d_s2.numpy()[:ps, 0, :]
        s2_err = float(np.max(np.abs(got - w2[:, 0, 0, :]) / np.maximum(np.abs(w2[:, 0, 0, :]), 1.0)))
        cpu_s = D * t_vm + P * t_pr
        out["cpu_baseline"] = {
            "value": 1.0 / t_pr, "unit": "neighbour pair entropies/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle vmKDE on {sample} torsions ({t_vm * 1e3:.1f} ms each) and "
                      f"{ps} neighbour pairs; whole workload estimated {cpu_s:.1f} s on one core",
            "end_to_end_s_estimate": cpu_s}
        out["parity_sample"] = {
            "grid_max_rel_err": grid_err, "codes_equal_given_device_grid": same_given_grid,
            "frames_differing_with_oracle_grid": frames_diff, "frames_checked": sample * F,
            "s2_max_rel_err": s2_err, "pairs_checked": ps}
        out["speedup_end_to_end_vs_cpu_1core"] = cpu_s / (td + tp + tm)
    print(json.dumps(out), flush=True)


def host_workload(args, hip):
    """The C++ host (centre_amd/centre_hip) end to end on the headline
    workload: a CDF-5 BAT file of 600 DOF x 100,000 frames is written, then
    one run discretises it, computes every 1-D and 2-D contribution (all
    179,700 pairs, 10 steps) and the MIE / MIST tables, reading and writing
    files as centre_OMP does (one JSON line)."""
    import shutil
    import subprocess
    import tempfile
    from centre_amd import cdf
    host = os.path.join(os.path.dirname(os.path.abspath(__file__)), "centre_amd", "centre_hip")
    if not os.path.exists(host):
        raise SystemExit("centre_amd/centre_hip is not built (run __graft_entry__.build())")
    nbd, nan, ndh, F, nb, S = args.n_bond, args.n_angle, args.n_dih, args.frames, args.nbins, args.nsteps
    D = nbd + nan + ndh
    P = D * (D - 1) // 2
    work = tempfile.mkdtemp(prefix="centre_host_")
    try:
        bat = hip.synth_bat(0xCE27E + 1, nbd, nan, ndh, F)
        t0 = time.perf_counter()
        cdf.write_dataset(
            os.path.join(work, "bat.nc"),
            [("frame", F), ("bond", nbd), ("angle", nan), ("dihedral", ndh)],
            [cdf.VarSpec("bond", ("frame", "bond"), np.ascontiguousarray(bat[:, :nbd]),
                         {"units": "angstrom"}),
             cdf.VarSpec("angle", ("frame", "angle"), np.ascontiguousarray(bat[:, nbd:nbd + nan]),
                         {"units": "radian"}),
             cdf.VarSpec("dihedral", ("frame", "dihedral"), np.ascontiguousarray(bat[:, nbd + nan:]),
                         {"units": "radian"})],
            attrs={"convention": "CENTRE"}, version=5, unlimited="frame")
        t_write = time.perf_counter() - t0
        del bat
        os.makedirs(os.path.join(work, "out"))
        cfg = os.path.join(work, "centre.cfg")
        with open(cfg, "w") as f:
            f.write(f"""[control]
    infilepath           = {work}/
    outfilepath          = {work}/out/
    infofile             = progress.info
    nsteps               = {S}
    discretize           = true
    calcentropy          = true
    genconvgdata         = true
[discretization]
    fname                = bat.nc
    pdfmethod            = histogram
[histogram]
    referencenbins       = {nb}
    nbins                = {nb}
[entropy]
    workset              = 2D
    usesubset            = false
    useneighbor          = false
    jacobian             = true
    scoringmethod        = MIST
    estimator            = ML
""")
        t0 = time.perf_counter()
        r = subprocess.run([host, "-i", cfg, "-O"], capture_output=True, text=True)
        t_run = time.perf_counter() - t0
        if r.returncode != 0:
            raise SystemExit(f"host failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}")
        stages = {}
        for line in r.stdout.splitlines():
            if line.startswith("TIME: "):
                k, v = line[6:].rsplit(" ", 2)[:2]
                stages[k] = float(v)
        outs = sorted(os.listdir(os.path.join(work, "out")))
        sizes = sum(os.path.getsize(os.path.join(work, "out", o)) for o in outs)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    out = {
        "metric": "pair entropies/sec end to end through the C++ host, file I/O included",
        "value": P / t_run, "unit": "pair entropies/s", "higher_is_better": True,
        "n_gpus": 1, "scaling": "weak", "vs_baseline": None, "dtype": mfma_form(F),
        "data": "synthetic (seeded Philox BAT generator, latent-coupled), CDF-5 BAT file",
        "config": {"workload": f"centre_hip -i centre.cfg: discretise + entropy + MIE/MIST, {D} DOF "
                               f"({nbd}B/{nan}A/{ndh}T) x {F} frames, {nb} bins, {S} steps, ML",
                   "n_pair": P},
        "seconds": t_run, "stages_s": stages, "bat_file_write_s": t_write,
        "output_files": len(outs), "output_bytes": sizes,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
