"""End to end through the C++ host on one GPU, against the reference's stored
example-1 run (histogram, 30 bins, 4 steps, ML, MIST).

Inputs are the reference's own stored bin-code files (codes and extrema, as
fixtures), a subset file listing the DOFs its 1-D files hold and a neighbour
file listing the pairs its 2-D files hold; the host runs the entropy and
scoring stages (``discretize = false``, as centre_OMP does when restarted
from its bin files) and its outputs are compared with the stored ones:
every 1-D and 2-D entropy contribution (1e-10 relative), the MIE and MIST
convergence tables (printed with 4 decimals) and the MIST trees of every step.
"""
import csv
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from centre_amd import cdf, hip

import h5read  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "centre_amd", "centre_hip")
TYPES = (("b", "bin_bonds.nc", "B", "bond", "bnd"), ("a", "bin_angles.nc", "A", "angle", "ang"),
         ("t", "bin_torsions.nc", "T", "torsion", "tor"))
KIND_WORD = {"bnd": "bond", "ang": "angle", "tor": "torsion",
             "ba": "bacross", "bd": "bdcross", "ad": "adcross"}


def write_inputs(d, g, hist=""):
    out = os.path.join(d, "out")
    os.makedirs(out)
    for t, fname, content, _, _ in TYPES:
        codes = g[f"ex1_codes_{t}"]
        D, F = codes.shape
        sch = g[f"ex1_binschemes_{t}"].astype(np.uint8)
        cdf.write_dataset(
            os.path.join(out, fname),
            [("dims", D), ("extremes", 4), ("bin_scheme_count", len(sch)), ("frames", F)],
            [cdf.VarSpec("dimid", ("dims",), g[f"ex1_dimid_{t}"].astype(np.uint32)),
             cdf.VarSpec("extreme", ("dims", "extremes"), g[f"ex1_extreme_{t}"]),
             cdf.VarSpec("coord", ("dims", "bin_scheme_count", "frames"), codes[:, None, :])],
            attrs={"title": f"{content} Integer Vectors", "application": "CENTRE",
                   "program": "Real2Int", "programVersion": "1.0.0",
                   "convention": "CENTRE_Real", "conventionVersion": "1.0", "content": content,
                   "ncoordims": np.array([D], np.uint32),
                   "nbinSchemes": np.array([len(sch)], np.uint8), "binSchemes": sch},
            version=5, unlimited="dims")
    sub = []
    for t, _, _, word, tag in TYPES:
        ids = g[f"ex1_id_{tag}1d"][:, 0]
        sub.append(f"{word} {g[f'ex1_codes_{t}'].shape[0]} {len(ids)} " + " ".join(map(str, ids)) + ";")
    with open(os.path.join(d, "subset.txt"), "w") as f:
        f.write("\n".join(sub) + "\n")
    neigh = []
    for tag, word in KIND_WORD.items():
        ids = g[f"ex1_id_{tag}2d"]
        rows = {}
        for i, j in ids:
            rows.setdefault(int(i), []).append(int(j))
        for i in sorted(rows):
            neigh.append(f"{word} {i} {len(rows[i])} " + " ".join(map(str, sorted(rows[i]))) + ";")
    with open(os.path.join(d, "neigh.txt"), "w") as f:
        f.write("\n".join(neigh) + "\n")
    cfg = os.path.join(d, "centre.cfg")
    with open(cfg, "w") as f:
        f.write(f"""[control]
    infilepath           = {d}/
    outfilepath          = {out}/
    infofile             = progress-1.info
    nsteps               = 4
    discretize           = false
    calcentropy          = true
    genconvgdata         = true
    cacheentcdimspercpu  = 10
[discretization]
    fname                = 0O.lig-s500_dfs_11-8-7_bat.nc
    pdfmethod            = histogram
[histogram]
    referencenbins       = 30
    nbins                = 30
{hist}[entropy]
    workset              = 2D
    usesubset            = true
    useneighbor          = true
    jacobian             = true
    subsetfile           = subset.txt
    neighborfile         = neigh.txt
    scoringmethod        = MIST
    estimator            = ML
[mi355x]
    procsfororder        = 2
""")
    return cfg, out


def read_table(path):
    with open(path) as f:
        rows = [[c.strip() for c in r] for r in csv.reader(f)]
    return rows[0], rows[1:]


@pytest.mark.gpu
def test_host_example1_entropy_and_scoring(tmp_path, golden):
    assert os.path.exists(HOST), "host binary missing: run __graft_entry__.build()"
    cfg, out = write_inputs(str(tmp_path), golden)
    print("running host on example-1", flush=True)
    r = subprocess.run([HOST, "-i", cfg], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr

    # entropy contributions: NetCDF-4 as the reference writes them, same
    # layout (the reference's stored headers), same records, same values
    for tag in ("ang", "ba"):
        p = os.path.join(out, f"entcontri_{tag}-2d.nc")
        assert h5read.is_hdf5(p), p
        got = h5read.header(p)
        if got is not None:
            with open(os.path.join(ROOT, "tests", "golden", f"ex1h_entcontri_{tag}-2d_header.txt")) as f:
                assert got == f.read().rstrip("\n"), f"{tag}-2d header differs from the reference's"
    for tag in ("bnd", "ang", "tor"):
        ds = h5read.open_any(os.path.join(out, f"entcontri_{tag}-1d.nc"))
        assert ds.attrs["convention"] == "EntropyCont" and ds.attrs["order"][0] == 1
        ids, S = ds.read("id")[:, 0], ds.read("contri_S")[:, 0, 0, :]
        np.testing.assert_array_equal(ids, golden[f"ex1_id_{tag}1d"][:, 0])
        np.testing.assert_allclose(S, golden[f"ex1_S_{tag}1d"], rtol=1e-10, atol=1e-12)
    for tag in KIND_WORD:
        ds = h5read.open_any(os.path.join(out, f"entcontri_{tag}-2d.nc"))
        same = tag in ("bnd", "ang", "tor")
        assert ds.attrs["storagetype"][0] == (3 if same else 0)
        assert ds.attrs["title"][0] == (1 if same else 0)
        ids, S = ds.read("id"), ds.read("contri_S")[:, 0, 0, :]
        # the stored files' own record order (the MPI build's, two ranks)
        np.testing.assert_array_equal(ids, golden[f"ex1_id_{tag}2d"], err_msg=tag)
        np.testing.assert_allclose(S, golden[f"ex1_S_{tag}2d"], rtol=1e-10, atol=1e-12, err_msg=tag)

    # convergence tables (4 decimals as printed)
    for meth in ("MIE", "MIST"):
        hdr, rows = read_table(os.path.join(out, f"{meth}-ML_convergence_b-30.csv"))
        assert hdr == list(golden[f"ex1h_{meth}_hdr"])
        got = np.array([[float(v) for v in row] for row in rows])
        np.testing.assert_allclose(got, golden[f"ex1h_{meth}_csv"], rtol=0, atol=1.5e-4, err_msg=meth)

    # MIST trees: same edges, same MI, largest MI first
    tcode = {"B": 0, "A": 1, "T": 2}
    for s in range(4):
        hdr, rows = read_table(os.path.join(out, f"Tree-MIST-ML_b-30_s-{s}.csv"))
        assert hdr == ["TD1", "TD2", "IndexDim1", "IndexDim2", "MI"]
        got = np.array([[tcode[r[0]], tcode[r[1]], int(r[2]), int(r[3]), float(r[4])] for r in rows])
        want = golden[f"ex1h_tree{s}"]
        assert got.shape == want.shape
        key = lambda m: sorted(tuple(int(x) for x in e[:4]) for e in m)
        assert key(got) == key(want), f"tree edges differ at step {s}"
        assert np.all(np.diff(got[:, 4]) <= 0)
        np.testing.assert_allclose(np.sort(got[:, 4]), np.sort(want[:, 4]), rtol=0, atol=1.5e-6)


def accumulated_mids(mn, mx, nb):
    """binData1D's bin mids: the offset grows by one width per bin."""
    w = (mx - mn) / nb
    o, out = mn, []
    for _ in range(nb):
        out.append(o + w / 2.0)
        o += w
    return out


@pytest.mark.gpu
def test_host_example1_count_files(tmp_path, golden):
    """histogram.writefreq: hist_*-{1d,2d}.nc hold, per record, the ids,
    extrema, accumulated bin mids and the prefix counts of steps firstStep,
    firstStep + stepStride, ... (here 1 and 3 of 4).  The layout follows the
    reference's stored hist_tor-1d.nc header; counts are bit-exact against
    the oracle's binData2D restatement, in each record's own row/column order."""
    cfg, out = write_inputs(str(tmp_path), golden,
                            "    writefreq            = true\n"
                            "    writeset             = B1D, A1D, D1D, BB2D, AA2D, DD2D, BA2D, BD2D, AD2D\n"
                            "    writestepstart       = 1\n"
                            "    writestepstride      = 2\n")
    print("running host on example-1 with count files", flush=True)
    r = subprocess.run([HOST, "-i", cfg], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(os.path.join(os.path.dirname(__file__), "golden", "ex2h_hist_tor-1d_header.json")) as f:
        ref = json.load(f)
    steps = [1, 3]
    codes = {t: golden[f"ex1_codes_{t}"] for t in "bat"}
    ext = {t: golden[f"ex1_extreme_{t}"] for t in "bat"}
    F = codes["b"].shape[1]
    ss = F // 4
    for t, _, _, _, tag in TYPES:
        ds = cdf.open_dataset(os.path.join(out, f"hist_{tag}-1d.nc"))
        assert ds.version == ref["version"] and ds.unlimited == ref["unlimited"]
        assert [k for k, _ in ref["dims"]] == list(ds.dims)
        assert [a[0] for a in ref["attrs"]] == list(ds.attrs)
        assert [(v[0], v[1], v[2]) for v in ref["vars"]] == \
            [(v.name, list(v.dims), v.nc_type) for v in ds.variables.values()]
        for k, typ, val in ref["attrs"]:
            got = ds.attrs[k]
            if typ != "text":
                assert np.asarray(got).dtype.name == typ, k
        assert ds.attrs["content"] == {"bnd": "B", "ang": "A", "tor": "D"}[tag]
        assert list(ds.attrs["firstStep"]) == [1] and list(ds.attrs["stepStride"]) == [2]
        ids = ds.read("id")[:, 0]
        np.testing.assert_array_equal(ids, golden[f"ex1_id_{tag}1d"][:, 0])
        e = ext[t][ids]
        np.testing.assert_array_equal(ds.read("extrema")[:, 0, :], e)
        mids = np.array([accumulated_mids(a, b, 30) for a, b in e[:, :2]])
        np.testing.assert_array_equal(ds.read("mid"), mids)
        freq = ds.read("freq")
        for k, d in enumerate(ids):
            for q, s in enumerate(steps):
                want = np.bincount(codes[t][d, :(s + 1) * ss], minlength=30)
                np.testing.assert_array_equal(freq[k, q], want)
    kinds = {"bnd": ("b", "b", "B/B"), "ang": ("a", "a", "A/A"), "tor": ("t", "t", "D/D"),
             "ba": ("b", "a", "B/A"), "bd": ("b", "t", "B/D"), "ad": ("a", "t", "A/D")}
    for tag, (x, y, content) in kinds.items():
        ds = cdf.open_dataset(os.path.join(out, f"hist_{tag}-2d.nc"))
        assert ds.attrs["content"] == content
        assert list(ds.attrs["storagetype"]) == [0 if x == y else 3]
        assert dict(ds.dims)["bin_pow_dims_sum"] == 900 and dict(ds.dims)["dimsXbin_sum"] == 60
        ent = h5read.open_any(os.path.join(out, f"entcontri_{tag}-2d.nc"))
        ids = ds.read("id")
        np.testing.assert_array_equal(ids, ent.read("id"))
        ex = ds.read("extrema")
        np.testing.assert_array_equal(ex[:, 0], ext[x][ids[:, 0]])
        np.testing.assert_array_equal(ex[:, 1], ext[y][ids[:, 1]])
        mids = ds.read("mid")
        for k in range(0, len(ids), 97):
            want = accumulated_mids(*ext[x][ids[k, 0], :2], 30) + accumulated_mids(*ext[y][ids[k, 1], :2], 30)
            np.testing.assert_array_equal(mids[k], want)
        freq = ds.read("freq")
        # the oracle on one concatenated matrix, rows first DOF, columns second
        both = np.concatenate([codes[x], codes[y]])
        pr = np.stack([ids[:, 0], ids[:, 1] + len(codes[x])], 1)
        want = O.pair_counts(both, 30, pr, 4)[:, steps].reshape(len(ids), 2, 900)
        np.testing.assert_array_equal(freq, want)


@pytest.mark.gpu
def test_host_all_pairs_from_bat_file(tmp_path):
    """Discretise a BAT file and compute every pair (no subset, no neighbour
    list: the host's single all-pairs launch), with ML and MM; every 1-D and
    2-D record agrees with the oracle run on the host's own bin files."""
    nbd, nan, ndh, F, nb, S = 9, 8, 10, 6000, 30, 3
    bat = hip.synth_bat(0xCE27E + 77, nbd, nan, ndh, F)
    d = str(tmp_path)
    out = os.path.join(d, "out")
    os.makedirs(out)
    cdf.write_dataset(os.path.join(d, "bat.nc"),
                      [("frame", F), ("bond", nbd), ("angle", nan), ("dihedral", ndh)],
                      [cdf.VarSpec("bond", ("frame", "bond"), np.ascontiguousarray(bat[:, :nbd])),
                       cdf.VarSpec("angle", ("frame", "angle"),
                                   np.ascontiguousarray(bat[:, nbd:nbd + nan])),
                       cdf.VarSpec("dihedral", ("frame", "dihedral"),
                                   np.ascontiguousarray(bat[:, nbd + nan:]))],
                      attrs={"convention": "CENTRE"}, version=5, unlimited="frame")
    cfg = os.path.join(d, "centre.cfg")
    with open(cfg, "w") as f:
        f.write(f"""[control]
    infilepath           = {d}/
    outfilepath          = {out}/
    nsteps               = {S}
    discretize           = true
    calcentropy          = true
    genconvgdata         = false
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
    estimator            = ML, MM
""")
    print("running host all-pairs", flush=True)
    r = subprocess.run([HOST, "-i", cfg], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    codes, ext = {}, {}
    # bin files: the layout of the reference's stored example-1 bin_bonds.nc
    # (CDF-5; dims, attribute order and types, fixed texts, variables)
    with open(os.path.join(ROOT, "tests", "golden", "ex1h_bin_bonds_header.json")) as f:
        ref = json.load(f)
    sizes = {"b": nbd, "a": nan, "t": ndh}
    for t, fname, content, _, _ in TYPES:
        ds = cdf.open_dataset(os.path.join(out, fname))
        assert ds.version == ref["version"] and ds.unlimited == ref["unlimited"]
        assert list(ds.dims) == [k for k, _ in ref["dims"]]
        assert dict(ds.dims) == {"dims": sizes[t], "extremes": 4, "bin_scheme_count": 1, "frames": F}
        assert list(ds.attrs) == [a[0] for a in ref["attrs"]]
        for k, typ, val in ref["attrs"]:
            got = ds.attrs[k]
            if typ == "text":
                want = {"title": f"{content} Integer Vectors", "content": content}.get(k, val)
                assert got == want, k
            else:
                assert np.asarray(got).dtype.name == typ, k
        assert list(ds.attrs["ncoordims"]) == [sizes[t]] and list(ds.attrs["binSchemes"]) == [nb]
        assert [(v.name, list(v.dims), v.nc_type) for v in ds.variables.values()] == \
            [(v[0], v[1], v[2]) for v in ref["vars"]]
        codes[t] = ds.read("coord")[:, 0, :]
        ext[t] = ds.read("extreme")
    tnum = {"b": 1, "a": 2, "t": 3}
    for tag, t in (("bnd", "b"), ("ang", "a"), ("tor", "t")):
        ds = h5read.open_any(os.path.join(out, f"entcontri_{tag}-1d.nc"))
        want = O.dof_entropies(codes[t], np.full(len(codes[t]), tnum[t], np.uint8), ext[t], S,
                               [nb], (0, 1))
        np.testing.assert_allclose(ds.read("contri_S")[:, :, 0, :], want[:, :, 0, :],
                                   rtol=1e-10, atol=1e-12)
    kinds = {"bnd": "bb", "ang": "aa", "tor": "tt", "ba": "ba", "bd": "bt", "ad": "at"}
    for tag, (x, y) in kinds.items():
        ds = h5read.open_any(os.path.join(out, f"entcontri_{tag}-2d.nc"))
        ids = ds.read("id").astype(np.int64)
        n_x, n_y = len(codes[x]), len(codes[y])
        want_n = n_x * (n_x - 1) // 2 if x == y else n_x * n_y
        assert len(ids) == want_n, tag
        both = np.concatenate([codes[x], codes[y]])
        types = np.array([tnum[x]] * n_x + [tnum[y]] * n_y, np.uint8)
        e = np.concatenate([ext[x], ext[y]])
        pr = np.stack([ids[:, 0], ids[:, 1] + n_x], 1)
        want = O.pair_entropies(both, types, e, pr, S, [nb], (0, 1))
        np.testing.assert_allclose(ds.read("contri_S")[:, :, 0, :], want[:, :, 0, :],
                                   rtol=1e-10, atol=1e-12, err_msg=tag)


@pytest.mark.gpu
def test_host_netcdf4_bat_shuffled_and_strided(tmp_path):
    """The input side: a NetCDF-4 BAT file in the reference's layout
    (Netcdf_BAT.cpp:440-670: float variables bond/angle/dihedral[frame][n],
    frame unlimited, dimension scales named _nc4_non_coord_* where a variable
    shares the dimension's name), discretised with frame shuffling
    (main_smp.cpp:312-357: std::shuffle with mt19937(randseed + 1000), then
    Real2Int::shuffleArray's swap loop, Real2Int.cpp:12-25) and a start /
    stride / nframe entropy range (Netcdf_TrjInt.cpp:353-359).  The bin files
    equal the oracle's codes of the same float32 columns passed through the
    same swap loop, and every 1-D and 2-D contribution equals the oracle's
    on the selected frames."""
    nbd, nan, ndh, F, nb, S = 6, 5, 7, 5000, 30, 3
    start, stride, nframe = 3, 2, 4501
    bat = hip.synth_bat(0xCE27E + 91, nbd, nan, ndh, F).astype(np.float32)
    d = str(tmp_path)
    out = os.path.join(d, "out")
    os.makedirs(out)
    cols = {"bond": bat[:, :nbd], "angle": bat[:, nbd:nbd + nan], "dihedral": bat[:, nbd + nan:]}
    h5read.write_netcdf4(
        os.path.join(d, "bat.nc"),
        [("tuple2", 2, False), ("pseudos", 1, False), ("frame", F, True), ("bond", nbd, False),
         ("angle", nan, False), ("dihedral", ndh, False)],
        [("pseudoBonds", ("pseudos", "tuple2"), np.zeros((1, 2), np.int32), None),
         ("bond", ("frame", "bond"), cols["bond"], (4096, 4)),
         ("angle", ("frame", "angle"), cols["angle"], None),
         ("dihedral", ("frame", "dihedral"), cols["dihedral"], None)],
        {"title": "synthetic", "application": "CENTRE", "program": "Cart2BAT",
         "programVersion": "1.0", "convention": "CENTRE", "conventionVersion": "1.0"})
    cfg = os.path.join(d, "centre.cfg")
    with open(cfg, "w") as f:
        f.write(f"""[control]
    infilepath           = {d}/
    outfilepath          = {out}/
    nsteps               = {S}
    discretize           = true
    calcentropy          = true
    genconvgdata         = false
[discretization]
    fname                = bat.nc
    pdfmethod            = histogram
    shuffleframes        = true
    randseed             = 7
[histogram]
    referencenbins       = {nb}
    nbins                = {nb}
[entropy]
    workset              = 2D
    jacobian             = true
    estimator            = ML, MM
    startframe           = {start}
    strideframe          = {stride}
    nframe               = {nframe}
""")
    print("running host on a NetCDF-4 BAT file (shuffled, strided)", flush=True)
    r = subprocess.run([HOST, "-i", cfg], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    idx = np.fromfile(os.path.join(out, "shuffle_indices.bin"), np.uint64)
    assert np.array_equal(np.sort(idx), np.arange(F)) and not np.array_equal(idx, np.arange(F))
    tnum = {"bond": 1, "angle": 2, "dihedral": 3}
    codes, ext = {}, {}
    for (var, x), (t, fname, _, _, _) in zip(cols.items(), TYPES):
        ds = cdf.open_dataset(os.path.join(out, fname))
        got = ds.read("coord")[:, 0, :]
        want_c, want_e = [], []
        for k in range(x.shape[1]):
            c, e = O.real2int(x[:, k], tnum[var], [nb], nb, True)
            want_c.append(O.shuffle(idx, c[0]))
            want_e.append(e)
        np.testing.assert_array_equal(got, np.array(want_c), err_msg=var)
        np.testing.assert_array_equal(ds.read("extreme"), np.array(want_e), err_msg=var)
        codes[t], ext[t] = got[:, start:nframe:stride], ds.read("extreme")
    for tag, t in (("bnd", "b"), ("ang", "a"), ("tor", "t")):
        ds = h5read.open_any(os.path.join(out, f"entcontri_{tag}-1d.nc"))
        want = O.dof_entropies(np.ascontiguousarray(codes[t]),
                               np.full(len(codes[t]), "bat".index(t) + 1, np.uint8), ext[t], S, [nb], (0, 1))
        np.testing.assert_allclose(ds.read("contri_S")[:, :, 0, :], want[:, :, 0, :],
                                   rtol=1e-10, atol=1e-12, err_msg=tag)
    kinds = {"bnd": "bb", "ang": "aa", "tor": "tt", "ba": "ba", "bd": "bt", "ad": "at"}
    for tag, (x, y) in kinds.items():
        ds = h5read.open_any(os.path.join(out, f"entcontri_{tag}-2d.nc"))
        ids = ds.read("id").astype(np.int64)
        n_x = len(codes[x])
        both = np.ascontiguousarray(np.concatenate([codes[x], codes[y]]))
        types = np.array([" bat".index(x)] * n_x + [" bat".index(y)] * len(codes[y]), np.uint8)
        e = np.concatenate([ext[x], ext[y]])
        pr = np.stack([ids[:, 0], ids[:, 1] + n_x], 1)
        want = O.pair_entropies(both, types, e, pr, S, [nb], (0, 1))
        np.testing.assert_allclose(ds.read("contri_S")[:, :, 0, :], want[:, :, 0, :],
                                   rtol=1e-10, atol=1e-12, err_msg=tag)
