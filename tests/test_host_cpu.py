"""The C++ host (``centre_amd/centre_hip``) without a GPU: input-file parsing
(Utils::parseConfiguration / Inputs, Utils.cpp:198-270, Inputs.cpp:14-751)
and the NetCDF readers for both file families the reference uses: classic
CDF-5 (its bin files) and NetCDF-4/HDF5 (its entropy-contribution files; the
fixture is example-1's ``entcontri_ang-1d.nc`` as the reference stored it)."""
import os
import subprocess

import numpy as np
import pytest

from centre_amd import cdf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "centre_amd", "centre_hip")
NC4 = os.path.join(ROOT, "tests", "golden", "ex1h_entcontri_ang-1d.nc")

pytestmark = pytest.mark.skipif(not os.path.exists(HOST), reason="host not built (run build())")


def run(*args):
    r = subprocess.run([HOST, *args], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout, r.stderr


def kv(out):
    d = {}
    for line in out.splitlines():
        k, _, v = line.partition(" ")
        d[k] = v
    return d


def test_print_config_reference_dialect(tmp_path):
    cfg = tmp_path / "in.cfg"
    cfg.write_text("""# comment line
[control]
    nsteps = 4   # inline comment
    discretize = false
    outfilepath = "out dir/"
[discretization]
    fname = bat.nc
    pdfmethod = histogram
[histogram]
    referencenbins = 30
    nbins = 30, 50
[entropy]
    workset = BB2D, DD2D, B1D
    usesubset = true
    estimator = MM, ML, MM
    scoringmethod = mie
""")
    rc, out, err = run("--print-config", str(cfg))
    assert rc == 0, err
    d = kv(out)
    assert d["nsteps"] == "4" and d["discretize"] == "0" and d["calcentropy"] == "1"
    assert d["outfilepath"] == "out dir/"
    assert d["nbins"] == "30 50" and d["referencenbins"] == "30"
    assert d["workset"] == str(8 | 32 | 1)
    assert d["estimators"] == "MM ML"          # duplicates dropped, order kept
    assert d["method"] == "0" and d["usesubset"] == "1" and d["jacobian"] == "1"
    assert d["cacheentcdimspercpu"] == "10"    # default (Inputs.cpp defaults)


def test_print_config_missing_required(tmp_path):
    cfg = tmp_path / "in.cfg"
    cfg.write_text("[control]\n nsteps = 3\n[discretization]\n fname = a.nc\n")
    rc, out, _ = run("--print-config", str(cfg))
    assert rc == 1
    assert "error histogram.nbins" in out


def test_inspect_classic(tmp_path):
    codes = (np.arange(3 * 2 * 7, dtype=np.uint8) % 5).reshape(3, 2, 7)
    ext = np.arange(12, dtype=np.float64).reshape(3, 4) * 0.5
    p = tmp_path / "bin_bonds.nc"
    cdf.write_dataset(str(p), [("dims", 3), ("extremes", 4), ("bin_scheme_count", 2), ("frames", 7)],
                      [cdf.VarSpec("extreme", ("dims", "extremes"), ext),
                       cdf.VarSpec("coord", ("dims", "bin_scheme_count", "frames"), codes)],
                      attrs={"convention": "CENTRE_Real", "content": "B"}, version=5, unlimited="dims")
    rc, out, err = run("--inspect", str(p), "coord", "extreme", "nope")
    assert rc == 0, err
    lines = out.splitlines()
    assert lines[0] == "format classic"
    assert "attr convention CENTRE_Real" in lines and "attr content B" in lines
    assert f"var coord shape 3,2,7 sum {int(codes.sum())}" in lines
    assert "var extreme shape 3,4 sum 33" in lines
    assert "var nope absent" in lines


def test_inspect_netcdf4_reference_file(golden):
    if not os.path.exists(NC4):
        pytest.skip("fixture missing")
    rc, out, err = run("--inspect", NC4, "id", "contri_S")
    if rc != 0 and "needs libhdf5" in err:
        pytest.skip("no libhdf5 on this machine")
    assert rc == 0, err
    d = {l.split(" shape")[0]: l for l in out.splitlines() if l.startswith("var ")}
    assert out.splitlines()[0] == "format netcdf4-hdf5"
    assert "attr convention EntropyCont" in out and "attr content A" in out
    ids, S = golden["ex1_id_ang1d"], golden["ex1_S_ang1d"]
    assert d["var id"].startswith(f"var id shape {ids.shape[0]},1 sum {int(ids.sum())}")
    line = d["var contri_S"]
    assert line.startswith(f"var contri_S shape {S.shape[0]},1,1,{S.shape[1]} sum ")
    assert float(line.split(" sum ")[1]) == pytest.approx(float(S.sum()), rel=1e-13)


def test_refuses_non_netcdf(tmp_path):
    p = tmp_path / "x.nc"
    p.write_bytes(b"hello world, not netcdf")
    rc, _, err = run("--inspect", str(p))
    assert rc == 2 and "not a NetCDF file" in err


@pytest.mark.parametrize("start,stride,nframe,F,want", [
    (0, 1, 0, 1000, (0, 1, 1000)),          # defaults: every frame
    (10, 3, 0, 1000, (10, 3, 1000)),
    (1000, 2, 0, 1000, (0, 2, 1000)),       # start at the end reads from 0
    (50, 5000, 0, 1000, (50, 1, 1000)),     # stride past the end reads every frame
    (700, 2, 600, 1000, (0, 2, 600)),       # start past nframe
    (0, 1, 5000, 1000, (0, 1, 1000)),       # nframe beyond the file
])
def test_frame_range_clamps(tmp_path, start, stride, nframe, F, want):
    """The frame selection follows the reference's reader clamps
    (Netcdf_TrjInt.cpp:120-121) and nframe default (Inputs.cpp:632-634)."""
    cfg = tmp_path / "in.cfg"
    cfg.write_text(f"[control]\n nsteps = 3\n[discretization]\n fname = a.nc\n"
                   f"[histogram]\n referencenbins = 30\n nbins = 30\n"
                   f"[entropy]\n startframe = {start}\n strideframe = {stride}\n"
                   f" nframe = {nframe}\n")
    rc, out, err = run("--print-config", str(cfg), str(F))
    assert rc == 0, out + err
    d = kv(out)
    assert (int(d["startframe"]), int(d["strideframe"]), int(d["nframe"])) == want


def test_zero_frame_stride_is_refused(tmp_path):
    cfg = tmp_path / "in.cfg"
    cfg.write_text("[control]\n nsteps = 3\n[discretization]\n fname = a.nc\n"
                   "[histogram]\n referencenbins = 30\n nbins = 30\n"
                   "[entropy]\n strideframe = 0\n")
    rc, out, _ = run("--print-config", str(cfg), "1000")
    assert rc == 1
    assert "error entropy.strideframe must be at least 1" in out


# ---- NetCDF-4 entropy-contribution writer (no GPU: the host's --convert) ----

import h5read  # noqa: E402

H5OK = h5read.lib() is not None and os.path.exists(h5read.H5DUMP)
needs_h5 = pytest.mark.skipif(not H5OK, reason="libhdf5 or h5dump absent")


@needs_h5
def test_netcdf4_writer_reproduces_the_reference_file(tmp_path):
    """The reference's stored example-1 entcontri_ang-1d.nc (NetCDF-4 from
    netcdf-c), rewritten by the host's NetCDF-4 writer: the h5dump header
    (creation order of every object and attribute, dimension scales and
    their links, chunking, fill values, attribute types and values) is the
    same apart from the _NCProperties provenance text, and so is the data."""
    dst = str(tmp_path / "a.nc")
    rc, _, err = run("--convert", NC4, dst)
    assert rc == 0, err
    assert h5read.header(dst) == h5read.header(NC4)
    with h5read.File(NC4) as a, h5read.File(dst) as b:
        for v in ("id", "contri_S"):
            np.testing.assert_array_equal(a.read(v), b.read(v))
        assert b.attr("_NCProperties").startswith("version=2,")


def entcontri_cdf5(path, ids, S, content, upper, dim_lens):
    """An entropy-contribution file in CDF-5 with the reference's dims,
    attribute order and variables (one estimator, one 30-bin scheme)."""
    n, steps = S.shape
    cdf.write_dataset(
        path,
        [("records", n), ("estimators", 1), ("nbin_schemes", 1), ("nsteps", steps),
         ("ids", ids.shape[1])],
        [cdf.VarSpec("id", ("records", "ids"), ids.astype(np.uint32)),
         cdf.VarSpec("contri_S", ("records", "estimators", "nbin_schemes", "nsteps"),
                     S[:, None, None, :])],
        attrs={"title": np.array([1 if upper else 0], np.uint8), "application": "CENTRE",
               "program": "Real2Int", "programVersion": "1.0.0", "convention": "EntropyCont",
               "conventionVersion": "1.0", "content": content,
               "order": np.array([ids.shape[1]], np.uint8), "nsteps": np.array([steps], np.uint8),
               "storagetype": np.array([3 if upper else 0], np.uint8),
               "ndims": np.array([ids.shape[1]], np.uint8),
               "dimLengths": np.array(dim_lens, np.uint64),
               "nbinSchemes": np.array([1], np.uint8), "binSchemes": np.array([30], np.uint8)},
        version=5, unlimited="records")


@needs_h5
@pytest.mark.parametrize("tag,content,upper,lens", [("ang", "A/A", True, ("ang", "ang")),
                                                     ("ba", "B/A", False, ("bnd", "ang"))])
def test_netcdf4_writer_2d_layout_matches_reference(tmp_path, golden, tag, content, upper, lens):
    """Example-1's 2-D records (same-type UPPER and cross-type FULL) written
    through the host's NetCDF-4 writer give the h5dump header of the
    reference's stored entcontri_{ang,ba}-2d.nc, and the same values."""
    ids, S = golden[f"ex1_id_{tag}2d"], golden[f"ex1_S_{tag}2d"]
    dims = [len(golden[f"ex1_id_{t}1d"]) for t in lens]
    src, dst = str(tmp_path / "in.nc"), str(tmp_path / "out.nc")
    entcontri_cdf5(src, ids, S, content, upper, dims)
    rc, _, err = run("--convert", src, dst, "netcdf4")
    assert rc == 0, err
    with open(os.path.join(ROOT, "tests", "golden", f"ex1h_entcontri_{tag}-2d_header.txt")) as f:
        assert h5read.header(dst) == f.read().rstrip("\n")
    with h5read.File(dst) as b:
        np.testing.assert_array_equal(b.read("id"), ids)
        np.testing.assert_array_equal(b.read("contri_S")[:, 0, 0, :], S)


@needs_h5
def test_cdf5_output_of_a_netcdf4_file(tmp_path):
    """mi355x.entcontriformat = cdf5 keeps the same dims, attributes and
    values in the classic format."""
    dst = str(tmp_path / "a.nc")
    rc, _, err = run("--convert", NC4, dst, "cdf5")
    assert rc == 0, err
    ds = cdf.open_dataset(dst)
    assert ds.version == 5 and ds.unlimited == "records"
    assert list(ds.dims) == ["records", "estimators", "nbin_schemes", "nsteps", "ids"]
    assert list(ds.attrs)[:3] == ["title", "application", "program"]
    with h5read.File(NC4) as a:
        np.testing.assert_array_equal(ds.read("contri_S"), a.read("contri_S"))
        np.testing.assert_array_equal(ds.read("id"), a.read("id"))
        for k in ("order", "storagetype", "dimLengths", "binSchemes", "nsteps"):
            np.testing.assert_array_equal(np.asarray(ds.attrs[k]), a.attr(k), err_msg=k)


def test_convert_refuses_other_files(tmp_path):
    p = tmp_path / "b.nc"
    cdf.write_dataset(str(p), [("dims", 2)], [cdf.VarSpec("dimid", ("dims",), np.arange(2, dtype=np.uint32))],
                      attrs={"convention": "CENTRE_Real"}, version=5)
    rc, _, err = run("--convert", str(p), str(tmp_path / "o.nc"))
    assert rc == 2 and "not an entropy-contribution file" in err


def test_entcontri_format_option(tmp_path):
    cfg = tmp_path / "in.cfg"
    base = ("[control]\n nsteps = 3\n[discretization]\n fname = a.nc\n"
            "[histogram]\n referencenbins = 30\n nbins = 30\n[mi355x]\n")
    cfg.write_text(base + " entcontriformat = hdf9\n")
    rc, out, _ = run("--print-config", str(cfg))
    assert rc == 1 and "error mi355x.entcontriformat must be netcdf4 or cdf5" in out
