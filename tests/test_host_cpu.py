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
