"""Host-side checks that need no GPU: the C-ABI library loads and exports
every function include/centre_hip.h declares; the classic-format reader and
writer round-trip; the synthetic generator is deterministic."""
import os
import re

import numpy as np
import pytest

from centre_amd import cdf
from centre_amd import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "centre_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(centre_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = hip.lib()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(hip.exported_symbols())


def test_no_gpu_init_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(hip.CentreError):
        hip.Context(0)


def test_synth_deterministic_across_threads():
    a = hip.synth_bat(123, 3, 4, 5, 777, n_threads=1)
    b = hip.synth_bat(123, 3, 4, 5, 777, n_threads=7)
    assert np.array_equal(a, b)
    c = hip.synth_bat(124, 3, 4, 5, 777)
    assert not np.array_equal(a, c)
    t = a[:, 7:]
    assert t.min() >= 0 and (t.astype(np.float32).astype(np.float64) < 2 * np.pi).all()
    assert abs(a[:, :3].mean() - 1.5) < 0.01 and abs(a[:, 3:7].mean() - 1.91) < 0.02


def test_cdf5_roundtrip(tmp_path):
    p = str(tmp_path / "x.nc")
    codes = np.arange(3 * 2 * 10, dtype=np.uint8).reshape(3, 2, 10)
    ext = np.arange(12, dtype=np.float64).reshape(3, 4)
    ids = np.array([5, 6, 7], np.uint32)
    cdf.write_dataset(
        p, [("dims", 3), ("extremes", 4), ("bin_scheme_count", 2), ("frames", 10)],
        [cdf.VarSpec("dimid", ("dims",), ids),
         cdf.VarSpec("extreme", ("dims", "extremes"), ext),
         cdf.VarSpec("coord", ("dims", "bin_scheme_count", "frames"), codes)],
        attrs={"convention": "CENTRE_Real", "binSchemes": np.array([30, 40], np.uint8)},
        version=5, unlimited="dims")
    ds = cdf.open_dataset(p)
    assert ds.version == 5 and ds.numrecs == 3
    assert np.array_equal(ds.read("coord"), codes)
    assert np.array_equal(ds.read("extreme"), ext)
    assert np.array_equal(ds.read("dimid"), ids)
    assert ds.attrs["convention"] == "CENTRE_Real"
    assert list(ds.attrs["binSchemes"]) == [30, 40]
    assert np.array_equal(ds.read("coord", [1, 0, 2], [2, 1, 5]), codes[1:3, 0:1, 2:7])


@pytest.mark.parametrize("version", [1, 2, 5])
def test_classic_versions_single_record_var(tmp_path, version):
    p = str(tmp_path / "b.nc")
    bat = np.random.default_rng(0).normal(size=(7, 3))
    cdf.write_dataset(p, [("frame", 7), ("bond", 3)],
                      [cdf.VarSpec("bond", ("frame", "bond"), bat)],
                      attrs={"convention": "CENTRE"}, version=version,
                      unlimited="frame")
    ds = cdf.open_dataset(p)
    assert np.array_equal(ds.read("bond"), bat)
