"""Extract the reference's stored example outputs into small numpy fixtures.

Run once, in the build container only (it reads the reference tree and uses
the HDF5 command-line dumper, neither of which exists where the GPU tests
run).  The fixtures it writes are data: the reference's own stored inputs
and outputs for ``test_examples/example-{1,2}``:

* ``bin_{bonds,angles,torsions}.nc`` (CDF-5): uint8 codes + extrema, read
  with this project's own classic-format reader;
* ``entcontri_*.nc`` (NetCDF-4/HDF5): ``id`` and ``contri_S``, dumped with
  ``h5dump -m %.17g`` (17 significant digits round-trip a double exactly);
* ``hist_tor-1d.nc`` (CDF-5): uint64 counts, extrema and bin mids;
* the convergence CSVs, the MIST tree CSVs and ``vmkde_torsions_minimas.out``;
* one stored NetCDF-4 file copied as is (``ex1h_entcontri_ang-1d.nc``, 18 KB),
  the format fixture of the host's NetCDF-4 reader;
* the header of example-2's ``hist_tor-1d.nc`` and of example-1's
  ``bin_bonds.nc`` as JSON, and the h5dump headers of example-1's
  ``entcontri_{ang,ba}-2d.nc``: the format fixtures of the host's writers.

Usage:  python tests/golden/make_golden.py [/root/reference]
"""
from __future__ import annotations

import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from centre_amd.cdf import open_dataset  # noqa: E402

H5DUMP = "/opt/conda/bin/h5dump"
TYPES_1D = ["bnd", "ang", "tor"]
TYPES_2D = ["bnd", "ang", "tor", "ba", "bd", "ad"]


def h5_dataset(path: str, name: str) -> np.ndarray:
    """Dump one dataset with h5dump and parse its values and shape."""
    out = subprocess.run([H5DUMP, "-m", "%.17g", "-y", "-d", name, "-w", "0", path],
                         check=True, capture_output=True, text=True).stdout
    m = re.search(r"DATASPACE\s+SIMPLE\s*\{\s*\(\s*([0-9, ]+)\)", out)
    shape = tuple(int(x) for x in m.group(1).split(","))
    body = out.split("DATA {", 1)[1]
    body = body.split("}", 1)[0]
    vals = re.findall(r"[-+]?(?:\d+\.?\d*(?:[eE][-+]?\d+)?|nan|inf)", body)
    arr = np.array([float(v) for v in vals])
    if arr.size != int(np.prod(shape)):
        raise RuntimeError(f"{path}:{name}: parsed {arr.size} values, "
                           f"shape {shape}")
    return arr.reshape(shape)


def read_csv_table(path: str):
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines() if ln.strip()]
    header = [h.strip() for h in lines[0].split(",")]
    rows = [[c.strip() for c in ln.split(",")] for ln in lines[1:]]
    return header, rows


def main(ref: str) -> None:
    ex1 = os.path.join(ref, "test_examples", "example-1")
    ex2 = os.path.join(ref, "test_examples", "example-2")
    out = {}

    # ---- example-1 bin codes (histogram run) -------------------------------
    for tag, fname in (("b", "bin_bonds.nc"), ("a", "bin_angles.nc"),
                       ("t", "bin_torsions.nc")):
        ds = open_dataset(os.path.join(ex1, "outfiles-h-ref", fname))
        out[f"ex1_codes_{tag}"] = ds.read("coord")[:, 0, :].astype(np.uint8)
        out[f"ex1_extreme_{tag}"] = ds.read("extreme").astype(np.float64)
        out[f"ex1_dimid_{tag}"] = ds.read("dimid").astype(np.uint32)
        out[f"ex1_binschemes_{tag}"] = np.asarray(ds.attrs["binSchemes"],
                                                  dtype=np.uint8)

    # ---- entropy contributions (both examples, histogram runs) --------------
    for exn, exdir in (("ex1", ex1), ("ex2", ex2)):
        d = os.path.join(exdir, "outfiles-h-ref")
        for t in TYPES_1D:
            p = os.path.join(d, f"entcontri_{t}-1d.nc")
            out[f"{exn}_id_{t}1d"] = h5_dataset(p, "/id").astype(np.uint32)
            out[f"{exn}_S_{t}1d"] = h5_dataset(p, "/contri_S")[:, 0, 0, :]
        for t in TYPES_2D:
            p = os.path.join(d, f"entcontri_{t}-2d.nc")
            out[f"{exn}_id_{t}2d"] = h5_dataset(p, "/id").astype(np.uint32)
            out[f"{exn}_S_{t}2d"] = h5_dataset(p, "/contri_S")[:, 0, 0, :]

    # ---- example-2 stored 1-D torsion counts ---------------------------------
    ds = open_dataset(os.path.join(ex2, "outfiles-h-ref", "hist_tor-1d.nc"))
    out["ex2_hist_id"] = ds.read("id")[:, 0].astype(np.uint32)
    out["ex2_hist_freq"] = ds.read("freq")[:, 0, :].astype(np.uint64)
    out["ex2_hist_extrema"] = ds.read("extrema")[:, 0, :].astype(np.float64)
    out["ex2_hist_mid"] = ds.read("mid").astype(np.float64)
    out["ex2_hist_firststep"] = np.asarray(ds.attrs["firstStep"], np.uint8)

    # ---- CSV outputs: convergence tables and MIST trees -----------------------
    for exn, exdir in (("ex1", ex1), ("ex2", ex2)):
        for run, b in (("h", 30), ("v", 3)):
            d = os.path.join(exdir, f"outfiles-{run}-ref")
            for meth in ("MIE", "MIST"):
                hdr, rows = read_csv_table(
                    os.path.join(d, f"{meth}-ML_convergence_b-{b}.csv"))
                out[f"{exn}{run}_{meth}_csv"] = np.array(
                    [[float(c) for c in r] for r in rows])
                out[f"{exn}{run}_{meth}_hdr"] = np.array(hdr)
            for s in range(4):
                hdr, rows = read_csv_table(
                    os.path.join(d, f"Tree-MIST-ML_b-{b}_s-{s}.csv"))
                tcode = {"B": 0, "A": 1, "T": 2}
                out[f"{exn}{run}_tree{s}"] = np.array(
                    [[tcode[r[0]], tcode[r[1]], int(r[2]), int(r[3]),
                      float(r[4])] for r in rows])
            # vmKDE minima (only the v runs, and example-2 h run, list them)
            mp = os.path.join(d, "vmkde_torsions_minimas.out")
            if os.path.getsize(mp) > 0:
                recs = []
                with open(mp) as f:
                    for ln in f:
                        ln = ln.strip().rstrip(";").split()
                        if not ln:
                            continue
                        tid, n = int(ln[1]), int(ln[2])
                        mins = [float(x) for x in ln[3:3 + n]]
                        recs.append([tid, n] + mins + [np.nan] * (9 - n))
                out[f"{exn}{run}_minima"] = np.array(recs)

    dst = os.path.join(HERE, "reference_examples.npz")
    np.savez_compressed(dst, **out)
    import shutil
    shutil.copyfile(os.path.join(ex1, "outfiles-h-ref", "entcontri_ang-1d.nc"),
                    os.path.join(HERE, "ex1h_entcontri_ang-1d.nc"))
    print(f"wrote {dst} ({os.path.getsize(dst) / 1e6:.2f} MB, "
          f"{len(out)} arrays)")


def write_classic_header(src: str, dst: str) -> None:
    """Header of a stored classic-format file (dimension order and sizes,
    attribute order, types and values, variable dimensions and types) as
    JSON."""
    import json
    ds = open_dataset(src)
    attrs = []
    for k, v in ds.attrs.items():
        if isinstance(v, str):
            attrs.append([k, "text", v])
        else:
            a = np.asarray(v)
            attrs.append([k, a.dtype.name, a.tolist()])
    hdr = {"version": ds.version, "unlimited": ds.unlimited,
           "dims": [[k, n] for k, n in ds.dims.items()], "attrs": attrs,
           "vars": [[v.name, list(v.dims), v.nc_type] for v in ds.variables.values()]}
    with open(dst, "w") as f:
        json.dump(hdr, f, indent=1)
        f.write("\n")


def write_headers(ref: str) -> None:
    """Format fixtures of the host's writers:

    * example-2's ``hist_tor-1d.nc`` (CDF-5 count file) and example-1's
      ``bin_bonds.nc`` (CDF-5 bin-code file) headers as JSON;
    * the h5dump headers (creation order, storage properties, attribute
      values; object addresses and the _NCProperties text taken out) of
      example-1's ``entcontri_ang-2d.nc`` (same-type pairs, UPPER) and
      ``entcontri_ba-2d.nc`` (cross-type pairs), both NetCDF-4.
    """
    sys.path.insert(0, os.path.dirname(HERE))
    import h5read
    ex1 = os.path.join(ref, "test_examples", "example-1", "outfiles-h-ref")
    write_classic_header(os.path.join(ref, "test_examples", "example-2", "outfiles-h-ref",
                                      "hist_tor-1d.nc"),
                         os.path.join(HERE, "ex2h_hist_tor-1d_header.json"))
    write_classic_header(os.path.join(ex1, "bin_bonds.nc"),
                         os.path.join(HERE, "ex1h_bin_bonds_header.json"))
    for tag in ("ang", "ba"):
        with open(os.path.join(HERE, f"ex1h_entcontri_{tag}-2d_header.txt"), "w") as f:
            f.write(h5read.header(os.path.join(ex1, f"entcontri_{tag}-2d.nc")) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    write_headers(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
