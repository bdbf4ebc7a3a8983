"""ctypes binding of the C oracle (``centre_oracle.c``) with numpy helpers.

TEST INFRASTRUCTURE ONLY (see the package docstring).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BOND, ANGLE, DIHEDRAL = 1, 2, 3
ML, MM, CS, JS = 0, 1, 2, 3

_lib = None


# timing-only builds of the same source (CPU baseline of the benchmark)
VARIANTS = {"O2": LIB_PATH,
            "O3": os.path.join(HERE, "liboracle_O3.so"),
            "Ofast": os.path.join(HERE, "liboracle_Ofast.so")}
_variants = {}


def build(force: bool = False) -> str:
    """Compile the oracle shared libraries in place."""
    src = os.path.join(HERE, "centre_oracle.c")
    if force or any(not os.path.exists(p) or os.path.getmtime(p) < os.path.getmtime(src)
                    for p in VARIANTS.values()):
        subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", HERE], check=True)
    return LIB_PATH


def lib(variant: str = "O2"):
    """The oracle library; variant "O3" / "Ofast" selects a timing-only build."""
    global _lib
    if variant != "O2":
        if variant not in _variants:
            if not os.path.exists(VARIANTS[variant]):
                build()
            L = C.CDLL(VARIANTS[variant])
            _declare(L)
            _variants[variant] = L
        return _variants[variant]
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = C.c_void_p
U8, U32, U64, I32, F64 = C.c_uint8, C.c_uint32, C.c_uint64, C.c_int, C.c_double


def _declare(L):
    sig = {
        "or_shuffle": [P, P, U64],
        "or_bnd_real2int": [P, U64, P, I32, P, P],
        "or_ang_real2int": [P, U64, P, I32, P, P],
        "or_dih_real2int": [P, U64, P, I32, U8, I32, P, P],
        "or_bin_data_1d": [U8, P, I32, U64, F64, F64, P, P, P],
        "or_bin_data_2d": [U8, P, I32, U64, F64, F64, F64, F64, P, P, P, P],
        "or_entropy_1d": [I32, P, I32, U8, U64, F64, F64, I32, I32, P, I32, P, P],
        "or_entropy_2d": [I32, I32, P, I32, U8, U64, F64, F64, F64, F64, I32,
                          I32, P, I32, P, P],
        "or_dof_entropies": [P, U32, U64, P, P, U8, P, I32, P, I32, I32, I32,
                             P, I32],
        "or_pair_entropies": [P, U32, U64, P, P, P, U64, U8, P, I32, P, I32,
                              I32, I32, P, I32],
        "or_pair_counts": [P, U32, U64, I32, P, U64, U8, P],
        "or_boruvka": [U32, P, P, P, U64, P, P],
        "or_vmkde_grid": [P, U64, F64, U32, P],
        "or_vmkde_states": [P, U64, U8, U32, F64, U32, U32, P, P, P, P],
        "or_vmkde": [P, U64, U8, F64, U32, F64, U32, U32, P, P, P],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    for name, args in {"or_bessel_i": [I32, F64],
                       "or_vmkde_bandwidth": [U64, F64],
                       "or_dih_gap_shift": [P, U64, I32]}.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_double


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ---- discretizer -----------------------------------------------------------
def real2int(x, dtype: int, schemes, nbins_ref: int = 0, optimize: bool = True):
    """One DOF column (float32) -> (codes[nsch, F] uint8, extrema[4])."""
    x = _c(x, np.float32)
    sch = _c(np.atleast_1d(schemes), np.uint8)
    codes = np.zeros((len(sch), x.size), np.uint8)
    ext = np.zeros(4)
    L = lib()
    if dtype == BOND:
        r = L.or_bnd_real2int(_p(x), x.size, _p(sch), len(sch), _p(codes), _p(ext))
    elif dtype == ANGLE:
        r = L.or_ang_real2int(_p(x), x.size, _p(sch), len(sch), _p(codes), _p(ext))
    else:
        r = L.or_dih_real2int(_p(x), x.size, _p(sch), len(sch),
                              nbins_ref or int(sch[0]), int(optimize),
                              _p(codes), _p(ext))
    if r != 0:
        raise ValueError(f"oracle real2int failed ({r})")
    return codes, ext


def dih_gap_shift(x, m: int) -> float:
    x = _c(x, np.float32)
    return lib().or_dih_gap_shift(_p(x), x.size, m)


def bessel_i(n: int, x: float) -> float:
    return lib().or_bessel_i(n, x)


def vmkde_bandwidth(nframes: int, kappa: float) -> float:
    return lib().or_vmkde_bandwidth(nframes, kappa)


def vmkde_grid(x, kappa: float, grid_intervals: int = 360) -> np.ndarray:
    x = _c(x, np.float32)
    out = np.zeros(2 * (grid_intervals + 1))
    lib().or_vmkde_grid(_p(x), x.size, kappa, grid_intervals, _p(out))
    return out


def vmkde_states(x, grid, nmaxconf=5, sdo_step=5, conv=1e-4, max_iter=1000,
                 grid_intervals=360):
    x = _c(x, np.float32)
    grid = _c(grid, np.float64)
    codes = np.zeros(x.size, np.uint8)
    mins = np.zeros(9)
    nf = np.zeros(1, np.uint8)
    r = lib().or_vmkde_states(_p(x), x.size, nmaxconf, sdo_step, conv,
                              max_iter, grid_intervals, _p(grid), _p(mins),
                              _p(nf), _p(codes))
    if r != 0:
        raise ValueError(f"oracle vmkde_states failed ({r})")
    return codes, mins[:nf[0]], int(nf[0])


def vmkde(x, nmaxconf=5, kappa=1.0, sdo_step=5, conv=1e-4, max_iter=1000,
          grid_intervals=360):
    grid = vmkde_grid(x, kappa, grid_intervals)
    return vmkde_states(x, grid, nmaxconf, sdo_step, conv, max_iter,
                        grid_intervals)


# ---- counts and entropies ----------------------------------------------------
def bin_data_1d(codes, nsteps, schemes, mn, mx):
    codes = _c(codes, np.uint8)
    sch = _c(np.atleast_1d(schemes), np.uint8)
    F = codes.size // len(sch)
    freqs = np.zeros(nsteps * int(sch.astype(int).sum()), np.uint64)
    mids = np.zeros(int(sch.astype(int).sum()))
    lib().or_bin_data_1d(nsteps, _p(sch), len(sch), F, mn, mx, _p(codes),
                         _p(freqs), _p(mids))
    return freqs, mids


def bin_data_2d(c1, c2, nsteps, schemes, mn1, mx1, mn2, mx2):
    c1 = _c(c1, np.uint8)
    c2 = _c(c2, np.uint8)
    sch = _c(np.atleast_1d(schemes), np.uint8)
    F = c1.size // len(sch)
    bsq = int((sch.astype(int) ** 2).sum())
    freqs = np.zeros(nsteps * bsq, np.uint64)
    mids = np.zeros(2 * int(sch.astype(int).sum()))
    lib().or_bin_data_2d(nsteps, _p(sch), len(sch), F, mn1, mx1, mn2, mx2,
                         _p(c1), _p(c2), _p(freqs), _p(mids))
    return freqs, mids


def entropy_1d(dtype, ests, nsteps, step_size, mn, mx, jacobian, kde,
               schemes, freqs):
    ests = _c(np.atleast_1d(ests), np.int32)
    sch = _c(np.atleast_1d(schemes), np.uint8)
    freqs = _c(freqs, np.uint64)
    out = np.zeros(len(ests) * len(sch) * nsteps)
    lib().or_entropy_1d(dtype, _p(ests), len(ests), nsteps, step_size, mn, mx,
                        int(jacobian), int(kde), _p(sch), len(sch),
                        _p(freqs), _p(out))
    return out


def entropy_2d(t1, t2, ests, nsteps, step_size, mn1, mx1, mn2, mx2, jacobian,
               kde, schemes, freqs):
    ests = _c(np.atleast_1d(ests), np.int32)
    sch = _c(np.atleast_1d(schemes), np.uint8)
    freqs = _c(freqs, np.uint64)
    out = np.zeros(len(ests) * len(sch) * nsteps)
    lib().or_entropy_2d(t1, t2, _p(ests), len(ests), nsteps, step_size, mn1,
                        mx1, mn2, mx2, int(jacobian), int(kde), _p(sch),
                        len(sch), _p(freqs), _p(out))
    return out


def _check_steps(nsteps):
    # the C side takes nsteps as uint8_t, the reference's hbin_t
    # (configcentre.h:24, Estimators.h:54); 256 would wrap to 0
    if not 1 <= int(nsteps) <= 255:
        raise ValueError("nsteps must be in 1..255")


def dof_entropies(codes, types, ext, nsteps, schemes=(30,), ests=(ML,),
                  jacobian=True, kde=False, nthreads=0):
    """codes [D, nsch, F] (or [D, F]) -> S1 [D, nest, nsch, nsteps]."""
    _check_steps(nsteps)
    sch = _c(np.atleast_1d(schemes), np.uint8)
    codes = _c(codes, np.uint8)
    D = codes.shape[0]
    F = codes.size // (D * len(sch))
    ests = _c(np.atleast_1d(ests), np.int32)
    types = _c(types, np.uint8)
    ext = _c(ext, np.float64)
    out = np.zeros((D, len(ests), len(sch), nsteps))
    lib().or_dof_entropies(_p(codes), D, F, _p(types), _p(ext), nsteps,
                           _p(sch), len(sch), _p(ests), len(ests),
                           int(jacobian), int(kde), _p(out), nthreads)
    return out


def pair_entropies(codes, types, ext, pairs, nsteps, schemes=(30,),
                   ests=(ML,), jacobian=True, kde=False, nthreads=0, variant="O2"):
    """codes [D, nsch, F] -> S2 [npair, nest, nsch, nsteps].  variant picks
    the build ("O3" / "Ofast" for timing only)."""
    _check_steps(nsteps)
    sch = _c(np.atleast_1d(schemes), np.uint8)
    codes = _c(codes, np.uint8)
    D = codes.shape[0]
    F = codes.size // (D * len(sch))
    ests = _c(np.atleast_1d(ests), np.int32)
    pairs = _c(pairs, np.uint32).reshape(-1, 2)
    types = _c(types, np.uint8)
    ext = _c(ext, np.float64)
    out = np.zeros((len(pairs), len(ests), len(sch), nsteps))
    r = lib(variant).or_pair_entropies(_p(codes), D, F, _p(types), _p(ext),
                                _p(pairs), len(pairs), nsteps, _p(sch),
                                len(sch), _p(ests), len(ests), int(jacobian),
                                int(kde), _p(out), nthreads)
    if r != 0:
        raise ValueError("pair index out of range")
    return out


def pair_counts(codes, nb, pairs, nsteps):
    _check_steps(nsteps)
    codes = _c(codes, np.uint8)
    D = codes.shape[0]
    F = codes.shape[-1]
    pairs = _c(pairs, np.uint32).reshape(-1, 2)
    out = np.zeros((len(pairs), nsteps, nb, nb), np.uint64)
    lib().or_pair_counts(_p(codes), D, F, nb, _p(pairs), len(pairs), nsteps,
                         _p(out))
    return out


def boruvka(n_nodes, src, dst, w):
    src = _c(src, np.uint32)
    dst = _c(dst, np.uint32)
    w = _c(w, np.float64)
    tree = np.zeros(max(n_nodes, 1), np.uint32)
    nt = np.zeros(1, np.uint32)
    r = lib().or_boruvka(n_nodes, _p(src), _p(dst), _p(w), len(w), _p(tree),
                         _p(nt))
    if r != 0:
        raise ValueError("edge endpoint out of range")
    return tree[:nt[0]]


def shuffle(rand_idx, data):
    rand_idx = _c(rand_idx, np.uint64)
    data = np.array(data, dtype=np.uint8, copy=True)
    lib().or_shuffle(_p(rand_idx), _p(data), data.size)
    return data
