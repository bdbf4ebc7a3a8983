"""ctypes binding of ``libcentre_hip.so`` (the C ABI in include/centre_hip.h).

The library is the only compute path: if it is missing or the GPU is not
usable, every entry point raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcentre_hip.so")

ML, MM, CS, JS = 0, 1, 2, 3
BOND, ANGLE, DIHEDRAL = 1, 2, 3

P = C.c_void_p
U8, U32, U64, I32, F64 = C.c_uint8, C.c_uint32, C.c_uint64, C.c_int, C.c_double

_SIGS = {
    "centre_init": ([I32, P], I32),
    "centre_destroy": ([P], None),
    "centre_last_error": ([P], C.c_char_p),
    "centre_sync": ([P], I32),
    "centre_stream": ([P], P),
    "centre_last_pair_kernel_ms": ([P], C.c_float),
    "centre_set_mfma_format": ([P, I32], I32),
    "centre_batch_pair_entropy": ([P, P, I32, I32, I32, P, I32, I32, I32, I32, I32, P, P, P, P],
                                  I32),
    "centre_screen_n_params": ([], I32),
    "centre_screen_tiles": ([P, P, U32, I32, C.c_float, I32, P, P], I32),
    "centre_pair_plan_from_tiles": ([P, U32, I32, P, U64, P], I32),
    "centre_pair_dofs_per_tile": ([I32], I32),
    "centre_screen_features": ([P, P, U64, U32, U64, P, P, I32, U32, P], I32),
    "centre_mlp_forward": ([P, P, P, U64, P], I32),
    "centre_mlp_loss_grad": ([P, P, P, P, U64, P, P], I32),
    "centre_mlp_train": ([P, P, P, P, U64, I32, I32, C.c_float, U64, P], I32),
    "centre_dev_discretize_hist": ([P, P, U64, U32, U64, I32, P, I32, I32, I32, P, P], I32),
    "centre_discretize_hist": ([P, P, U64, U32, U64, I32, P, I32, I32, I32, P, P], I32),
    "centre_dev_discretize_vmkde": ([P, P, U64, U32, U64, I32, F64, U32, F64, U32, U32,
                                     P, P, P, P], I32),
    "centre_discretize_vmkde": ([P, P, U64, U32, U64, I32, F64, U32, F64, U32, U32,
                                 P, P, P, P], I32),
    "centre_dev_dof_entropy": ([P, P, U64, U32, U64, P, P, I32, I32, P, I32, I32, I32,
                                P, P], I32),
    "centre_pair_plan_create": ([P, U32, I32, P, U64, P], I32),
    "centre_pair_plan_destroy": ([P], None),
    "centre_pair_plan_records": ([P], U64),
    "centre_pair_plan_tiles": ([P], I32),
    "centre_dev_pair_entropy": ([P, P, P, U64, U64, P, P, I32, P, I32, I32, I32, P, P, P,
                                 P], I32),
    "centre_dev_pair_entropy_schemes": ([P, P, I32, P, U64, U64, P, P, I32, P, I32, I32, I32,
                                         P, P, P], I32),
    "centre_dev_pair_entropy_noalloc": ([P, P, P, U64, U64, P, P, I32, P, I32, I32, I32, P,
                                         P, P], I32),
    "centre_pair_entropy": ([P, P, U32, U64, P, P, I32, I32, P, U64, P, I32, I32, I32,
                             P, P, P], I32),
    "centre_dev_mst": ([P, P, P, P, U64, U32, P, P], I32),
    "centre_mst": ([P, P, P, P, U64, U32, P, P], I32),
    "centre_dev_segment_sum": ([P, P, P, I32, P], I32),
    "centre_synth_bat": ([U64, U32, U32, U32, U64, P, I32], I32),
    "centre_dev_alloc": ([P, C.c_size_t, P], I32),
    "centre_dev_free": ([P, P], I32),
    "centre_memcpy_h2d": ([P, P, P, C.c_size_t], I32),
    "centre_memcpy_d2h": ([P, P, P, C.c_size_t], I32),
    "centre_memcpy_d2d": ([P, P, P, C.c_size_t], I32),
    "centre_memcpy_d2d_2d": ([P, P, C.c_size_t, P, C.c_size_t, C.c_size_t, C.c_size_t], I32),
    "centre_device_sync": ([P], I32),
}

_lib = None


def exported_symbols() -> Sequence[str]:
    return list(_SIGS)


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with centre_amd/build.py "
                f"(the HIP library is the only compute path)")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


class BatchSystem(C.Structure):
    """centre_batch_system (device pointers as integers)."""
    _fields_ = [("d_codes", C.c_void_p), ("code_stride", C.c_uint64), ("n_dof", C.c_uint32),
                ("n_frame", C.c_uint64), ("d_types", C.c_void_p), ("d_extrema", C.c_void_p),
                ("d_s1", C.c_void_p), ("d_s2", C.c_void_p), ("d_mi", C.c_void_p)]


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def synth_bat(seed: int, n_bond: int, n_angle: int, n_dih: int,
              n_frame: int, n_threads: int = 0) -> np.ndarray:
    """Seeded synthetic BAT trajectory [n_frame, D] float64 (host)."""
    out = np.empty((n_frame, n_bond + n_angle + n_dih), np.float64)
    r = lib().centre_synth_bat(seed, n_bond, n_angle, n_dih, n_frame,
                               _p(out), n_threads)
    if r != 0:
        raise CentreError(r, "centre_synth_bat failed")
    return out


class CentreError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"centre error {code}: {msg}")
        self.code = code


class Context:
    """One GPU, one HIP stream (centre_init / centre_destroy)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = C.c_void_p()
        r = self._L.centre_init(device, C.byref(h))
        if r != 0:
            raise CentreError(r, "centre_init failed (no usable GPU?)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self._L.centre_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def check(self, r: int):
        if r != 0:
            msg = self._L.centre_last_error(self.h)
            raise CentreError(r, msg.decode() if msg else "")

    @property
    def stream(self) -> int:
        return self._L.centre_stream(self.h) or 0

    def sync(self):
        self.check(self._L.centre_sync(self.h))

    def device_sync(self):
        self.check(self._L.centre_device_sync(self.h))

    # ---- device memory -------------------------------------------------------
    def alloc(self, shape, dtype) -> "DevArray":
        return DevArray(self, shape, dtype)

    def to_device(self, x: np.ndarray) -> "DevArray":
        x = np.ascontiguousarray(x)
        d = DevArray(self, x.shape, x.dtype)
        d.upload(x)
        return d

    def batch_pair_entropy(self, systems, nbins, nsteps, ests=(0,), jacobian=True, kde=False,
                           order=0, tile_dofs=0, shape_used=False):
        """systems: list of BatchSystem -> (per-system ms, total ms, order
        used[, DOFs per tile side used when shape_used])."""
        arr = (BatchSystem * len(systems))(*systems)
        ests = _c(np.atleast_1d(ests), np.int32)
        sys_ms = np.zeros(len(systems), np.float64)
        tot = C.c_double(0.0)
        used = C.c_int(0)
        rused = C.c_int(0)
        self.check(self._L.centre_batch_pair_entropy(
            self.h, C.cast(arr, C.c_void_p), len(systems), nbins, nsteps, _p(ests), len(ests),
            int(jacobian), int(kde), order, tile_dofs, _p(sys_ms), C.byref(tot), C.byref(used),
            C.byref(rused)))
        if shape_used:
            return sys_ms, tot.value, used.value, rused.value
        return sys_ms, tot.value, used.value

    # -- learned pair screen (device pointers) ---------------------------------
    def screen_features(self, d_codes, stride, n_dof, n_frame, d_types, d_ext, nbins,
                        n_sub, d_feat):
        self.check(self._L.centre_screen_features(self.h, d_codes, stride, n_dof, n_frame,
                                                  d_types, d_ext, nbins, n_sub, d_feat))

    def screen_tiles(self, d_pred, n_dof, nbins, cutoff, topk) -> np.ndarray:
        R = int(self._L.centre_pair_dofs_per_tile(nbins))
        nblk = (n_dof + R - 1) // R
        buf = np.zeros((nblk * (nblk + 1) // 2, 2), np.int32)
        n = C.c_uint64(0)
        self.check(self._L.centre_screen_tiles(self.h, d_pred, n_dof, nbins, cutoff, topk, _p(buf),
                                               C.byref(n)))
        return buf[:n.value].copy()

    def mlp_forward(self, d_params, d_feat, n, d_out):
        self.check(self._L.centre_mlp_forward(self.h, d_params, d_feat, n, d_out))

    def mlp_loss_grad(self, d_params, d_feat, d_target, n, d_grad) -> float:
        loss = C.c_double(0.0)
        self.check(self._L.centre_mlp_loss_grad(self.h, d_params, d_feat, d_target, n, d_grad,
                                                C.byref(loss)))
        return loss.value

    def mlp_train(self, d_params, d_feat, d_target, n, epochs, batch, lr, seed) -> np.ndarray:
        hist = np.zeros(epochs, np.float64)
        self.check(self._L.centre_mlp_train(self.h, d_params, d_feat, d_target, n, epochs,
                                            batch, lr, seed, _p(hist)))
        return hist

    def copy_d2d(self, dst: int, src: int, nbytes: int):
        self.check(self._L.centre_memcpy_d2d(self.h, dst, src, nbytes))

    def copy_2d(self, dst: int, dpitch: int, src: int, spitch: int, width: int, height: int):
        self.check(self._L.centre_memcpy_d2d_2d(self.h, dst, dpitch, src, spitch, width, height))

    def last_pair_kernel_ms(self) -> float:
        return float(self._L.centre_last_pair_kernel_ms(self.h))

    def set_mfma_format(self, fmt: int) -> None:
        """Pair-tile operands: 0 = FP4 where the counts stay exact, 1 = int8."""
        self.check(self._L.centre_set_mfma_format(self.h, int(fmt)))

    # ---- host-buffer entry points ------------------------------------------
    def discretize_hist(self, bat: np.ndarray, dof_type: int, schemes,
                        nbins_ref: int = 0, optimize: bool = True):
        """bat [F, D] float64 -> codes [D, nsch, F] uint8, extrema [D, 4]."""
        bat = _c(bat, np.float64)
        F, D = bat.shape
        sch = _c(np.atleast_1d(schemes), np.uint8)
        codes = np.zeros((D, len(sch), F), np.uint8)
        ext = np.zeros((D, 4))
        self.check(self._L.centre_discretize_hist(
            self.h, _p(bat), D, D, F, dof_type, _p(sch), len(sch),
            int(nbins_ref or sch[0]), int(optimize), _p(codes), _p(ext)))
        return codes, ext

    def discretize_vmkde(self, bat: np.ndarray, nmaxconf=5, kappa=1.0,
                         sdo_step=5, sdo_conv=1e-4, sdo_iter=1000,
                         grid_intervals=360, want_grid=False):
        bat = _c(bat, np.float64)
        F, D = bat.shape
        codes = np.zeros((D, F), np.uint8)
        mins = np.zeros((D, 9))
        nf = np.zeros(D, np.uint8)
        grid = np.zeros((D, 2, grid_intervals + 1)) if want_grid else None
        self.check(self._L.centre_discretize_vmkde(
            self.h, _p(bat), D, D, F, nmaxconf, kappa, sdo_step, sdo_conv,
            sdo_iter, grid_intervals, _p(codes), _p(mins), _p(nf), _p(grid)))
        return codes, mins, nf, grid

    def pair_entropy(self, codes: np.ndarray, types, extrema, nbins: int,
                     nsteps: int, pairs=None, ests=(ML,), jacobian=True,
                     kde=False, want_s1=True, want_mi=True):
        """codes [D, F] uint8 -> (S1 [D,nest,S], S2 [P,nest,S], MI [P,S])."""
        codes = _c(codes, np.uint8)
        D, F = codes.shape
        types = _c(types, np.uint8)
        ext = _c(extrema, np.float64)
        ests = _c(np.atleast_1d(ests), np.int32)
        nest = len(ests)
        if pairs is None:
            npair = D * (D - 1) // 2
            pp = None
        else:
            pp = _c(pairs, np.uint32).reshape(-1, 2)
            npair = len(pp)
        s1 = np.zeros((D, nest, nsteps)) if want_s1 else None
        s2 = np.zeros((npair, nest, nsteps))
        mi = np.zeros((npair, nsteps)) if want_mi else None
        self.check(self._L.centre_pair_entropy(
            self.h, _p(codes), D, F, _p(types), _p(ext), nbins, nsteps, _p(pp),
            npair, _p(ests), nest, int(jacobian), int(kde), _p(s1), _p(s2),
            _p(mi)))
        return s1, s2, mi

    def mst(self, n_node: int, src, dst, w):
        src = _c(src, np.uint32)
        dst = _c(dst, np.uint32)
        w = _c(w, np.float64)
        out = np.zeros(max(n_node, 1), np.uint32)
        nt = C.c_uint32(0)
        self.check(self._L.centre_mst(self.h, _p(src), _p(dst), _p(w), len(w),
                                      n_node, _p(out), C.byref(nt)))
        return out[:nt.value]

    # ---- device-pointer entry points (pointers as ints) ---------------------
    def dev_discretize_hist(self, d_bat: int, ld: int, n_dof: int, n_frame: int,
                            dof_type: int, schemes, nbins_ref: int,
                            optimize: bool, d_codes: int, d_ext: int):
        sch = _c(np.atleast_1d(schemes), np.uint8)
        self.check(self._L.centre_dev_discretize_hist(
            self.h, d_bat, ld, n_dof, n_frame, dof_type, _p(sch), len(sch),
            nbins_ref, int(optimize), d_codes, d_ext))

    def dev_discretize_vmkde(self, d_bat: int, ld: int, n_dof: int, n_frame: int,
                             nmaxconf: int, kappa: float, d_codes: int, d_minima: int,
                             d_nfound: int, sdo_step=5, sdo_conv=1e-4, sdo_iter=1000,
                             grid_intervals=360, d_grid: int = 0):
        self.check(self._L.centre_dev_discretize_vmkde(
            self.h, d_bat, ld, n_dof, n_frame, nmaxconf, kappa, sdo_step, sdo_conv,
            sdo_iter, grid_intervals, d_codes, d_minima, d_nfound, d_grid or None))

    def dev_dof_entropy(self, d_codes, stride, n_dof, n_frame, d_types, d_ext,
                        nbins, nsteps, ests, jacobian, kde, d_s1, d_counts=None):
        ests = _c(np.atleast_1d(ests), np.int32)
        self.check(self._L.centre_dev_dof_entropy(
            self.h, d_codes, stride, n_dof, n_frame, d_types, d_ext, nbins,
            nsteps, _p(ests), len(ests), int(jacobian), int(kde), d_s1,
            d_counts))

    def dev_pair_entropy(self, plan: "PairPlan", d_codes, stride, n_frame,
                         d_types, d_ext, nsteps, ests, jacobian, kde, d_s1,
                         d_s2, d_mi, d_counts=None):
        ests = _c(np.atleast_1d(ests), np.int32)
        self.check(self._L.centre_dev_pair_entropy(
            self.h, plan.h, d_codes, stride, n_frame, d_types, d_ext, nsteps,
            _p(ests), len(ests), int(jacobian), int(kde), d_s1, d_s2, d_mi,
            d_counts))

    def dev_pair_entropy_schemes(self, plans, d_codes, stride, n_frame, d_types, d_ext,
                                 nsteps, ests, jacobian, kde, d_s1, d_s2, d_mi):
        """plans: one PairPlan per scheme; codes [D][nsch][stride]."""
        ests = _c(np.atleast_1d(ests), np.int32)
        arr = (C.c_void_p * len(plans))(*[pl.h.value for pl in plans])
        self.check(self._L.centre_dev_pair_entropy_schemes(
            self.h, C.cast(arr, C.c_void_p), len(plans), d_codes, stride, n_frame, d_types,
            d_ext, nsteps, _p(ests), len(ests), int(jacobian), int(kde), d_s1, d_s2, d_mi))

    def dev_pair_entropy_noalloc(self, plan: "PairPlan", d_codes, stride, n_frame, d_types,
                                 d_ext, nsteps, ests, jacobian, kde, d_s1, d_s2, d_mi):
        ests = _c(np.atleast_1d(ests), np.int32)
        self.check(self._L.centre_dev_pair_entropy_noalloc(
            self.h, plan.h, d_codes, stride, n_frame, d_types, d_ext, nsteps,
            _p(ests), len(ests), int(jacobian), int(kde), d_s1, d_s2, d_mi))

    def dev_mst(self, d_src, d_dst, d_w, n_edge, n_node):
        out = np.zeros(max(n_node, 1), np.uint32)
        nt = C.c_uint32(0)
        self.check(self._L.centre_dev_mst(self.h, d_src, d_dst, d_w, n_edge,
                                          n_node, _p(out), C.byref(nt)))
        return out[:nt.value]

    def dev_segment_sum(self, d_x, seg_begin):
        seg = _c(seg_begin, np.uint64)
        out = np.zeros(len(seg) - 1)
        self.check(self._L.centre_dev_segment_sum(self.h, d_x, _p(seg),
                                                  len(seg) - 1, _p(out)))
        return out


class DevArray:
    """A device buffer owned through the C ABI (shape and dtype kept here)."""

    def __init__(self, ctx: Context, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(v) for v in np.atleast_1d(shape)) if np.ndim(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        ctx.check(ctx._L.centre_dev_alloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value or 0

    def upload(self, x: np.ndarray):
        x = np.ascontiguousarray(x, dtype=self.dtype)
        if x.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        self.ctx.check(self.ctx._L.centre_memcpy_h2d(self.ctx.h, self.ptr, _p(x), x.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        self.ctx.check(self.ctx._L.centre_memcpy_d2h(self.ctx.h, _p(out), self.ptr, self.nbytes))
        return out

    def offset(self, nbytes: int) -> int:
        return self.ptr + nbytes

    def free(self):
        if getattr(self, "ptr", 0) and self.ctx.h:
            self.ctx._L.centre_dev_free(self.ctx.h, self.ptr)
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PairPlan:
    """centre_pair_plan: the pair set and tiling of one pair-stage call."""

    def __init__(self, ctx: Context, n_dof: int, nbins: int, pairs=None, tiles=None):
        self.ctx = ctx
        h = C.c_void_p()
        if tiles is not None:
            tl = _c(tiles, np.int32).reshape(-1, 2)
            ctx.check(ctx._L.centre_pair_plan_from_tiles(ctx.h, n_dof, nbins, _p(tl), len(tl),
                                                         C.byref(h)))
            self.h = h
            self._keep = tl
            return
        pp = None if pairs is None else _c(pairs, np.uint32).reshape(-1, 2)
        ctx.check(ctx._L.centre_pair_plan_create(
            ctx.h, n_dof, nbins, _p(pp), 0 if pp is None else len(pp),
            C.byref(h)))
        self.h = h
        self._keep = pp

    @property
    def records(self) -> int:
        return int(self.ctx._L.centre_pair_plan_records(self.h))

    @property
    def tiles(self) -> int:
        return int(self.ctx._L.centre_pair_plan_tiles(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.ctx._L.centre_pair_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
