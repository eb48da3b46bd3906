"""Native tier: an rfd_system (include/rfd.h, tier 2) driven from Python.

The system owns SoA particle and edge-list bond storage in HBM and runs whole
explicit steps on the device (forces, platens, integration, cut box, loading
control), so a step needs no host synchronisation.  This module builds
specimens into it, maps the reference slot layout onto it (for parity tests)
and reads results back.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import RfdError, lib
from .params import Compression

_u32p = C.POINTER(C.c_uint32)


class SystemParams(C.Structure):
    _fields_ = [("E", C.c_double), ("nu", C.c_double), ("rho", C.c_double),
                ("mu", C.c_double), ("mur", C.c_double), ("e", C.c_double),
                ("dt", C.c_double), ("farcut", C.c_double), ("skin", C.c_double),
                ("rebuild_every", C.c_uint32), ("contact_capacity", C.c_uint32),
                ("refmode", C.c_int), ("fp64", C.c_int)]


class Loading(C.Structure):
    _fields_ = [("mode", C.c_int), ("cut", C.c_int),
                ("cx", C.c_double), ("cy", C.c_double), ("cz", C.c_double),
                ("RR", C.c_double), ("zcut", C.c_double),
                ("top", C.c_double), ("bottom", C.c_double), ("top_r", C.c_double),
                ("bottom_r", C.c_double), ("band", C.c_double),
                ("v", C.c_double), ("v_load", C.c_double),
                ("size", C.c_double), ("size0", C.c_double),
                ("inv_area", C.c_double), ("min_stress", C.c_double),
                ("zero_stress", C.c_double), ("max_strain", C.c_double),
                ("stress", C.c_double), ("strain", C.c_double),
                ("fracture_stress", C.c_double), ("fracture_strain", C.c_double),
                ("sum_bottom", C.c_double), ("sum_top", C.c_double),
                ("loading", C.c_int), ("fracture", C.c_int),
                ("steps", C.c_uint32), ("hist_len", C.c_uint32),
                ("box_c", C.c_double * 3), ("box_h", C.c_double * 3),
                ("hidden", C.c_double * 3)]


class FireParams(C.Structure):
    _fields_ = [("ftol", C.c_double), ("dt0", C.c_double), ("dt_max", C.c_double),
                ("alpha0", C.c_double), ("f_inc", C.c_double), ("f_dec", C.c_double),
                ("f_alpha", C.c_double), ("n_delay", C.c_uint32), ("max_iter", C.c_uint32),
                ("check_every", C.c_uint32)]


class FireResult(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("converged", C.c_int), ("fmax0", C.c_double),
                ("fmax", C.c_double), ("mmax", C.c_double), ("sum_bottom", C.c_double),
                ("sum_top", C.c_double), ("strain", C.c_double), ("seconds", C.c_double),
                ("warm_used", C.c_int), ("fmax0_plain", C.c_double), ("fmax0_pred", C.c_double),
                ("breaks", C.c_uint32), ("break_rounds", C.c_uint32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneResult(C.Structure):
    _fields_ = [("force_tile", C.c_int), ("nbr_block", C.c_int), ("single_eval", C.c_int),
                ("cached", C.c_int), ("bucket", C.c_int), ("force_ms", C.c_double * 4),
                ("nbr_ms", C.c_double * 4)]

    FORCE_VARIANTS = ("64", "32", "128", "64 every half")

    def as_dict(self) -> dict:
        return {"force_tile": self.force_tile, "nbr_block": self.nbr_block,
                "single_eval": bool(self.single_eval), "cached": bool(self.cached),
                "bucket": self.bucket,
                "force_ms": {t: self.force_ms[k] for k, t in enumerate(self.FORCE_VARIANTS)},
                "nbr_ms": {b: self.nbr_ms[k] for k, b in enumerate((128, 64, 256, 0))}}


class PackParams(C.Structure):
    _fields_ = [("cx", C.c_double), ("cy", C.c_double), ("radius", C.c_double),
                ("z0", C.c_double), ("z1", C.c_double), ("s0", C.c_double),
                ("growth", C.c_double), ("beta", C.c_double), ("momentum", C.c_double),
                ("tol", C.c_double),
                ("stage_iters", C.c_uint32), ("final_iters", C.c_uint32)]


class PackResult(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("stages", C.c_uint32), ("converged", C.c_int),
                ("row_overflow", C.c_int), ("max_overlap", C.c_double), ("seconds", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_vp, _u32, _dbl = C.c_void_p, C.c_uint32, C.c_double
NATIVE_SIGNATURES = {
    "rfd_system_create": [_vp, C.POINTER(SystemParams), _u32, C.POINTER(_vp)],
    "rfd_system_destroy": [_vp],
    "rfd_system_set_particles": [_vp, _vp, _vp],
    "rfd_system_find_pairs": [_vp, _dbl, _vp, _u32, _u32p],
    "rfd_system_set_bonds": [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp],
    "rfd_system_set_bond_state": [_vp, _vp, _vp],
    "rfd_system_get_bond_state": [_vp, _vp, _vp],
    "rfd_system_set_velocities": [_vp, _vp, _vp],
    "rfd_system_import_reference_layout": [_vp] + [_vp] * 10 + [C.c_uint32] * 2
                                          + [C.c_float] * 4,
    "rfd_system_export_reference_layout": [_vp] + [_vp] * 10 + [C.c_uint32],
    "rfd_system_set_loading": [_vp, C.POINTER(Loading)],
    "rfd_system_get_loading": [_vp, C.POINTER(Loading)],
    "rfd_system_get_state": [_vp, _vp, _vp, _vp, _vp, _vp],
    "rfd_system_get_state64": [_vp, _vp, _vp, _vp, _vp, _vp],
    "rfd_system_compute_forces": [_vp],
    "rfd_system_step": [_vp, _u32],
    "rfd_system_get_history": [_vp, _vp, _u32, _u32p],
    "rfd_system_bond_counts": [_vp, _u32p],
    "rfd_system_contact_entries": [_vp, C.POINTER(C.c_uint64)],
    "rfd_system_sizes": [_vp, _u32p, _u32p],
    "rfd_system_device_arrays": [_vp, _vp, _vp, _vp, _vp, _vp],
    "rfd_system_timing": [_vp, C.c_int],
    "rfd_system_fire_load_step": [_vp, _dbl, C.POINTER(FireParams), _vp, _vp,
                                  C.POINTER(FireResult)],
    "rfd_system_residual": [_vp, C.POINTER(C.c_double)],
    "rfd_system_half_mirror": [_vp, C.POINTER(_vp)],
    "rfd_system_mpnn_inputs": [_vp, _dbl, _dbl, _vp, _vp],
    "rfd_system_fire_buffers": [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp),
                                C.POINTER(_vp)],
    "rfd_system_bond_graph": [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp),
                              C.POINTER(_vp), _u32p],
    "rfd_system_timing_read": [_vp, C.POINTER(C.c_double), _u32p],
    "rfd_system_set_specimens": [_vp, _u32, _vp, C.POINTER(Loading), _u32],
    "rfd_system_get_specimen_loading": [_vp, _u32, C.POINTER(Loading)],
    "rfd_system_get_specimen_history": [_vp, _u32, _vp, _u32, _u32p],
    "rfd_system_set_grid": [_vp, C.POINTER(C.c_double), _dbl, _u32p],
    "rfd_system_get_grid": [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double), _u32p],
    "rfd_system_step_graph": [_vp, _u32],
    "rfd_system_autotune": [_vp, C.c_int, C.POINTER(TuneResult)],
    "rfd_system_set_tiles": [_vp, C.c_int, C.c_int, C.c_int],
    "rfd_system_get_tiles": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "rfd_system_set_fusion": [_vp, C.c_int],
    "rfd_system_get_fusion": [_vp, C.POINTER(C.c_int)],
    "rfd_system_set_persistent": [_vp, C.c_int],
    "rfd_system_get_persistent": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "rfd_tune_cache_save": [C.c_char_p],
    "rfd_system_pack": [_vp, C.POINTER(PackParams), C.POINTER(PackResult)],
    "rfd_tune_cache_load": [C.c_char_p],
}


def _bind():
    L = lib()
    if not getattr(L, "_native_bound", False):
        for name, args in NATIVE_SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L._native_bound = True
    return L


def _np(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data_as(C.c_void_p)


def morton_order(R: np.ndarray, cell: float) -> np.ndarray:
    """Permutation sorting points by the Morton (Z-order) code of their cell,
    so that particles close in space are close in memory."""
    q = np.floor((R - R.min(0)) / cell).astype(np.uint64)
    q = np.minimum(q, (1 << 21) - 1)

    def spread(x):
        x = x & np.uint64(0x1FFFFF)
        x = (x | (x << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
        x = (x | (x << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
        x = (x | (x << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
        x = (x | (x << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
        x = (x | (x << np.uint64(2))) & np.uint64(0x1249249249249249)
        return x

    code = spread(q[:, 0]) | (spread(q[:, 1]) << np.uint64(1)) | (spread(q[:, 2]) << np.uint64(2))
    return np.argsort(code, kind="stable")


@dataclass
class Bonds:
    bi: np.ndarray
    bj: np.ndarray
    br: np.ndarray
    sc: np.ndarray
    tc: np.ndarray
    L0: np.ndarray | None = None  # rest lengths (None: current distances)

    @property
    def n(self) -> int:
        return int(self.bi.size)


class NativeSystem:
    """rfd_system handle.  Positions are kept in the order given (callers that
    want spatial locality reorder first, see ``morton_order``)."""

    def __init__(self, ctx, params: SystemParams, R: np.ndarray, radius: np.ndarray):
        self.L = _bind()
        self.ctx = ctx
        self.params = params
        R = np.asarray(R, np.float64)
        self.N = N = R.shape[0]
        h = C.c_void_p()
        self._call("rfd_system_create", ctx.h, C.byref(params), N, C.byref(h), sys=False)
        self.h = h
        self.radius = np.ascontiguousarray(radius, np.float32)
        soa, p = _np(R.T.reshape(-1), np.float64)
        _, pr = _np(self.radius, np.float32)
        self._call("rfd_system_set_particles", self.h, p, pr)
        self.bonds: Bonds | None = None
        self.S = 0

    def _call(self, name, *args, sys=True):
        st = getattr(self.L, name)(*args)
        if st != 0:
            msg = self.L.rfd_last_error(self.ctx.h)
            raise RfdError(name, st, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            # a system outliving its context (e.g. held by a traceback until
            # interpreter exit) is left to the process end: its destroy call
            # would use the freed context
            if getattr(self.ctx, "h", None):
                self.L.rfd_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- setup ---------------------------------------------------------
    def set_positions(self, R: np.ndarray):
        """New positions (N, 3) for the same particles; velocities, loads and
        contact rows are cleared, bonds and their state are kept."""
        R = np.asarray(R, np.float64)
        assert R.shape == (self.N, 3)
        _, p = _np(R.T.reshape(-1), np.float64)
        _, pr = _np(self.radius, np.float32)
        self._call("rfd_system_set_particles", self.h, p, pr)

    def find_pairs(self, gap: float) -> np.ndarray:
        cnt = C.c_uint32()
        self._call("rfd_system_find_pairs", self.h, gap, None, 0, C.byref(cnt))
        n = cnt.value
        out = np.zeros(2 * max(n, 1), np.uint32)
        self._call("rfd_system_find_pairs", self.h, gap, out.ctypes.data_as(C.c_void_p), n,
                   C.byref(cnt))
        return out[: 2 * n].reshape(-1, 2)

    def set_bonds(self, b: Bonds):
        order = np.lexsort((b.bj, b.bi))
        L0 = None if b.L0 is None else np.ascontiguousarray(b.L0[order], np.float32)
        b = Bonds(*(np.ascontiguousarray(x[order]) for x in (b.bi, b.bj, b.br, b.sc, b.tc)), L0)
        arrs = [_np(b.bi, np.uint32), _np(b.bj, np.uint32), _np(b.br, np.float32),
                _np(b.sc, np.float32), _np(b.tc, np.float32)]
        pl = L0.ctypes.data_as(C.c_void_p) if L0 is not None else None
        self._call("rfd_system_set_bonds", self.h, b.n, *[p for _, p in arrs], pl)
        self.bonds = b

    def set_bond_state(self, state: np.ndarray, flag: np.ndarray):
        """state[Nb, 12], flag[Nb] in the system's bond order (sorted by (i, j))."""
        _, ps = _np(state, np.float32)
        _, pf = _np(flag, np.uint8)
        self._call("rfd_system_set_bond_state", self.h, ps, pf)

    def bond_state(self):
        nb = self.bonds.n
        st = np.zeros(12 * nb, np.float32)
        fl = np.zeros(nb, np.uint8)
        self._call("rfd_system_get_bond_state", self.h, st.ctypes.data_as(C.c_void_p),
                   fl.ctypes.data_as(C.c_void_p))
        return st.reshape(nb, 12), fl

    # -- the reference slot layout ---------------------------------------
    _SLOT_ARGS = ("rij", "Oijt", "Mijn", "Mijt")

    def import_reference_layout(self, lay, mat):
        """Particles, bonds with their state, and contact springs from a
        reference-layout system (rfd.reflayout.RefLayout, device tensors);
        radius, bond radius and strengths from the material.  The bond list
        and order are those of bonds_from_slots."""
        d = [t.data_ptr() for t in (lay.R, lay.V, lay.W, lay.IL, lay.ILtype, lay.iL)]
        d += [lay.slot[k].data_ptr() for k in self._SLOT_ARGS]
        self._call("rfd_system_import_reference_layout", self.h, *d, lay.N, lay.K,
                   mat.p_r, mat.b_r, mat.sc, mat.tc)
        # the host copy of the bond list (the library selects and orders the
        # same way)
        self.bonds, _, _ = bonds_from_slots(lay.host(), lay.N, lay.K, mat)
        self.radius = np.full(self.N, mat.p_r, np.float32)

    def export_reference_layout(self, lay):
        """Write the state into a reference-layout system's device tensors
        (the same N; rows of lay.K slots)."""
        assert lay.N == self.N
        d = [t.data_ptr() for t in (lay.R, lay.V, lay.W, lay.IL, lay.ILtype, lay.iL)]
        d += [lay.slot[k].data_ptr() for k in self._SLOT_ARGS]
        self._call("rfd_system_export_reference_layout", self.h, *d, lay.K)

    def set_velocities(self, V=None, W=None):
        a = _np(V, np.float32) if V is not None else (None, None)
        b = _np(W, np.float32) if W is not None else (None, None)
        self._call("rfd_system_set_velocities", self.h, a[1], b[1])

    def set_loading(self, L: Loading):
        self._call("rfd_system_set_loading", self.h, C.byref(L))

    def loading(self) -> Loading:
        L = Loading()
        self._call("rfd_system_get_loading", self.h, C.byref(L))
        return L

    # -- run -----------------------------------------------------------
    def compute_forces(self):
        self._call("rfd_system_compute_forces", self.h)

    def step(self, n: int = 1):
        self._call("rfd_system_step", self.h, n)

    def state(self):
        n3 = 3 * self.N
        R = np.zeros(n3, np.float64)
        V, W, F, M = (np.zeros(n3, np.float32) for _ in range(4))
        self._call("rfd_system_get_state", self.h, *(a.ctypes.data_as(C.c_void_p)
                                                      for a in (R, V, W, F, M)))
        return {"R": R, "V": V, "W": W, "F": F, "M": M}

    def state64(self):
        """R, V, W, F, M as (N, 3) float64 arrays: the FP64 path's own values
        (the FP32 path's widened)."""
        out = {k: np.zeros(3 * self.N, np.float64) for k in ("R", "V", "W", "F", "M")}
        self._call("rfd_system_get_state64", self.h,
                   *(out[k].ctypes.data_as(C.c_void_p) for k in ("R", "V", "W", "F", "M")))
        return {k: v.reshape(3, self.N).T.copy() for k, v in out.items()}

    def history(self) -> np.ndarray:
        n = C.c_uint32()
        self._call("rfd_system_get_history", self.h, None, 0, C.byref(n))
        h = np.zeros(3 * max(n.value, 1), np.float64)
        self._call("rfd_system_get_history", self.h, h.ctypes.data_as(C.c_void_p), n.value,
                   C.byref(n))
        return h[: 3 * n.value].reshape(-1, 3)

    # -- quasi-static load steps -----------------------------------------
    def fire_load_step(self, dstrain: float, params: FireParams, dU=None, dTh=None) -> FireResult:
        """One load step; dU / dTh are device tensors (N, 3) float32 or None."""
        r = FireResult()
        pu = C.c_void_p(dU.data_ptr()) if dU is not None else None
        pt = C.c_void_p(dTh.data_ptr()) if dTh is not None else None
        self._call("rfd_system_fire_load_step", self.h, dstrain, C.byref(params), pu, pt,
                   C.byref(r))
        return r

    def half_mirror(self):
        """Device view hm[H] int32: the other half of each half's bond."""
        a = C.c_void_p()
        self._call("rfd_system_half_mirror", self.h, C.byref(a))
        return device_view(a.value, (2 * self.bonds.n,), "<i4", self)

    def mpnn_inputs(self, su: float, sth: float, X, E):
        """Fill the warm-start network inputs X [N x 32], E [H x 4] (device)."""
        assert X.shape == (self.N, 32) and E.shape == (2 * self.bonds.n, 4)
        self._call("rfd_system_mpnn_inputs", self.h, su, sth, C.c_void_p(X.data_ptr()),
                   C.c_void_p(E.data_ptr()))

    def mean_radius(self) -> float:
        return float(self.radius.astype(np.float64).mean())

    def residual(self) -> tuple[float, float]:
        out = (C.c_double * 2)()
        self._call("rfd_system_residual", self.h, out)
        return out[0], out[1]

    def fire_buffers(self):
        """(du_last, du_prev, dth_last, dth_prev) as torch views (N, 3) of the
        device arrays."""
        ptrs = [C.c_void_p() for _ in range(4)]
        self._call("rfd_system_fire_buffers", self.h, *[C.byref(q) for q in ptrs])
        return tuple(device_view(q.value, (self.N, 3), "<f4", self) for q in ptrs)

    def bond_graph(self):
        """Device views: hoff[N+1] int32, half[H, 2] int32 (other end, bond),
        bond[Nb, 4] float32 (1/L0, br, sc, tc), state[Nb, 8] float32 (O, mn, Mt,
        flag as uint32 bits: read column 7 through .view(torch.int32)).  The
        state view is the buffer current at the time of the call."""
        a, b, c, d, H = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_uint32()
        self._call("rfd_system_bond_graph", self.h, C.byref(a), C.byref(b), C.byref(c),
                   C.byref(d), C.byref(H))
        nb = H.value // 2
        return (device_view(a.value, (self.N + 1,), "<i4", self),
                device_view(b.value, (H.value, 2), "<i4", self),
                device_view(c.value, (nb, 4), "<f4", self),
                device_view(d.value, (nb, 8), "<f4", self))

    def seat_platens(self, preload: float = 0.0, L: Loading | None = None):
        """Move the platens onto the specimen: top at the highest and bottom at
        the lowest particle centre inside the platen footprint (load axis of
        the loading mode), each pushed `preload` further in so that the whole
        end layer is held; size0 becomes their distance."""
        L = L or self.loading()
        R = self.state()["R"].reshape(3, -1)
        ax = 2 if L.mode == 1 else 1
        if L.mode == 1:
            inside = (R[0] - L.cx) ** 2 + (R[1] - L.cy) ** 2 < L.RR
        else:
            inside = (R[0] - L.cx) ** 2 + (R[2] - L.cz) ** 2 < L.RR
        h = R[ax][inside]
        L.top, L.bottom = float(h.max()) - preload, float(h.min()) + preload
        L.top_r = L.top - L.band
        L.bottom_r = L.bottom + 0.5 * L.band
        L.size = L.size0 = L.top - L.bottom
        L.strain = 0.0
        L.v = 0.0
        self.set_loading(L)
        return L

    def timing(self, enable: bool = True):
        self._call("rfd_system_timing", self.h, int(enable))

    # -- batched specimens, grid, graphs, launch shapes --------------------
    def set_specimens(self, first, loadings, hist_cap: int = 1 << 14):
        """Specimen k owns particles [first[k], first[k+1]) and is driven by
        loadings[k]; an empty list returns to the single loading control."""
        S = len(loadings)
        if S == 0:
            self._call("rfd_system_set_specimens", self.h, 0, None, None, 0)
            self.S = 0
            return
        f = np.ascontiguousarray(first, np.uint32)
        arr = (Loading * S)(*loadings)
        self._call("rfd_system_set_specimens", self.h, S, f.ctypes.data_as(C.c_void_p), arr,
                   hist_cap)
        self.S = S
        self.spec_first = f.copy()

    def specimen_loadings(self) -> list[Loading]:
        arr = (Loading * self.S)()
        self._call("rfd_system_get_specimen_loading", self.h, self.S, arr)
        return list(arr)

    def specimen_history(self, k: int) -> np.ndarray:
        n = C.c_uint32()
        self._call("rfd_system_get_specimen_history", self.h, k, None, 0, C.byref(n))
        h = np.zeros(3 * max(n.value, 1), np.float64)
        self._call("rfd_system_get_specimen_history", self.h, k, h.ctypes.data_as(C.c_void_p),
                   n.value, C.byref(n))
        return h[: 3 * min(n.value, len(h) // 3)].reshape(-1, 3)

    def set_grid(self, origin, cell: float, dims):
        o = (C.c_double * 3)(*[float(x) for x in origin])
        d = (C.c_uint32 * 3)(*[int(x) for x in dims])
        self._call("rfd_system_set_grid", self.h, o, float(cell), d)

    def grid(self):
        o = (C.c_double * 3)()
        c = C.c_double()
        d = (C.c_uint32 * 3)()
        self._call("rfd_system_get_grid", self.h, o, C.byref(c), d)
        return tuple(o), c.value, tuple(d)

    def step_graph(self, n: int = 1):
        self._call("rfd_system_step_graph", self.h, n)

    def pack(self, p: PackParams) -> PackResult:
        r = PackResult()
        self._call("rfd_system_pack", self.h, C.byref(p), C.byref(r))
        return r

    def set_tiles(self, force_tile: int, nbr_block: int, single_eval: bool = True):
        self._call("rfd_system_set_tiles", self.h, force_tile, nbr_block, int(single_eval))

    def tiles(self) -> tuple[int, int, bool]:
        """(force tile, row-search block, single-evaluation pass running)."""
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._call("rfd_system_get_tiles", self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, bool(c.value)

    def set_fusion(self, on: bool = True):
        """Fuse the particle update into the force pass of explicit steps
        (on by default; the same bits either way) and the FIRE measure into
        the quasi-static force pass (there only the order of the FP64 partial
        sums differs)."""
        self._call("rfd_system_set_fusion", self.h, int(on))

    def fused(self) -> bool:
        """Whether explicit steps currently run the fused pass."""
        f = C.c_int()
        self._call("rfd_system_get_fusion", self.h, C.byref(f))
        return bool(f.value)

    def set_persistent(self, on: bool = True):
        """Run the fused steps between two contact-row rebuilds as one
        cooperative launch where the system qualifies (off by default; the
        same bits either way)."""
        self._call("rfd_system_set_persistent", self.h, int(on))

    def persistent(self) -> tuple:
        """(multi-step launches requested, the current system qualifies)."""
        a, b = C.c_int(), C.c_int()
        self._call("rfd_system_get_persistent", self.h, C.byref(a), C.byref(b))
        return bool(a.value), bool(b.value)

    def force_kernel(self) -> str:
        """Name of the force-pass kernel the next explicit step launches."""
        if self.params.fp64:
            return "k_f64_forces"
        if self.fused():
            return "k_nat_forces1<fused>"  # the particle update runs in the same launch
        return "k_nat_forces1" if self.tiles()[2] else "k_nat_forces"

    def autotune(self, use_cache: bool = True) -> TuneResult:
        r = TuneResult()
        self._call("rfd_system_autotune", self.h, int(use_cache), C.byref(r))
        return r

    def timing_read(self) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_uint32()
        self._call("rfd_system_timing_read", self.h, C.byref(ms), C.byref(n))
        return ms.value, n.value

    def bond_counts(self) -> tuple[int, int, int]:
        out = (C.c_uint32 * 3)()
        self._call("rfd_system_bond_counts", self.h, out)
        return tuple(out)

    def contact_entries(self) -> int:
        """Entries in the current contact rows (never-bonded candidates)."""
        n = C.c_uint64()
        self._call("rfd_system_contact_entries", self.h, C.byref(n))
        return n.value


def tune_cache_save(path: str) -> None:
    if _bind().rfd_tune_cache_save(str(path).encode()) != 0:
        raise RfdError("rfd_tune_cache_save", -1, f"cannot write {path}")


def tune_cache_load(path: str) -> None:
    if _bind().rfd_tune_cache_load(str(path).encode()) != 0:
        raise RfdError("rfd_tune_cache_load", -1, f"cannot read {path}")


# ------------------------------------------------------------- loading
class _DevArray:
    """Minimal __cuda_array_interface__ wrapper so torch can view device memory
    owned by the library.  It holds a reference to the owning object (the
    NativeSystem), and the tensor torch makes from it holds this wrapper, so
    the memory outlives every view of it."""

    def __init__(self, ptr: int, shape, typestr: str, owner=None):
        self.owner = owner
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr,
                                         "data": (ptr, False), "version": 3, "strides": None}


def device_view(ptr: int, shape, typestr: str, owner=None):
    import torch

    return torch.as_tensor(_DevArray(ptr, shape, typestr, owner), device="cuda")


def fire_params(mat_E: float, r_mean: float, dt: float, rel_tol: float = 1e-6,
                max_iter: int = 20000, check_every: int = 100) -> FireParams:
    """FIRE settings; the tolerance is rel_tol * F_ref with F_ref = E*pi*r^2,
    the bond force at unit strain for a bond of the mean radius."""
    p = FireParams()
    p.ftol = rel_tol * mat_E * 3.141592653589793 * r_mean * r_mean
    p.dt0 = dt
    p.dt_max = 10.0 * dt
    p.alpha0, p.f_inc, p.f_dec, p.f_alpha = 0.1, 1.1, 0.5, 0.99
    p.n_delay = 5
    p.max_iter = max_iter
    p.check_every = check_every
    return p


def loading_from_compression(cmp: Compression, zcut: float, box=None) -> Loading:
    """Device loading control from the host compression_data mirror."""
    L = Loading()
    L.mode = 1 if cmp.mode == "uniaxial" else 2
    L.cx, L.cy, L.cz = cmp.center_d
    L.RR = cmp.RR
    L.zcut = zcut
    L.top, L.bottom, L.top_r, L.bottom_r = cmp.top, cmp.bottom, cmp.top_r, cmp.bottom_r
    L.band = 2.0 * cmp.p_r
    L.v, L.v_load = cmp.V, cmp.v_load
    L.size = L.size0 = cmp.size0
    L.inv_area = cmp.inv_area
    L.min_stress, L.zero_stress, L.max_strain = cmp.min_stress, cmp.zero_stress, cmp.max_strain
    if box is not None:
        L.cut = 1
        for d in range(3):
            L.box_c[d] = box.center[d]
            L.box_h[d] = box.spacesized2[d]
            L.hidden[d] = box.hidden_point[d]
    return L


def params_from_material(mat, refmode: int = 1, skin: float = 0.05, rebuild_every: int = 20,
                         contact_capacity: int = 16, fp64: bool = False) -> SystemParams:
    p = SystemParams()
    p.E, p.nu, p.rho = mat.E, mat.nu, mat.rho
    p.mu, p.mur, p.e = mat.p_mu, mat.p_mur, mat.p_e
    p.dt = mat.dt
    p.farcut = mat.a_farcut / (2.0 * mat.p_r)
    p.skin = skin
    p.rebuild_every = rebuild_every
    p.contact_capacity = contact_capacity
    p.refmode = refmode
    p.fp64 = 1 if fp64 else 0
    return p


# ----------------------------------------------- reference slot layout map
def bonds_from_slots(h: dict, N: int, K: int, mat) -> tuple[Bonds, np.ndarray, np.ndarray]:
    """Bond records (i<j) from the reference slot arrays: a pair is a bond
    record when slot i->j has ILtype 1; its state is i's slot state
    (rij -> n_prev, Oijt, Mijn, Mijt)."""
    IL = h["IL"].reshape(N, K)
    ty = h["ILtype"].reshape(N, K)
    ii, kk = np.nonzero((IL != 0xFFFFFFFF) & (ty == 1))
    jj = IL[ii, kk].astype(np.int64)
    m = ii < jj
    ii, kk, jj = ii[m], kk[m], jj[m]
    order = np.lexsort((jj, ii))  # the system's bond order: sorted by (i, j)
    ii, kk, jj = ii[order], kk[order], jj[order]
    s = ii * K + kk
    st = np.concatenate([h[k].reshape(-1, 3)[s] for k in ("rij", "Oijt", "Mijn", "Mijt")], 1)
    n = ii.size
    # the bond law uses 1/iL as the rest length (md_calculateforcesdem3.cu:96)
    L0 = (np.float32(1.0) / h["iL"][s].astype(np.float32)).astype(np.float32)
    b = Bonds(ii.astype(np.uint32), jj.astype(np.uint32), np.full(n, mat.b_r, np.float32),
              np.full(n, mat.sc, np.float32), np.full(n, mat.tc, np.float32), L0)
    return b, st.astype(np.float32), np.ones(n, np.uint8)


def strengths(g: np.random.Generator, n: int, mean_t=20e6, mean_s=40e6, cv=0.2):
    """Bond strengths N(mu, cv*mu) truncated > 0, in stress units (1e9 Pa)."""
    from .params import STRESS
    from .specimen import truncated_normal

    sc = truncated_normal(g, mean_t / STRESS, cv * mean_t / STRESS, n)
    tc = truncated_normal(g, mean_s / STRESS, cv * mean_s / STRESS, n)
    return sc.astype(np.float32), tc.astype(np.float32)


def lattice_rest_gap(d_nn: float, r_lo: float) -> float:
    """Gap below which every nearest lattice neighbour of an FCC lattice of
    spacing d_nn is joined (radii >= r_lo)."""
    return d_nn - 2.0 * r_lo + 1e-3 * d_nn
