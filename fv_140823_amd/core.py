"""Python face of the C ABI: the host forest, the flattened mesh and the device
solver context.  Arrays cross the boundary as contiguous float64 numpy arrays in
the var-major layouts of include/fvgpu.h ([neq][ncell], [6][neq][ncell], ...).
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, Sequence

import numpy as np

from . import _capi
from ._capi import (FIELD_FACEFLUX, FIELD_RHS, FIELD_S, FIELD_SF, FIELD_SG, FIELD_SN, FIELD_STAGE_IN, FVG_MINMOD,
                    FVG_MIXED, FVG_PLAIN, FVG_VANALBADA, FvgParams, check, lib)

# reference bc order of the six sides (inputset.c:99-104)
SIDES = ("-x", "+x", "-y", "+y", "-z", "+z")
BC_INLET, BC_OUTFLOW = 1, 2


def neq_of(model: int) -> int:
    n = lib().fvg_neq(model)
    if n < 0:
        raise ValueError(f"model {model} is not supported")
    return n


def params(model: int = 1, **kw) -> FvgParams:
    """fvg_params with the reference's defaults, overridden by keyword."""
    p = FvgParams()
    lib().fvg_params_default(C.byref(p), model)
    for k, v in kw.items():
        if k == "precision" and isinstance(v, str):
            v = {"plain": FVG_PLAIN, "mixed": FVG_MIXED}[v]
        if k == "limiter" and isinstance(v, str):
            v = {"minmod": FVG_MINMOD, "vanalbada": FVG_VANALBADA}[v]
        if k == "wave_speeds" and isinstance(v, str):
            v = {"davis": 0, "pvrs": 1}[v]
        if k == "inlet":
            arr = (C.c_double * 5)(*([float(x) for x in v] + [0.0] * (5 - len(v))))
            p.inlet = arr
            continue
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Forest:
    """Host forest of octrees (reference BRANCH/CELL/LEAF structs)."""

    def __init__(self, nroot: Sequence[int] = (1, 1, 1), x0: Sequence[float] = (0, 0, 0),
                 x1: Sequence[float] = (1, 1, 1), level: int = 3,
                 bc: Sequence[int] = (2, 2, 2, 2, 2, 2), periodic: Sequence[int] = (0, 0, 0),
                 neq: int = 5):
        L = lib()
        self.neq = int(neq)
        h = C.c_void_p()
        ni = (C.c_int * 3)(*nroot)
        a0 = (C.c_double * 3)(*map(float, x0))
        a1 = (C.c_double * 3)(*map(float, x1))
        b6 = (C.c_int * 6)(*bc)
        pe = (C.c_int * 3)(*periodic)
        check(L.fvh_forest_create(ni, a0, a1, int(level), b6, pe, self.neq, C.byref(h)),
              "fvh_forest_create")
        self._h = h
        self.box = (np.array(x0, float), np.array(x1, float))
        self.periodic = tuple(int(p) for p in periodic)

    @classmethod
    def _adopt(cls, h: C.c_void_p) -> "Forest":
        self = cls.__new__(cls)
        self._h = h
        nr, bc, pe, neq = (C.c_int * 3)(), (C.c_int * 6)(), (C.c_int * 3)(), C.c_int()
        a0, a1 = (C.c_double * 3)(), (C.c_double * 3)()
        check(lib().fvh_forest_geometry(h, nr, a0, a1, bc, pe, C.byref(neq)), "fvh_forest_geometry")
        self.neq = neq.value
        self.box = (np.array(a0[:], float), np.array(a1[:], float))
        self.periodic = tuple(pe[:])
        return self

    def geometry(self) -> dict:
        nr, bc, pe, neq = (C.c_int * 3)(), (C.c_int * 6)(), (C.c_int * 3)(), C.c_int()
        a0, a1 = (C.c_double * 3)(), (C.c_double * 3)()
        check(lib().fvh_forest_geometry(self._h, nr, a0, a1, bc, pe, C.byref(neq)), "fvh_forest_geometry")
        return dict(nroot=tuple(nr[:]), x0=tuple(a0[:]), x1=tuple(a1[:]), bc=tuple(bc[:]),
                    periodic=tuple(pe[:]), neq=neq.value)

    # ---- restart / export
    def save(self, path: str, tstep: int = 0, time: float = 0.0) -> None:
        """Binary restart (exact): geometry, leaf levels in list order, S in f64."""
        check(lib().fvh_forest_save(self._h, str(path).encode(), int(tstep), float(time)), "fvh_forest_save")

    @classmethod
    def load(cls, path: str):
        """Forest from a binary restart; returns (forest, tstep, time)."""
        h, ts, tm = C.c_void_p(), C.c_int32(), C.c_double()
        check(lib().fvh_forest_load(str(path).encode(), C.byref(h), C.byref(ts), C.byref(tm)), "fvh_forest_load")
        return cls._adopt(h), ts.value, tm.value

    def save_text(self, path: str, tstep: int = 0, time: float = 0.0) -> None:
        """The reference's text restart layout (restart_save.c)."""
        check(lib().fvh_forest_save_text(self._h, str(path).encode(), int(tstep), float(time)),
              "fvh_forest_save_text")

    def load_text(self, path: str):
        """Spawn the leaves of a reference text restart into this forest of bare roots
        (restart_load.c); returns (tstep, time)."""
        ts, tm = C.c_int32(), C.c_double()
        check(lib().fvh_forest_load_text(self._h, str(path).encode(), C.byref(ts), C.byref(tm)),
              "fvh_forest_load_text")
        return ts.value, tm.value

    def export_tecplot(self, path: str, model: int, p: np.ndarray) -> None:
        """Tecplot FEPOINT brick zone (exportfield_linked_V3.c) from the leaves' S and SG and
        the cell pressure p."""
        p = np.ascontiguousarray(p, np.float64)
        if p.shape != (self.nleaves,):
            raise ValueError("one pressure per leaf")
        check(lib().fvh_export_tecplot(self.run, str(path).encode(), int(model), _dptr(p)),
              "fvh_export_tecplot")

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().fvh_forest_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def run(self) -> C.c_void_p:
        """The RUN* of this forest (for the reference-signature entry points)."""
        return C.c_void_p(lib().fvh_forest_run(self._h))

    @property
    def nleaves(self) -> int:
        return lib().fvh_forest_nleaves(self._h)

    @property
    def max_level(self) -> int:
        return lib().fvh_forest_max_level(self._h)

    def leaf_info(self):
        n = self.nleaves
        xc = np.empty((n, 3), np.float64)
        lev = np.empty(n, np.int32)
        check(lib().fvh_forest_leaf_info(self._h, _dptr(xc), lev.ctypes.data_as(C.POINTER(C.c_int32))),
              "fvh_forest_leaf_info")
        return xc, lev

    def _shape(self, which: int):
        n, q = self.nleaves, self.neq
        return {FIELD_SF: (6, q, n), FIELD_SG: (q + 1, 3, n)}.get(which, (q, n))

    def set_field(self, which: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, np.float64).reshape(self._shape(which))
        check(lib().fvh_forest_set_field(self._h, which, _dptr(a)), "fvh_forest_set_field")

    def get_field(self, which: int) -> np.ndarray:
        a = np.empty(self._shape(which), np.float64)
        check(lib().fvh_forest_get_field(self._h, which, _dptr(a)), "fvh_forest_get_field")
        return a

    def set_state(self, S: np.ndarray) -> None:
        self.set_field(FIELD_S, S)

    def get_state(self) -> np.ndarray:
        return self.get_field(FIELD_S)

    def adapt(self, flags: np.ndarray, max_level: int) -> int:
        f = np.ascontiguousarray(flags, np.int8)
        if f.shape != (self.nleaves,):
            raise ValueError("one flag per leaf")
        n = lib().fvh_forest_adapt(self._h, f.ctypes.data_as(C.POINTER(C.c_int8)), int(max_level))
        if n < 0:
            check(n, "fvh_forest_adapt")
        return n


class FlatMesh:
    """The flattened struct-of-arrays mesh of a RUN (owned native arrays)."""

    def __init__(self, run: C.c_void_p):
        h = C.c_void_p()
        check(lib().fvh_flatten_run(run, C.byref(h)), "fvh_flatten_run")
        self._h = h
        self.mesh = lib().fvh_flat_mesh(h).contents

    def close(self):
        if getattr(self, "_h", None):
            lib().fvh_flat_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ncell(self) -> int:
        return self.mesh.ncell

    @property
    def nface(self) -> int:
        return self.mesh.nface

    def tile_table(self) -> np.ndarray:
        """The interior 8^3 tiles fvg_set_mesh finds on this mesh (host only): rows of
        (first leaf, first leaves of the neighbour tiles across faces 0..5, level)."""
        n = C.c_int()
        check(lib().fvh_tile_table(C.byref(self.mesh), None, 0, C.byref(n)), "fvh_tile_table")
        out = np.zeros((n.value, 8), np.int32)
        if n.value:
            check(lib().fvh_tile_table(C.byref(self.mesh), out.ctypes.data_as(C.POINTER(C.c_int32)),
                                        n.value, C.byref(n)), "fvh_tile_table")
        return out

    def arrays(self) -> dict:
        m = self.mesh
        N, F, Q, nl = m.ncell, m.nface, m.nquad, m.nlevel

        def arr(ptr, n, dt):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

        return dict(
            cflag=arr(m.cflag, N, np.uint32), nbr=arr(m.nbr, 6 * N, np.int32).reshape(6, N),
            quad=arr(m.quad_nbr, 4 * Q, np.int32).reshape(Q, 4) if Q else np.zeros((0, 4), np.int32),
            face_L=arr(m.face_L, F, np.int32), face_R=arr(m.face_R, F, np.int32),
            face_meta=arr(m.face_meta, F, np.uint32), gslot=arr(m.gslot, 6 * N, np.int32).reshape(6, N),
            h=arr(m.h, 3 * nl, np.float64).reshape(nl, 3),
            dist=arr(m.dist, 18 * nl, np.float64).reshape(nl, 3, 3, 2),
            area=arr(m.area, 3 * nl, np.float64).reshape(nl, 3),
            vol=arr(m.vol, nl, np.float64), nface_axis=tuple(m.nface_axis),
            nface_reg=tuple(m.nface_reg), face_LL=arr(m.face_LL, F, np.int32),
            face_RR=arr(m.face_RR, F, np.int32))


class Solver:
    """A device context (fvg_ctx) on one mesh."""

    def __init__(self, p: FvgParams, device: int = 0):
        h = C.c_void_p()
        self.params = p
        check(lib().fvg_create(int(device), C.byref(p), C.byref(h)), "fvg_create")
        self._h = h
        self.neq = neq_of(p.model)
        self.ncell = 0
        self.nface = 0

    @classmethod
    def from_forest(cls, forest: Forest, p: FvgParams, device: int = 0, upload: bool = True,
                    device_flatten: bool = False):
        """A context on the forest's mesh, flattened on the host (fvh_flatten_run) or,
        with device_flatten, built on the device from the leaf keys (fvg_remesh_*), which
        lets later re-flattens carry the state over on the device."""
        s = cls(p, device)
        if device_flatten:
            s.remesh(forest, carry_state=False)
        else:
            s.set_mesh(FlatMesh(forest.run))
        if upload:
            s.set_state(forest.get_state())
        return s

    def remesh(self, forest: Forest, carry_state: bool = True) -> None:
        """Rebuild the mesh on the device from the (adapted) forest's leaf keys.  With
        carry_state the state comes from the previous device mesh by key (kept, copied
        from the split parent, or the mean of the 8 merged kids); otherwise S is zero."""
        self._ck(lib().fvh_forest_remesh_device(forest._h, self._h, int(bool(carry_state))),
                 "fvh_forest_remesh_device")
        sz = self.mesh_sizes()
        self.ncell, self.nface = sz["ncell"], sz["nface"]

    @property
    def has_keys(self) -> bool:
        return lib().fvg_has_keys(self._h) == 1

    def mesh_sizes(self) -> dict:
        s = (C.c_int32 * 12)()
        n = None
        self._ck(lib().fvg_get_mesh(self._h, s, n, n, n, n, n, n, n, n, n, n), "fvg_get_mesh")
        return dict(ncell=s[0], nface=s[1], nface_axis=tuple(s[2:5]), nquad=s[5], nlevel=s[6],
                    nface_reg=tuple(s[7:10]), has_keys=bool(s[10]))

    def mesh_arrays(self) -> dict:
        """The device mesh copied back (same keys and shapes as FlatMesh.arrays())."""
        z = self.mesh_sizes()
        N, F, Q, nl = z["ncell"], z["nface"], z["nquad"], z["nlevel"]
        a = dict(cflag=np.empty(N, np.uint32), nbr=np.empty((6, N), np.int32),
                 quad=np.zeros((Q, 4), np.int32), face_L=np.empty(F, np.int32),
                 face_R=np.empty(F, np.int32), face_meta=np.empty(F, np.uint32),
                 face_LL=np.empty(F, np.int32), face_RR=np.empty(F, np.int32),
                 gslot=np.empty((6, N), np.int32), tab=np.empty(34 * nl, np.float64))
        u32 = lambda x: x.ctypes.data_as(C.POINTER(C.c_uint32))
        i32 = lambda x: x.ctypes.data_as(C.POINTER(C.c_int32))
        self._ck(lib().fvg_get_mesh(self._h, None, u32(a["cflag"]), i32(a["nbr"]),
                                    i32(a["quad"]) if Q else None, i32(a["face_L"]),
                                    i32(a["face_R"]), u32(a["face_meta"]), i32(a["face_LL"]),
                                    i32(a["face_RR"]), i32(a["gslot"]), _dptr(a["tab"])),
                 "fvg_get_mesh")
        t = a.pop("tab")
        a.update(h=t[:3 * nl].reshape(nl, 3), dist=t[3 * nl:21 * nl].reshape(nl, 3, 3, 2),
                 area=t[21 * nl:24 * nl].reshape(nl, 3), vol=t[24 * nl:25 * nl].copy(),
                 nface_axis=z["nface_axis"], nface_reg=z["nface_reg"])
        return a

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().fvg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def _ck(self, rc: int, where: str) -> None:
        check(rc, where, self._h)

    def set_params(self, p: FvgParams) -> None:
        self._ck(lib().fvg_set_params(self._h, C.byref(p)), "fvg_set_params")
        self.params = p

    def set_mesh(self, flat: FlatMesh) -> None:
        self._ck(lib().fvg_set_mesh(self._h, C.byref(flat.mesh)), "fvg_set_mesh")
        self.ncell = flat.ncell
        self.nface = flat.nface

    def _shape(self, which: int):
        n, q = self.ncell, self.neq
        return {FIELD_SF: (6, q, n), FIELD_SG: (q + 1, 3, n), FIELD_FACEFLUX: (q, self.nface)}.get(which, (q, n))

    def set_field(self, which: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, np.float64).reshape(self._shape(which))
        self._ck(lib().fvg_set_field(self._h, which, _dptr(a)), "fvg_set_field")

    def get_field(self, which: int) -> np.ndarray:
        a = np.empty(self._shape(which), np.float64)
        self._ck(lib().fvg_get_field(self._h, which, _dptr(a)), "fvg_get_field")
        return a

    def set_state(self, S: np.ndarray) -> None:
        self.set_field(FIELD_S, S)

    def get_state(self) -> np.ndarray:
        return self.get_field(FIELD_S)

    def device_ptr(self, which: int) -> int:
        p = C.c_void_p()
        self._ck(lib().fvg_device_ptr(self._h, which, C.byref(p)), "fvg_device_ptr")
        return p.value

    def faceinterp(self) -> None:
        self._ck(lib().fvg_faceinterp(self._h), "fvg_faceinterp")

    def gradcalc(self) -> None:
        self._ck(lib().fvg_gradcalc(self._h), "fvg_gradcalc")

    def rhs(self) -> None:
        self._ck(lib().fvg_rhs(self._h), "fvg_rhs")

    def rk(self, irk: int, dt: float) -> None:
        self._ck(lib().fvg_rk(self._h, int(irk), float(dt)), "fvg_rk")

    def step(self, dt: float, method: int = 3) -> None:
        self._ck(lib().fvg_step(self._h, int(method), float(dt)), "fvg_step")

    def steps(self, n: int, dt: float, method: int = 3) -> None:
        self._ck(lib().fvg_steps(self._h, int(n), int(method), float(dt)), "fvg_steps")

    def adapt_flags(self, refine: float, coarsen: float, base_level: int, max_level: int,
                    kind: int = 0) -> np.ndarray:
        """Split/merge flags (+1/-1/0 per cell) from the density gradient, on the device."""
        f = np.empty(self.ncell, np.int8)
        self._ck(lib().fvg_adapt_flags(self._h, int(kind), float(refine), float(coarsen), int(base_level),
                                       int(max_level), f.ctypes.data_as(C.POINTER(C.c_int8))),
                 "fvg_adapt_flags")
        return f

    def pressure(self) -> np.ndarray:
        """Cell pressure from the device state (FP64 EOS)."""
        p = np.empty(self.ncell, np.float64)
        self._ck(lib().fvg_pressure(self._h, _dptr(p)), "fvg_pressure")
        return p

    def cfl_dt(self, cfl: float) -> float:
        d = C.c_double()
        self._ck(lib().fvg_cfl_dt(self._h, float(cfl), C.byref(d)), "fvg_cfl_dt")
        return d.value

    def totals(self) -> np.ndarray:
        out = np.empty(self.neq + 1, np.float64)
        self._ck(lib().fvg_totals(self._h, _dptr(out)), "fvg_totals")
        return out

    def check_state(self) -> int:
        b = C.c_int32()
        self._ck(lib().fvg_check_state(self._h, C.byref(b)), "fvg_check_state")
        return b.value

    def sync(self) -> None:
        self._ck(lib().fvg_sync(self._h), "fvg_sync")

    def set_tiling(self, mode) -> None:
        """LDS-tiled stages for the interior 8^3 tiles: True / False force them on / off,
        "auto" (the default) lets the library choose (fvg_set_tiling)."""
        m = {"auto": 2, True: 1, False: 0}[mode] if isinstance(mode, (bool, str)) else int(mode)
        self._ck(lib().fvg_set_tiling(self._h, m), "fvg_set_tiling")

    def set_graphs(self, on: bool) -> None:
        """Steps six at a time as one captured HIP graph (off by default)."""
        self._ck(lib().fvg_set_graphs(self._h, 1 if on else 0), "fvg_set_graphs")

    @property
    def tile_count(self) -> int:
        return int(lib().fvg_tile_count(self._h))

    def time_steps(self, n: int, dt: float, method: int = 3):
        tot = C.c_double()
        k = (C.c_double * 8)()
        self._ck(lib().fvg_time_steps(self._h, int(n), int(method), float(dt), C.byref(tot), k),
                 "fvg_time_steps")
        return tot.value, list(k)

    # ---- device-pointer state copies (torch interop, stream-synchronous)
    def set_state_device(self, ptr: int) -> None:
        self._ck(lib().fvg_set_state_device(self._h, C.c_void_p(ptr)), "fvg_set_state_device")

    def get_state_device(self, ptr: int) -> None:
        self._ck(lib().fvg_get_state_device(self._h, C.c_void_p(ptr)), "fvg_get_state_device")

    # ---- adjoint / tangent of rollouts
    def adj_set_probe(self, mask: np.ndarray) -> None:
        m = np.ascontiguousarray(mask, np.uint8)
        if m.shape != (self.ncell,):
            raise ValueError("one mask byte per cell")
        self._ck(lib().fvg_adj_set_probe(self._h, m.ctypes.data_as(C.POINTER(C.c_uint8))),
                 "fvg_adj_set_probe")

    def adj_forward(self, nsteps: int, ckpt: int, dt: float) -> float:
        """The rollout with S kept every ckpt steps; ckpt = 0 keeps the whole rollout
        where it fits in half the free device memory (else the widest spacing that does)."""
        J = C.c_double()
        self._ck(lib().fvg_adj_forward(self._h, int(nsteps), int(ckpt), float(dt), C.byref(J)),
                 "fvg_adj_forward")
        return J.value

    @property
    def adj_checkpoint_spacing(self) -> int:
        """The checkpoint spacing of the last adj_forward (0 before one)."""
        return int(lib().fvg_adj_checkpoint_spacing(self._h))

    def adj_objective(self):
        """(hi, lo) of the last rollout's objective, J = hi + lo unevaluated."""
        out = np.empty(2, np.float64)
        self._ck(lib().fvg_adj_objective(self._h, _dptr(out)), "fvg_adj_objective")
        return float(out[0]), float(out[1])

    def adj_backward(self) -> np.ndarray:
        g = np.empty((self.neq, self.ncell), np.float64)
        self._ck(lib().fvg_adj_backward(self._h, _dptr(g), None), "fvg_adj_backward")
        return g

    def adj_backward_device(self, ptr: int) -> None:
        self._ck(lib().fvg_adj_backward(self._h, None, C.c_void_p(ptr)), "fvg_adj_backward")

    def tlm(self, direction: np.ndarray, nsteps: int, dt: float):
        d = np.ascontiguousarray(direction, np.float64).reshape(self.neq, self.ncell)
        J, dJ = C.c_double(), C.c_double()
        self._ck(lib().fvg_tlm(self._h, _dptr(d), int(nsteps), float(dt), C.byref(J), C.byref(dJ)),
                 "fvg_tlm")
        return J.value, dJ.value

    def full_rhs(self) -> np.ndarray:
        """faceinterp (MUSCL) + gradcalc (viscous) + RHS of the current S."""
        if self.params.ischeme == 2:
            self.faceinterp()
        if self.params.ns:
            self.gradcalc()
        self.rhs()
        return self.get_field(FIELD_RHS)
