"""ctypes declarations of libfv140823.so (include/fvgpu.h).

The library is loaded from the package's _lib/ directory.  There is no
fallback: if the library is missing the import fails with a message naming the
build step.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

_LIB_PATH = pathlib.Path(__file__).resolve().parent / "_lib" / "libfv140823.so"

FVG_OK = 0
FVG_ERR_ARG = -1
FVG_ERR_HIP = -2
FVG_ERR_STATE = -3
FVG_ERR_NONFINITE = -4
FVG_ERR_UNSUPPORTED = -5
FVG_ERR_NOMEM = -6
FVG_ERR_IO = -7

FVG_PLAIN, FVG_MIXED = 0, 1
FVG_MINMOD, FVG_VANALBADA = 0, 1
FVG_WAVES_DAVIS, FVG_WAVES_PVRS = 0, 1
FIELD_S, FIELD_SN, FIELD_RHS, FIELD_SF, FIELD_SG, FIELD_FACEFLUX, FIELD_STAGE_IN = range(7)
FVG_TILING_OFF, FVG_TILING_ON, FVG_TILING_AUTO = range(3)

ERR_NAMES = {
    FVG_ERR_ARG: "bad argument",
    FVG_ERR_HIP: "HIP runtime error",
    FVG_ERR_STATE: "call out of order",
    FVG_ERR_NONFINITE: "non-finite or non-physical state",
    FVG_ERR_UNSUPPORTED: "unsupported option",
    FVG_ERR_NOMEM: "out of memory",
    FVG_ERR_IO: "file error",
}


class FvgParams(C.Structure):
    _fields_ = [
        ("model", C.c_int),
        ("ns", C.c_int),
        ("ischeme", C.c_int),
        ("limiter", C.c_int),
        ("precision", C.c_int),
        ("visc_work", C.c_int),
        ("gamma", C.c_double),
        ("pinf", C.c_double),
        ("mu", C.c_double),
        ("ymin", C.c_double),
        ("baro_nexp", C.c_double),
        ("baro_beta", C.c_double),
        ("baro_cbar", C.c_double),
        ("baro_psat", C.c_double),
        ("baro_rsat", C.c_double),
        ("mix_RG", C.c_double),
        ("mix_Tref", C.c_double),
        ("mix_rsat", C.c_double),
        ("mix_psat", C.c_double),
        ("mix_cliq", C.c_double),
        ("mix_cmix", C.c_double),
        ("inlet", C.c_double * 5),
        ("vanalbada_eps", C.c_double),
        ("iprimtv", C.c_int),
        ("wave_speeds", C.c_int),
        ("prandtl", C.c_double),
    ]


class FvgMesh(C.Structure):
    _fields_ = [
        ("ncell", C.c_int32),
        ("nface", C.c_int32),
        ("nface_axis", C.c_int32 * 3),
        ("nquad", C.c_int32),
        ("nlevel", C.c_int32),
        ("neq", C.c_int32),
        ("cflag", C.POINTER(C.c_uint32)),
        ("nbr", C.POINTER(C.c_int32)),
        ("quad_nbr", C.POINTER(C.c_int32)),
        ("face_L", C.POINTER(C.c_int32)),
        ("face_R", C.POINTER(C.c_int32)),
        ("face_meta", C.POINTER(C.c_uint32)),
        ("gslot", C.POINTER(C.c_int32)),
        ("h", C.POINTER(C.c_double)),
        ("dist", C.POINTER(C.c_double)),
        ("area", C.POINTER(C.c_double)),
        ("vol", C.POINTER(C.c_double)),
        ("nface_reg", C.c_int32 * 3),
        ("face_LL", C.POINTER(C.c_int32)),
        ("face_RR", C.POINTER(C.c_int32)),
    ]


_dp = C.POINTER(C.c_double)
_vp = C.c_void_p
_ip = C.POINTER(C.c_int)

# (name, restype, argtypes)
_DECLS = [
    ("fvg_params_default", None, [C.POINTER(FvgParams), C.c_int]),
    ("fvg_neq", C.c_int, [C.c_int]),
    ("fvh_forest_create", C.c_int, [_ip, _dp, _dp, C.c_int, _ip, _ip, C.c_int, C.POINTER(_vp)]),
    ("fvh_forest_destroy", None, [_vp]),
    ("fvh_forest_run", _vp, [_vp]),
    ("fvh_forest_nleaves", C.c_int, [_vp]),
    ("fvh_forest_max_level", C.c_int, [_vp]),
    ("fvh_forest_geometry", C.c_int, [_vp, _ip, _dp, _dp, _ip, _ip, _ip]),
    ("fvh_forest_leaf_info", C.c_int, [_vp, _dp, C.POINTER(C.c_int32)]),
    ("fvh_forest_set_field", C.c_int, [_vp, C.c_int, _dp]),
    ("fvh_forest_get_field", C.c_int, [_vp, C.c_int, _dp]),
    ("fvh_forest_adapt", C.c_int, [_vp, C.POINTER(C.c_int8), C.c_int]),
    ("fvh_forest_save", C.c_int, [_vp, C.c_char_p, C.c_int32, C.c_double]),
    ("fvh_forest_load", C.c_int, [C.c_char_p, C.POINTER(_vp), C.POINTER(C.c_int32), _dp]),
    ("fvh_forest_save_text", C.c_int, [_vp, C.c_char_p, C.c_int32, C.c_double]),
    ("fvh_forest_load_text", C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_int32), _dp]),
    ("fvh_export_tecplot", C.c_int, [_vp, C.c_char_p, C.c_int, _dp]),
    ("fvh_flatten_run", C.c_int, [_vp, C.POINTER(_vp)]),
    ("fvh_flat_mesh", C.POINTER(FvgMesh), [_vp]),
    ("fvh_flat_destroy", None, [_vp]),
    ("fvg_create", C.c_int, [C.c_int, C.POINTER(FvgParams), C.POINTER(_vp)]),
    ("fvg_destroy", C.c_int, [_vp]),
    ("fvg_set_params", C.c_int, [_vp, C.POINTER(FvgParams)]),
    ("fvg_set_mesh", C.c_int, [_vp, C.POINTER(FvgMesh)]),
    ("fvg_ncell", C.c_int, [_vp]),
    ("fvg_set_tiling", C.c_int, [_vp, C.c_int]),
    ("fvg_tile_count", C.c_int, [_vp]),
    ("fvg_set_graphs", C.c_int, [_vp, C.c_int]),
    ("fvh_tile_table", C.c_int, [C.POINTER(FvgMesh), C.POINTER(C.c_int32), C.c_int, C.POINTER(C.c_int)]),
    ("fvg_set_state", C.c_int, [_vp, _dp]),
    ("fvg_get_state", C.c_int, [_vp, _dp]),
    ("fvg_set_field", C.c_int, [_vp, C.c_int, _dp]),
    ("fvg_get_field", C.c_int, [_vp, C.c_int, _dp]),
    ("fvg_device_ptr", C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    ("fvg_stream", _vp, [_vp]),
    ("fvg_faceinterp", C.c_int, [_vp]),
    ("fvg_gradcalc", C.c_int, [_vp]),
    ("fvg_rhs", C.c_int, [_vp]),
    ("fvg_rk", C.c_int, [_vp, C.c_int, C.c_double]),
    ("fvg_step", C.c_int, [_vp, C.c_int, C.c_double]),
    ("fvg_steps", C.c_int, [_vp, C.c_int, C.c_int, C.c_double]),
    ("fvg_cfl_dt", C.c_int, [_vp, C.c_double, _dp]),
    ("fvg_totals", C.c_int, [_vp, _dp]),
    ("fvg_check_state", C.c_int, [_vp, C.POINTER(C.c_int32)]),
    ("fvg_sync", C.c_int, [_vp]),
    ("fvg_last_error", C.c_char_p, [_vp]),
    ("fvg_time_steps", C.c_int, [_vp, C.c_int, C.c_int, C.c_double, _dp, _dp]),
    ("fvg_adj_set_probe", C.c_int, [_vp, C.POINTER(C.c_uint8)]),
    ("fvg_adapt_flags", C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                  C.POINTER(C.c_int8)]),
    ("fvg_adapt_flags_static", C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                         C.POINTER(C.c_int8)]),
    ("fvg_pressure", C.c_int, [_vp, _dp]),
    ("fvg_remesh_begin", C.c_int, [_vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32)]),
    ("fvg_remesh_end", C.c_int, [_vp, C.c_int32, _dp, _dp, _dp, _dp, C.c_int]),
    ("fvg_has_keys", C.c_int, [_vp]),
    ("fvh_forest_remesh_device", C.c_int, [_vp, _vp, C.c_int]),
    ("fvh_forest_leaves", _vp, [_vp]),
    ("fvh_tables_from_owners", C.c_int, [_vp, C.c_int, C.POINTER(C.c_int32), _dp, _dp, _dp, _dp,
                                         C.POINTER(C.c_int32)]),
    ("fvg_get_mesh", C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_uint32), C.POINTER(C.c_int32),
                               C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                               C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                               C.POINTER(C.c_int32), _dp]),
    ("fvg_adj_forward", C.c_int, [_vp, C.c_int, C.c_int, C.c_double, _dp]),
    ("fvg_adj_checkpoint_spacing", C.c_int, [_vp]),
    ("fvg_adj_objective", C.c_int, [_vp, _dp]),
    ("fvg_adj_backward", C.c_int, [_vp, _dp, _vp]),
    ("fvg_tlm", C.c_int, [_vp, _dp, C.c_int, C.c_double, _dp, _dp]),
    ("fvg_eos_host", C.c_int, [C.POINTER(FvgParams), _dp, _dp]),
    ("fvg_face_flux_host", C.c_int, [C.POINTER(FvgParams), C.c_int, C.c_int, _dp, _dp, _dp]),
    ("fvg_face_vjp_host", C.c_int, [C.POINTER(FvgParams), C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]),
    ("fvg_face_flux_ns_host", C.c_int, [C.POINTER(FvgParams), C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]),
    ("fvg_face_vjp_ns_host", C.c_int, [C.POINTER(FvgParams), C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp,
                                       _dp, _dp, _dp]),
    ("fvg_face_jvp_host", C.c_int, [C.POINTER(FvgParams), C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]),
    ("fvg_limiter_host", C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double, _dp, _dp, _dp]),
    ("fvg_limiter_axis_host", C.c_int, [C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.POINTER(C.c_double)]),
    ("fvg_stream_copy", C.c_int, [C.c_int, _vp, _vp, C.c_size_t, C.c_int, _dp]),
    ("fvg_set_state_device", C.c_int, [_vp, _vp]),
    ("fvg_get_state_device", C.c_int, [_vp, _vp]),
    ("fv_gpu_attach", C.c_int, [_vp, C.c_int, C.POINTER(FvgParams), C.c_int]),
    ("fv_gpu_detach", C.c_int, [_vp]),
    ("fv_gpu_remesh", C.c_int, [_vp]),
    ("fv_gpu_upload", C.c_int, [_vp]),
    ("fv_gpu_download", C.c_int, [_vp]),
    ("fv_gpu_context", _vp, [_vp]),
    ("fv_gpu_last_status", C.c_int, [_vp]),
    ("RHS", None, [_vp]),
    ("RK", None, [C.c_int, _vp]),
    ("faceinterp", None, [_vp]),
    ("gradcalc", None, [_vp]),
]

EXPORTED = [d[0] for d in _DECLS]

_lib = None


def lib() -> C.CDLL:
    """The loaded native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise ImportError(
                f"{_LIB_PATH} is missing: build it with `python -m fv_140823_amd.build` "
                "(or __graft_entry__.build()); the HIP path has no fallback"
            )
        # torch ships its own HIP runtime under the same soname (libamdhip64.so.7)
        # but asks for it by its unversioned name, so if this library were loaded
        # first the process would end up with two HIP runtimes and torch would see
        # no GPU.  Loading torch first makes this library bind to torch's runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(os.fspath(_LIB_PATH))
        for name, res, args in _DECLS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def lib_path() -> pathlib.Path:
    return _LIB_PATH


class FvError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str = ""):
        self.code = code
        super().__init__(f"{where}: {ERR_NAMES.get(code, code)} ({code}) {detail}".rstrip())


def check(code: int, where: str, ctx=None) -> None:
    if code != FVG_OK:
        detail = ""
        if ctx is not None:
            msg = lib().fvg_last_error(ctx)
            detail = msg.decode() if msg else ""
        raise FvError(code, where, detail)
