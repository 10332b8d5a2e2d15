"""ctypes binding of ``lib/libpyhype_hip.so`` (C ABI declared in include/pyhype_hip.h).

Loading fails loudly: there is no CPU fallback for any operation of this package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libpyhype_hip.so"

PHX_ABI_VERSION = 3
PHX_OK, PHX_E_INVALID, PHX_E_HIP, PHX_E_UNREALIZABLE, PHX_E_STATE = 0, -1, -2, -3, -4
E, W, N, S = 0, 1, 2, 3
SIDE_NEIGHBOR, SIDE_OPEN, SIDE_WALL, SIDE_OUTLET = 0, 1, 2, 3
FLUX_ROE, FLUX_HLLL = 0, 1
RECON_CONSERVATIVE, RECON_PRIMITIVE = 0, 1
LIMITER_VENKATAKRISHNAN = 0
LIMITER_BARTH_JESPERSEN = 1
LIMITER_VAN_ALBADA = 2
LIMITER_VAN_LEER = 3
LIMITERS = {"Venkatakrishnan": LIMITER_VENKATAKRISHNAN, "BarthJespersen": LIMITER_BARTH_JESPERSEN,
            "VanAlbada": LIMITER_VAN_ALBADA, "VanLeer": LIMITER_VAN_LEER}
GEOM_GENERAL, GEOM_CARTESIAN = 0, 1

EXPORTS = (
    "phx_plan_create", "phx_plan_destroy", "phx_state_len", "phx_reset", "phx_run", "phx_run_frozen",
    "phx_run_record", "phx_stage", "phx_residual", "phx_dt", "phx_ghosts", "phx_status", "phx_residual_vjp",
    "phx_step_vjp", "phx_set_launch", "phx_set_variant", "phx_set_xcd_order", "phx_launch_info", "phx_last_error",
    "phx_residual_vjp_noalloc", "phx_residual_jvp", "phx_step_jvp",
)


class BlockDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("case_id", C.c_int32), ("geom_kind", C.c_int32),
        ("side_type", C.c_int32 * 4), ("neighbor", C.c_int32 * 4), ("state_offset", C.c_int64),
        ("geom", C.POINTER(C.c_double)), ("geom_len", C.c_int64),
    ]


class PlanDesc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("nblocks", C.c_int32), ("blocks", C.POINTER(BlockDesc)),
        ("ncases", C.c_int32), ("flux", C.c_int32), ("limiter", C.c_int32), ("nstages", C.c_int32),
        ("butcher_a", C.POINTER(C.c_double)), ("gamma", C.c_double), ("gm1", C.c_double),
        ("one_over_gm1", C.c_double), ("g_over_gm1", C.c_double), ("cfl", C.c_double),
        ("t_final", C.POINTER(C.c_double)), ("rows_per_item", C.c_int32), ("grid_waves", C.c_int32),
        ("reconstruction", C.c_int32),
        ("quadrature_points", C.c_int32),
    ]


class BackendError(RuntimeError):
    """A call into the HIP library failed."""


_lib = None


def load(path: Path = LIB_PATH):
    """Load the HIP library (once).  Raises ImportError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not path.exists():
        raise ImportError(f"{path} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(str(path))
    vp, i32, i64, dp = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double)
    sig = {
        "phx_plan_create": ([C.POINTER(PlanDesc), C.POINTER(vp)], C.c_int),
        "phx_plan_destroy": ([vp], C.c_int),
        "phx_state_len": ([vp], i64),
        "phx_reset": ([vp, dp, i32, vp], C.c_int),
        "phx_run": ([vp, vp, i32, i32, vp], C.c_int),
        "phx_run_frozen": ([vp, vp, i32, vp, vp], C.c_int),
        "phx_run_record": ([vp, vp, i32, vp, vp, vp], C.c_int),
        "phx_stage": ([vp, vp, i32, vp], C.c_int),
        "phx_residual": ([vp, vp, vp, vp], C.c_int),
        "phx_dt": ([vp, vp, vp, vp], C.c_int),
        "phx_ghosts": ([vp, vp, vp, vp], C.c_int),
        "phx_status": ([vp, C.POINTER(i32), C.POINTER(i32), dp, dp], C.c_int),
        "phx_residual_vjp": ([vp, vp, vp, vp, vp], C.c_int),
        "phx_residual_vjp_noalloc": ([vp, vp, vp, vp, vp], C.c_int),
        "phx_residual_jvp": ([vp, vp, vp, vp, vp], C.c_int),
        "phx_step_jvp": ([vp, vp, vp, vp, vp], C.c_int),
        "phx_step_vjp": ([vp, vp, vp, vp, vp], C.c_int),
        "phx_set_launch": ([vp, i32, i32, i32], C.c_int),
        "phx_set_variant": ([vp, i32, i32], C.c_int),
        "phx_set_xcd_order": ([vp, i32], C.c_int),
        "phx_launch_info": ([vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)], C.c_int),
        "phx_last_error": ([], C.c_char_p),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> None:
    if rc != PHX_OK:
        msg = _lib.phx_last_error().decode() if _lib is not None else ""
        if rc == PHX_E_INVALID:
            raise ValueError(f"{what}: {msg}")
        raise BackendError(f"{what} failed ({rc}): {msg}")
