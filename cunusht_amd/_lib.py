"""ctypes binding of libnusht_amd.so (the C ABI declared in include/nusht_amd.h).

The library is built in-tree (cunusht_amd/libnusht_amd.so) by cunusht_amd/csrc/Makefile.
Loading fails loudly if it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnusht_amd.so")

NUSHT_OK = 0
_ERRNAMES = {1: "argument", 2: "HIP", 3: "unsupported", 4: "state", 5: "out of memory"}

_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double


class NuShtError(RuntimeError):
    """Raised when a C-ABI call returns a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"nusht {_ERRNAMES.get(code, code)} error: {msg}")
        self.code = code


class RingDesc(ctypes.Structure):
    _fields_ = [("lmax", _i), ("spin", _i), ("nring", _i), ("theta", _vp), ("phi0", _vp), ("nph", _vp),
                ("ofs", _vp), ("weight", _vp)]


class PlanDesc(ctypes.Structure):
    _fields_ = [("lmax", _i), ("mmax", _i), ("spin", _i), ("ncomp", _i), ("epsilon", _d), ("npts_max", _i64),
                ("sigma", _d), ("max_maps", _i)]


JOB_GCLM2LENMAP, JOB_LENMAP2GCLM, JOB_NUSHT2D2, JOB_NUSHT2D1 = 0, 1, 2, 3


class Job(ctypes.Structure):
    _fields_ = [("kind", _i), ("plan", _vp), ("ring", _vp), ("dlm_scaled", _vp), ("polrot", _i), ("gamma", _vp),
                ("alm_in", _vp), ("alm_out", _vp), ("map_in0", _vp), ("map_in1", _vp), ("map_out0", _vp),
                ("map_out1", _vp)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("ntheta", _i), ("ntheta_dfs", _i), ("nfine", _i), ("nrow_fine", _i), ("kernel_width", _i),
                ("single_precision", _i), ("sigma", _d), ("beta", _d), ("workspace_bytes", _i64),
                ("legendre_pair_steps", _i64), ("legendre_variant", _i), ("legendre_ms", ctypes.c_float),
                ("nufft_tile_rows", _i), ("nufft_tile_cols", _i), ("nufft_candidates", _i),
                ("nufft_type2_ms", ctypes.c_float), ("nufft_type1_ms", ctypes.c_float), ("nufft_spread_rows", _i)]


class RingInfo(ctypes.Structure):
    _fields_ = [("npix", _i64), ("nring", _i), ("ring_pairs", _i), ("ring_dft", _i), ("legendre_pair_steps", _i64),
                ("legendre_variant", _i), ("legendre_ms", ctypes.c_float), ("ring_lengths", _i),
                ("workspace_bytes", _i64)]


# name -> argtypes (all return int unless listed in _RESTYPES)
SIGNATURES = {
    "nusht_last_error": [],
    "nusht_version": [],
    "nusht_doubling_f64": [_vp, _i, _i, _i, _vp, _vp],
    "nusht_doubling_f32": [_vp, _i, _i, _i, _vp, _vp],
    "nusht_doubling_c128": [_vp, _i, _i, _i, _vp, _vp],
    "nusht_adjoint_doubling_f64": [_vp, _i, _i, _i, _vp, _vp],
    "nusht_adjoint_doubling_f32": [_vp, _i, _i, _i, _vp, _vp],
    "nusht_adjoint_doubling_c128": [_vp, _i, _i, _i, _vp, _vp],
    "nusht_pointing_f64": [_vp, _vp, _vp, _vp, _i, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "nusht_magnification_f64": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i, _vp, _vp],
    "nusht_almxfl_c128": [_vp, _i, _i, _vp, _vp, _vp],
    "nusht_ringplan_create": [ctypes.POINTER(RingDesc), _vp, ctypes.POINTER(_vp)],
    "nusht_ringplan_destroy": [_vp],
    "nusht_ringplan_npix": [_vp],
    "nusht_ringplan_get_info": [_vp, ctypes.POINTER(RingInfo)],
    "nusht_legendre_variants": [],
    "nusht_ringplan_synthesis": [_vp, _vp, _i, _vp, _vp, _vp],
    "nusht_ringplan_adjoint": [_vp, _vp, _vp, _vp, _i, _vp],
    "nusht_ringplan_analysis": [_vp, _vp, _vp, _vp, _i, _vp],
    "nusht_ringplan_dlm2pointing": [_vp, _vp, _vp, _vp, _vp, _vp],
    "nusht_pointplan_create": [_i, _i, _i64, _vp, _vp, ctypes.POINTER(_vp)],
    "nusht_pointplan_destroy": [_vp],
    "nusht_pointplan_synthesis": [_vp, _vp, _i, _vp, _vp, _vp, _vp, _vp],
    "nusht_pointplan_synthesis_noalloc": [_i, _i, _i64, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp],
    "nusht_plan_create": [ctypes.POINTER(PlanDesc), _vp, ctypes.POINTER(_vp)],
    "nusht_plan_destroy": [_vp],
    "nusht_plan_get_info": [_vp, ctypes.POINTER(PlanInfo)],
    "nusht_plan_workspace_bytes": [_vp],
    "nusht_sht_synth_grad_gl": [_vp, _vp, _vp, _vp, _vp],
    "nusht_fft2d_c2c_f64": [_vp, _i, _i, _i, _i, _vp],
    "nusht_fft2d_c2c_f32": [_vp, _i, _i, _i, _i, _vp],
    "nusht_sht_synth_cc": [_vp, _vp, _vp, _vp],
    "nusht_sht_adjoint_cc": [_vp, _vp, _vp, _vp],
    "nusht_nufft_setpts": [_vp, _vp, _vp, _i64, _vp],
    "nusht_nufft_interp": [_vp, _vp, _vp, _vp, _vp, _vp],
    "nusht_nufft_spread": [_vp, _vp, _vp, _vp, _vp, _vp],
    "nusht_batch_execute": [ctypes.POINTER(Job), _i, _vp],
    "nusht_plan_set_pts": [_vp, _vp, _vp, _i64, _vp],
    "nusht_plan_set_pts_async": [_vp, _vp, _vp, _i64, _vp],
    "nusht_nusht2d2": [_vp, _vp, _vp, _vp, _vp, _vp],
    "nusht_nusht2d1": [_vp, _vp, _vp, _vp, _vp, _vp],
    "nusht_gclm2lenmap": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _vp],
    "nusht_lenmap2gclm": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _vp],
    "nusht_nusht2d2_batch": [_vp, _i, _vp, _i64, _vp, _vp, _i64, _vp, _vp],
    "nusht_gclm2lenmap_batch": [_vp, _vp, _i, _vp, _i64, _vp, _vp, _vp, _i64, _i, _vp],
    "nusht_timing_enable": [_i],
    "nusht_timing_reset": [],
    "nusht_timing_count": [],
    "nusht_timing_get": [_i, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(_d), ctypes.POINTER(_i64)],
}
_RESTYPES = {"nusht_last_error": ctypes.c_char_p, "nusht_ringplan_npix": _i64, "nusht_plan_workspace_bytes": _i64}

_LIB = None


def load():
    """Load the library (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `make -C cunusht_amd/csrc` (or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, _i)
    _LIB = lib
    return lib


def check(code: int) -> None:
    if code != NUSHT_OK:
        msg = load().nusht_last_error()
        raise NuShtError(code, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def timing_enable(on: bool = True) -> None:
    call("nusht_timing_enable", 1 if on else 0)


def timing_reset() -> None:
    call("nusht_timing_reset")


def timing_read() -> dict:
    """{stage: (total_ms, calls)} accumulated since the last reset (synchronizes on the timing events)."""
    lib = load()
    n = lib.nusht_timing_count()
    if n < 0:
        check(-n)
    out = {}
    for i in range(n):
        nm = ctypes.c_char_p()
        ms = _d()
        c = _i64()
        call("nusht_timing_get", i, ctypes.byref(nm), ctypes.byref(ms), ctypes.byref(c))
        out[nm.value.decode()] = (ms.value, c.value)
    return out
