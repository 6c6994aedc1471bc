"""ctypes bindings to oracle/liboracle.so (oracle; test infrastructure only)."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.oracle_sht_synth_rings.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
        L.oracle_sht_adjoint_rings.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
        L.oracle_es_interp.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_long, _dp, _dp, ctypes.c_int,
                                       ctypes.c_double, _dp]
        L.oracle_es_spread.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_long, _dp, _dp, _dp, ctypes.c_int,
                                       ctypes.c_double]
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def sht_synth_rings(alm, lmax: int, spin: int, theta) -> np.ndarray:
    """Ring coefficients G_m(theta_i).  spin 0: (lmax+1, nring), m = 0..lmax.
    spin > 0: (2 lmax + 1, nring), row m + lmax.  alm: (nalm,) or (ncomp, nalm)."""
    alm = np.ascontiguousarray(np.atleast_2d(alm), dtype=np.complex128)
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    nm = lmax + 1 if spin == 0 else 2 * lmax + 1
    G = np.zeros((nm, theta.size), dtype=np.complex128)
    rc = lib().oracle_sht_synth_rings(lmax, spin, alm.shape[0], _p(alm.view(np.float64)), theta.size, _p(theta),
                                      _p(G.view(np.float64)))
    assert rc == 0
    return G


def sht_adjoint_rings(G, lmax: int, spin: int, ncomp: int, theta) -> np.ndarray:
    """Exact adjoint of sht_synth_rings; returns (ncomp, nalm)."""
    G = np.ascontiguousarray(G, dtype=np.complex128)
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    n = (lmax + 1) * (lmax + 2) // 2
    alm = np.zeros((ncomp, n), dtype=np.complex128)
    rc = lib().oracle_sht_adjoint_rings(lmax, spin, ncomp, _p(G.view(np.float64)), theta.size, _p(theta),
                                        _p(alm.view(np.float64)))
    assert rc == 0
    return alm


def es_interp(grid, x1, x2, w: int, beta: float) -> np.ndarray:
    grid = np.ascontiguousarray(grid, dtype=np.complex128)
    x1 = np.ascontiguousarray(x1, dtype=np.float64)
    x2 = np.ascontiguousarray(x2, dtype=np.float64)
    out = np.zeros(x1.size, dtype=np.complex128)
    rc = lib().oracle_es_interp(grid.shape[0], grid.shape[1], _p(grid.view(np.float64)), x1.size, _p(x1), _p(x2),
                                w, beta, _p(out.view(np.float64)))
    assert rc == 0
    return out


def es_spread(shape, x1, x2, c, w: int, beta: float) -> np.ndarray:
    grid = np.zeros(shape, dtype=np.complex128)
    x1 = np.ascontiguousarray(x1, dtype=np.float64)
    x2 = np.ascontiguousarray(x2, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.complex128)
    rc = lib().oracle_es_spread(shape[0], shape[1], _p(grid.view(np.float64)), x1.size, _p(x1), _p(x2),
                                _p(c.view(np.float64)), w, beta)
    assert rc == 0
    return grid
