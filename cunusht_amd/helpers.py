"""Lensing helpers on the pointing grid (cunusht/helper_GPU.py:88-161).

get_spin_raise / get_spin_lower are the spin-raising / lowering response coefficients;
dlm2A is the determinant of the magnification matrix of a deflection field, computed on the GPU:
three syntheses on the pointing grid (spin 0, 1, 2; plus spin 0 for a curl part) and one
pointwise kernel of libnusht_amd.so.  It is used for inverse lensing: lenmap2gclm applied to a
lensed map multiplied by A (GPU_nufft_transformer.py:584-586).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .utils import Alm


def get_spin_raise(s: int, lmax: int) -> np.ndarray:
    """sqrt((l - s)(l + s + 1)) for |s| <= l <= lmax, else 0 (helper_GPU.py:88-97)."""
    ret = np.zeros(lmax + 1, dtype=float)
    ll = np.arange(abs(s), lmax + 1, dtype=np.float64)
    ret[abs(s):] = np.sqrt((ll - s) * (ll + s + 1))
    return ret


def get_spin_lower(s: int, lmax: int) -> np.ndarray:
    """-sqrt((l + s)(l - s + 1)) for |s| <= l <= lmax, else 0 (helper_GPU.py:99-110)."""
    ret = np.zeros(lmax + 1, dtype=float)
    ll = np.arange(abs(s), lmax + 1, dtype=np.float64)
    ret[abs(s):] = -np.sqrt((ll + s) * (ll - s + 1))
    return ret


def _almxfl(alm: torch.Tensor, fl: torch.Tensor, lmax: int, out: torch.Tensor) -> torch.Tensor:
    nc = 1 if alm.dim() == 1 else alm.shape[0]
    _lib.call("nusht_almxfl_c128", _ptr(alm), lmax, nc, _ptr(fl), _ptr(out), _stream())
    return out


def _ptr(t):
    import ctypes
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream():
    import ctypes
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def dlm2A(dlm, lmax_dlm: int, mmax_dlm: int, geom, nthreads=None, dclm=None, exact: bool = True) -> torch.Tensor:
    """Determinant of the magnification matrix on the pointing grid (helper_GPU.py:113-161).

    geom: a GPU_nusht_transformer (its pointing grid and ring plans are used) or a geominfo tuple.
    exact=True gives the solid-angle ratio dOmega'/dOmega of the deflection map; exact=False the
    reference's formula, whose shear sign in the |d|-dependent term (helper_GPU.py:126) agrees only
    through second order in |d|.  nthreads is accepted for signature compatibility and ignored.
    """
    from .transformer import GPU_nusht_transformer
    T = geom if isinstance(geom, GPU_nusht_transformer) else GPU_nusht_transformer(geom)
    assert lmax_dlm == T.lmax and mmax_dlm == T.lmax, "dlm2A: lmax/mmax must match the transformer's lmax"
    return T.dlm2A(dlm, dclm=dclm, exact=exact)


def magnification_from_alm(T, dlm: torch.Tensor, dclm, exact: bool) -> torch.Tensor:
    L = T.lmax
    n = Alm.getsize(L, L)
    dev = dict(dtype=torch.complex128, device="cuda")
    d2k = torch.as_tensor(-0.5 * get_spin_lower(1, L), device="cuda")
    d2g = torch.as_tensor(-0.5 * get_spin_raise(1, L), device="cuda")
    dg = torch.zeros((2, n), **dev)
    dg[0] = dlm
    if dclm is not None:
        dg[1] = dclm
    klm = _almxfl(dg[0], d2k, L, torch.empty(n, **dev))
    glm = _almxfl(dg, d2g, L, torch.empty((2, n), **dev))
    k = T.synthesis(klm, spin=0)
    g = T.synthesis(glm, spin=2)
    d = T.synthesis(dg, spin=1)
    w = None
    if dclm is not None:
        w = T.synthesis(_almxfl(dg[1], d2k, L, torch.empty(n, **dev)), spin=0)
    A = torch.empty(k.numel(), dtype=torch.float64, device="cuda")
    _lib.call("nusht_magnification_f64", _ptr(k), _ptr(g[0]), _ptr(g[1]), _ptr(d[0]), _ptr(d[1]), _ptr(w),
              A.numel(), 1 if exact else 0, _ptr(A), _stream())
    return A
