"""cunusht_amd: non-uniform spherical harmonic transforms on AMD MI355X (gfx950).

Drop-in for cunuSHT's GPU backend: get_transformer(backend='GPU') returns the operator class with
nusht2d2, nusht2d1, gclm2lenmap, lenmap2gclm and dlm2pointing (cunusht/cunusht.py:63-98).
"""
from __future__ import annotations

from .geometry import Geom, get_geom, good_size
from .helpers import dlm2A, get_spin_lower, get_spin_raise
from .utils import Alm, almxfl, synalm

__all__ = ["get_transformer", "get_geom", "Geom", "Alm", "almxfl", "synalm", "good_size", "dlm2A",
           "get_spin_raise", "get_spin_lower"]


def get_transformer(backend: str, solver: str = "default", mode: str = "nuFFT"):
    """Returns the transformer class for `backend` (cunusht/cunusht.py:63-82).

    backend 'GPU' (solver 'default' / 'cufinufft' / 'amd'), mode 'nuFFT'.  The reference's CPU
    backend (ducc0 / lenspyx) is not part of this package.
    """
    if backend not in ("GPU",):
        raise AssertionError(f"{backend} backend not found")
    if mode not in ("nuFFT",):
        raise AssertionError(f"{mode} mode not found")
    if solver not in ("default", "cufinufft", "amd"):
        raise AssertionError(f"Solver not found: {solver}")
    from .transformer import GPU_nusht_transformer
    return GPU_nusht_transformer
