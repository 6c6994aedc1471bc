"""cunusht_amd: non-uniform spherical harmonic transforms on AMD MI355X (gfx950).

Drop-in for cunuSHT's GPU backend: get_transformer(backend='GPU') returns the operator class with
nusht2d2, nusht2d1, gclm2lenmap, lenmap2gclm and dlm2pointing (cunusht/cunusht.py:63-98).
get_transformer(backend='CPU') returns the host operator class with the signatures of the reference's
CPU_DUCCnufft_transformer (cunusht/deflection/CPU_nufft_transformer.py:181-558).
"""
from __future__ import annotations

from .geometry import Geom, get_geom, good_size
from .helpers import dlm2A, get_spin_lower, get_spin_raise
from .utils import Alm, almxfl, synalm

__all__ = ["get_transformer", "get_geom", "Geom", "Alm", "almxfl", "synalm", "good_size", "dlm2A",
           "get_spin_raise", "get_spin_lower"]


def get_transformer(backend: str, solver: str = "default", mode: str = "nuFFT"):
    """Returns the transformer class for `backend` (cunusht/cunusht.py:63-82), mode 'nuFFT'.

    backend 'GPU': solver 'default' / 'cufinufft' / 'amd' -> GPU_nusht_transformer (libnusht_amd.so).
    backend 'CPU' (libnusht_cpu.so): solver 'duccnufft' / 'amd' -> CPU_nusht_transformer, the algorithm and
    signatures of the reference's CPU_DUCCnufft_transformer; solver 'default' / 'lenspyx' ->
    CPU_nusht_lenspyx_transformer, the same operators with the signatures of the reference's default CPU solver
    CPU_Lenspyx_transformer.  The reference's 'finufft' CPU solver (broken there) is not provided.
    """
    if backend not in ("GPU", "CPU"):
        raise AssertionError(f"{backend} backend not found")
    if mode not in ("nuFFT",):
        raise AssertionError(f"{mode} mode not found")
    if backend == "CPU":
        from .cpu_transformer import CPU_nusht_lenspyx_transformer, CPU_nusht_transformer
        if solver in ("default", "lenspyx"):
            return CPU_nusht_lenspyx_transformer
        if solver in ("duccnufft", "amd"):
            return CPU_nusht_transformer
        raise AssertionError(f"Solver not found: {solver}")
    if solver not in ("default", "cufinufft", "amd"):
        raise AssertionError(f"Solver not found: {solver}")
    from .transformer import GPU_nusht_transformer
    return GPU_nusht_transformer
