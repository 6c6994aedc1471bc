"""alm layout helpers and seeded synthetic inputs (oracle; test infrastructure only).

m-major healpy layout: idx = m (2 lmax + 1 - m) / 2 + l  (cunusht/utils.py:31,57).
"""
from __future__ import annotations

import numpy as np


def nalm(lmax: int, mmax: int | None = None) -> int:
    """Number of alm (cunusht/utils.py:31)."""
    mmax = lmax if mmax is None else mmax
    return ((mmax + 1) * (mmax + 2)) // 2 + (mmax + 1) * (lmax - mmax)


def almidx(lmax: int, l, m):
    """Index of (l, m) in the m-major layout (cunusht/utils.py:57)."""
    return m * (2 * lmax + 1 - m) // 2 + l


def almxfl(alm: np.ndarray, fl: np.ndarray, lmax: int) -> np.ndarray:
    """alm * fl[l] (cunusht/utils.py:84-119)."""
    out = np.array(alm, copy=True)
    for m in range(lmax + 1):
        b = almidx(lmax, m, m)
        out[..., b:b + lmax - m + 1] *= fl[m:lmax + 1]
    return out


def synalm(cl: np.ndarray, lmax: int, rng: np.random.Generator) -> np.ndarray:
    """Gaussian alm with spectrum cl, seeded (cunusht/utils.py:121-149 semantics).

    a_lm = sqrt(cl/2) (N + iN); m = 0 entries real and scaled by sqrt(2).
    """
    n = nalm(lmax)
    alm = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    alm = almxfl(alm, np.sqrt(cl[:lmax + 1] * 0.5), lmax)
    alm[:lmax + 1] = alm[:lmax + 1].real * np.sqrt(2.0)
    return alm


def alm2cl(alm: np.ndarray, lmax: int) -> np.ndarray:
    """Auto power spectrum (cunusht/utils.py:152-198 semantics)."""
    cl = np.zeros(lmax + 1)
    for m in range(lmax + 1):
        b = almidx(lmax, m, m)
        a = alm[b:b + lmax - m + 1]
        w = 1.0 if m == 0 else 2.0
        cl[m:] += w * np.abs(a) ** 2
    return cl / (2 * np.arange(lmax + 1) + 1)


def random_alm(lmax: int, spin: int, ncomp: int, rng: np.random.Generator) -> np.ndarray:
    """Uniform(-1,1) alm with real m=0 and zero l<spin (test/test_adjointOperator.py:27-36)."""
    n = nalm(lmax)
    res = rng.uniform(-1.0, 1.0, (ncomp, n)) + 1j * rng.uniform(-1.0, 1.0, (ncomp, n))
    res[:, 0:lmax + 1].imag = 0.0
    ofs = 0
    for s in range(spin):
        res[:, ofs:ofs + spin - s] = 0.0
        ofs += lmax + 1 - s
    return res


def myalmdot(a1: np.ndarray, a2: np.ndarray, lmax: int) -> float:
    """Real inner product of real-field alm (test/test_adjointOperator.py:39-48)."""
    a1 = np.asarray(a1)
    a2 = np.asarray(a2)
    v = np.sum(a1[:lmax + 1].real * a2[:lmax + 1].real)
    v += 2.0 * np.sum((a1[lmax + 1:] * np.conj(a2[lmax + 1:])).real)
    return float(v)
