"""alm layout helpers (cunusht/utils.py:10-149 semantics, healpy m-major layout)."""
from __future__ import annotations

import numpy as np


class Alm:
    @staticmethod
    def getsize(lmax: int, mmax: int) -> int:
        """cunusht/utils.py:31."""
        return ((mmax + 1) * (mmax + 2)) // 2 + (mmax + 1) * (lmax - mmax)

    @staticmethod
    def getidx(lmax: int, l, m):
        """cunusht/utils.py:57."""
        return m * (2 * lmax + 1 - m) // 2 + l

    @staticmethod
    def getlmax(s: int, mmax: int | None) -> int:
        """cunusht/utils.py:60-82."""
        if mmax is not None and mmax >= 0:
            x = (2 * s + mmax ** 2 - mmax - 2) / (2 * mmax + 2)
        else:
            x = (-3 + np.sqrt(1 + 8 * s)) / 2
        return int(x) if x == np.floor(x) else -1


def almxfl(alm: np.ndarray, fl: np.ndarray, mmax: int | None = None, inplace: bool = False):
    """alm * fl[l] (cunusht/utils.py:84-119)."""
    lmax = Alm.getlmax(alm.shape[-1], mmax)
    mmax = lmax if mmax is None or mmax < 0 else mmax
    out = alm if inplace else np.array(alm, copy=True)
    for m in range(mmax + 1):
        b = Alm.getidx(lmax, m, m)
        out[..., b:b + lmax - m + 1] *= fl[m:lmax + 1]
    return None if inplace else out


def synalm(cl: np.ndarray, lmax: int, mmax: int | None = None, rng: np.random.Generator | None = None):
    """Gaussian alm with spectrum cl (cunusht/utils.py:121-149), seeded through `rng`."""
    rng = np.random.default_rng() if rng is None else rng
    mmax = lmax if mmax is None or mmax < 0 else mmax
    n = Alm.getsize(lmax, mmax)
    alm = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    almxfl(alm, np.sqrt(cl[:lmax + 1] * 0.5), mmax, True)
    alm[:lmax + 1] = alm[:lmax + 1].real * np.sqrt(2.0)
    return alm
