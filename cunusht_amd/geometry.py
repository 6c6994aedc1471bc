"""Iso-latitude pointing geometries (Gauss-Legendre, Clenshaw-Curtis, Fejer-1) and FFT sizes.

Mirrors cunusht/geometry.py: Geom (:29-57), get_gl_geometry (:266-284), get_cc_geometry
(:228-245), get_f1_geometry (:247-264), get_geom (:19-27) and the ducc0 good_size rule used at
cunusht/deflection/GPU_nufft_transformer.py:138-141 (2,3,5,7,11-smooth for complex transforms,
2,3,5-smooth for real ones).  Gauss-Legendre nodes are computed here (ducc0 is not a dependency):
Newton's iteration in theta on P_n(cos theta), evaluated with the difference form of the
Legendre recurrence so that nodes and weights keep full relative accuracy near the poles.

Pixel weights follow the ducc0 convention the reference's geometries are built on (ducc0.misc
GL_weights, ducc0.sht.experimental.get_gridweights): ring quadrature weight x 2 pi / nphi, so the
weights of all pixels sum to 4 pi and sum(weight * map) integrates the map over the sphere.
Clenshaw-Curtis and Fejer-1 weights are the closed forms of the two rules (Trefethen, Spectral
Methods in MATLAB, clencurt; Waldvogel, BIT 46 (2006) 195).
"""
from __future__ import annotations

import inspect

import numpy as np


def good_size(n: int, real: bool = False) -> int:
    primes = (2, 3, 5) if real else (2, 3, 5, 7, 11)
    k = max(int(n), 1)
    while True:
        m = k
        for p in primes:
            while m % p == 0:
                m //= p
        if m == 1:
            return k
        k += 1


def _pn_diff(n: int, theta: np.ndarray):
    y = 2.0 * np.sin(0.5 * theta) ** 2     # 1 - cos(theta)
    p = 1.0 - y                             # P_1
    d = -y                                  # D_1 = P_1 - P_0
    pm = np.ones_like(theta)
    for k in range(1, n):
        d = (k * d - (2 * k + 1) * y * p) / (k + 1)
        pm = p
        p = p + d
    if n == 1:
        pm = np.ones_like(theta)
    return p, pm, d, y


def gauss_legendre_thetas(n: int):
    """(theta, weights) of the n-point Gauss-Legendre rule, theta increasing, sum(weights) = 2."""
    nh = (n + 1) // 2
    k = np.arange(1, nh + 1)
    th = np.pi * (k - 0.25) / (n + 0.5)
    odd_mid = n % 2 == 1
    if odd_mid:
        th[-1] = 0.5 * np.pi
    for _ in range(100):
        p, pm, d, y = _pn_diff(n, th)
        step = p / (n * (d - y * p) / np.sin(th))
        if odd_mid:
            step[-1] = 0.0
        th = th - step
        if np.max(np.abs(step)) < 1e-17:
            break
    p, pm, d, y = _pn_diff(n, th)
    w = 2.0 * np.sin(th) ** 2 / (n * pm) ** 2
    ns = n // 2
    return np.concatenate([th, np.pi - th[:ns][::-1]]), np.concatenate([w, w[:ns][::-1]])


def cc_weights(n: int) -> np.ndarray:
    """Clenshaw-Curtis weights on cos(theta_j), theta_j = j pi / (n - 1), j = 0..n-1 (sum 2)."""
    if n == 1:
        return np.array([2.0])
    N = n - 1
    th = np.pi * np.arange(n) / N
    k = np.arange(1, N // 2 + 1)
    b = np.where(2 * k == N, 1.0, 2.0)
    v = 1.0 - np.sum(b[None, :] * np.cos(2.0 * k[None, :] * th[:, None]) / (4.0 * k[None, :] ** 2 - 1.0), axis=1)
    c = np.full(n, 2.0)
    c[0] = c[-1] = 1.0
    return c * v / N


def f1_weights(n: int) -> np.ndarray:
    """Fejer's first rule on cos(theta_j), theta_j = (j + 1/2) pi / n (sum 2)."""
    th = np.pi * (np.arange(n) + 0.5) / n
    k = np.arange(1, n // 2 + 1)
    v = 1.0 - 2.0 * np.sum(np.cos(2.0 * k[None, :] * th[:, None]) / (4.0 * k[None, :] ** 2 - 1.0), axis=1)
    return 2.0 * v / n


class Geom:
    """Iso-latitude pixelisation (cunusht/geometry.py:29-57): rings sorted by ofs."""

    def __init__(self, theta, phi0, nphi, ringstart, w):
        theta = np.asarray(theta, dtype=np.float64)
        order = np.argsort(np.asarray(ringstart))
        self.theta = theta[order]
        self.phi0 = np.asarray(phi0, dtype=np.float64)[order]
        self.nph = np.asarray(nphi, dtype=np.int64)[order]
        self.ofs = np.asarray(ringstart, dtype=np.int64)[order]
        self.weight = np.asarray(w, dtype=np.float64)[order]

    def npix(self) -> int:
        return int(np.sum(self.nph))

    def pixel_angles(self):
        th = np.repeat(self.theta, self.nph)
        ph = np.concatenate([self.phi0[i] + 2 * np.pi * np.arange(self.nph[i]) / self.nph[i]
                             for i in range(self.theta.size)])
        return th, ph

    @staticmethod
    def get_gl_geometry(lmax: int, good_size_real: bool = True, nphi: int | None = None):
        n = lmax + 1
        if nphi is None:
            nphi = good_size(2 * lmax + 1, good_size_real)
        th, wt = gauss_legendre_thetas(n)
        nph = np.full(n, nphi, dtype=np.int64)
        ofs = np.arange(n, dtype=np.int64) * nphi
        return Geom(th, np.zeros(n), nph, ofs, 2.0 * np.pi * wt / nph)

    @staticmethod
    def get_cc_geometry(ntheta: int, nphi: int):
        th = np.linspace(0, np.pi, ntheta)
        nph = np.full(ntheta, nphi, dtype=np.int64)
        ofs = np.arange(ntheta, dtype=np.int64) * nphi
        return Geom(th, np.zeros(ntheta), nph, ofs, 2.0 * np.pi * cc_weights(ntheta) / nphi)

    @staticmethod
    def get_f1_geometry(ntheta: int, nphi: int):
        th = np.pi * (np.arange(ntheta) + 0.5) / ntheta
        nph = np.full(ntheta, nphi, dtype=np.int64)
        ofs = np.arange(ntheta, dtype=np.int64) * nphi
        return Geom(th, np.zeros(ntheta), nph, ofs, 2.0 * np.pi * f1_weights(ntheta) / nphi)


def get_geom(geometry=("gl", {"lmax": 255})) -> Geom:
    """cunusht/geometry.py:19-27."""
    geo = getattr(Geom, "get_" + geometry[0] + "_geometry", None)
    if geo is None:
        raise ValueError(f"Geometry {geometry[0]} not supported (gl, cc, f1)")
    # lmax / mmax travel in the same dict for the transformer; geometries that do not take them ignore them
    params = inspect.signature(geo).parameters
    kw = {k: v for k, v in geometry[1].items() if k in params or k not in ("lmax", "mmax")}
    return geo(**kw)
