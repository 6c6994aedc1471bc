"""Ring geometries and FFT sizes (oracle; test infrastructure only).

Follows cunusht/geometry.py:228-245 (CC), :266-284 (GL) and the ducc0 good_size
rule used at cunusht/deflection/GPU_nufft_transformer.py:138-141 (ducc0 itself is
absent here: the rule is restated -- smallest 2^a 3^b 5^c 7^d 11^e >= n for complex
transforms, smallest 2^a 3^b 5^c >= n for real ones).

Parity: good_size pinned against the sizes the reference logged
(first_steps/quickstart.ipynb:209 ntheta=1584, nphi=3168 at lmax 1575; :239 GL nphi 3200).
GL nodes/weights pinned against numpy.polynomial.legendre.leggauss and the exactness of
the quadrature on polynomials of degree <= 2n-1.  Pixel weights carry ducc0's factor
2 pi / nphi (ducc0.misc.GL_weights, used at cunusht/geometry.py:280), pinned by their sum
(4 pi) and by analysis(synthesis(alm)) = alm.
"""
from __future__ import annotations

import numpy as np


def _smooth(n: int, primes) -> bool:
    for p in primes:
        while n % p == 0:
            n //= p
    return n == 1


def good_size(n: int, real: bool = False) -> int:
    primes = (2, 3, 5) if real else (2, 3, 5, 7, 11)
    k = max(int(n), 1)
    while not _smooth(k, primes):
        k += 1
    return k


def _legendre_pd(n: int, theta: np.ndarray):
    """P_n, P_{n-1} and D_n = P_n - P_{n-1} at cos(theta), from y = 1 - cos(theta) = 2 sin^2(theta/2).

    Difference form of the three-term recurrence: (k+1) D_{k+1} = k D_k - (2k+1) y P_k,
    P_{k+1} = P_k + D_{k+1}; it never forms 1 - y, so it keeps full relative accuracy
    near the poles where x = cos(theta) cannot resolve the nodes."""
    y = 2.0 * np.sin(0.5 * theta) ** 2
    p = np.ones_like(theta)
    d = -y  # D_1 = P_1 - P_0 = -y
    pm = p
    p = p + d
    for k in range(1, n):
        d = (k * d - (2 * k + 1) * y * p) / (k + 1)
        pm = p
        p = p + d
    return p, pm, d, y


def gl_nodes(n: int):
    """Gauss-Legendre nodes theta (increasing) and weights (sum 2), Newton iteration in theta."""
    # northern half only (y <= 1 keeps the difference recurrence well conditioned); mirror the rest
    nh = (n + 1) // 2
    k = np.arange(1, nh + 1)
    theta = np.pi * (k - 0.25) / (n + 0.5)
    if n % 2 == 1:
        theta[-1] = 0.5 * np.pi
    for _ in range(100):
        p, pm, d, y = _legendre_pd(n, theta)
        # d/dtheta P_n(cos theta) = n (D_n - y P_n) / sin(theta)
        dp = n * (d - y * p) / np.sin(theta)
        step = p / dp
        if n % 2 == 1:
            step[-1] = 0.0
        theta = theta - step
        if np.max(np.abs(step)) < 1e-17:
            break
    p, pm, d, y = _legendre_pd(n, theta)
    w = 2.0 * np.sin(theta) ** 2 / (n * pm) ** 2
    ns = n // 2
    theta = np.concatenate([theta, np.pi - theta[:ns][::-1]])
    w = np.concatenate([w, w[:ns][::-1]])
    return theta, w


class Rings:
    """Iso-latitude rings: theta, phi0, nph, ofs (ring-major pixel order), weights."""

    def __init__(self, theta, phi0, nph, ofs, w):
        self.theta = np.asarray(theta, dtype=np.float64)
        self.phi0 = np.asarray(phi0, dtype=np.float64)
        self.nph = np.asarray(nph, dtype=np.int64)
        self.ofs = np.asarray(ofs, dtype=np.int64)
        self.weight = np.asarray(w, dtype=np.float64)

    def npix(self) -> int:
        return int(np.sum(self.nph))

    def pixel_angles(self):
        th = np.repeat(self.theta, self.nph)
        ph = np.concatenate([self.phi0[i] + 2 * np.pi * np.arange(self.nph[i]) / self.nph[i]
                             for i in range(self.theta.size)])
        return th, ph


def gl_geometry(lmax: int, nphi: int | None = None) -> Rings:
    """cunusht/geometry.py:266-284 (default nphi = good_size(2 lmax + 1, real=True))."""
    n = lmax + 1
    if nphi is None:
        nphi = good_size(2 * lmax + 1, True)
    th, w = gl_nodes(n)
    nph = np.full(n, nphi, dtype=np.int64)
    ofs = np.arange(n, dtype=np.int64) * nphi
    return Rings(th, np.zeros(n), nph, ofs, 2.0 * np.pi * w / nph)


def cc_thetas(ntheta: int) -> np.ndarray:
    """cunusht/geometry.py:240: np.linspace(0, pi, ntheta)."""
    return np.linspace(0.0, np.pi, ntheta)


def reference_cc_sizes(lmax: int):
    """(ntheta, nphi) of the reference's CC grid (GPU_nufft_transformer.py:138-140)."""
    nt = (good_size(lmax + 2) + 3) // 4 * 4
    nph = 2 * good_size(lmax + 1)
    return nt, nph
