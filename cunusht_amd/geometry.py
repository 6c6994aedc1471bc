"""Iso-latitude pointing geometries (Gauss-Legendre, thinned Gauss-Legendre, HEALPix, Clenshaw-Curtis,
Fejer-1) and FFT sizes.

Mirrors cunusht/geometry.py: st2mmax (:9-17), Geom and its helpers (:29-192), the geometries
(:194-297), pbounds (:300-326), get_geom (:19-27) and the ducc0 good_size rule used at
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

import functools
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
    th, w = _gauss_legendre_cached(n)
    return th.copy(), w.copy()


@functools.lru_cache(maxsize=32)
def _gauss_legendre_cached(n: int):
    nh = (n + 1) // 2
    k = np.arange(1, nh + 1)
    th = np.pi * (k - 0.25) / (n + 0.5)
    odd_mid = n % 2 == 1
    if odd_mid:
        th[-1] = 0.5 * np.pi
    # Newton converges quadratically: once every step is below 1e-14 of its node, one more step leaves the
    # nodes at rounding level (steps then stall at ulp size instead of reaching zero)
    last = False
    for _ in range(100):
        p, pm, d, y = _pn_diff(n, th)
        step = p / (n * (d - y * p) / np.sin(th))
        if odd_mid:
            step[-1] = 0.0
        th = th - step
        if last:
            break
        last = bool(np.max(np.abs(step) / th) < 1e-14)
    p, pm, d, y = _pn_diff(n, th)
    w = 2.0 * np.sin(th) ** 2 / (n * pm) ** 2
    ns = n // 2
    th, w = np.concatenate([th, np.pi - th[:ns][::-1]]), np.concatenate([w, w[:ns][::-1]])
    th.setflags(write=False)
    w.setflags(write=False)
    return th, w


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


def st2mmax(spin, tht, lmax):
    """Largest m that contributes at colatitude tht (libsharp polar optimisation, Reinecke & Seljebotn 2013
    Eqs. 7-8; cunusht/geometry.py:9-17): the positive root of m^2 + 2 s cos(tht) m - (T + lmax sin(tht))^2 + s^2
    with T = max(0.01 lmax, 100)."""
    T = max(0.01 * lmax, 100)
    b = -2 * spin * np.cos(tht)
    c = -(T + lmax * np.sin(tht)) ** 2 + spin ** 2
    return 0.5 * (-b + np.sqrt(b * b - 4 * c))


def healpix_rings(nside: int):
    """(theta, phi0, nph, ofs) of the HEALPix RING scheme (Gorski et al. 2005, ApJ 622, 759, Sec. 4):
    polar-cap ring i < nside: z = 1 - i^2 / (3 nside^2), 4 i pixels from phi0 = pi / (4 i); equatorial ring
    nside <= i <= 3 nside: z = 4/3 - 2 i / (3 nside), 4 nside pixels from phi0 = pi / (4 nside) when i + nside is
    even, else 0; the south cap mirrors the north.  Cap colatitudes use theta = 2 asin(i / (sqrt(6) nside)),
    which keeps full relative accuracy at the poles."""
    nside = int(nside)
    assert nside >= 1, nside
    nr = 4 * nside - 1
    i = np.arange(1, nr + 1)
    inorth = np.minimum(i, 4 * nside - i)
    cap = inorth < nside
    th = np.empty(nr)
    th[cap] = 2.0 * np.arcsin(inorth[cap] / (np.sqrt(6.0) * nside))
    eq = ~cap
    th[eq] = np.arccos((2 * nside - i[eq]) * (2.0 / (3.0 * nside)))
    south = i > 2 * nside
    th[south & cap] = np.pi - th[south & cap]
    nph = np.where(cap, 4 * inorth, 4 * nside).astype(np.int64)
    phi0 = np.where(cap, np.pi / (4.0 * np.maximum(inorth, 1)),
                    np.where((i + nside) % 2 == 0, np.pi / (4.0 * nside), 0.0))
    ofs = np.concatenate([[0], np.cumsum(nph[:-1])]).astype(np.int64)
    return th, phi0, nph, ofs


class Geom:
    """Iso-latitude pixelisation (cunusht/geometry.py:29-57): rings sorted by ofs."""

    def __init__(self, theta, phi0, nphi, ringstart, w):
        theta = np.asarray(theta, dtype=np.float64).reshape(-1)
        for arr in (phi0, nphi, ringstart, w):
            assert np.size(arr) == theta.size, (np.size(arr), theta.size)
        order = np.argsort(np.asarray(ringstart), kind="stable")
        self.theta = theta[order]
        self.phi0 = np.asarray(phi0, dtype=np.float64).reshape(-1)[order]
        self.nph = np.asarray(nphi, dtype=np.int64).reshape(-1)[order]
        self.ofs = np.asarray(ringstart, dtype=np.int64).reshape(-1)[order]
        self.weight = np.asarray(w, dtype=np.float64).reshape(-1)[order]

    def npix(self) -> int:
        return int(np.sum(self.nph))

    def fsky(self) -> float:
        """Fraction of the sphere covered (cunusht/geometry.py:65-69)."""
        return float(np.sum(self.weight * self.nph) / (4 * np.pi))

    def sort(self, arr, inplace: bool):
        """Rings reordered by argsort(arr) (cunusht/geometry.py:71-82)."""
        arr = np.asarray(arr)
        assert arr.size == self.theta.size, (arr.size, self.theta.size)
        asort = np.argsort(arr, kind="stable")
        if inplace:
            for name in ("theta", "weight", "phi0", "nph", "ofs"):
                setattr(self, name, getattr(self, name)[asort])
            return self
        g = Geom.__new__(Geom)
        g.theta, g.phi0, g.nph, g.ofs, g.weight = (self.theta[asort], self.phi0[asort], self.nph[asort],
                                                   self.ofs[asort], self.weight[asort])
        return g

    @staticmethod
    def nalm(lmax: int, mmax: int) -> int:
        return ((mmax + 1) * (mmax + 2)) // 2 + (mmax + 1) * (lmax - mmax)

    def restrict(self, tht_min: float, tht_max: float, northsouth_sym: bool, update_ringstart=False):
        """Rings with tht_min <= theta <= tht_max (and the mirrored band if northsouth_sym); ring starts refer to
        the original map unless update_ringstart (cunusht/geometry.py:90-108)."""
        cond = (self.theta <= tht_max) & (self.theta >= tht_min)
        if northsouth_sym:
            cond = cond | ((self.theta <= np.pi - tht_min) & (self.theta >= np.pi - tht_max))
        band = np.where(cond)[0]
        ofs = (np.concatenate([[0], np.cumsum(self.nph[band][:-1])]).astype(np.int64) if update_ringstart
               else self.ofs[band])
        return Geom(self.theta[band], self.phi0[band], self.nph[band], ofs, self.weight[band])

    def thinout(self, spin, good_size_real=True):
        """Fewer longitude points at high latitude, never coarser than the equator (cunusht/geometry.py:110-124)."""
        ntht, tht = self.theta.size, self.theta
        st = np.sin(tht)
        lmax = ntht - 1
        mmax = np.minimum(np.maximum(st2mmax(spin, tht, lmax), st2mmax(-spin, tht, lmax)), np.ones(ntht) * lmax)
        nph = np.array([good_size(int(np.ceil(2 * m + 1)), good_size_real) for m in mmax])
        dph_eq = 1.0 / nph[np.argmax(st)]
        nphi_eq = np.array([good_size(int(np.ceil(x / dph_eq)), good_size_real) for x in st])
        nph = np.where((st / nph) > dph_eq, nphi_eq, nph)
        ofs = np.concatenate([[0], np.cumsum(nph[:-1])]).astype(np.int64)
        return Geom(tht, self.phi0, nph, ofs, self.weight / nph * self.nph)

    def split(self, nbands: int, verbose=False):
        """nbands colatitude bands, each with its mirrored band; ring starts refer to the full map
        (cunusht/geometry.py:126-153)."""
        thts = np.linspace(0.0, np.pi * 0.5, nbands + 1)
        geoms = [self.restrict(lo, hi, True) for lo, hi in zip(thts[:-1], thts[1:])]
        npix, tot = self.npix(), sum(g.npix() for g in geoms)
        assert tot >= npix, (npix, tot)
        if tot > npix and verbose:
            print(f"(split with overlap, {tot - npix} additional pixels out of {npix})")
        return geoms

    @staticmethod
    def rings2pix(geom, rings):
        return np.concatenate([geom.ofs[ir] + np.arange(geom.nph[ir], dtype=np.int64) for ir in rings])

    @staticmethod
    def phis(geom, ir):
        assert ir < geom.theta.size
        nph = geom.nph[ir]
        return (geom.phi0[ir] + np.arange(nph) * (2 * np.pi / nph)) % (2.0 * np.pi)

    @staticmethod
    def rings2phi(geom, rings):
        return np.concatenate([Geom.phis(geom, ir) for ir in rings])

    @staticmethod
    def get_supported_geometries() -> str:
        return "".join(" " + k.split("_")[1] for k in vars(Geom)
                       if len(k.split("_")) == 3 and k.startswith("get_") and k.endswith("_geometry"))

    @staticmethod
    def show_supported_geometries():
        print("supported geometries:" + Geom.get_supported_geometries())

    def pixel_angles(self):
        """(theta, phi) of every pixel in map order (phi not reduced mod 2 pi)."""
        th = np.repeat(self.theta, self.nph)
        ph = np.concatenate([self.phi0[i] + 2 * np.pi * np.arange(self.nph[i]) / self.nph[i]
                             for i in range(self.theta.size)])
        return th, ph

    @staticmethod
    def get_thingauss_geometry(lmax: int, smax: int, good_size_real: bool = True):
        """Gauss-Legendre rings with nphi = good_size(2 mmax(theta) + 1) (cunusht/geometry.py:194-215)."""
        th, wt = gauss_legendre_thetas(lmax + 1)
        mmax = np.minimum(np.maximum(st2mmax(smax, th, lmax), st2mmax(-smax, th, lmax)), np.ones(th.size) * lmax)
        nph = np.array([good_size(int(np.ceil(2 * m + 1)), good_size_real) for m in mmax], dtype=np.int64)
        ofs = np.concatenate([[0], np.cumsum(nph[:-1])]).astype(np.int64)
        return Geom(th, np.zeros(th.size), nph, ofs, 2.0 * np.pi * wt / nph)

    @staticmethod
    def get_tgl_geometry(lmax: int, smax: int, good_size_real: bool = True):
        return Geom.get_thingauss_geometry(lmax, smax, good_size_real=good_size_real)

    @staticmethod
    def get_healpix_geometry(nside: int):
        """HEALPix RING scheme, every pixel weighted 4 pi / (12 nside^2) (cunusht/geometry.py:216-226)."""
        th, phi0, nph, ofs = healpix_rings(nside)
        return Geom(th, phi0, nph, ofs, np.full(th.size, 4.0 * np.pi / (12 * nside ** 2)))

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


class pbounds:
    """Longitude interval [pctr - prange / 2, pctr + prange / 2] (mod 2 pi) (cunusht/geometry.py:300-326)."""

    def __init__(self, pctr: float, prange: float):
        assert prange >= 0.0, prange
        self.pctr = pctr % (2.0 * np.pi)
        self.hext = min(prange * 0.5, np.pi)

    def __repr__(self):
        return "ctr:%.2f range:%.2f" % (self.pctr, self.hext * 2)

    def __eq__(self, other):
        return self.pctr == other.pctr and self.hext == other.hext

    def get_range(self):
        return 2 * self.hext

    def get_ctr(self):
        return self.pctr

    def contains(self, phs):
        dph = (np.asarray(phs) - self.pctr) % (2 * np.pi)
        return (dph <= self.hext) | ((2 * np.pi - dph) <= self.hext)


class pbdGeometry:
    """A geometry with a longitude cut (cunusht/geometry.py:329-334): the Geom and its pbounds."""

    def __init__(self, geom: Geom, pbound: pbounds):
        self.geom = geom
        self.pbound = pbound

    def __getattr__(self, name):
        return getattr(self.geom, name)


def get_geom(geometry=("gl", {"lmax": 255})) -> Geom:
    """cunusht/geometry.py:19-27."""
    geo = getattr(Geom, "get_" + geometry[0] + "_geometry", None)
    if geo is None:
        raise ValueError(f"Geometry {geometry[0]} not found, available geometries:" + Geom.get_supported_geometries())
    # lmax / mmax travel in the same dict for the transformer; geometries that do not take them ignore them
    params = inspect.signature(geo).parameters
    kw = {k: v for k, v in geometry[1].items() if k in params or k not in ("lmax", "mmax")}
    return geo(**kw)
