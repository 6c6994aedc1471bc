"""CPU nuSHT operators: get_transformer(backend='CPU').

Host counterpart of the reference's CPU_DUCCnufft_transformer (cunusht/deflection/CPU_nufft_transformer.py:181-558)
and of its `deflection` helper (:57-178), with the same constructor and call signatures: the Clenshaw-Curtis
synthesis, the double-Fourier-sphere doubling, the 2D FFT, the ES-kernel nuFFT of type 2 and the spin
rotation, and the exact adjoints of all of them.  The numerical core is libnusht_cpu.so (include/nusht_cpu.h:
Legendre recursion, ES interpolation and spreading, pointing; C++ with OpenMP); the FFTs are scipy.fft with
`workers=nthreads`.  ducc0 is neither needed nor used.  Arithmetic is FP64; for epsilon > 1e-6 the results are
returned as float32 / complex64, the dtypes of the reference's single-precision mode (:235-257).

The stages are arranged for fewer operations than the reference's array layout, with the same mathematics:
  * the doubling and the theta FFT act on the ring coefficients G_m(theta) of each m (the phi FFT of the CC
    map and the phi half of the reference's 2D FFT cancel each other, so neither is formed);
  * for spin 0 only m >= 0 is carried (the field is real) and the phi modes are centred by e^{-i c phi},
    c = lmax // 2;
  * the nuFFT reads only the fine rows that points with theta in [0, pi] touch (about half of the torus).
execmode 'debug' also forms the reference's own intermediates (CAR map, doubled map, its normalised FFT) from
the same ring coefficients; execmode 'timing' records per-stage wall times under the reference's timer keys.

Differences from the reference, all deliberate:
  * lenmap2gclm applies e^{-i s gamma} (polrot=True), so that it is the exact adjoint of gclm2lenmap; the
    reference's CPU adjoint omits the rotation (:487-529).
  * execmode defaults to 'normal' (the reference's gclm2lenmap default 0 fails its own assert, :455-458).
  * mmax < lmax alm layouts are accepted (they are prefixes of the mmax = lmax layout).
Errors: AssertionError for the reference's own checks, ValueError for invalid arguments.
"""
from __future__ import annotations

import ctypes
import math
import os
import time

import numpy as np
import scipy.fft as sfft

from . import diagnostics
from .geometry import get_geom, good_size
from .utils import Alm

_dp = ctypes.POINTER(ctypes.c_double)
_LIB = None
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libnusht_cpu.so")

SYMBOLS = ("nusht_cpu_version", "nusht_cpu_synth_rings", "nusht_cpu_adjoint_rings", "nusht_cpu_es_kernel_info",
           "nusht_cpu_interp", "nusht_cpu_spread", "nusht_cpu_pointing")


def load():
    """The host core library (ImportError when it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `make -C cunusht_amd/csrc/host`")
        lib = ctypes.CDLL(LIB_PATH)
        i, d, i64 = ctypes.c_int, ctypes.c_double, ctypes.c_int64
        lib.nusht_cpu_version.argtypes = []
        lib.nusht_cpu_synth_rings.argtypes = [i, i, i, _dp, i, _dp, _dp, i]
        lib.nusht_cpu_adjoint_rings.argtypes = [i, i, i, _dp, i, _dp, _dp, i]
        lib.nusht_cpu_es_kernel_info.argtypes = [i, d, d, ctypes.POINTER(i), _dp]
        lib.nusht_cpu_interp.argtypes = [ctypes.c_void_p, _dp, i64, _dp, _dp, i, _dp, i, _dp, i]
        lib.nusht_cpu_spread.argtypes = [ctypes.c_void_p, _dp, i64, _dp, _dp, i, _dp, i, _dp, i]
        lib.nusht_cpu_pointing.argtypes = [i64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, i]
        for name in SYMBOLS:
            getattr(lib, name).restype = ctypes.c_int
        _LIB = lib
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _call(name, *args):
    rc = getattr(load(), name)(*args)
    if rc != 0:
        raise ValueError(f"{name}: invalid argument")


def _nthreads(nthreads) -> int:
    return int(nthreads) if nthreads else len(os.sched_getaffinity(0))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def cc_sizes(lmax: int):
    """(ntheta, nphi) of the reference's CAR grid (CPU_nufft_transformer.py:208-211)."""
    return (good_size(lmax + 2) + 3) // 4 * 4, 2 * good_size(lmax + 1)


def _alm_full(alm, lmax: int, mmax: int, ncomp_expected=None) -> np.ndarray:
    """(ncomp, nalm(lmax, lmax)) complex128; an mmax < lmax layout is the prefix of the full one."""
    a = np.atleast_2d(np.asarray(alm))
    n = Alm.getsize(lmax, mmax)
    if a.shape[-1] != n:
        raise ValueError(f"alm has {a.shape[-1]} coefficients, expected {n} for lmax={lmax}, mmax={mmax}")
    full = Alm.getsize(lmax, lmax)
    if n == full:
        return _c128(a)
    out = np.zeros((a.shape[0], full), dtype=np.complex128)
    out[:, :n] = a
    return out


def _mode(alm, spin: int) -> str:
    """ducc_sht_mode (CPU_nufft_transformer.py:52-55): GRAD_ONLY for a single component of spin > 0."""
    a = np.atleast_2d(alm)
    return "GRAD_ONLY" if (a.shape[0] == 1 and abs(spin) > 0) else "STANDARD"


# ---------------------------------------------------------------------------------------- ring stages
def synth_rings(alm2d: np.ndarray, lmax: int, spin: int, theta, nthreads=None) -> np.ndarray:
    """Ring coefficients G (rows m = 0..lmax for spin 0, m = -lmax..lmax for spin > 0; columns rings)."""
    alm2d = _c128(alm2d)
    theta = _f64(theta)
    nrows = lmax + 1 if spin == 0 else 2 * lmax + 1
    G = np.empty((nrows, theta.size), dtype=np.complex128)
    _call("nusht_cpu_synth_rings", lmax, spin, alm2d.shape[0], _p(alm2d.view(np.float64)), theta.size, _p(theta),
          _p(G.view(np.float64)), _nthreads(nthreads))
    return G


def adjoint_rings(G: np.ndarray, lmax: int, spin: int, ncomp: int, theta, nthreads=None) -> np.ndarray:
    G = _c128(G)
    theta = _f64(theta)
    alm = np.empty((ncomp, Alm.getsize(lmax, lmax)), dtype=np.complex128)
    _call("nusht_cpu_adjoint_rings", lmax, spin, ncomp, _p(G.view(np.float64)), theta.size, _p(theta),
          _p(alm.view(np.float64)), _nthreads(nthreads))
    return alm


def _ms(lmax: int, spin: int) -> np.ndarray:
    return np.arange(lmax + 1) if spin == 0 else np.arange(-lmax, lmax + 1)


def phi_synth_rings(G: np.ndarray, lmax: int, spin: int, nph: int, phi0, nthreads=None) -> np.ndarray:
    """Values on nring rings of nph equidistant points from phi0: (nring, nph), real for spin 0."""
    ms = _ms(lmax, spin)
    coef = G.T * np.exp(1j * np.outer(phi0, ms))
    if spin == 0:
        coef[:, 1:] *= 2.0
    A = np.zeros((G.shape[1], nph), dtype=np.complex128)
    if nph >= ms.max() - ms.min() + 1:
        A[:, ms % nph] = coef
    else:  # rings with fewer points than modes: fold (alias) the coefficients
        np.add.at(A, (slice(None), ms % nph), coef)
    v = sfft.ifft(A, axis=1, norm="forward", workers=_nthreads(nthreads))
    return v.real if spin == 0 else v


def phi_adjoint_rings(M: np.ndarray, lmax: int, spin: int, phi0, nthreads=None) -> np.ndarray:
    """Exact adjoint of phi_synth_rings: G'_m(ring) = sum_k M_k e^{-i m phi_k}, (rows, nring)."""
    ms = _ms(lmax, spin)
    nph = M.shape[1]
    F = sfft.fft(_c128(M), axis=1, workers=_nthreads(nthreads))
    return (F[:, ms % nph] * np.exp(-1j * np.outer(phi0, ms))).T


# ---------------------------------------------------------------------------------------- ES nuFFT
class _Grid(ctypes.Structure):
    """nusht_cpu_grid (include/nusht_cpu.h)."""
    _fields_ = [("nf1", ctypes.c_int), ("nf2", ctypes.c_int), ("row_lo", ctypes.c_int), ("nrow", ctypes.c_int),
                ("w", ctypes.c_int), ("beta", ctypes.c_double), ("ktol", ctypes.c_double)]


def _runs(dst):
    """(src, dst, length) runs of consecutive entries of the index map src -> dst[src]."""
    dst = np.asarray(dst)
    cut = np.nonzero(np.diff(dst) != 1)[0] + 1
    starts = np.concatenate([[0], cut])
    ends = np.concatenate([cut, [dst.size]])
    return [(int(a), int(dst[a]), int(b - a)) for a, b in zip(starts, ends)]


class EsGrid:
    """ES-kernel nuFFT between 2D Fourier modes (theta modes k = -lmax..lmax, phi modes of the spin) and points.

    Kernel parameters by the published ES method (Barnett, Magland & af Klinteberg, SIAM J. Sci. Comput. 41
    (2019)): upsampling sigma = 2 for eps < 1e-8 else 1.25, width w = ceil(ln(1/eps) / (pi sqrt(1 - 1/sigma))),
    beta = 0.97 pi w (1 - 1/(2 sigma)); fine sizes even and 2,3,5,7,11-smooth, at least 2 w; deconvolution by
    the kernel's Fourier transform, integrated by Gauss-Legendre quadrature; kernel values from per-cell
    polynomials accurate to max(eps / 100, 1e-15).  Spin 0 carries the phi modes m = 0..lmax shifted by
    c = lmax // 2 (m - c is centred), the phase e^{i c phi} is applied at the points."""

    def __init__(self, lmax: int, spin: int, eps: float):
        if not (0.0 < eps < 1.0):
            raise ValueError("epsilon must be in (0, 1)")
        self.lmax, self.spin, self.eps = lmax, spin, eps
        sigma = 2.0 if eps < 1e-8 else 1.25
        w = max(2, int(math.ceil(math.log(1.0 / eps) / (math.pi * math.sqrt(1.0 - 1.0 / sigma)))))
        if w > 32:
            raise ValueError(f"epsilon {eps} needs an ES kernel wider than 32 cells")
        self.w, self.beta = w, 0.97 * math.pi * w * (1.0 - 1.0 / (2.0 * sigma))
        self.ks = np.arange(-lmax, lmax + 1)
        self.ms = _ms(lmax, spin)
        self.shift = lmax // 2 if spin == 0 else 0
        mm = self.ms - self.shift

        def fine(n):
            nf = good_size(max(int(math.ceil(sigma * n)), 2 * w))
            while nf % 2:
                nf = good_size(nf + 1)
            return nf

        self.nf1, self.nf2 = fine(self.ks.size), fine(self.ms.size)
        lo = -(w // 2) - 1
        hi = int(math.ceil(0.5 * self.nf1 - 0.5 * w)) + w
        self.nrow = min(self.nf1, hi - lo + 1)
        self.row_lo = 0 if self.nrow == self.nf1 else lo
        self.grid = _Grid(self.nf1, self.nf2, self.row_lo, self.nrow, w, self.beta, max(eps / 100.0, 1e-15))
        # index maps as contiguous runs: theta modes -> torus rows, window rows -> torus rows, phi modes -> columns
        self.k_runs = _runs(self.ks % self.nf1)
        self.row_runs = _runs((self.row_lo + np.arange(self.nrow)) % self.nf1)
        self.m_runs = _runs(mm % self.nf2)
        xq, wq = np.polynomial.legendre.leggauss(4 * w + 60)
        psi = wq * np.exp(self.beta * (np.sqrt(np.maximum(0.0, 1.0 - xq * xq)) - 1.0))

        def inv_hat(k, nf):
            return 1.0 / (0.5 * w * (psi @ np.cos(np.pi * w * np.outer(xq, k) / nf)))

        self.dec = np.outer(inv_hat(self.ks, self.nf1), inv_hat(mm, self.nf2))

    def kernel_info(self):
        """(degree, max error) of the kernel polynomials."""
        deg, err = ctypes.c_int(), ctypes.c_double()
        _call("nusht_cpu_es_kernel_info", self.w, self.beta, self.grid.ktol, ctypes.byref(deg), ctypes.byref(err))
        return deg.value, err.value

    def check_points(self, theta, phi):
        theta, phi = _f64(theta), _f64(phi)
        if theta.shape != phi.shape or theta.ndim != 1:
            raise ValueError("theta and phi must be 1-D arrays of one length")
        if theta.size and not (np.all(theta >= 0.0) and np.all(theta <= np.pi)):
            raise ValueError("point colatitudes must lie in [0, pi]")
        if theta.size and not np.all(np.isfinite(phi)):
            raise ValueError("point longitudes must be finite")
        return theta, phi

    def _window(self, A, nth):
        """Torus rows (nf1, nmodes) of the theta transform -> the window (nrow, nf2) over phi."""
        W = np.zeros((self.nrow, self.nf2), dtype=np.complex128)
        for rs, rd, rn in self.row_runs:
            for ms, md, mn in self.m_runs:
                W[rs:rs + rn, md:md + mn] = A[rd:rd + rn, ms:ms + mn]
        return sfft.ifft(W, axis=1, norm="forward", workers=nth, overwrite_x=True)

    def u2nu(self, F: np.ndarray, theta, phi, nthreads=None, gamma=None, spin_rot: int = 0) -> np.ndarray:
        """e^{i (c phi + spin_rot gamma)} sum_{k,m} F[k, m] e^{i (k theta + (m - c) phi)} at the points (within
        eps).  F is (2 lmax + 1, nmodes) and may be overwritten."""
        nth = _nthreads(nthreads)
        F *= self.dec
        A = np.zeros((self.nf1, self.ms.size), dtype=np.complex128)
        for s0, d0, n in self.k_runs:
            A[d0:d0 + n] = F[s0:s0 + n]
        A = sfft.ifft(A, axis=0, norm="forward", workers=nth, overwrite_x=True)
        W = self._window(A, nth)
        del A
        out = np.empty(theta.size, dtype=np.complex128)
        g = None if gamma is None or not spin_rot else _f64(gamma)
        _call("nusht_cpu_interp", ctypes.byref(self.grid), _p(W.view(np.float64)), theta.size, _p(theta), _p(phi),
              self.shift, _p(g), int(spin_rot), _p(out.view(np.float64)), nth)
        return out

    def nu2u(self, values: np.ndarray, theta, phi, nthreads=None, gamma=None, spin_rot: int = 0) -> np.ndarray:
        """Exact adjoint of u2nu: F'[k, m] = sum_p e^{-i (k theta_p + (m - c) phi_p)} e^{-i (c phi_p + spin_rot
        gamma_p)} values_p (within eps)."""
        nth = _nthreads(nthreads)
        v = _c128(values)
        W = np.zeros((self.nrow, self.nf2), dtype=np.complex128)
        g = None if gamma is None or not spin_rot else _f64(gamma)
        _call("nusht_cpu_spread", ctypes.byref(self.grid), _p(v.view(np.float64)), v.size, _p(theta), _p(phi),
              self.shift, _p(g), int(spin_rot), _p(W.view(np.float64)), nth)
        W = sfft.fft(W, axis=1, workers=nth, overwrite_x=True)
        T = np.zeros((self.nf1, self.ms.size), dtype=np.complex128)
        for rs, rd, rn in self.row_runs:  # window rows are distinct torus rows
            for ms, md, mn in self.m_runs:
                T[rd:rd + rn, ms:ms + mn] = W[rs:rs + rn, md:md + mn]
        del W
        T = sfft.fft(T, axis=0, workers=nth, overwrite_x=True)
        out = np.empty((self.ks.size, self.ms.size), dtype=np.complex128)
        for s0, d0, n in self.k_runs:
            out[s0:s0 + n] = T[d0:d0 + n]
        out *= self.dec
        return out


# ---------------------------------------------------------------------------------------- the operator
class CPU_nusht_transformer:
    """nuSHT operators on the host (cunusht/deflection/CPU_nufft_transformer.py:181-558)."""

    def __init__(self, geominfo_deflection, shttransformer_desc="ducc", nuFFTtype=None, epsilon=None):
        if nuFFTtype:
            assert False, "nuFFTtype mode not supported"
        if shttransformer_desc not in ("ducc", "shtns", "cunusht", "amd"):
            raise ValueError('shttransformer_desc must be either "ducc" or "shtns"')
        load()
        self.backend = "CPU"
        self.shttransformer_desc = shttransformer_desc
        self.nuFFTtype = nuFFTtype
        self.epsilon = epsilon
        self.execmode = "normal"
        self.single_prec = False
        self.nthreads = None
        self.ret = {}
        self.timer = diagnostics.Timer()
        assert "lmax" in geominfo_deflection[1], "geominfo_deflection must give lmax (sizes the CC grid)"
        self.lmax = int(geominfo_deflection[1]["lmax"])
        self.ntheta_CAR, self.nphi_CAR = cc_sizes(self.lmax)
        self.nphihalf_CAR = self.nphi_CAR // 2
        self.ntheta_dCAR, self.nphi_dCAR = 2 * self.ntheta_CAR - 2, self.nphi_CAR
        self.deflectionlib = deflection(geominfo_deflection, shttransformer_desc, timer_instance=self.timer)
        self.geom = self.deflectionlib.geom
        self._grids = {}

    # ------------------------------------------------------------------ helpers
    @staticmethod
    def hashdict():
        return "CPU_nusht_transformer"

    def _tick(self, key, t0):
        t1 = time.perf_counter()
        self.timer.keys[key] = self.timer.keys.get(key, 0.0) + (t1 - t0)
        return t1

    def _grid(self, lmax, spin, eps) -> EsGrid:
        key = (lmax, spin, float(eps))
        g = self._grids.get(key)
        if g is None:
            g = EsGrid(lmax, spin, eps)
            self._grids[key] = g
        return g

    def _out_real(self, a):
        return a.astype(np.float32) if self.single_prec else a

    def _out_complex(self, a):
        return a.astype(np.complex64) if self.single_prec else a

    @staticmethod
    def _loc(loc):
        loc = np.asarray(loc, dtype=np.float64)
        if loc.ndim != 2 or loc.shape[1] != 2:
            raise ValueError("loc must be (npts, 2) = (theta, phi)")
        return _f64(loc[:, 0]), _f64(loc[:, 1])

    def _theta_coeffs(self, G, lmax, spin, nthreads):
        """Doubling over theta of every m row and its normalised theta FFT: F[k, m], k = -lmax..lmax, from the
        CC ring coefficients G (rows m, nt rings): D[r >= nt] = (-1)^(s + m) G[2 nt - 2 - r]."""
        nt = G.shape[1]
        n1 = 2 * nt - 2
        ms = _ms(lmax, spin)
        sg = np.where((spin + ms) % 2 == 0, 1.0, -1.0)[None, :]
        D = np.empty((n1, G.shape[0]), dtype=np.complex128)
        D[:nt] = G.T
        np.multiply(D[nt - 2:0:-1], sg, out=D[nt:])
        D = sfft.fft(D, axis=0, norm="forward", workers=_nthreads(nthreads), overwrite_x=True)
        F = np.empty((2 * lmax + 1, G.shape[0]), dtype=np.complex128)
        F[:lmax] = D[n1 - lmax:]
        F[lmax:] = D[:lmax + 1]
        return F

    def _theta_coeffs_adjoint(self, Fp, lmax, spin, nt, nthreads):
        n1 = 2 * nt - 2
        ms = _ms(lmax, spin)
        sg = np.where((spin + ms) % 2 == 0, 1.0, -1.0)[None, :]
        D = np.zeros((n1, ms.size), dtype=np.complex128)
        D[n1 - lmax:] = Fp[:lmax]
        D[:lmax + 1] = Fp[lmax:]
        D = sfft.ifft(D, axis=0, workers=_nthreads(nthreads), overwrite_x=True)
        Gp = D[:nt].copy()
        Gp[1:nt - 1] += sg * D[n1 - 1:nt - 1:-1]
        return Gp.T

    def _cc_theta(self, lmax):
        nt = cc_sizes(lmax)[0]
        return np.pi * np.arange(nt) / (nt - 1)

    def _debug_intermediates(self, G, lmax, spin):
        """The reference's CAR map, doubled map and its normalised 2D FFT (CPU_nufft_transformer.py:266-325)."""
        nt, nph = cc_sizes(lmax)
        car = phi_synth_rings(G, lmax, spin, nph, np.zeros(nt), self.nthreads)
        h = nph // 2
        d = np.empty((2 * nt - 2, nph), dtype=car.dtype)
        d[:nt] = car
        d[nt:, :h] = d[nt - 2:0:-1, h:]
        d[nt:, h:] = d[nt - 2:0:-1, :h]
        if spin % 2:
            d[nt:] *= -1
        self.ret["synthesis"] = np.array([car]) if spin == 0 else np.array([car.real, car.imag])
        self.ret["doubling"] = d
        self.ret["C2C"] = sfft.fft2(d, norm="forward", workers=_nthreads(self.nthreads))

    # ------------------------------------------------------------------ SHT on the pointing grid
    def synthesis(self, alm, out=None, lmax=None, mmax=None, nthreads=None, spin=0, mode=None):
        """alm -> map on the pointing geometry (CPU_nufft_transformer.py:373-380)."""
        lmax = self.lmax if lmax is None else lmax
        return self.deflectionlib.synthesis(alm, spin, lmax, lmax if mmax is None else mmax, nthreads, map=out,
                                            mode=mode)

    def adjoint_synthesis(self, synthmap, lmax=None, mmax=None, nthreads=None, out=None, spin=0, apply_weights=True,
                          mode=None):
        """Adjoint synthesis with the pixel weights, i.e. quadrature analysis (CPU_nufft_transformer.py:382-388,
        CPU_sht_transformer.py:32-48); apply_weights=False gives the exact transpose of synthesis."""
        lmax = self.lmax if lmax is None else lmax
        return self.deflectionlib.adjoint_synthesis(synthmap, spin, lmax, lmax if mmax is None else mmax, nthreads,
                                                    alm=out, apply_weights=apply_weights, mode=mode)

    # ------------------------------------------------------------------ general (nu) synthesis
    def _synth_general(self, a, lmax, spin, theta, phi, eps, nthreads, gamma=None):
        """alm (ncomp, nalm(lmax, lmax)) -> complex values at the points, times e^{i spin gamma} when gamma is
        given (spin 0: real field with imaginary part 0)."""
        if not (a.shape[0] == 1 or (spin > 0 and a.shape[0] == 2)):
            raise ValueError(f"spin {spin} takes 1 or 2 alm components, got {a.shape[0]}")
        grid = self._grid(lmax, spin, eps)
        theta, phi = grid.check_points(theta, phi)
        t = time.perf_counter()
        G = synth_rings(a, lmax, spin, self._cc_theta(lmax), nthreads)
        t = self._tick("synthesis", t)
        if self.execmode == "debug":
            self._debug_intermediates(G, lmax, spin)
            t = time.perf_counter()
        if spin == 0:
            G[1:] *= 2.0
        F = self._theta_coeffs(G, lmax, spin, nthreads)
        del G
        t = self._tick("doubling", t)
        u = grid.u2nu(F, theta, phi, nthreads, gamma=gamma, spin_rot=spin if gamma is not None else 0)
        if spin == 0:
            u = u.real.astype(np.complex128)
        self._tick("nuFFT2d2", t)
        return u

    def _adjoint_general(self, v, lmax, spin, ncomp, theta, phi, eps, nthreads, gamma=None):
        """Exact adjoint of _synth_general: complex values (spin 0: real part used) -> (ncomp, nalm(lmax, lmax))."""
        grid = self._grid(lmax, spin, eps)
        theta, phi = grid.check_points(theta, phi)
        v = np.asarray(v).reshape(-1)
        if v.size != theta.size:
            raise ValueError(f"pointmap has {v.size} values for {theta.size} points")
        t = time.perf_counter()
        u = np.real(v).astype(np.complex128) if spin == 0 else _c128(v)
        Fp = grid.nu2u(u, theta, phi, nthreads, gamma=gamma, spin_rot=spin if gamma is not None else 0)
        t = self._tick("nuFFT2d1", t)
        th = self._cc_theta(lmax)
        Gp = self._theta_coeffs_adjoint(Fp, lmax, spin, th.size, nthreads)
        del Fp
        t = self._tick("adjointdoubling", t)
        res = adjoint_rings(Gp, lmax, spin, ncomp, th, nthreads)
        self._tick("adjointsynthesis", t)
        return res

    def synthesis_general(self, pointmap, lmax, mmax, alm, loc, spin, epsilon, nthreads, mode, verbose):
        """alm -> complex values at loc (npts, 2) (CPU_nufft_transformer.py:390-405): spin 0 gives the real
        field (imaginary part 0), spin > 0 gives Q + iU.  mode 'GRAD_ONLY' or a single component is E only."""
        theta, phi = self._loc(loc)
        a = _alm_full(alm, lmax, mmax)
        if mode == "GRAD_ONLY" and spin > 0:
            a = a[:1]
        u = self._synth_general(a, lmax, spin, theta, phi, epsilon, nthreads)
        if self.execmode == "debug":
            self.ret["nuFFT2d2"] = u.copy()
        return u

    def adjoint_synthesis_general(self, lmax, mmax, pointmap, loc, spin, epsilon, nthreads, mode, alm, verbose):
        """Exact adjoint of synthesis_general (CPU_nufft_transformer.py:410-427): complex values -> alm
        (ncomp, nalm(lmax, mmax)); ncomp = 1 for spin 0 or mode 'GRAD_ONLY', else 2."""
        theta, phi = self._loc(loc)
        ncomp = 1 if (spin == 0 or mode == "GRAD_ONLY") else 2
        res = self._adjoint_general(pointmap, lmax, spin, ncomp, theta, phi, epsilon, nthreads)
        res = res[:, :Alm.getsize(lmax, mmax)]
        if alm is not None:
            alm = np.asarray(alm)
            if alm.size == res.size and alm.dtype in (np.complex128, np.complex64) and alm.flags.writeable:
                alm.reshape(res.shape)[...] = res
        return res

    # GPU-class names for the same two operators (values and loc as numpy arrays)
    def nusht2d2(self, lmax, mmax, alm, loc, pointmap=None, epsilon=None, verbose=0, spin=0, nthreads=None):
        eps = self.epsilon if epsilon is None else epsilon
        u = self.synthesis_general(pointmap, lmax, mmax, alm, loc, spin, eps, nthreads, _mode(alm, spin), verbose)
        return u.real if spin == 0 else np.array([u.real, u.imag])

    def nusht2d1(self, lmax, mmax, pointmap, loc, epsilon=None, alm=None, verbose=0, spin=0, nthreads=None):
        eps = self.epsilon if epsilon is None else epsilon
        v = np.asarray(pointmap)
        if spin > 0 and not np.iscomplexobj(v):
            v = v[0] + 1j * v[1]
        mode = "GRAD_ONLY" if (spin > 0 and alm is not None and np.atleast_2d(alm).shape[0] == 1) else "STANDARD"
        res = self.adjoint_synthesis_general(lmax, mmax, v, loc, spin, eps, nthreads, mode, alm, verbose)
        return res[0] if spin == 0 else res

    # ------------------------------------------------------------------ lensing
    def dlm2pointing(self, dlm, mmax_dlm=None, single_prec=False, nthreads=None):
        """Deflected (theta, phi) of every pointing pixel, (2, npix) (CPU_nufft_transformer.py:175-178).
        `dlm` is the gradient deflection alm (not divided by sqrt(l(l+1)))."""
        return self.deflectionlib.dlm2pointing(dlm, self.lmax if mmax_dlm is None else mmax_dlm, single_prec,
                                               nthreads)

    def _gamma_for(self, polrot, spin):
        """The rotation angle when the spin rotation applies, else None."""
        if not (polrot and spin):
            return None
        g = self.deflectionlib.gamma()
        if g is None:
            raise ValueError("polrot needs the rotation angle: pass dlm (not only ptg)")
        return g

    def _setup(self, execmode, epsilon, nthreads):
        if execmode in (0, None):
            execmode = "normal"
        assert execmode in ["normal", "debug", "timing"]
        assert epsilon is not None, "epsilon must be given"
        self.execmode = execmode
        self.epsilon = epsilon
        self.single_prec = epsilon > 1e-6
        self.nthreads = nthreads if nthreads is not None else self.nthreads
        self.timer.keys = {}
        self.ret = {}

    def _pointing(self, dlm, ptg, lmax, mmax, nthreads, verbose):
        if ptg is not None:
            ptg = np.asarray(ptg, dtype=np.float64)
            if ptg.ndim != 2 or ptg.shape[1] != 2:
                raise ValueError("ptg must be (npix, 2) = (theta, phi)")
            self.deflectionlib.set_ptg(ptg)
            return _f64(ptg[:, 0]), _f64(ptg[:, 1])
        if dlm is None:
            raise ValueError("dlm or ptg must be given")
        d = np.atleast_2d(np.asarray(dlm))
        if verbose:
            s2 = sum(float(np.sum(_alm2cl(d[i], lmax, lmax) * (2 * np.arange(lmax + 1) + 1))) for i in range(len(d)))
            s2 /= 4 * np.pi
            if len(d) > 1:
                s2 /= np.sqrt(2.0)
            sig = np.sqrt(s2 / self.geom.fsky()) / np.pi * 180 * 60
            print("deflection std is %.2e amin" % sig)
        t = time.perf_counter()
        # the deflection alm has mmax_dlm = lmax (CPU_nufft_transformer.py:467)
        tp, pp = self.deflectionlib.dlm2pointing(d[0], lmax, self.single_prec, nthreads,
                                                 dclm=d[1] if len(d) > 1 else None)
        self._tick("pointing", t)
        if self.execmode == "debug":
            self.ret["dlm2pointing"] = np.array([tp, pp])
        return _f64(tp), _f64(pp)

    def gclm2lenmap(self, gclm, dlm, lmax, mmax, spin, epsilon, nthreads, polrot=True, ptg=None, lenmap=None,
                    verbose=1, execmode="normal"):
        """Lensed map(s) on the pointing geometry (CPU_nufft_transformer.py:429-484): real (npix,) for spin 0,
        (2, npix) = (Q, U) for spin > 0.  gclm (ncomp, nalm): ncomp 1 is gradient only."""
        self._setup(execmode, epsilon, nthreads)
        t0 = time.perf_counter()
        gclm = np.atleast_2d(np.asarray(gclm))
        tp, pp = self._pointing(dlm, ptg, lmax, mmax, nthreads, verbose)
        a = _alm_full(gclm, lmax, mmax)
        gam = self._gamma_for(polrot, spin)
        if self.execmode == "debug":  # the unrotated values too, as the reference's stages return them
            u = self._synth_general(a, lmax, spin, tp, pp, epsilon, nthreads)
            self.ret["nuFFT2d2"] = u.copy()
            if gam is not None:
                u = u * np.exp(1j * spin * gam)
            self.ret["rotate"] = u.copy()
        else:
            u = self._synth_general(a, lmax, spin, tp, pp, epsilon, nthreads, gamma=gam)
        self._tick("gclm2lenmap", t0)
        res = self._out_real(u.real if spin == 0 else np.array([u.real, u.imag]))
        if lenmap is not None:
            lm = np.asarray(lenmap)
            if lm.size == res.size and lm.dtype == res.dtype and lm.flags.writeable:
                lm.reshape(res.shape)[...] = res
        if self.execmode == "timing":
            diagnostics.dump_timer(self, "gclm2lenmap", "lmax{}_epsilon{}".format(lmax, epsilon))
            print(self.timer)
        if self.execmode == "debug":
            print("::debug:: Returned component results")
            return self.ret
        return res

    def lenmap2gclm(self, lenmap, dlm, gclm_out, spin, lmax, mmax, nthreads, epsilon=None, ptg=None, verbose=1,
                    execmode="normal", polrot=True):
        """Exact adjoint of gclm2lenmap (CPU_nufft_transformer.py:487-529): lenmap real (npix,) for spin 0;
        (2, npix) = (Q, U) or complex Q + iU for spin > 0.  Returns (ncomp, nalm), ncomp from gclm_out
        (1 for spin 0 or gradient only, else 2)."""
        self._setup(execmode, epsilon, nthreads)
        t0 = time.perf_counter()
        tp, pp = self._pointing(dlm, ptg, lmax, mmax, nthreads, verbose)
        lm = np.asarray(lenmap)
        if spin == 0:
            v = np.real(lm).reshape(-1).astype(np.complex128)
        elif np.iscomplexobj(lm):
            v = _c128(lm.reshape(-1))
        else:
            lm = lm.reshape(2, -1)
            v = lm[0] + 1j * lm[1]
        ncomp = 1 if spin == 0 else (np.atleast_2d(gclm_out).shape[0] if gclm_out is not None else 2)
        res = self._adjoint_general(v, lmax, spin, ncomp, tp, pp, epsilon, nthreads,
                                    gamma=self._gamma_for(polrot, spin))[:, :Alm.getsize(lmax, mmax)]
        res = self._out_complex(res)
        if gclm_out is not None:
            g = np.asarray(gclm_out)
            if g.size == res.size and g.dtype == res.dtype and g.flags.writeable:
                g.reshape(res.shape)[...] = res
        self._tick("lenmap2gclm", t0)
        if self.execmode == "timing":
            diagnostics.dump_timer(self, "lenmap2gclm", "lmax{}_epsilon{}".format(lmax, epsilon))
            print(self.timer)
        if self.execmode == "debug":
            self.ret["adjointsynthesis"] = res
            print("::debug:: Returned component results")
            return self.ret
        return res

    def gclm2lenpixs(self, gclm, mmax, spin, pixs, polrot=True, ptg=None, nthreads=10):
        """Lensed values at the pointing pixels `pixs` by exact synthesis at their deflected positions
        (CPU_nufft_transformer.py:531-558).  ptg: (2, npix) or (npix, 2); default the last pointing."""
        assert spin >= 0, spin
        gclm = np.atleast_2d(np.asarray(gclm))
        lmax = Alm.getlmax(gclm[0].size, mmax)
        mmax = lmax if mmax is None else mmax
        if ptg is None:
            ptg = self.deflectionlib.ptg()
            if ptg is None:
                raise ValueError("no pointing yet: pass ptg, or run gclm2lenmap / dlm2pointing first")
        ptg = np.asarray(ptg, dtype=np.float64)
        if ptg.ndim == 2 and ptg.shape[1] == 2 and ptg.shape[0] != 2:
            ptg = ptg.T
        pixs = np.asarray(pixs, dtype=np.int64).reshape(-1)
        if pixs.size == 0 or pixs.min() < 0 or pixs.max() >= ptg.shape[1]:
            raise IndexError(f"pixs must be non-empty indices into the {ptg.shape[1]} pointing pixels")
        th, ph = ptg[0, pixs], ptg[1, pixs]
        a = _alm_full(gclm, lmax, mmax)
        G = synth_rings(a, lmax, spin, th, nthreads)
        ms = _ms(lmax, spin)
        e = np.exp(1j * np.outer(ms, ph))
        if spin == 0:
            e[1:] *= 2.0
            return np.einsum("mp,mp->p", G, e).real
        v = np.einsum("mp,mp->p", G, e)
        if polrot:
            g = self.deflectionlib.gamma()
            if g is None:
                raise ValueError("polrot needs the rotation angle: run gclm2lenmap / dlm2pointing with dlm first")
            v = np.exp(1j * spin * g[pixs]) * v
            return v.real, v.imag
        return np.array([v.real, v.imag])

    def lensgclm(self, gclm, dlm, spin, lmax_out, nthreads=None, mmax=None, mmax_out=None, gclm_out=None,
                 polrot=True, out_sht_mode="STANDARD", epsilon=None):
        """alm -> lensed map -> adjoint synthesis with weights on the pointing geometry
        (CPU_nufft_transformer.py:948-971)."""
        if out_sht_mode not in ("STANDARD", "GRAD_ONLY"):
            raise ValueError(f"unknown out_sht_mode {out_sht_mode}")
        gclm = np.atleast_2d(np.asarray(gclm))
        lmax = Alm.getlmax(gclm[0].size, mmax)
        mmax = lmax if mmax is None else mmax
        eps = epsilon if epsilon is not None else (self.epsilon if self.epsilon is not None else 1e-10)
        m = self.gclm2lenmap(gclm, dlm, lmax, mmax, spin, eps, nthreads, polrot=polrot, verbose=0)
        mode = "GRAD_ONLY" if (spin > 0 and out_sht_mode == "GRAD_ONLY") else "STANDARD"
        mmax_out = lmax_out if mmax_out is None else mmax_out
        return self.deflectionlib.adjoint_synthesis(np.atleast_2d(m).astype(np.float64), spin, lmax_out, mmax_out,
                                                    nthreads, alm=gclm_out, apply_weights=True, mode=mode)


def _alm2cl(a, lmax, mmax):
    a = np.asarray(a)
    cl = np.zeros(lmax + 1)
    for m in range(mmax + 1):
        i0 = Alm.getidx(lmax, m, m)
        seg = np.abs(a[i0:i0 + lmax + 1 - m]) ** 2
        cl[m:] += seg if m == 0 else 2.0 * seg
    return cl / (2 * np.arange(lmax + 1) + 1)


class deflection:
    """Pointing-geometry SHT and deflected angles (CPU_nufft_transformer.py:57-178)."""

    def __init__(self, geominfo, shttransformer_desc="ducc", timer_instance=None):
        self.geominfo = geominfo
        self.geom = get_geom(geominfo)
        self.shttransformer_desc = shttransformer_desc
        self.timer = timer_instance if timer_instance is not None else diagnostics.Timer()
        self._ptg = None
        self._gamma = None
        self._key = None

    # per-ring phi transforms over the geometry, rings grouped by their number of points
    def _groups(self):
        g = self.geom
        nph = np.asarray(g.nph, dtype=np.int64)
        for n in np.unique(nph):
            rs = np.nonzero(nph == n)[0]
            idx = np.asarray(g.ofs, dtype=np.int64)[rs][:, None] + np.arange(n)[None, :]
            yield int(n), rs, idx

    def _map_from_G(self, G, lmax, spin, nthreads):
        out = np.empty(self.geom.npix(), dtype=np.float64 if spin == 0 else np.complex128)
        for n, rs, idx in self._groups():
            out[idx] = phi_synth_rings(G[:, rs], lmax, spin, n, np.asarray(self.geom.phi0)[rs], nthreads)
        return out

    def _G_from_map(self, m, lmax, spin, nthreads):
        nrows = lmax + 1 if spin == 0 else 2 * lmax + 1
        G = np.empty((nrows, len(self.geom.theta)), dtype=np.complex128)
        for n, rs, idx in self._groups():
            G[:, rs] = phi_adjoint_rings(m[idx], lmax, spin, np.asarray(self.geom.phi0)[rs], nthreads)
        return G

    def synthesis(self, gclm, spin, lmax, mmax, nthreads, map=None, mode=None, **kwargs):
        """Maps on the geometry, (1, npix) for spin 0 and (2, npix) for spin > 0 (CPU_sht_transformer.py:22-30)."""
        a = _alm_full(gclm, lmax, mmax)
        if mode == "GRAD_ONLY" and spin > 0:
            a = a[:1]
        G = synth_rings(a, lmax, spin, self.geom.theta, nthreads)
        v = self._map_from_G(G, lmax, spin, nthreads)
        res = v[None, :] if spin == 0 else np.array([v.real, v.imag])
        if map is not None:
            mp = np.asarray(map)
            if mp.size == res.size and mp.dtype == res.dtype and mp.flags.writeable:
                mp.reshape(res.shape)[...] = res
        return res

    def adjoint_synthesis(self, m, spin, lmax, mmax, nthreads, alm=None, apply_weights=True, mode=None, **kwargs):
        """(ncomp, nalm(lmax, mmax)) from maps (1 or 2, npix); weighted when apply_weights
        (CPU_sht_transformer.py:32-48).  Unlike the reference, the input map is not modified."""
        m = np.atleast_2d(np.asarray(m, dtype=np.float64))
        npix = self.geom.npix()
        if m.shape[-1] != npix:
            raise ValueError(f"map has {m.shape[-1]} pixels, the geometry {npix}")
        if apply_weights:
            w = np.empty(npix)
            for n, rs, idx in self._groups():
                w[idx] = np.asarray(self.geom.weight)[rs][:, None]
            m = m * w
        if spin == 0:
            v = m[0]
            ncomp = 1
        else:
            v = m[0] + 1j * m[1]
            ncomp = 1 if mode == "GRAD_ONLY" else 2
        G = self._G_from_map(v, lmax, spin, nthreads)
        res = adjoint_rings(G, lmax, spin, ncomp, self.geom.theta, nthreads)[:, :Alm.getsize(lmax, mmax)]
        if alm is not None:
            a = np.asarray(alm)
            if a.size == res.size and a.dtype == res.dtype and a.flags.writeable:
                a.reshape(res.shape)[...] = res
        return res

    def dlm2pointing(self, dlm, mmax_dlm, single_prec, nthreads, dclm=None):
        """(theta', phi') of every pixel after the deflection with gradient alm dlm (and curl alm dclm)
        (CPU_nufft_transformer.py:86-178); the rotation angle gamma is kept for the spin rotation."""
        lmax = Alm.getlmax(np.asarray(dlm).size, mmax_dlm)
        d = np.asarray(dlm).reshape(1, -1) if dclm is None else np.array([np.asarray(dlm).reshape(-1),
                                                                            np.asarray(dclm).reshape(-1)])
        key = (lmax, mmax_dlm, d.tobytes())
        if key != self._key:
            a = _alm_full(d, lmax, mmax_dlm if mmax_dlm is not None else lmax)
            G = synth_rings(a, lmax, 1, self.geom.theta, nthreads)
            dv = self._map_from_G(G, lmax, 1, nthreads)
            th, ph = self.geom.pixel_angles()
            th, ph = _f64(th), _f64(ph)
            dt, dp = _f64(dv.real), _f64(dv.imag)
            n = th.size
            tp, pp, gam = np.empty(n), np.empty(n), np.empty(n)
            _call("nusht_cpu_pointing", n, _p(th), _p(ph), _p(dt), _p(dp), _p(tp), _p(pp), _p(gam),
                  _nthreads(nthreads))
            self._ptg = np.array([tp, pp])
            self._gamma = gam
            self._key = key
        out = self._ptg
        return out.astype(np.float32) if single_prec else out.copy()

    def set_ptg(self, ptg_npix2):
        """A given pointing (npix, 2); the rotation angle is kept when it is the pointing computed last."""
        p = np.ascontiguousarray(np.asarray(ptg_npix2, dtype=np.float64).T)
        if self._ptg is not None and p.shape == self._ptg.shape and np.array_equal(p, self._ptg):
            return
        self._ptg = p
        self._gamma = None
        self._key = None

    def ptg(self):
        return self._ptg

    def gamma(self):
        return self._gamma


class CPU_nusht_lenspyx_transformer(CPU_nusht_transformer):
    """The same CPU operators with the signatures of the reference's default CPU solver, CPU_Lenspyx_transformer
    (cunusht/deflection/CPU_nufft_transformer.py:560-762), which wraps lenspyx / ducc0 synthesis_general.

    Differences, on purpose: the transforms are this package's CC + ES-nuFFT route (within epsilon of the exact
    synthesis_general the reference calls); gclm2lenmap honours `polrot` (the reference passes it but never
    rotates) and returns the maps of the DUCC-signature class (real (npix,) for spin 0, (2, npix) for spin > 0);
    `backwards` is accepted and, as in the reference, unused."""

    def __init__(self, geominfo_deflection, dglm=None, mmax_dlm=None, nthreads=None, verbose=None, epsilon=None,
                 single_prec=None):
        super().__init__(geominfo_deflection, epsilon=epsilon)
        self.dglm = None if dglm is None else np.asarray(dglm)
        self.mmax_dlm = mmax_dlm
        self.nthreads = nthreads
        self.verbosity = 0 if verbose is None else verbose
        if single_prec is not None:
            self.single_prec = bool(single_prec)

    def _dlm(self, dlm, dclm=None):
        d = self.dglm if dlm is None else np.asarray(dlm)
        if d is None:
            raise ValueError("no deflection: pass dlm (or dglm at construction)")
        if dclm is not None:
            d = np.array([np.asarray(d).reshape(-1), np.asarray(dclm).reshape(-1)])
        return d

    def synthesis(self, alm, spin, out=None, lmax=None, mmax=None, nthreads=None):
        """(ncomp, npix) maps on the pointing geometry (:629-635)."""
        lmax = self.lmax if lmax is None else lmax
        return self.deflectionlib.synthesis(alm, spin, lmax, lmax if mmax is None else mmax, nthreads, map=out,
                                            mode=_mode(alm, spin))

    def synthesis_general(self, lmax, mmax, alm, loc, spin, epsilon, nthreads, mode, verbose):
        """Real maps (ncomp, npts) at loc (npts, 2) (:643-647, ducc0 synthesis_general's layout)."""
        u = CPU_nusht_transformer.synthesis_general(self, None, lmax, mmax, alm, loc, spin, epsilon, nthreads, mode,
                                                    verbose)
        return u.real[None, :] if spin == 0 else np.array([u.real, u.imag])

    def adjoint_synthesis_general(self, lmax, mmax, pointmap, loc, spin, epsilon, mode, nthreads, alm, verbose):
        """Exact adjoint of synthesis_general: real maps (ncomp, npts) -> alm (:649-653)."""
        m = np.atleast_2d(np.asarray(pointmap))
        v = m[0] if spin == 0 else (m[0] + 1j * m[1] if m.shape[0] == 2 else m[0])
        return CPU_nusht_transformer.adjoint_synthesis_general(self, lmax, mmax, v, loc, spin, epsilon, nthreads,
                                                               mode, alm, verbose)

    def dlm2pointing(self, dlm=None, mmax_dlm=None, single_prec=None, nthreads=None, epsilon=1e-14, verbose=1):
        """Deflected (theta, phi) as (npix, 2) (:655-663)."""
        d = self._dlm(dlm)
        lmax = Alm.getlmax(np.atleast_2d(d)[0].size, mmax_dlm)
        ptg = self.deflectionlib.dlm2pointing(np.atleast_2d(d)[0], lmax, bool(single_prec), nthreads)
        return np.ascontiguousarray(ptg.T)

    def gclm2lenmap(self, gclm, lmax, mmax, spin, nthreads, dlm=None, epsilon=None, backwards=False, polrot=True,
                    ptg=None, dclm=None, lenmap=None, verbose=None, execmode="normal"):
        """Lensed map(s) (:664-696)."""
        eps = self.epsilon if epsilon is None else epsilon
        assert eps is not None
        return CPU_nusht_transformer.gclm2lenmap(self, gclm, None if ptg is not None else self._dlm(dlm, dclm),
                                                 lmax, mmax, spin, eps, nthreads, polrot=polrot, ptg=ptg,
                                                 lenmap=lenmap, verbose=self.verbosity if verbose is None else verbose,
                                                 execmode=execmode)

    def lenmap2gclm(self, lenmap, dlm, spin, lmax, mmax, nthreads, epsilon=None, gclm_out=None, ptg=None,
                    verbose=None, execmode="normal", polrot=True):
        """Exact adjoint of gclm2lenmap (:698-731)."""
        eps = self.epsilon if epsilon is None else epsilon
        assert eps is not None
        self.epsilon = eps
        return CPU_nusht_transformer.lenmap2gclm(self, lenmap, None if ptg is not None else self._dlm(dlm),
                                                 gclm_out, spin, lmax, mmax, nthreads, epsilon=eps, ptg=ptg,
                                                 verbose=self.verbosity if verbose is None else verbose,
                                                 execmode=execmode, polrot=polrot)

    def gclm2lenpixs(self, gclm, mmax, spin, pixs, polrot=True, ptg=None):
        """As the DUCC-signature method; ptg is (npix, 2) here (:734-762)."""
        return CPU_nusht_transformer.gclm2lenpixs(self, gclm, mmax, spin, pixs, polrot=polrot,
                                                  ptg=None if ptg is None else np.asarray(ptg).T)
