"""CPU restatement of the reference nuSHT pipeline, FP64 (oracle; test infrastructure only).

Follows CPU_DUCCnufft_transformer (cunusht/deflection/CPU_nufft_transformer.py:181-529) stage by stage,
on the reference's own CC sizes (GPU_nufft_transformer.py:138-141):

  synthesis_general          :390-405   CC synthesis -> doubling (:284-296) -> c2c/N (:266-273) -> u2nu (:346-351)
  adjoint_synthesis_general  :410-427   nu2u (:336-344) -> c2c backward /N (:278-279) -> adjoint doubling
                                        (:298-318) -> adjoint CC synthesis (:327-331)
  gclm2lenmap / lenmap2gclm  :429-529   with the pointing from dlm (:118-166) and the rotation (:356-368)

The CC synthesis is oracle/src/oracle_sht.c followed by a phi-FFT; the nuFFT is oracle/nufft.py.
The adjoint applies exp(-i s gamma), the exact adjoint of the forward rotation.

Parity: pinned against direct FP64 Y_lm summation at the deflected points (oracle/sht.py) at small
lmax, and by the dot-product adjointness of each pair (test/test_adjointOperator.py:247-277).
"""
from __future__ import annotations

import numpy as np
from scipy import fft as sfft

from . import clib, nufft
from .almutil import almxfl
from .geometry import Rings, cc_thetas, reference_cc_sizes
from .pointing import deflected_angles, gamma as gamma_fn


def phi_synth(G: np.ndarray, lmax: int, spin: int, nphi: int, phi0=None) -> np.ndarray:
    """Rows of G (m = 0..lmax for spin 0, -lmax..lmax otherwise) -> ring values on nphi points."""
    nring = G.shape[1]
    arr = np.zeros((nring, nphi), dtype=np.complex128)
    ms = np.arange(lmax + 1) if spin == 0 else np.arange(-lmax, lmax + 1)
    ph = np.ones((ms.size, nring)) if phi0 is None else np.exp(1j * np.outer(ms, phi0))
    Gs = G * ph
    if spin == 0:
        np.add.at(arr, (slice(None), ms % nphi), Gs.T)
        neg = ms[1:]
        np.add.at(arr, (slice(None), (-neg) % nphi), np.conj(Gs[1:]).T)
    else:
        np.add.at(arr, (slice(None), ms % nphi), Gs.T)
    out = sfft.ifft(arr, axis=1, workers=nufft.fft_workers()) * nphi
    return out.real if spin == 0 else out


def phi_adjoint(M: np.ndarray, lmax: int, spin: int, phi0=None) -> np.ndarray:
    """Exact adjoint of phi_synth (per ring): G'_m = sum_k M_k e^{-i m phi_k}."""
    nring, nphi = M.shape
    F = sfft.fft(np.asarray(M, dtype=np.complex128), axis=1, workers=nufft.fft_workers())
    ms = np.arange(lmax + 1) if spin == 0 else np.arange(-lmax, lmax + 1)
    G = F[:, ms % nphi].T
    if phi0 is not None:
        G = G * np.exp(-1j * np.outer(ms, phi0))
    return G


class NuSHT:
    """The reference CPU algorithm for one lmax (mmax = lmax)."""

    def __init__(self, lmax: int):
        self.lmax = lmax
        self.ntheta, self.nphi = reference_cc_sizes(lmax)
        self.thetas = cc_thetas(self.ntheta)

    def _doubling(self, cmap: np.ndarray, spin: int) -> np.ndarray:
        nt, nph = self.ntheta, self.nphi
        h = nph // 2
        d = np.empty((2 * nt - 2, nph), dtype=np.complex128)
        d[:nt] = cmap
        d[nt:, :h] = d[nt - 2:0:-1, h:]
        d[nt:, h:] = d[nt - 2:0:-1, :h]
        if spin % 2:
            d[nt:] *= -1
        return d

    def _adjoint_doubling(self, d: np.ndarray, spin: int) -> np.ndarray:
        nt, nph = self.ntheta, self.nphi
        h = nph // 2
        d = d.copy()
        sg = -1.0 if spin % 2 else 1.0
        d[1:nt - 1, :h] += sg * d[-1:nt - 1:-1, h:]
        d[1:nt - 1, h:] += sg * d[-1:nt - 1:-1, :h]
        return d[:nt]

    def synthesis_general(self, alm, spin: int, theta, phi, eps: float) -> np.ndarray:
        L = self.lmax
        G = clib.sht_synth_rings(alm, L, spin, self.thetas)
        cmap = phi_synth(G, L, spin, self.nphi)
        d = self._doubling(cmap, spin)
        F = sfft.fft2(d, workers=nufft.fft_workers()) / d.size
        c = nufft.nufft2d2(F, theta, phi, eps)
        return c.real if spin == 0 else c

    def adjoint_synthesis_general(self, values, spin: int, ncomp: int, theta, phi, eps: float) -> np.ndarray:
        L = self.lmax
        shape = (2 * self.ntheta - 2, self.nphi)
        F = nufft.nufft2d1(np.asarray(values, dtype=np.complex128), theta, phi, shape, eps)
        d = sfft.ifft2(F, workers=nufft.fft_workers())
        cmap = self._adjoint_doubling(d, spin)
        if spin == 0:
            cmap = cmap.real
        G = phi_adjoint(cmap, L, spin)
        return clib.sht_adjoint_rings(G, L, spin, ncomp, self.thetas)


def dlm2pointing(dlm_scaled, lmax: int, rings: Rings, with_gamma: bool = True):
    """Spin-1 gradient synthesis of dlm = sqrt(l(l+1)) dlm_scaled on the rings, then the pointing.

    Mirrors GPU_nufft_transformer.py:64-104 (synth_grad + pointing kernel) and
    CPU_nufft_transformer.py:118-166 (gamma)."""
    ll = np.arange(lmax + 1, dtype=np.float64)
    dlm = almxfl(np.asarray(dlm_scaled, dtype=np.complex128), np.sqrt(ll * (ll + 1)), lmax)
    G = clib.sht_synth_rings(dlm[None, :], lmax, 1, rings.theta)
    th, ph = rings.pixel_angles()
    dt = np.empty(th.size)
    dp = np.empty(th.size)
    for n in np.unique(rings.nph):  # rings of one length at once
        sel = np.flatnonzero(rings.nph == n)
        v = phi_synth(G[:, sel], lmax, 1, int(n), phi0=rings.phi0[sel])
        pix = (rings.ofs[sel][:, None] + np.arange(n)[None, :]).reshape(-1)
        dt[pix] = v.real.reshape(-1)
        dp[pix] = v.imag.reshape(-1)
    tp, pp = deflected_angles(th, ph, dt, dp)
    g = gamma_fn(th, dt, dp) if with_gamma else None
    return tp, pp, g, (dt, dp)


def gclm2lenmap(gclm, dlm_scaled, lmax: int, spin: int, rings: Rings, eps: float, polrot: bool = True):
    """CPU_nufft_transformer.py:429-484: returns real map (spin 0) or (Q, U)."""
    tp, pp, g, _ = dlm2pointing(dlm_scaled, lmax, rings, with_gamma=spin > 0)
    op = NuSHT(lmax)
    v = op.synthesis_general(gclm, spin, tp, pp, eps)
    if spin == 0:
        return v
    if polrot:
        v = v * np.exp(1j * spin * g)
    return np.array([v.real, v.imag])


def lenmap2gclm(lenmap, dlm_scaled, lmax: int, spin: int, ncomp: int, rings: Rings, eps: float,
                polrot: bool = True):
    """Exact adjoint of gclm2lenmap (CPU_nufft_transformer.py:486-529)."""
    tp, pp, g, _ = dlm2pointing(dlm_scaled, lmax, rings, with_gamma=spin > 0)
    op = NuSHT(lmax)
    lm = np.asarray(lenmap)
    if spin == 0:
        v = np.real(lm).reshape(-1).astype(np.complex128)
    else:
        v = lm[0] + 1j * lm[1]
        if polrot:
            v = v * np.exp(-1j * spin * g)
    return op.adjoint_synthesis_general(v, spin, ncomp, tp, pp, eps)
