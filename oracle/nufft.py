"""2D type-1 / type-2 nuFFT with the ES kernel, FP64 (oracle; test infrastructure only).

Semantics follow the reference's nuFFT calls (cunusht/deflection/GPU_nufft_transformer.py:242-246
with modeord=1, i.e. modes in FFT order; CPU_nufft_transformer.py:341-351 with fft_order=True):

  type 2:  c_j = sum_{k1,k2} F[k1 mod n1, k2 mod n2] e^{+i (k1 x1_j + k2 x2_j)}
  type 1:  F[k1, k2] = sum_j c_j e^{-i (k1 x1_j + k2 x2_j)}

with k in [-n/2, n/2).  Method: Barnett, Magland & af Klinteberg (2019): upsampling
sigma = 2, width w = ceil(log10(1/eps)) + 2, beta = 2.30 w, deconvolution by the kernel's
Fourier transform computed with Gauss-Legendre quadrature.

Parity: pinned against the brute-force NUDFT (nudft2d1/nudft2d2 below) and by the
type-1/type-2 dot-product adjointness.
"""
from __future__ import annotations

import os

import numpy as np
from scipy import fft as sfft

from . import clib


def fft_workers() -> int:
    """Threads for the oracle's FFTs (scipy.fft; the transform itself is the textbook DFT): the OpenMP
    thread count when set, else the cores this process may use."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = len(os.sched_getaffinity(0))
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def es_params(eps: float):
    w = int(np.ceil(np.log10(1.0 / eps))) + 2
    w = max(w, 3)
    return w, 2.30 * w


def es_fourier(k: np.ndarray, nf: int, w: int, beta: float) -> np.ndarray:
    """int_{-w/2}^{w/2} psi(2t/w) cos(2 pi k t / nf) dt."""
    xq, wq = np.polynomial.legendre.leggauss(4 * w + 60)
    t = 0.5 * w * xq
    psi = np.exp(beta * (np.sqrt(1.0 - (2 * t / w) ** 2) - 1.0))
    return 0.5 * w * (np.cos(2 * np.pi * np.outer(k, t) / nf) @ (wq * psi))


def _fine(n: int) -> int:
    return 2 * n


def _kidx(n: int) -> np.ndarray:
    return np.fft.fftfreq(n, 1.0 / n).astype(np.int64)


def nufft2d2(F: np.ndarray, x1, x2, eps: float) -> np.ndarray:
    n1, n2 = F.shape
    w, beta = es_params(eps)
    nf1, nf2 = _fine(n1), _fine(n2)
    k1, k2 = _kidx(n1), _kidx(n2)
    d1 = es_fourier(k1, nf1, w, beta)
    d2 = es_fourier(k2, nf2, w, beta)
    G = np.zeros((nf1, nf2), dtype=np.complex128)
    G[np.ix_(k1 % nf1, k2 % nf2)] = F / np.outer(d1, d2)
    g = sfft.ifft2(G, workers=fft_workers()) * (nf1 * nf2)
    return clib.es_interp(g, x1, x2, w, beta)


def nufft2d1(c, x1, x2, shape, eps: float) -> np.ndarray:
    n1, n2 = shape
    w, beta = es_params(eps)
    nf1, nf2 = _fine(n1), _fine(n2)
    k1, k2 = _kidx(n1), _kidx(n2)
    d1 = es_fourier(k1, nf1, w, beta)
    d2 = es_fourier(k2, nf2, w, beta)
    g = clib.es_spread((nf1, nf2), x1, x2, np.asarray(c, dtype=np.complex128), w, beta)
    G = sfft.fft2(g, workers=fft_workers())
    return G[np.ix_(k1 % nf1, k2 % nf2)] / np.outer(d1, d2)


def nudft2d2(F: np.ndarray, x1, x2) -> np.ndarray:
    """Brute force type 2 (small inputs only)."""
    n1, n2 = F.shape
    k1, k2 = _kidx(n1), _kidx(n2)
    e1 = np.exp(1j * np.outer(np.asarray(x1), k1))
    e2 = np.exp(1j * np.outer(np.asarray(x2), k2))
    return np.einsum("ja,ab,jb->j", e1, F, e2)


def nudft2d1(c, x1, x2, shape) -> np.ndarray:
    """Brute force type 1 (small inputs only)."""
    n1, n2 = shape
    k1, k2 = _kidx(n1), _kidx(n2)
    e1 = np.exp(-1j * np.outer(np.asarray(x1), k1))
    e2 = np.exp(-1j * np.outer(np.asarray(x2), k2))
    return np.einsum("j,ja,jb->ab", np.asarray(c), e1, e2)
