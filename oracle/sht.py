"""Spin-weighted spherical harmonics in FP64 numpy (oracle; test infrastructure only).

Convention (orthonormal, Condon-Shortley phase, healpy/libsharp spin convention):

    sY_lm(theta, phi) = (-1)^s sqrt((2l+1)/(4 pi)) d^l_{m,-s}(theta) e^{i m phi}

with d^j_{m'm} the Wigner small-d matrix.  For s = 1 this is
sY_lm = -(d_theta + i/sin(theta) d_phi) Y_lm / sqrt(l(l+1)), so spin-1 GRAD_ONLY
synthesis of dlm equals (d_theta, d_phi/sin(theta)) of sum_lm dlm/sqrt(l(l+1)) Y_lm --
the identity the reference's pointing test relies on
(test/test_pointingGPU.py:93-101, SHTns synth_grad vs ducc spin-1 GRAD_ONLY).

Spin-s fields (s > 0) follow the convention of the reference CPU backend's
ducc spin SHT (cunusht/deflection/CPU_nufft_transformer.py:320-325):

    Q + iU = - sum_lm (E_lm + i B_lm) sY_lm

with E, B the alm of two real fields (a_{l,-m} = (-1)^m conj(a_lm)).

The l-recursion is the standard three-term Wigner-d recursion at fixed (m, m'),
restated here from the textbook relation

    l sqrt(((l+1)^2-m^2)((l+1)^2-m'^2)) d^{l+1}
        = (2l+1) (l(l+1) cos(theta) - m m') d^l - (l+1) sqrt((l^2-m^2)(l^2-m'^2)) d^{l-1}

started from the closed-form d^{l0}_{m,m'} at l0 = max(|m|, |m'|).  Underflow
near the poles is avoided with log-space starting values and a running binary
scale exponent.
"""
from __future__ import annotations

import math
import numpy as np

from .almutil import nalm, almidx

_SCALE_BITS = 600
_BIG = 2.0 ** _SCALE_BITS
_SMALL = 2.0 ** (-_SCALE_BITS)
_THRESH = 2.0 ** (_SCALE_BITS // 2)


def _log_start(l0: int, m: int, mprime: int, theta: np.ndarray):
    """log|d^{l0}_{m,mprime}(theta)| and its sign, l0 = max(|m|, |mprime|), m >= 0."""
    c = np.cos(0.5 * theta)
    s = np.sin(0.5 * theta)
    with np.errstate(divide="ignore"):
        lc = np.log(np.abs(c))
        ls = np.log(np.abs(s))
    if l0 == m:  # first index equals j: d^j_{j,mu} = (-1)^{j-mu} sqrt((2j)!/((j+mu)!(j-mu)!)) c^{j+mu} s^{j-mu}
        mu = mprime
        lg = 0.5 * (math.lgamma(2 * l0 + 1) - math.lgamma(l0 + mu + 1) - math.lgamma(l0 - mu + 1))
        pc, ps = l0 + mu, l0 - mu
        sign = (-1.0) ** (l0 - mu)
    elif mprime == l0:  # d^j_{m',j} = sqrt((2j)!/((j+m')!(j-m')!)) c^{j+m'} s^{j-m'}
        lg = 0.5 * (math.lgamma(2 * l0 + 1) - math.lgamma(l0 + m + 1) - math.lgamma(l0 - m + 1))
        pc, ps = l0 + m, l0 - m
        sign = 1.0
    else:  # mprime == -l0: d^j_{m',-j} = (-1)^{m'+j} sqrt((2j)!/((j+m')!(j-m')!)) c^{j-m'} s^{j+m'}
        lg = 0.5 * (math.lgamma(2 * l0 + 1) - math.lgamma(l0 + m + 1) - math.lgamma(l0 - m + 1))
        pc, ps = l0 - m, l0 + m
        sign = (-1.0) ** (m + l0)
    with np.errstate(invalid="ignore"):
        logv = lg + np.where(pc > 0, pc * lc, 0.0) + np.where(ps > 0, ps * ls, 0.0)
    sgn = sign * np.where(pc % 2 == 1, np.sign(c), 1.0) * np.where(ps % 2 == 1, np.sign(s), 1.0)
    return logv, sgn


def _split_scaled(logv, sgn):
    """Represent exp(logv)*sgn as v * 2^(600*k) with |v| in a safe range."""
    # k = 0 whenever |value| > 2^-300; otherwise v lands in (2^-300, 2^300].
    k = np.ceil((logv / math.log(2.0) - _SCALE_BITS // 2) / _SCALE_BITS)
    k = np.where(np.isfinite(k), k, -10**6)
    k = np.minimum(k, 0)
    v = sgn * np.exp(np.where(np.isfinite(logv), logv - k * _SCALE_BITS * math.log(2.0), -np.inf))
    return v, k.astype(np.int64)


def wigner_lambda(lmax: int, m: int, mprime: int, theta: np.ndarray) -> np.ndarray:
    """sqrt((2l+1)/(4 pi)) d^l_{m,mprime}(theta) for l = 0..lmax (zero for l < l0).

    Returns array of shape (lmax+1, ntheta).  m >= 0, |mprime| <= lmax.
    """
    theta = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    out = np.zeros((lmax + 1, theta.size))
    l0 = max(abs(m), abs(mprime))
    if l0 > lmax:
        return out
    x = np.cos(theta)
    logv, sgn = _log_start(l0, m, mprime, theta)
    logv = logv + 0.5 * math.log((2 * l0 + 1) / (4 * math.pi))
    cur, k = _split_scaled(logv, sgn)
    prev = np.zeros_like(cur)
    out[l0] = np.where(k == 0, cur, 0.0)
    mm = float(m * mprime)
    for l in range(l0, lmax):
        if l == 0:
            nxt = math.sqrt(3.0) * x * cur
        else:
            d1 = math.sqrt(((l + 1) ** 2 - m * m) * ((l + 1) ** 2 - mprime * mprime))
            d0 = math.sqrt((l * l - m * m) * (l * l - mprime * mprime))
            f = math.sqrt((2 * l + 3) * (2 * l + 1))
            alpha = f * (l + 1) / d1
            beta = f * mm / (l * d1)
            gamma = math.sqrt((2 * l + 3) / (2 * l - 1)) * (l + 1) * d0 / (l * d1)
            nxt = (alpha * x - beta) * cur - gamma * prev
        prev, cur = cur, nxt
        big = (k < 0) & (np.abs(cur) > _THRESH)
        cur = np.where(big, cur * _SMALL, cur)
        prev = np.where(big, prev * _SMALL, prev)
        k = k + big
        out[l + 1] = np.where(k == 0, cur, 0.0)
    return out


def spin_lambda(lmax: int, spin: int, m: int, theta: np.ndarray) -> np.ndarray:
    """sY_lm(theta, 0) for l = 0..lmax at fixed m (any sign of m and spin)."""
    if m >= 0:
        return (-1.0) ** spin * wigner_lambda(lmax, m, -spin, theta)
    # sY_{l,-|m|} = (-1)^(s+m) conj(-sY_{l,|m|}); lambda is real
    return (-1.0) ** (spin + m) * spin_lambda(lmax, -spin, -m, theta)


def synth_ring_coeffs(alm: np.ndarray, lmax: int, spin: int, theta: np.ndarray) -> np.ndarray:
    """Fourier coefficients G_m(theta) of the synthesized field, m = -lmax..lmax.

    spin == 0: alm shape (nalm,), field = sum_lm a_lm Y_lm (real).
    spin > 0 : alm shape (2, nalm) = (E, B), field Q+iU = -sum (E+iB) sY_lm.
    Returns complex array (2*lmax+1, ntheta), row index m + lmax.
    """
    theta = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    G = np.zeros((2 * lmax + 1, theta.size), dtype=np.complex128)
    if spin == 0:
        a = np.asarray(alm).reshape(-1)
        for m in range(lmax + 1):
            lam = spin_lambda(lmax, 0, m, theta)[m:]
            g = a[almidx(lmax, np.arange(m, lmax + 1), m)] @ lam
            G[lmax + m] = g
            if m > 0:
                G[lmax - m] = np.conj(g)
        return G
    E = np.asarray(alm[0]).reshape(-1)
    B = np.asarray(alm[1]).reshape(-1) if len(alm) > 1 else np.zeros_like(E)
    for m in range(-lmax, lmax + 1):
        am = abs(m)
        idx = almidx(lmax, np.arange(am, lmax + 1), am)
        e, b = E[idx], B[idx]
        if m < 0:
            e = (-1.0) ** am * np.conj(e)
            b = (-1.0) ** am * np.conj(b)
        lam = spin_lambda(lmax, spin, m, theta)[am:]
        G[lmax + m] = -((e + 1j * b) @ lam)
    return G


def synth_direct(alm, lmax: int, spin: int, theta, phi) -> np.ndarray:
    """Direct FP64 summation at arbitrary points.

    spin 0 -> real array; spin > 0 -> complex Q+iU.
    """
    theta = np.asarray(theta, dtype=np.float64).reshape(-1)
    phi = np.asarray(phi, dtype=np.float64).reshape(-1)
    G = synth_ring_coeffs(alm, lmax, spin, theta)
    ms = np.arange(-lmax, lmax + 1)
    val = np.einsum("mp,mp->p", G, np.exp(1j * np.outer(ms, phi)))
    return val.real if spin == 0 else val


def adjoint_direct(values, lmax: int, spin: int, theta, phi) -> np.ndarray:
    """Exact adjoint of synth_direct w.r.t. the real-field alm inner product.

    <a, b> = sum_l a_l0 b_l0 + 2 sum_{m>0} Re(a_lm conj(b_lm)) per component, and
    <x, y> = sum Re(conj(x) y) on point values (test/test_adjointOperator.py:39-48,265-268).
    """
    theta = np.asarray(theta, dtype=np.float64).reshape(-1)
    phi = np.asarray(phi, dtype=np.float64).reshape(-1)
    ms = np.arange(-lmax, lmax + 1)
    ph = np.exp(-1j * np.outer(ms, phi))
    v = np.asarray(values)
    n = nalm(lmax, lmax)
    if spin == 0:
        v = np.real(v).astype(np.float64)
        out = np.zeros(n, dtype=np.complex128)
        for m in range(lmax + 1):
            lam = spin_lambda(lmax, 0, m, theta)[m:]
            Gp = (ph[lmax + m] * v)  # sum_p v_p e^{-i m phi_p} lambda
            b = lam @ Gp
            if m == 0:
                b = b.real.astype(np.complex128)
            out[almidx(lmax, np.arange(m, lmax + 1), m)] = b
        return out
    v = v.astype(np.complex128)
    Pm = ph * v[None, :]  # P'_m(p) = v_p e^{-i m phi_p}
    E = np.zeros(n, dtype=np.complex128)
    B = np.zeros(n, dtype=np.complex128)
    sgn = (-1.0) ** spin
    for m in range(lmax + 1):
        idx = almidx(lmax, np.arange(m, lmax + 1), m)
        u = spin_lambda(lmax, spin, m, theta)[m:]      # sY_lm
        w = spin_lambda(lmax, -spin, m, theta)[m:]     # (-s)Y_lm
        X = u @ Pm[lmax + m]                                 # sum_p P'_m u
        Y = w @ np.conj(Pm[lmax - m])                        # sum_p conj(P'_{-m}) v
        E[idx] = -0.5 * (X + sgn * Y)
        B[idx] = 0.5j * (X - sgn * Y)
    return np.array([E, B])
