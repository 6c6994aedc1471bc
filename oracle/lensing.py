"""Magnification determinant of the lensing deflection (oracle; test infrastructure only).

magnification() restates cunusht/helper_GPU.py:113-161 (dlm2A):
    kappa = spin-0 synthesis of  0.5 sqrt(l(l+1)) dlm           (-1/2 times spin lowering, :88-110)
    gamma = spin-2 synthesis of -0.5 sqrt((l-1)(l+2)) [dlm, dclm] (-1/2 times spin raising)
    d     = spin-1 synthesis of [dlm, dclm]
    omega = spin-0 synthesis of  0.5 sqrt(l(l+1)) dclm
    A = f0 ((1 - k)^2 - g1^2 - g2^2 + w^2) + f1 (1 - k - ((d1^2 - d2^2) g1 + 2 d1 d2 g2) / |d|^2),
    f0 = sin|d| / |d|, f1 = cos|d| - f0.
The reference divides by |d| where |d| = 0 (NaN there); here f0 and f1 / |d|^2 use their
series below |d| = 1e-4.

Parity: pinned by solid_angle_ratio(), an independent construction: the Jacobian determinant
dOmega'/dOmega of the deflection map n -> n' = cos|d| n + sin|d| d_hat (d = d1 e_theta +
d2 e_phi, the map of cunusht/c/pointing.cu:205-223), by 4th-order central differences of the
deflected 3-vector.  With the spin-s convention of this package (Q + iU = -sum (E + iB) sY),
the formula equals that Jacobian exactly once the sign of the shear in the |d|-dependent term
is reversed (exact=True); the reference's sign (helper_GPU.py:126, marked "check the sign")
agrees only through second order in |d| (exact=False reproduces it).
"""
from __future__ import annotations

import numpy as np

from . import almutil, sht


def spin_raise(s: int, lmax: int) -> np.ndarray:
    """sqrt((l - s)(l + s + 1)) for |s| <= l <= lmax (cunusht/helper_GPU.py:88-97)."""
    ret = np.zeros(lmax + 1)
    ll = np.arange(abs(s), lmax + 1, dtype=np.float64)
    ret[abs(s):] = np.sqrt((ll - s) * (ll + s + 1))
    return ret


def spin_lower(s: int, lmax: int) -> np.ndarray:
    """-sqrt((l + s)(l - s + 1)) for |s| <= l <= lmax (cunusht/helper_GPU.py:99-110)."""
    ret = np.zeros(lmax + 1)
    ll = np.arange(abs(s), lmax + 1, dtype=np.float64)
    ret[abs(s):] = -np.sqrt((ll + s) * (ll - s + 1))
    return ret


def combine(k, g1, g2, d1, d2, w, exact: bool = True):
    """The pointwise formula of cunusht/helper_GPU.py:150-160 (exact=False: the reference's shear sign
    in the |d|-dependent term)."""
    d2sum = d1 * d1 + d2 * d2
    d = np.sqrt(d2sum)
    small = d < 1e-4
    with np.errstate(invalid="ignore", divide="ignore"):
        f0 = np.where(small, 1.0 - d2sum / 6.0, np.sin(d) / d)
        f1 = np.cos(d) - f0
        f1od2 = np.where(small, -1.0 / 3.0 + d2sum / 30.0, f1 / d2sum)
    sg = 1.0 if exact else -1.0
    return (f0 * ((1.0 - k) ** 2 - g1 * g1 - g2 * g2 + w * w) + f1 * (1.0 - k)
            + sg * f1od2 * ((d1 * d1 - d2 * d2) * g1 + 2.0 * d1 * d2 * g2))


def fields(dlm, lmax: int, theta, phi, dclm=None):
    """(kappa, g1, g2, d1, d2, omega) at the points by direct summation."""
    d2k = -0.5 * spin_lower(1, lmax)
    d2g = -0.5 * spin_raise(1, lmax)
    dcl = np.zeros_like(dlm) if dclm is None else dclm
    k = sht.synth_direct(almutil.almxfl(dlm, d2k, lmax), lmax, 0, theta, phi)
    g = sht.synth_direct(np.array([almutil.almxfl(dlm, d2g, lmax), almutil.almxfl(dcl, d2g, lmax)]), lmax, 2,
                         theta, phi)
    d = sht.synth_direct(np.array([dlm, dcl]), lmax, 1, theta, phi)
    w = sht.synth_direct(almutil.almxfl(dcl, d2k, lmax), lmax, 0, theta, phi) if dclm is not None \
        else np.zeros_like(k)
    return k, g.real, g.imag, d.real, d.imag, w


def magnification(dlm, lmax: int, theta, phi, dclm=None, exact: bool = True):
    return combine(*fields(dlm, lmax, theta, phi, dclm), exact=exact)


def _deflected_vector(dlm, lmax, theta, phi, dclm):
    dcl = np.zeros_like(dlm) if dclm is None else dclm
    d = sht.synth_direct(np.array([dlm, dcl]), lmax, 1, theta, phi)
    d1, d2 = d.real, d.imag
    a = np.sqrt(d1 * d1 + d2 * d2)
    with np.errstate(invalid="ignore", divide="ignore"):
        sa = np.where(a > 0, np.sin(a) / np.where(a > 0, a, 1.0), 1.0)
    st, ct, sp, cp = np.sin(theta), np.cos(theta), np.sin(phi), np.cos(phi)
    n = np.stack([st * cp, st * sp, ct])
    et = np.stack([ct * cp, ct * sp, -st])
    ep = np.stack([-sp, cp, np.zeros_like(sp)])
    return np.cos(a) * n + sa * (d1 * et + d2 * ep)


def solid_angle_ratio(dlm, lmax: int, theta, phi, dclm=None, h: float = 1e-3):
    """dOmega'/dOmega of the deflection map at the points (central differences, error O(h^4))."""
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)

    def deriv(dth, dph):
        f = lambda s: _deflected_vector(dlm, lmax, theta + s * dth, phi + s * dph, dclm)
        return (-f(2 * h) + 8 * f(h) - 8 * f(-h) + f(-2 * h)) / (12 * h)

    nt = deriv(1.0, 0.0)
    nph = deriv(0.0, 1.0)
    n1 = _deflected_vector(dlm, lmax, theta, phi, dclm)
    return np.einsum("ip,ip->p", np.cross(nt, nph, axis=0), n1) / np.sin(theta)
