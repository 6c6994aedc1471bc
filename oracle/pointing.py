"""Deflected angles and polarisation rotation angle, FP64 (oracle; test infrastructure only).

deflected_angles restates cunusht/c/pointing.cu:205-223 (per pixel: a = |d|,
n' from sin/cos theta, cos a, sin(a)/a; theta' = atan2(sqrt(n1^2+n2^2), n3),
phi' = phi - atan2(n2, n1), wrapped only when >= 2 pi).  The reference's branch choice
between the exact and the Taylor sin(a)/a is driven by a defective test
(pointing.cu:89 reads element [1]); here the exact branch is used unless a < 1e-3,
where the series is accurate to 1e-19.

gamma restates cunusht/deflection/CPU_nufft_transformer.py:158-160:
    gamma = atan2(d_phi, d_theta) - atan2(d_phi, |d| sin|d| cot(theta) + d_theta cos|d|)

Parity: both are pinned by rotation_3d below, an independent construction with 3-vectors
(n' = cos a n + sin a d_hat; gamma = angle of the geodesic's direction at n minus its
angle at n', each measured from e_theta towards e_phi), and by the special cases
d = 0 (identity) and pure d_theta (theta' = theta + d_theta).
"""
from __future__ import annotations

import numpy as np


def _sinc(a):
    a2 = a * a
    series = 1.0 - a2 / 6.0 * (1.0 - a2 / 20.0 * (1.0 - a2 / 42.0))
    with np.errstate(invalid="ignore", divide="ignore"):
        exact = np.sin(a) / a
    return np.where(a < 1e-3, series, exact)


def deflected_angles(theta, phi, dt, dp):
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    dt = np.asarray(dt, dtype=np.float64)
    dp = np.asarray(dp, dtype=np.float64)
    a = np.sqrt(dt * dt + dp * dp)
    sa = _sinc(a)
    ca = np.cos(a)
    st, ct = np.sin(theta), np.cos(theta)
    n1 = st * ca + ct * dt * sa
    n2 = -dp * sa
    n3 = ct * ca - st * dt * sa
    tp = np.arctan2(np.sqrt(n1 * n1 + n2 * n2), n3)
    pp = phi - np.arctan2(n2, n1)
    pp = np.where(pp >= 2 * np.pi, pp - 2 * np.pi, pp)
    return tp, pp


def gamma(theta, dt, dp):
    theta = np.asarray(theta, dtype=np.float64)
    d = np.sqrt(dt * dt + dp * dp)
    cot = np.cos(theta) / np.sin(theta)
    return np.arctan2(dp, dt) - np.arctan2(dp, d * np.sin(d) * cot + dt * np.cos(d))


def rotation_3d(theta, phi, dt, dp):
    """Independent 3-vector construction of (theta', phi', gamma)."""
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    st, ct, sp, cp = np.sin(theta), np.cos(theta), np.sin(phi), np.cos(phi)
    n = np.stack([st * cp, st * sp, ct])
    et = np.stack([ct * cp, ct * sp, -st])
    ep = np.stack([-sp, cp, np.zeros_like(sp)])
    a = np.sqrt(dt * dt + dp * dp)
    dhat = (dt * et + dp * ep) / np.where(a > 0, a, 1.0)
    n2 = np.cos(a) * n + np.sin(a) * dhat
    tdir = -np.sin(a) * n + np.cos(a) * dhat  # geodesic direction at n'
    tp = np.arctan2(np.hypot(n2[0], n2[1]), n2[2])
    pp = np.arctan2(n2[1], n2[0])
    stp, ctp, spp, cpp = np.sin(tp), np.cos(tp), np.sin(pp), np.cos(pp)
    et2 = np.stack([ctp * cpp, ctp * spp, -stp])
    ep2 = np.stack([-spp, cpp, np.zeros_like(spp)])
    beta = np.arctan2(np.sum(tdir * ep2, 0), np.sum(tdir * et2, 0))
    alpha = np.arctan2(dp, dt)
    g = alpha - beta
    g = (g + np.pi) % (2 * np.pi) - np.pi
    return tp, pp % (2 * np.pi), g
