"""Spin-s fields of constant ambient tensors and their exact lensing by parallel transport (oracle; test
infrastructure only).

This is the independent check of the polarisation rotation that the reference applies through ducc0
`lensing_rotate` (cunusht/deflection/CPU_nufft_transformer.py:356-368, angle gamma from :158-160): both
its angle and its direction.  It uses no spherical harmonics, no gamma formula and no pointing formula:

* a constant real 3-vector v gives the spin-1 field  f(n) = v . m(n),
  a constant real symmetric traceless 3x3 tensor T gives the spin-2 field  f(n) = m(n)^T T m(n) / 2,
  with m = e_theta + i e_phi.  With the spin-weighted harmonics of this oracle (sht.py, pinned by the eth
  operator) these are band-limited spin +s fields at l = s; the tests check that, since with
  m = e_theta - i e_phi the field would be spin -s instead.
* The lensed field at n takes the field at n' = cos|d| n + sin|d| d_hat (the end of the geodesic leaving n
  along the deflection d = d_theta e_theta + d_phi e_phi) with the polarisation parallel-transported along
  that geodesic back to n.  Parallel transport along a great circle keeps the binormal b = n x d_hat and
  turns the tangent d_hat into the geodesic's tangent at n', so the lensed value is the tensor contracted
  with m(n) carried to n':  m' = (m . d_hat) t' + (m . b) b,  t' = -sin|d| n + cos|d| d_hat.

Parity: this construction pins the e^{+i s gamma} of the oracle pipeline (pipeline.gclm2lenmap) to rounding
at deflections up to 1.4 rad, and the opposite direction fails it by O(1).
"""
from __future__ import annotations

import numpy as np


def frame(theta, phi):
    """n, e_theta, e_phi as (3, N) arrays."""
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    st, ct, sp, cp = np.sin(theta), np.cos(theta), np.sin(phi), np.cos(phi)
    n = np.stack([st * cp, st * sp, ct])
    et = np.stack([ct * cp, ct * sp, -st])
    ep = np.stack([-sp, cp, np.zeros_like(sp)])
    return n, et, ep


def _contract(spin: int, tensor, m):
    if spin == 1:
        return np.einsum("i,ip->p", np.asarray(tensor), m)
    if spin == 2:
        return 0.5 * np.einsum("ip,ij,jp->p", m, np.asarray(tensor), m)
    raise ValueError("spin must be 1 or 2")


def ambient_field(spin: int, tensor, theta, phi) -> np.ndarray:
    """Complex spin-s field (Q + iU for s = 2) of the constant ambient vector (s = 1) or symmetric
    traceless tensor (s = 2)."""
    _, et, ep = frame(theta, phi)
    return _contract(spin, tensor, et + 1j * ep)


def lensed_ambient_field(spin: int, tensor, theta, phi, d_theta, d_phi) -> np.ndarray:
    """The same field lensed by the deflection (d_theta, d_phi) at each point, by parallel transport."""
    n, et, ep = frame(theta, phi)
    dt = np.asarray(d_theta, dtype=np.float64)
    dp = np.asarray(d_phi, dtype=np.float64)
    a = np.hypot(dt, dp)
    if np.any(a == 0):
        raise ValueError("a zero deflection has no geodesic direction; the field is then unchanged")
    dhat = (dt * et + dp * ep) / a
    b = np.cross(n.T, dhat.T).T
    tprime = -np.sin(a) * n + np.cos(a) * dhat
    m = et + 1j * ep
    mprime = np.sum(m * dhat, 0) * tprime + np.sum(m * b, 0) * b
    return _contract(spin, tensor, mprime)


def random_tensor(spin: int, rng: np.random.Generator):
    """A random vector (s = 1) or symmetric traceless 3x3 tensor (s = 2)."""
    if spin == 1:
        return rng.standard_normal(3)
    A = rng.standard_normal((3, 3))
    T = A + A.T
    return T - np.trace(T) / 3.0 * np.eye(3)
