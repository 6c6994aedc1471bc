"""Writes the golden vectors in this directory (run from the repository root:
`python tests/golden/make_golden.py`).

The reference ships no data files for its tests (its tests compare live against ducc0, absent here), so the
vectors are computed from the oracle's exact routines, each pinned independently in tests/test_oracle.py:
FP64 direct Y_lm / sY_lm summation at the target points (no nuFFT), the closed-form pointing and gamma,
the doubling index map, and the parallel-transport lensing of a constant-ambient-tensor field.  Freezing
them stops a later change to the oracle from moving the goalposts silently: tests/test_golden.py checks the
oracle and the library against these files.

Seeds and sizes are fixed here; every file is a numpy .npz without pickled objects.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import almutil, geometry, pipeline, pointing, sht, transport  # noqa: E402


def _deflection(L, rng, amp):
    ll = np.arange(L + 1)
    return almutil.synalm(np.where(ll >= 2, amp / (ll + 1.0) ** 2, 0.0), L, rng)


def _ring_table(rings):
    return dict(r_theta=rings.theta, r_phi0=rings.phi0, r_nph=rings.nph, r_ofs=rings.ofs, r_weight=rings.weight)


def lensing_cases():
    """gclm2lenmap / lenmap2gclm, GL lmax 16 (17 rings x 33), spin 0 and 2, by direct summation at the
    deflected points (rotation e^{+i s gamma} forward, e^{-i s gamma} in the adjoint)."""
    L = 16
    rings = geometry.gl_geometry(L, nphi=2 * L + 1)
    for spin in (0, 2):
        rng = np.random.default_rng(1000 + spin)
        nc = 1 if spin == 0 else 2
        gclm = almutil.random_alm(L, spin, nc, rng)
        dlm = _deflection(L, rng, 3e-3)
        tp, pp, gam, (dt, dp) = pipeline.dlm2pointing(dlm, L, rings)
        if spin == 0:
            lenmap = sht.synth_direct(gclm[0], L, 0, tp, pp)[None, :]
        else:
            v = sht.synth_direct(gclm, L, spin, tp, pp) * np.exp(1j * spin * gam)
            lenmap = np.array([v.real, v.imag])
        x = rng.standard_normal(lenmap.shape)
        if spin == 0:
            adj = sht.adjoint_direct(x[0], L, 0, tp, pp)[None, :]
        else:
            adj = sht.adjoint_direct((x[0] + 1j * x[1]) * np.exp(-1j * spin * gam), L, spin, tp, pp)
        yield f"lensing_spin{spin}_gl16", dict(lmax=L, spin=spin, gclm=gclm, dlm_scaled=dlm, lenmap=lenmap,
                                              adj_in=x, adj_out=adj, ptg_theta=tp, ptg_phi=pp, gamma=gam,
                                              d_theta=dt, d_phi=dp, **_ring_table(rings))


def point_cases():
    """nusht2d2 / nusht2d1 at 300 random points, lmax 20, spin 0 and 2 (direct summation and its adjoint),
    including points on both poles and longitudes outside [0, 2 pi)."""
    L, n = 20, 300
    for spin in (0, 2):
        rng = np.random.default_rng(2000 + spin)
        nc = 1 if spin == 0 else 2
        th = np.arccos(rng.uniform(-1, 1, n))
        ph = rng.uniform(-np.pi, 3 * np.pi, n)
        th[:2] = [0.0, np.pi]
        alm = almutil.random_alm(L, spin, nc, rng)
        vals = sht.synth_direct(alm if spin else alm[0], L, spin, th, ph)
        x = rng.standard_normal(n) + (1j * rng.standard_normal(n) if spin else 0)
        adj = sht.adjoint_direct(x, L, spin, th, ph)
        yield f"points_spin{spin}_l20", dict(lmax=L, spin=spin, alm=alm, loc=np.stack([th, ph], 1), values=vals,
                                            adj_in=x, adj_out=adj if spin else adj[None, :])


def pointing_case():
    """Pointing and gamma on a GL lmax 31 grid, deflections of 1e-3 rad with every 7th pixel x 400
    (pointing.cu:205-223; CPU_nufft_transformer.py:158-160)."""
    rng = np.random.default_rng(3000)
    rings = geometry.gl_geometry(31)
    th, ph = rings.pixel_angles()
    dt = rng.normal(0, 1e-3, th.size)
    dp = rng.normal(0, 1e-3, th.size)
    dt[::7] *= 400
    tp, pp = pointing.deflected_angles(th, ph, dt, dp)
    yield "pointing_gl31", dict(d_theta=dt, d_phi=dp, ptg_theta=tp, ptg_phi=pp, gamma=pointing.gamma(th, dt, dp),
                                **_ring_table(rings))


def doubling_case():
    """Doubling onto the torus and its adjoint (doubling.cu:180-201, :53-72), (ntheta, nphi) = (7, 10), by the
    index map D[r >= nt][p] = (-1)^s M[2 nt - 2 - r][(p + nphi / 2) mod nphi]."""
    rng = np.random.default_rng(4000)
    nt, nph = 7, 10
    M = rng.standard_normal((nt, nph))
    Dm = rng.standard_normal((2 * nt - 2, nph))
    out = {"map": M, "dmap": Dm}
    for spin in (0, 1, 2):
        sg = -1.0 if spin % 2 else 1.0
        D = np.empty((2 * nt - 2, nph))
        A = np.empty((nt, nph))
        for r in range(2 * nt - 2):
            for p in range(nph):
                D[r, p] = M[r, p] if r < nt else sg * M[2 * nt - 2 - r, (p + nph // 2) % nph]
        for r in range(nt):
            for p in range(nph):
                A[r, p] = Dm[r, p] + (sg * Dm[2 * nt - 2 - r, (p + nph // 2) % nph] if 0 < r < nt - 1 else 0.0)
        out[f"doubled_s{spin}"] = D
        out[f"adjoint_s{spin}"] = A
    yield "doubling_7x10", out


def transport_case():
    """Spin-2 field of a constant ambient tensor on GL lmax 12, lensed by parallel transport."""
    rng = np.random.default_rng(5000)
    L = 12
    rings = geometry.gl_geometry(L)
    Tn = transport.random_tensor(2, rng)
    th, ph = rings.pixel_angles()
    w = np.repeat(rings.weight, rings.nph)
    EB = sht.adjoint_direct(transport.ambient_field(2, Tn, th, ph) * w, L, 2, th, ph)
    dlm = _deflection(L, rng, 3e-2)
    _, _, _, (dt, dp) = pipeline.dlm2pointing(dlm, L, rings)
    f = transport.lensed_ambient_field(2, Tn, th, ph, dt, dp)
    yield "transport_spin2_gl12", dict(lmax=L, spin=2, tensor=Tn, gclm=EB, dlm_scaled=dlm,
                                       lenmap=np.array([f.real, f.imag]), **_ring_table(rings))


def main():
    for gen in (lensing_cases, point_cases, pointing_case, doubling_case, transport_case):
        for name, arrays in gen():
            path = os.path.join(HERE, name + ".npz")
            np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrays.items()})
            print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
