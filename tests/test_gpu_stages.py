"""GPU parity of the stage-level entry points: 2D FFT, the plan's SHT / nuFFT stages, gradient synthesis.

Tolerances: the 2D FFT within 1e-13 (f64) / 1e-5 (f32) relative max error of numpy.fft; the Legendre stage
within 1e-12 of the oracle's direct ring coefficients; composed stages equal to the composed entry points
(1e-14 relative: same kernels); adjointness of stage pairs to 1e-12.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import almutil, sht as osht

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _p(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _s():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _relmax(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("shape", [(1, 1), (1, 40), (13, 7), (130, 96), (64, 4096), (2048, 26)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("direction", [-1, 1])
def test_fft2d_vs_numpy(shape, prec, direction):
    from cunusht_amd import _lib
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    batch = 2
    x = rng.standard_normal((batch,) + shape) + 1j * rng.standard_normal((batch,) + shape)
    dt = torch.complex128 if prec == "f64" else torch.complex64
    d = torch.as_tensor(x, device="cuda").to(dt).contiguous()
    _lib.call(f"nusht_fft2d_c2c_{prec}", _p(d), shape[0], shape[1], batch, direction, _s())
    got = d.cpu().numpy()
    ref = np.fft.fft2(x) if direction == -1 else np.fft.ifft2(x) * (shape[0] * shape[1])
    assert _relmax(got, ref) < (1e-13 if prec == "f64" else 1e-5)


def test_fft2d_rejects_unsupported_length():
    """4177 is prime with 4176 = 2^4 3^2 29 and too long for Bluestein in LDS; 8501 exceeds the any-length
    limit; both are refused on the host before any launch."""
    from cunusht_amd import _lib
    d = torch.zeros((2, 8501), dtype=torch.complex128, device="cuda")
    with pytest.raises(_lib.NuShtError, match="not supported"):
        _lib.call("nusht_fft2d_c2c_f64", _p(d), 1, 4177, 1, -1, _s())
    with pytest.raises(_lib.NuShtError, match="exceeds"):
        _lib.call("nusht_fft2d_c2c_f64", _p(d), 2, 8501, 1, -1, _s())


@pytest.mark.parametrize("shape", [(11, 17), (22, 89), (97, 44), (8191, 2), (3, 4124)])
def test_fft2d_any_length_vs_numpy(shape):
    """Sides by Rader (11, 17, 97, 8191), Bluestein (22, 44, 89) and a split transform (4124 = 2 x 2062)."""
    from cunusht_amd import _lib
    rng = np.random.default_rng(shape[0] + 3 * shape[1])
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    d = _t(x)
    _lib.call("nusht_fft2d_c2c_f64", _p(d), shape[0], shape[1], 1, -1, _s())
    assert _relmax(d.cpu().numpy(), np.fft.fft2(x)) < 1e-12
    f = torch.as_tensor(x, device="cuda").to(torch.complex64).contiguous()
    _lib.call("nusht_fft2d_c2c_f32", _p(f), shape[0], shape[1], 1, 1, _s())
    assert _relmax(f.cpu().numpy(), np.fft.ifft2(x) * x.size) < 2e-5


def _plan(L, spin, eps, npts):
    from cunusht_amd.transformer import NuPlan
    nc = 1 if spin == 0 else 2
    return NuPlan(L, spin, nc, eps, npts)


@pytest.mark.parametrize("spin", [0, 2])
def test_plan_stages_compose_to_operators(spin):
    """synth_cc + nufft_interp == nusht2d2, nufft_spread + adjoint_cc == nusht2d1; G == oracle ring coefficients
    on the plan's equiangular rings; both stage pairs are adjoint."""
    from cunusht_amd import _lib
    L, eps, npts = 24, 1e-10, 700
    rng = np.random.default_rng(90 + spin)
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    ain = a[0] if spin == 0 else a
    th = np.arccos(rng.uniform(-1, 1, npts))
    ph = rng.uniform(0, 2 * np.pi, npts)
    pl = _plan(L, spin, eps, npts)
    pl.set_pts(_t(th), _t(ph))
    info = pl.info()
    nt = info["ntheta"]
    nrows = 2 * L + 1 if spin else L + 1
    G = torch.empty((nrows, nt), dtype=torch.complex128, device="cuda")
    ad = _t(ain)
    _lib.call("nusht_sht_synth_cc", pl._h, _p(ad), _p(G), _s())
    out = torch.empty((2, npts), dtype=torch.float64, device="cuda")
    _lib.call("nusht_nufft_interp", pl._h, _p(G), None, _p(out[0]), _p(out[1]) if spin else None, _s())
    ref_out = torch.empty_like(out)
    _lib.call("nusht_nusht2d2", pl._h, _p(ad), _p(ref_out[0]), _p(ref_out[1]) if spin else None, None, _s())
    torch.cuda.synchronize()
    k = 1 if spin == 0 else 2
    assert _relmax(out[:k].cpu().numpy(), ref_out[:k].cpu().numpy()) < 1e-14
    # G against the oracle's ring coefficients (rows m = 0..L for spin 0, m = -L..L for spin s)
    cc = np.pi * np.arange(nt) / (nt - 1)
    Gref = osht.synth_ring_coeffs(ain, L, spin, cc)
    Gref = Gref[L:] if spin == 0 else Gref
    assert _relmax(G.cpu().numpy(), Gref) < 1e-12
    # adjoint direction
    x = rng.standard_normal((k, npts))
    xd = _t(x)
    G2 = torch.empty_like(G)
    _lib.call("nusht_nufft_spread", pl._h, _p(xd[0]), _p(xd[1]) if spin else None, None, _p(G2), _s())
    alm2 = torch.empty((nc, almutil.nalm(L)), dtype=torch.complex128, device="cuda")
    _lib.call("nusht_sht_adjoint_cc", pl._h, _p(G2), _p(alm2), _s())
    alm_ref = torch.empty_like(alm2)
    _lib.call("nusht_nusht2d1", pl._h, _p(xd[0]), _p(xd[1]) if spin else None, None, _p(alm_ref), _s())
    torch.cuda.synchronize()
    assert _relmax(alm2.cpu().numpy(), alm_ref.cpu().numpy()) < 1e-13
    # <interp G, x> == <G, spread x> with the real inner product of the m-expansion: rows m > 0 of a real
    # field count twice (spin 0); <synth a, G'> == <a, adjoint G'>
    Gh, G2h = G.cpu().numpy(), G2.cpu().numpy()
    lhs = float(np.sum(out[:k].cpu().numpy() * x))
    w = np.ones((nrows, 1))
    if spin == 0:
        w[1:] = 2.0
    rhs = float(np.sum(w * np.real(Gh * np.conj(G2h))))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    a2h = alm2.cpu().numpy()
    rhs2 = sum(almutil.myalmdot(a[c], a2h[c], L) for c in range(nc))
    assert abs(rhs - rhs2) < 1e-12 * abs(rhs)


def test_synth_grad_gl_vs_direct():
    """Gradient synthesis of a scalar field on the GL grid = spin-1 synthesis of sqrt(l(l+1)) alm."""
    from cunusht_amd import _lib
    from cunusht_amd.geometry import get_geom
    from cunusht_amd.transformer import RingSHT
    L = 30
    rng = np.random.default_rng(12)
    a = almutil.random_alm(L, 0, 1, rng)[0]
    g = get_geom(("gl", {"lmax": L}))
    rs = RingSHT(g, L, 1)
    out = torch.empty((2, g.npix()), dtype=torch.float64, device="cuda")
    _lib.call("nusht_sht_synth_grad_gl", rs._h, _p(_t(a)), _p(out[0]), _p(out[1]), _s())
    torch.cuda.synchronize()
    tt, pp = g.pixel_angles()
    ll = np.arange(L + 1)
    ref = osht.synth_direct(np.array([almutil.almxfl(a, np.sqrt(ll * (ll + 1.0)), L), np.zeros_like(a)]), L, 1,
                            tt, pp)
    assert _relmax(out.cpu().numpy(), np.array([ref.real, ref.imag])) < 1e-12
    # and it is the gradient: finite differences of the scalar synthesis in theta at a few pixels
    f = lambda t, p: osht.synth_direct(a, L, 0, t, p)
    sel = rng.choice(g.npix(), 20, replace=False)
    h = 1e-5
    dth = (f(tt[sel] + h, pp[sel]) - f(tt[sel] - h, pp[sel])) / (2 * h)
    dph = (f(tt[sel], pp[sel] + h) - f(tt[sel], pp[sel] - h)) / (2 * h) / np.sin(tt[sel])
    assert np.max(np.abs(out[0].cpu().numpy()[sel] - dth)) < 1e-6 * np.max(np.abs(dth))
    assert np.max(np.abs(out[1].cpu().numpy()[sel] - dph)) < 1e-6 * np.max(np.abs(dph))
