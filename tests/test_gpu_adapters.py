"""GPU parity of the lensing adapters: quadrature analysis on the pointing grid, exact synthesis at
arbitrary points (gclm2lenpixs) and the remapping alm -> alm (lensgclm).

Tolerances: exact FP64 paths (analysis, point synthesis) 1e-12 of the oracle's direct summation;
paths through the nuFFT at epsilon 1e-12 within 1e-9.
"""
import numpy as np
import pytest
import torch

from oracle import almutil, geometry as ogeom, pipeline, sht as osht

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _relmax(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _deflection(L, rng, amp=1e-4):
    ll = np.arange(L + 1)
    return almutil.synalm(np.where(ll >= 2, amp / (ll + 1.0) ** 2, 0.0), L, rng)


# ------------------------------------------------------------------ analysis on the pointing grid
@pytest.mark.parametrize("spin", [0, 1, 2])
@pytest.mark.parametrize("geo", [("gl", {"lmax": 12}), ("gl", {"lmax": 12, "nphi": 29}),
                                 ("cc", {"lmax": 12, "ntheta": 28, "nphi": 27}),
                                 ("f1", {"lmax": 12, "ntheta": 27, "nphi": 26})])
def test_analysis_vs_oracle_and_round_trip(spin, geo):
    """adjoint_synthesis (weights applied) == oracle adjoint of weight x map, and inverts synthesis on grids
    that resolve lmax 12 (GL: 13 rings; CC / F1: > 25 rings).  nphi 29 / 26 / 27 exercise the Rader /
    direct ring transforms."""
    import cunusht_amd
    L = 12
    rng = np.random.default_rng(100 + spin)
    T = cunusht_amd.get_transformer("GPU")(geo)
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    ain = a[0] if spin == 0 else a
    m = T.synthesis(_t(ain), lmax=L, mmax=L, spin=spin)
    back = T.adjoint_synthesis(m, lmax=L, mmax=L, spin=spin).cpu().numpy().reshape(nc, -1)
    assert np.max(np.abs(back - a)) < 1e-12
    g = T.geom
    th, ph = g.pixel_angles()
    w = np.repeat(g.weight, g.nph)
    x = rng.standard_normal((nc, th.size)) if spin else rng.standard_normal(th.size)
    got = T.adjoint_synthesis(_t(x), lmax=L, mmax=L, spin=spin).cpu().numpy().reshape(nc, -1)
    xv = x if spin == 0 else x[0] + 1j * x[1]
    ref = np.reshape(osht.adjoint_direct(w * xv, L, spin, th, ph), (nc, -1))
    assert _relmax(got, ref) < 1e-12


# ------------------------------------------------------------------ exact synthesis at points
@pytest.mark.parametrize("spin", [0, 1, 2])
@pytest.mark.parametrize("rotate", [False, True])
def test_point_synthesis_vs_direct(spin, rotate):
    """PointPlan (C ABI nusht_pointplan_*) at random points, the poles, the equator and a mirrored pair."""
    from cunusht_amd.transformer import PointPlan
    L = 40
    rng = np.random.default_rng(7 * spin + rotate)
    th = np.concatenate([rng.uniform(0, np.pi, 57), [0.0, np.pi, 0.5 * np.pi, 0.3, np.pi - 0.3, 1e-7]])
    ph = rng.uniform(0, 2 * np.pi, th.size)
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    ain = a[0] if spin == 0 else a
    gam = rng.uniform(-0.1, 0.1, th.size) if (rotate and spin) else None
    pp = PointPlan(L, spin, th)
    out = pp.synthesis(_t(ain), nc, _t(ph), None if gam is None else _t(gam)).cpu().numpy()
    ref = osht.synth_direct(ain, L, spin, th, ph)
    if spin:
        if gam is not None:
            ref = np.exp(1j * spin * gam) * ref
        ref = np.array([ref.real, ref.imag])
    assert _relmax(out, ref) < 1e-12


def test_point_synthesis_grad_only():
    """spin 2 with one (gradient) component equals two components with B = 0."""
    from cunusht_amd.transformer import PointPlan
    L = 24
    rng = np.random.default_rng(3)
    th = rng.uniform(0, np.pi, 31)
    ph = rng.uniform(0, 2 * np.pi, 31)
    e = almutil.random_alm(L, 2, 1, rng)
    pp = PointPlan(L, 2, th)
    one = pp.synthesis(_t(e), 1, _t(ph)).cpu().numpy()
    two = pp.synthesis(_t(np.concatenate([e, np.zeros_like(e)])), 2, _t(ph)).cpu().numpy()
    assert _relmax(one, two) < 1e-14


# ------------------------------------------------------------------ gclm2lenpixs / lensgclm
@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("polrot", [True, False])
def test_gclm2lenpixs_matches_lenmap(spin, polrot):
    """Exact values at chosen pixels == the nuFFT-based lensed map at those pixels (epsilon 1e-12)."""
    import cunusht_amd
    L = 48
    rng = np.random.default_rng(11 + spin + 2 * polrot)
    dlm = _deflection(L, rng, 1e-3)
    nc = 1 if spin == 0 else 2
    g = almutil.random_alm(L, spin, nc, rng)
    gin = g[0] if spin == 0 else g
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=2, epsilon=1e-12, spin=spin)
    full = T.gclm2lenmap(_t(gin), L, L, dlm_scaled=_t(dlm), spin=spin, polrot=polrot).cpu().numpy()
    pixs = rng.choice(T.npix(), 50, replace=False)
    got = T.gclm2lenpixs(_t(gin), L, spin, pixs, polrot=polrot)
    got = got.cpu().numpy() if spin == 0 else np.array([got[0].cpu().numpy(), got[1].cpu().numpy()])
    want = full[pixs] if spin == 0 else full[:, pixs]
    assert _relmax(got, want) < 1e-9
    # with an explicit pointing the deflected angles come from ptg
    if spin == 0:
        ptg = T.dlm2pointing(_t(dlm))
        again = T.gclm2lenpixs(_t(gin), L, 0, pixs, ptg=ptg).cpu().numpy()
        assert _relmax(again, got) < 1e-14


@pytest.mark.parametrize("spin", [0, 2])
def test_lensgclm_vs_oracle(spin):
    """lensgclm = quadrature analysis of the lensed map; oracle: lensed map by its FP64 pipeline, then the
    weighted direct adjoint."""
    import cunusht_amd
    L = 32
    nphi = 2 * L + 1
    rng = np.random.default_rng(21 + spin)
    dlm = _deflection(L, rng)
    nc = 1 if spin == 0 else 2
    g = almutil.random_alm(L, spin, nc, rng)
    gin = g[0] if spin == 0 else g
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L, "nphi": nphi}), nuFFTtype=2, epsilon=1e-12,
                                           spin=spin)
    got = T.lensgclm(_t(gin), _t(dlm), spin, L).cpu().numpy().reshape(nc, -1)
    rings = ogeom.gl_geometry(L, nphi=nphi)
    th, ph = rings.pixel_angles()
    lm = pipeline.gclm2lenmap(gin, dlm, L, spin, rings, 1e-13)
    w = np.repeat(rings.weight, rings.nph)
    lv = lm if spin == 0 else lm[0] + 1j * lm[1]
    ref = np.reshape(osht.adjoint_direct(w * lv, L, spin, th, ph), (nc, -1))
    assert _relmax(got, ref) < 1e-9
    if spin:
        e_only = T.lensgclm(_t(gin), _t(dlm), spin, L, out_sht_mode="GRAD_ONLY").cpu().numpy().reshape(-1)
        assert _relmax(e_only, got[0]) < 1e-13


def test_lensgclm_zero_deflection_is_identity():
    import cunusht_amd
    L = 40
    rng = np.random.default_rng(5)
    g = almutil.random_alm(L, 2, 2, rng)
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=2, epsilon=1e-12, spin=2)
    zero = np.zeros_like(g[0])
    got = T.lensgclm(_t(g), _t(zero), 2, L).cpu().numpy()
    assert np.max(np.abs(got - g)) < 1e-9


# ------------------------------------------------------------------ ragged ring geometries
def _ring_plan_calls(geom, L, spin, alm, x, weighted):
    """synthesis of alm and (weighted) adjoint of x on `geom` through the C ABI ring plan."""
    import ctypes
    from cunusht_amd import _lib
    from cunusht_amd.transformer import RingSHT
    rs = RingSHT(geom, L, spin)
    nc = alm.shape[0] if alm.ndim == 2 else 1
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: ctypes.c_void_p(0 if t is None else t.data_ptr())
    a = _t(alm)
    out = torch.empty((2, geom.npix()), dtype=torch.float64, device="cuda")
    _lib.call("nusht_ringplan_synthesis", rs._h, p(a), nc, p(out[0]), p(out[1] if spin else None), s)
    xd = _t(x)
    res = torch.empty((nc, almutil.nalm(L)), dtype=torch.complex128, device="cuda")
    _lib.call("nusht_ringplan_analysis" if weighted else "nusht_ringplan_adjoint", rs._h, p(xd[0]),
              p(xd[1] if spin else None), p(res), nc, s)
    torch.cuda.synchronize()
    info = rs.info()
    return (out[0].cpu().numpy() if spin == 0 else out.cpu().numpy()), res.cpu().numpy(), info


@pytest.mark.parametrize("spin", [0, 2])
def test_ragged_rings_every_dft_method(spin):
    """Rings of different lengths, shorter (aliased) and longer than 2 lmax + 1, through every ring DFT:
    direct (1, 2, 3, 64, 96), Rader (31, 97, 8191), Bluestein (22, 89), split (4124 = 2 x 2062 with Bluestein
    sub-transforms, 8189 = 19 x 431).  Synthesis and the unweighted adjoint vs direct summation."""
    from cunusht_amd.geometry import Geom
    L = 40
    rng = np.random.default_rng(30 + spin)
    nph = np.array([96, 1, 64, 31, 4124, 22, 2, 97, 89, 3, 8191, 8189], dtype=np.int64)
    th = np.sort(rng.uniform(0.05, np.pi - 0.05, nph.size))
    th[1] = np.pi - th[0]  # one mirrored pair shares a recursion
    phi0 = rng.uniform(0, 0.3, nph.size)
    ofs = np.concatenate([[0], np.cumsum(nph[:-1])])
    g = Geom(th, phi0, nph, ofs, np.ones(nph.size))
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    x = rng.standard_normal((2, g.npix()))
    got, ad, info = _ring_plan_calls(g, L, spin, a, x, weighted=False)
    assert info["ring_lengths"] == nph.size and info["ring_dft"] == 1  # the longest ring (8191) is Rader
    tt, pp = g.pixel_angles()
    ain = a[0] if spin == 0 else a
    ref = osht.synth_direct(ain, L, spin, tt, pp)
    if spin:
        ref = np.array([ref.real, ref.imag])
    assert _relmax(got, ref) < 1e-12
    xv = x[0] if spin == 0 else x[0] + 1j * x[1]
    refad = np.reshape(osht.adjoint_direct(xv, L, spin, tt, pp), (nc, -1))
    assert _relmax(ad, refad) < 1e-12


def test_unsupported_ring_length_is_rejected():
    """4177 is prime with 4176 = 2^4 3^2 29: no in-LDS DFT; the plan refuses it on the host."""
    from cunusht_amd._lib import NuShtError
    from cunusht_amd.geometry import Geom
    from cunusht_amd.transformer import RingSHT
    g = Geom(np.array([1.0]), np.zeros(1), np.array([4177]), np.zeros(1, dtype=np.int64), np.ones(1))
    with pytest.raises(NuShtError, match="not supported"):
        RingSHT(g, 8, 0)


@pytest.mark.parametrize("spin", [0, 1, 2])
def test_healpix_synthesis_and_weighted_adjoint(spin):
    """HEALPix nside 16 (polar rings of 4 .. 60 pixels, aliased for lmax 40; phi0 offsets) vs direct summation."""
    import cunusht_amd
    L, nside = 40, 16
    rng = np.random.default_rng(50 + spin)
    T = cunusht_amd.get_transformer("GPU")(("healpix", {"nside": nside, "lmax": L}))
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    ain = a[0] if spin == 0 else a
    got = T.synthesis(_t(ain), lmax=L, mmax=L, spin=spin).cpu().numpy()
    g = T.geom
    tt, pp = g.pixel_angles()
    ref = osht.synth_direct(ain, L, spin, tt, pp)
    if spin:
        ref = np.array([ref.real, ref.imag])
    assert _relmax(got, ref) < 1e-12
    x = rng.standard_normal((nc, g.npix())) if spin else rng.standard_normal(g.npix())
    ad = T.adjoint_synthesis(_t(x), lmax=L, mmax=L, spin=spin).cpu().numpy().reshape(nc, -1)
    xv = x if spin == 0 else x[0] + 1j * x[1]
    w = np.repeat(g.weight, g.nph)
    refad = np.reshape(osht.adjoint_direct(w * xv, L, spin, tt, pp), (nc, -1))
    assert _relmax(ad, refad) < 1e-12


@pytest.mark.parametrize("spin", [0, 2])
def test_thinned_gl_analysis_inverts_synthesis(spin):
    """Thinned GL at lmax 400: polar rings carry fewer than 2 lmax + 1 pixels; the aliased m are below the
    polar-optimisation threshold, so quadrature analysis still returns the alm."""
    import cunusht_amd
    L = 400
    rng = np.random.default_rng(60 + spin)
    T = cunusht_amd.get_transformer("GPU")(("tgl", {"lmax": L, "smax": 2}))
    assert T.geom.nph.min() < 2 * L + 1
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    ain = a[0] if spin == 0 else a
    m = T.synthesis(_t(ain), lmax=L, mmax=L, spin=spin)
    back = T.adjoint_synthesis(m, lmax=L, mmax=L, spin=spin).cpu().numpy().reshape(nc, -1)
    assert np.max(np.abs(back - a)) < 1e-10


def test_healpix_pointing_and_lensing():
    """dlm2pointing on HEALPix rings vs the oracle, and gclm2lenmap on that pointing vs the oracle pipeline."""
    import cunusht_amd
    from oracle import pointing as opointing
    L, nside = 32, 16
    rng = np.random.default_rng(70)
    dlm = _deflection(L, rng, 1e-3)
    T = cunusht_amd.get_transformer("GPU")(("healpix", {"nside": nside, "lmax": L}), nuFFTtype=2, epsilon=1e-12)
    pg = T.dlm2pointing(_t(dlm), gamma=True).cpu().numpy()
    g = T.geom
    tt, pp = g.pixel_angles()
    ll = np.arange(L + 1)
    d = osht.synth_direct(np.array([almutil.almxfl(dlm, np.sqrt(ll * (ll + 1.0)), L), np.zeros_like(dlm)]), L, 1,
                          tt, pp)
    tp, ph = opointing.deflected_angles(tt, pp, d.real, d.imag)
    assert np.max(np.abs(pg[0] - tp)) < 1e-12
    dphi = (pg[1] - ph + np.pi) % (2 * np.pi) - np.pi
    assert np.max(np.abs(dphi)) < 1e-12
    gl = almutil.random_alm(L, 0, 1, rng)[0]
    got = T.gclm2lenmap(_t(gl), L, L, dlm_scaled=_t(dlm)).cpu().numpy()
    ref = osht.synth_direct(gl, L, 0, tp, ph)
    assert _relmax(got, ref) < 1e-9
