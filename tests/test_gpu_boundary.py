"""GPU tests of the Python operator surface's caches and edge inputs, through the C ABI.

* Point and pointing caches compare contents, as the reference's full equality compares do
  (cunusht/deflection/GPU_nufft_transformer.py:417-420,470-473,550-556): host `loc` arrays whose device
  temporaries land on recycled memory, and a dlm_scaled buffer refilled in place through a raw pointer,
  must never reuse the previous points.  No test here keeps references to its temporaries.
* Points outside the formed fine-grid rows or not finite, let through by the unchecked set_pts, give NaN in
  nusht2d2 and are left out of nusht2d1.
* The pointing kernel skips pixels that a ring table places outside the pixel arrays.
* deflectionlib (GPU_nufft_transformer.py:164; first_steps/quickstart.ipynb:150, nusht2d1.ipynb:226,385,520),
  set_pts / gclm2lenmap(ptg=...) and get_grid_uniform on the GPU class.
Tolerances: FP64 paths at epsilon 1e-12 within 1e-9 relative max error of direct Y_lm summation.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import almutil, geometry as ogeom, pipeline, pointing as opointing, sht as osht

pytestmark = pytest.mark.gpu


def _relmax(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _deflection(L, rng, amp=1e-3):
    ll = np.arange(L + 1)
    return almutil.synalm(np.where(ll >= 2, amp / (ll + 1.0) ** 2, 0.0), L, rng)


def _loc(rng, n):
    return np.stack([np.arccos(rng.uniform(-1, 1, n)), rng.uniform(0, 2 * np.pi, n)], axis=1)


def _T(L, eps, spin=0):
    import cunusht_amd
    return cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=2, epsilon=eps, spin=spin)


@pytest.mark.filterwarnings("ignore::UserWarning")
def test_nusht_fresh_host_loc_every_call():
    """Two nusht2d2 calls with different host loc arrays of one size, then nusht2d1 with a third: each call
    must use its own points (the device copies of host arrays are temporaries on recycled memory)."""
    L, n = 24, 777
    rng = np.random.default_rng(301)
    T = _T(L, 1e-12)
    alm = almutil.random_alm(L, 0, 1, rng)[0]
    for k in range(2):
        loc = _loc(rng, n)
        print(f"case nusht2d2 host loc call {k}", flush=True)
        got = T.nusht2d2(L, L, alm, loc).cpu().numpy()
        ref = osht.synth_direct(alm, L, 0, loc[:, 0], loc[:, 1])
        assert _relmax(got, ref) < 1e-9, k
    loc = _loc(rng, n)
    x = rng.standard_normal(n)
    print("case nusht2d1 host loc", flush=True)
    got = T.nusht2d1(L, L, x, loc).cpu().numpy()
    ref = osht.adjoint_direct(x, L, 0, loc[:, 0], loc[:, 1])
    assert _relmax(got, ref) < 1e-9
    # the same values again, in a new host array: still correct
    got2 = T.nusht2d1(L, L, x, loc.copy()).cpu().numpy()
    assert _relmax(got2, ref) < 1e-9


@pytest.mark.filterwarnings("ignore::UserWarning")
def test_dlm_buffer_refilled_in_place():
    """A dlm_scaled buffer rewritten through the library's own out= path (a raw-pointer write that leaves
    torch's version counter alone) gives the pointing of its new contents on the next call."""
    L, spin = 20, 2
    rng = np.random.default_rng(302)
    T = _T(L, 1e-12, spin=spin)
    rings = ogeom.gl_geometry(L)
    gclm = almutil.random_alm(L, spin, 2, rng)
    g = torch.as_tensor(gclm, device="cuda")
    buf = torch.zeros(almutil.nalm(L), dtype=torch.complex128, device="cuda")
    print("case zero deflection", flush=True)
    m0 = T.gclm2lenmap(g, L, L, dlm_scaled=buf, spin=spin).cpu().numpy()
    ref0 = pipeline.gclm2lenmap(gclm, np.zeros(almutil.nalm(L), complex), L, spin, rings, 1e-12)
    assert _relmax(m0, ref0) < 1e-9
    dlm = _deflection(L, rng)
    dmap = T.synthesis(torch.as_tensor(dlm, device="cuda"))
    v0 = buf._version
    T.adjoint_synthesis(dmap, out=buf)  # quadrature analysis: exact on the GL grid, so buf = dlm
    torch.cuda.synchronize()
    assert buf._version == v0, "the library writes through a raw pointer"
    assert _relmax(buf.cpu().numpy(), dlm) < 1e-12
    print("case refilled deflection", flush=True)
    m1 = T.gclm2lenmap(g, L, L, dlm_scaled=buf, spin=spin).cpu().numpy()
    ref1 = pipeline.gclm2lenmap(gclm, dlm, L, spin, rings, 1e-12)
    assert _relmax(m1, ref1) < 1e-9
    assert _relmax(m1, ref0) > 1e-6  # the deflection moved the map


@pytest.mark.parametrize("spin", [0, 2])
def test_bad_points_unchecked(spin):
    """set_pts_async with a NaN colatitude, an infinite longitude and colatitudes outside the formed rows:
    nusht2d2 gives NaN there and the right values elsewhere; nusht2d1 stays finite and equals the adjoint
    of the good points alone."""
    from cunusht_amd import _lib
    from cunusht_amd.transformer import NuPlan
    L, n, eps = 16, 400, 1e-12
    rng = np.random.default_rng(303 + spin)
    loc = _loc(rng, n)
    bad = np.array([3, 50, 51, 200, 399])
    loc[3, 0] = np.nan
    loc[50, 1] = np.inf
    loc[51, 1] = np.nan
    loc[200, 0] = 6.0
    loc[399, 0] = -2.0
    good = np.setdiff1d(np.arange(n), bad)
    nc = 1 if spin == 0 else 2
    plan = NuPlan(L, spin, nc, eps, n)
    th = torch.as_tensor(np.ascontiguousarray(loc[:, 0]), device="cuda")
    ph = torch.as_tensor(np.ascontiguousarray(loc[:, 1]), device="cuda")
    print(f"case bad points spin={spin}", flush=True)
    plan.set_pts(th, ph, checked=False)
    alm = almutil.random_alm(L, spin, nc, rng)
    a = torch.as_tensor(alm if spin else alm[0], device="cuda")
    out = torch.zeros((2, n), dtype=torch.float64, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: ctypes.c_void_p(0 if t is None else t.data_ptr())
    _lib.call("nusht_nusht2d2", plan._h, p(a), p(out[0]), p(out[1]) if spin else None, None, s)
    x = rng.standard_normal((2, n))
    xt = torch.as_tensor(x, device="cuda")
    res = torch.zeros((nc, almutil.nalm(L)), dtype=torch.complex128, device="cuda")
    _lib.call("nusht_nusht2d1", plan._h, p(xt[0]), p(xt[1]) if spin else None, None, p(res), s)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    rows = [0] if spin == 0 else [0, 1]
    for r in rows:
        assert np.all(np.isnan(o[r, bad]))
        assert np.all(np.isfinite(o[r, good]))
    ref = osht.synth_direct(alm if spin else alm[0], L, spin, loc[good, 0], loc[good, 1])
    got = o[0, good] if spin == 0 else o[0, good] + 1j * o[1, good]
    assert _relmax(got, ref) < 1e-9
    r = res.cpu().numpy()
    assert np.all(np.isfinite(r))
    vals = x[0, good] if spin == 0 else x[0, good] + 1j * x[1, good]
    ref1 = osht.adjoint_direct(vals, L, spin, loc[good, 0], loc[good, 1])
    assert _relmax(r[0] if spin == 0 else r, ref1) < 1e-9
    # the checked entry point refuses the same points
    from cunusht_amd._lib import NuShtError
    with pytest.raises(NuShtError):
        plan.set_pts(th, ph, checked=True)


def test_pointing_ring_table_outside_pixels():
    """A ring table whose last ring starts 5 pixels before the end of the arrays: that ring's first 5 pixels
    are written, nothing past the arrays, and the pixels it no longer covers are left as they were."""
    from cunusht_amd import _lib
    rng = np.random.default_rng(304)
    L = 15
    geom = ogeom.gl_geometry(L)
    th, ph = geom.pixel_angles()
    npix = th.size
    dt = rng.normal(0, 1e-3, npix)
    dp = rng.normal(0, 1e-3, npix)
    ofs = geom.ofs.astype(np.int64).copy()
    last0 = int(ofs[-1])
    ofs[-1] = npix - 5
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    ring = [dev(geom.theta), dev(geom.phi0), dev(geom.nph.astype(np.int64)), dev(ofs)]
    d = [dev(dt), dev(dp)]
    o = [torch.full((npix,), 7.0, dtype=torch.float64, device="cuda") for _ in range(2)]
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    print("case pointing ring table past the arrays", flush=True)
    _lib.call("nusht_pointing_f64", p(ring[0]), p(ring[1]), p(ring[2]), p(ring[3]), geom.theta.size,
              int(geom.nph.max()), npix, p(d[0]), p(d[1]), p(o[0]), p(o[1]), None, s)
    torch.cuda.synchronize()
    pt, pp = o[0].cpu().numpy(), o[1].cpu().numpy()
    tp, ppr = opointing.deflected_angles(th, ph, dt, dp)
    assert np.max(np.abs(pt[:last0] - tp[:last0])) < 1e-14
    assert np.all(pt[last0:npix - 5] == 7.0) and np.all(pp[last0:npix - 5] == 7.0)
    # the last ring's pixels k = 0..4, at the shifted offset: its angles with the deflection found there
    k = np.arange(5)
    th_k = np.full(5, geom.theta[-1])
    ph_k = geom.phi0[-1] + k * 2 * np.pi / geom.nph[-1]
    tk, pk = opointing.deflected_angles(th_k, ph_k, dt[npix - 5:], dp[npix - 5:])
    assert np.max(np.abs(pt[npix - 5:] - tk)) < 1e-14
    assert np.max(np.abs((pp[npix - 5:] - pk + np.pi) % (2 * np.pi) - np.pi)) < 1e-12


@pytest.mark.filterwarnings("ignore::UserWarning")
def test_deflectionlib_surface():
    """deflectionlib as the reference's tutorials use it: geom.nalm / geom.npix / geom.theta / geom.nph,
    adjoint_synthesis of a map (quadrature analysis: the inverse of synthesis on the GL grid),
    dlm2pointing into given arrays, and the gradient synthesis."""
    L = 20
    rng = np.random.default_rng(305)
    T = _T(L, 1e-10)
    D = T.deflectionlib
    rings = ogeom.gl_geometry(L)
    assert D.geom.nalm(L, L) == almutil.nalm(L)
    assert D.geom.npix() == T.npix() == rings.npix()
    assert np.allclose(D.geom.theta, rings.theta, rtol=0, atol=1e-13)
    assert int(D.geom.nph[0]) == int(rings.nph[0])
    alm = almutil.random_alm(L, 0, 1, rng)[0]
    m = D.synthesis(torch.as_tensor(alm, device="cuda"))
    th, ph = D.geom.pixel_angles()
    assert _relmax(m.cpu().numpy(), osht.synth_direct(alm, L, 0, th, ph)) < 1e-10
    back = D.adjoint_synthesis(m.cpu().numpy())
    assert tuple(back.shape) == (1, almutil.nalm(L))
    assert _relmax(back[0].cpu().numpy(), alm) < 1e-10
    dlm = _deflection(L, rng)
    pt = torch.empty(T.npix(), dtype=torch.float64, device="cuda")
    pp = torch.empty(T.npix(), dtype=torch.float64, device="cuda")
    ptg = D.dlm2pointing(torch.as_tensor(dlm, device="cuda"), L, 0, pt, pp)
    tp, ppr, _, (dt, dp) = pipeline.dlm2pointing(dlm, L, rings)
    assert np.max(np.abs(ptg[0].cpu().numpy() - tp)) < 1e-12
    assert torch.equal(pt, ptg[0]) and torch.equal(pp, ptg[1])
    # the deflection of the pointing is the gradient of sum dlm_scaled Y_lm
    grad = D.synthesis_der1_cupy(torch.as_tensor(dlm, device="cuda"))
    assert _relmax(grad[0].cpu().numpy(), dt) < 1e-10
    assert _relmax(grad[1].cpu().numpy(), dp) < 1e-10


@pytest.mark.filterwarnings("ignore::UserWarning")
def test_set_pts_ptg_and_grid_uniform():
    """get_grid_uniform is the pointing grid; gclm2lenmap(ptg=...) and set_pts(ptg) + nusht2d2 at given
    angles agree with direct summation; a second ptg with new values in a fresh array is used."""
    L = 18
    rng = np.random.default_rng(306)
    T = _T(L, 1e-12)
    grid = T.get_grid_uniform().cpu().numpy()
    th, ph = ogeom.gl_geometry(L).pixel_angles()
    assert grid.shape == (2, th.size)
    assert np.max(np.abs(grid[0] - th)) < 1e-14 and np.max(np.abs(grid[1] - ph)) < 1e-14
    alm = almutil.random_alm(L, 0, 1, rng)[0]
    g = torch.as_tensor(alm, device="cuda")
    for k in range(2):
        ptg = grid + rng.normal(0, 2e-3, grid.shape)
        ptg[0] = np.clip(ptg[0], 0, np.pi)
        print(f"case gclm2lenmap ptg call {k}", flush=True)
        got = T.gclm2lenmap(g, L, L, ptg=ptg).cpu().numpy()
        assert _relmax(got, osht.synth_direct(alm, L, 0, ptg[0], ptg[1])) < 1e-9, k
    T.set_pts(ptg)
    assert T.ptg.shape == (2, th.size)
