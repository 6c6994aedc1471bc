"""GPU parity: the HIP library (through its C ABI) against the oracle on seeded inputs.

Tolerances: index/byte work (doubling) bit-exact; FP64 paths (epsilon <= 1e-10) within 1e-9
relative max error of the oracle's FP64 restatement or direct Y_lm summation; epsilon-controlled
paths within a small multiple of epsilon (relative L2).
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import almutil, clib, geometry as ogeom, pipeline, pointing as opointing, sht as osht, transport

pytestmark = pytest.mark.gpu


_KEEP = []


def _t(a):
    """Device copy of a host array, kept alive until _sync() (a temporary's memory would otherwise be
    recycled by the caching allocator before the kernel reads it)."""
    t = torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    _KEEP.append(t)
    return t


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _sync():
    torch.cuda.synchronize()
    _KEEP.clear()


def _lib():
    from cunusht_amd import _lib
    return _lib


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _relmax(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _rel2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


# ------------------------------------------------------------------ doubling (bit-exact)
@pytest.mark.parametrize("spin", [0, 1, 2])
@pytest.mark.parametrize("ntheta,nphi", [(6, 8), (34, 64), (67, 130)])
def test_doubling_exact(spin, ntheta, nphi):
    rng = np.random.default_rng(ntheta + spin)
    m = rng.standard_normal((nphi, ntheta))
    out = torch.empty((nphi, 2 * ntheta - 2), dtype=torch.float64, device="cuda")
    _lib().call("nusht_doubling_f64", _p(_t(m)), ntheta, nphi, spin, _p(out), _stream())
    _sync()
    # numpy restatement of doubling.cu:189-197 with the (-1)^s sign
    ref = np.empty((nphi, 2 * ntheta - 2))
    ref[:, :ntheta] = m
    for r in range(ntheta, 2 * ntheta - 2):
        ref[:, r] = np.roll(m, -nphi // 2, axis=0)[:, 2 * ntheta - 2 - r] * (-1) ** spin
    assert np.array_equal(out.cpu().numpy(), ref)
    mc = rng.standard_normal((nphi, ntheta)) + 1j * rng.standard_normal((nphi, ntheta))
    outc = torch.empty((nphi, 2 * ntheta - 2), dtype=torch.complex128, device="cuda")
    _lib().call("nusht_doubling_c128", _p(_t(mc)), ntheta, nphi, spin, _p(outc), _stream())
    _sync()
    refc = np.empty((nphi, 2 * ntheta - 2), dtype=complex)
    refc[:, :ntheta] = mc
    for r in range(ntheta, 2 * ntheta - 2):
        refc[:, r] = np.roll(mc, -nphi // 2, axis=0)[:, 2 * ntheta - 2 - r] * (-1) ** spin
    assert np.array_equal(outc.cpu().numpy(), refc)


@pytest.mark.parametrize("spin", [0, 1, 2])
@pytest.mark.parametrize("ntheta,nphi", [(6, 8), (34, 64), (67, 130)])
def test_adjoint_doubling_exact(spin, ntheta, nphi):
    rng = np.random.default_rng(100 + ntheta + spin)
    d = rng.standard_normal((2 * ntheta - 2, nphi)).astype(np.float32)
    out = torch.empty((ntheta, nphi), dtype=torch.float32, device="cuda")
    _lib().call("nusht_adjoint_doubling_f32", _p(_t(d)), ntheta, nphi, spin, _p(out), _stream())
    _sync()
    # restatement of doubling.cu:61-69 / CPU_nufft_transformer.py:303-308
    ref = d[:ntheta].copy()
    h = nphi // 2
    sg = np.float32(-1.0 if spin % 2 else 1.0)
    ref[1:ntheta - 1, :h] += sg * d[-1:ntheta - 1:-1, h:]
    ref[1:ntheta - 1, h:] += sg * d[-1:ntheta - 1:-1, :h]
    assert np.array_equal(out.cpu().numpy(), ref)


def test_doubling_pair_is_adjoint():
    rng = np.random.default_rng(5)
    nt, nph = 19, 36
    m = rng.standard_normal((nt, nph))       # ring-major map
    d = rng.standard_normal((2 * nt - 2, nph))
    # ring-major doubling == transpose of the theta-contiguous kernel
    dm = torch.empty((nph, 2 * nt - 2), dtype=torch.float64, device="cuda")
    _lib().call("nusht_doubling_f64", _p(_t(m.T)), nt, nph, 2, _p(dm), _stream())
    am = torch.empty((nt, nph), dtype=torch.float64, device="cuda")
    _lib().call("nusht_adjoint_doubling_f64", _p(_t(d)), nt, nph, 2, _p(am), _stream())
    _sync()
    lhs = np.sum(d * dm.cpu().numpy().T)
    rhs = np.sum(m * am.cpu().numpy())
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


# ------------------------------------------------------------------ pointing
def test_pointing_kernel():
    rng = np.random.default_rng(9)
    L = 31
    geom = ogeom.gl_geometry(L)
    th, ph = geom.pixel_angles()
    dt = rng.normal(0, 1e-3, th.size)
    dp = rng.normal(0, 1e-3, th.size)
    dt[::7] *= 400  # some large deflections take the exact sin(a)/a branch
    o = [torch.empty(th.size, dtype=torch.float64, device="cuda") for _ in range(3)]
    _lib().call("nusht_pointing_f64", _p(_t(geom.theta)), _p(_t(geom.phi0)), _p(_t(geom.nph)), _p(_t(geom.ofs)),
                geom.theta.size, int(geom.nph.max()), th.size, _p(_t(dt)), _p(_t(dp)), _p(o[0]), _p(o[1]), _p(o[2]),
                _stream())
    _sync()
    tp, pp = opointing.deflected_angles(th, ph, dt, dp)
    g = opointing.gamma(th, dt, dp)
    # test/test_pointingGPU.py:121-122 tolerances (std of differences)
    assert np.std(o[0].cpu().numpy() - tp) <= 1e-15
    dphi = (o[1].cpu().numpy() - pp + np.pi) % (2 * np.pi) - np.pi
    assert np.std(dphi) <= 1e-12
    assert np.max(np.abs(o[2].cpu().numpy() - g)) < 1e-12


# ------------------------------------------------------------------ ring SHT on the pointing grid
@pytest.mark.parametrize("L,nphi", [(16, 33), (12, 25), (20, 48), (11, 23), (24, 49), (8, 17)])
@pytest.mark.parametrize("spin", [0, 1, 2])
def test_ring_synthesis_and_adjoint(L, nphi, spin):
    """nphi covers the direct (25, 48, 49), Rader (17) and Bluestein (23, 33) ring DFTs."""
    import cunusht_amd
    rng = np.random.default_rng(L * 10 + spin)
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L, "nphi": nphi}))
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    ain = a[0] if spin == 0 else a
    out = T.synthesis(_t(ain), lmax=L, mmax=L, spin=spin)
    torch.cuda.synchronize()
    rings = ogeom.gl_geometry(L, nphi=nphi)
    th, ph = rings.pixel_angles()
    ref = osht.synth_direct(ain, L, spin, th, ph)
    got = out.cpu().numpy()
    if spin:
        ref = np.array([ref.real, ref.imag])
    assert _relmax(got, ref) < 1e-12
    x = rng.standard_normal(got.shape)
    ad = T.adjoint_synthesis(_t(x), lmax=L, mmax=L, spin=spin, ncomp=nc, apply_weights=False)
    ad = ad.cpu().numpy().reshape(nc, -1)
    lhs = float(np.sum(x * got))
    rhs = sum(almutil.myalmdot(a[c], ad[c], L) for c in range(nc))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


@pytest.mark.parametrize("L,nphi", [(40, 81), (33, 67), (31, 64)])
def test_dlm2pointing_vs_oracle(L, nphi):
    import cunusht_amd
    rng = np.random.default_rng(L)
    ll = np.arange(L + 1)
    dlm = almutil.synalm(np.where(ll >= 2, 1e-5 / (ll + 1.0) ** 2, 0.0), L, rng)
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L, "nphi": nphi}))
    got = T.dlm2pointing(_t(dlm), gamma=True).cpu().numpy()
    rings = ogeom.gl_geometry(L, nphi=nphi)
    tp, pp, g, _ = pipeline.dlm2pointing(dlm, L, rings)
    assert np.max(np.abs(got[0] - tp)) < 1e-14
    assert np.max(np.abs((got[1] - pp + np.pi) % (2 * np.pi) - np.pi)) < 1e-13
    assert np.max(np.abs(got[2] - g)) < 1e-10


@pytest.mark.parametrize("lmax", [255, 511, 1023])
def test_pointing_reference_configuration(lmax):
    """test/test_pointingGPU.py:83-122 at its sizes: alm of amplitude 1e-6 (U(-1, 1) re / im), the GPU
    gradient synthesis of alm sqrt(1/(l(l+1))) plus pointing against the CPU spin-1 synthesis of alm plus
    deflected angles (the oracle); std(dtheta) <= 1e-15 and std(dphi) <= 1e-12 with dphi > pi masked."""
    import cunusht_amd
    rng = np.random.default_rng(lmax)
    alm = 1e-6 * almutil.random_alm(lmax, 0, 1, rng)[0]
    ll = np.arange(lmax + 1, dtype=np.float64)
    with np.errstate(divide="ignore"):
        fl = np.where(ll > 0, 1.0 / np.sqrt(ll * (ll + 1.0)), 0.0)
    dlm_scaled = almutil.almxfl(alm, fl, lmax)
    print(f"case pointing lmax={lmax}", flush=True)
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": lmax}))
    got = T.dlm2pointing(_t(dlm_scaled)).cpu().numpy()
    rings = ogeom.gl_geometry(lmax, nphi=2 * lmax + 2)
    assert rings.npix() == got.shape[1]
    tp, pp, _, _ = pipeline.dlm2pointing(dlm_scaled, lmax, rings, with_gamma=False)
    dth = got[0] - tp
    dph = got[1] - pp
    dph = dph[np.abs(dph) <= np.pi]
    assert np.std(dth) <= 1e-15
    assert np.std(dph) <= 1e-12


@pytest.mark.parametrize("spin", [0, 1, 2])
@pytest.mark.parametrize("mcut", [0, 40, 120])
def test_ring_sht_polar_underflow(spin, mcut):
    _polar_underflow_case(spin, mcut)


@pytest.mark.parametrize("spin", [0, 2])
def test_ring_sht_polar_underflow_every_variant(spin, monkeypatch):
    """The same check with each Legendre synthesis kernel variant forced (plan creation otherwise keeps
    only the fastest one)."""
    nv = _lib().load().nusht_legendre_variants()
    assert nv >= 1
    for v in range(nv):
        monkeypatch.setenv("NUSHT_LEGENDRE_VARIANT", str(v))
        _polar_underflow_case(spin, 40, variant=v)


def _polar_underflow_case(spin, mcut, variant=None):
    """Rings 1e-4 .. 0.05 rad from both poles at lmax 320, alm zero for m < mcut: at such rings the
    Wigner-d values of large m start far below 2^-300 (scaled recursion, l-ranges that are skipped
    because a whole group of rings is still under range, partly masked groups).  Each ring's
    values and each m's alm are compared with the oracle relative to their own maximum; values
    below 1e-85 of the global maximum are exempt (the library drops terms below 2^-300 = 4.9e-91
    of the O(1) recursion values).  A southern ring theta = fl(pi - eps) fixes eps only to ulp(pi),
    and values ~ eps^m change by m ulp(theta) / sin(theta) relative to it; the library uses the
    exact mirror of the northern ring, so each ring's tolerance adds L ulp(theta) / sin(theta)."""
    from cunusht_amd.geometry import Geom
    from cunusht_amd.transformer import RingSHT
    L, nphi = 320, 648
    rng = np.random.default_rng(1000 + 10 * spin + mcut)
    polar = np.geomspace(1e-4, 0.05, 150)
    theta = np.concatenate([polar, np.pi - polar[::2], np.linspace(0.3, 2.8, 37)])
    nring = theta.size
    phi0 = rng.uniform(0, 2 * np.pi, nring)
    geom = Geom(theta, phi0, np.full(nring, nphi), np.arange(nring) * nphi, np.ones(nring))
    r = RingSHT(geom, L, spin)
    if variant is not None:
        assert r.info()["legendre_variant"] == variant
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    mofa = np.concatenate([np.full(L + 1 - m, m) for m in range(L + 1)])  # m of each alm index
    a[:, mofa < mcut] = 0
    npix = nring * nphi
    out = torch.zeros((2, npix), dtype=torch.float64, device="cuda")
    lib = _lib()
    ad = _t(a)
    lib.call("nusht_ringplan_synthesis", r._h, _p(ad), nc, _p(out[0]), _p(out[1]) if spin else None, _stream())
    torch.cuda.synchronize()
    G = clib.sht_synth_rings(a, L, spin, theta)
    ref = pipeline.phi_synth(G, L, spin, nphi, phi0)
    got = out.cpu().numpy()
    got = got[0].reshape(nring, nphi) if spin == 0 else (got[0] + 1j * got[1]).reshape(nring, nphi)
    floor = 1e-85 * np.max(np.abs(ref))
    rmax = np.max(np.abs(ref), axis=1)
    err = np.max(np.abs(got - ref), axis=1)
    cond = L * np.spacing(theta) / np.sin(theta)
    bad = np.nonzero(err > (1e-12 + cond) * rmax + floor)[0]
    assert bad.size == 0, [(int(i), theta[i], rmax[i], err[i]) for i in bad[:8]]
    # adjoint with values on the polar rings only (southern ones with eps >= 1e-2, where the
    # conditioning term stays below 1.5e-11)
    x = np.zeros((2, nring, nphi))
    sel = (np.arange(nring) < 225) & ((theta < 1.0) | (np.pi - theta >= 1e-2))
    x[:, sel] = rng.standard_normal((2, int(sel.sum()), nphi))
    if spin == 0:
        x[1] = 0
    xd = _t(x.reshape(2, -1))
    res = torch.zeros((nc, (L + 1) * (L + 2) // 2), dtype=torch.complex128, device="cuda")
    lib.call("nusht_ringplan_adjoint", r._h, _p(xd[0]), _p(xd[1]) if spin else None, _p(res), nc, _stream())
    _sync()
    xm = x[0] if spin == 0 else x[0] + 1j * x[1]
    Gt = pipeline.phi_adjoint(xm, L, spin, phi0)
    aref = clib.sht_adjoint_rings(Gt, L, spin, nc, theta)
    agot = res.cpu().numpy()
    floor = 1e-85 * np.max(np.abs(aref))
    for m in range(L + 1):
        sl = mofa == m
        am = np.max(np.abs(aref[:, sl]))
        e = np.max(np.abs(agot[:, sl] - aref[:, sl]))
        assert e <= 2e-11 * am + floor, (m, am, e)


# ------------------------------------------------------------------ nusht2d2 / nusht2d1
@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("eps", [1e-12, 1e-7, 1e-5])
def test_nusht2d2_vs_direct(spin, eps):
    import cunusht_amd
    rng = np.random.default_rng(int(-np.log10(eps)) + spin)
    L = 30
    npts = 3000
    theta = np.arccos(rng.uniform(-1, 1, npts))
    phi = rng.uniform(0, 2 * np.pi, npts)
    loc = np.stack([theta, phi], 1)
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    ain = a[0] if spin == 0 else a
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=2, epsilon=eps, spin=spin)
    got = T.nusht2d2(L, L, _t(ain), _t(loc), spin=spin).cpu().numpy()
    ref = osht.synth_direct(ain, L, spin, theta, phi)
    if spin:
        ref = np.array([ref.real, ref.imag])
    err = _rel2(got, ref)
    assert err < 10 * max(eps, 1e-13), err


@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("eps", [1e-10, 1e-7])
def test_nusht2d2_dense_tiles(spin, eps):
    """Tiles above the interpolation's per-tile ordering cap (1024 points, kept in binned order) beside ordinary
    tiles: 5000 points in a 0.05 rad patch (less than one fine cell at lmax 30) plus 1500 uniform ones, against
    direct summation."""
    import cunusht_amd
    rng = np.random.default_rng(97 + spin)
    L = 30
    nd, nu = 5000, 1500
    theta = np.concatenate([1.0 + 0.05 * rng.uniform(0, 1, nd), np.arccos(rng.uniform(-1, 1, nu))])
    phi = np.concatenate([2.0 + 0.05 * rng.uniform(0, 1, nd), rng.uniform(0, 2 * np.pi, nu)])
    perm = rng.permutation(nd + nu)
    theta, phi = theta[perm], phi[perm]
    loc = np.stack([theta, phi], 1)
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    ain = a[0] if spin == 0 else a
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=2, epsilon=eps, spin=spin)
    got = T.nusht2d2(L, L, _t(ain), _t(loc), spin=spin).cpu().numpy()
    ref = osht.synth_direct(ain, L, spin, theta, phi)
    if spin:
        ref = np.array([ref.real, ref.imag])
    err = _rel2(got, ref)
    assert err < 10 * max(eps, 1e-13), err


@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("eps", [1e-10, 1e-5])
def test_nusht2d1_adjointness(spin, eps):
    """Dot-product test <x, A y> = <A^H x, y> (test/test_adjointOperator.py:247-277, rtol 1e-9)."""
    import cunusht_amd
    rng = np.random.default_rng(48)
    L = 40
    npts = 4000
    loc = np.stack([np.pi * rng.uniform(0, 1, npts), 2 * np.pi * rng.uniform(0, 1, npts)], 1)
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    ain = a[0] if spin == 0 else a
    x = rng.uniform(-0.5, 0.5, (nc, npts))
    xin = x[0] if spin == 0 else x
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=1, epsilon=eps, spin=spin)
    y = T.nusht2d2(L, L, _t(ain), _t(loc), spin=spin).cpu().numpy().reshape(nc, -1)
    ad = T.nusht2d1(L, L, _t(xin), _t(loc), spin=spin).cpu().numpy().reshape(nc, -1)
    lhs = float(np.sum(x * y))
    rhs = sum(almutil.myalmdot(a[c], ad[c], L) for c in range(nc))
    tol = 1e-9 if eps < 1e-6 else 1e-5
    assert abs(lhs - rhs) < tol * abs(lhs), (lhs, rhs)
    # and against the oracle's adjoint
    vals = xin if spin == 0 else (x[0] + 1j * x[1])
    ref = osht.adjoint_direct(vals, L, spin, loc[:, 0], loc[:, 1]).reshape(nc, -1)
    assert _rel2(ad, ref) < 10 * max(eps, 1e-12)


@pytest.mark.parametrize("spin", [1, 0, 2])
@pytest.mark.parametrize("L", [33, 34])
@pytest.mark.parametrize("eps", [1e-10, 1e-5])
def test_nusht_pair_odd_spins_and_lmax(spin, L, eps):
    """nusht2d2 vs direct summation and nusht2d1 vs the direct adjoint at spin 1 (the parity (s + m) of the doubled
    theta columns flips with s, and the theta pass transforms columns of opposite parity together), spin 0 and
    spin 2, at odd and even lmax (spin 0 then has an even or odd number of m columns; with spin, 2 lmax + 1
    columns leave one unpaired)."""
    import cunusht_amd
    rng = np.random.default_rng(900 + 10 * spin + L + int(-np.log10(eps)))
    npts = 2500
    theta = np.arccos(rng.uniform(-1, 1, npts))
    phi = rng.uniform(0, 2 * np.pi, npts)
    theta[:2] = [0.0, np.pi]
    loc = np.stack([theta, phi], 1)
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    ain = a[0] if spin == 0 else a
    print(f"case pair spin={spin} L={L} eps={eps}", flush=True)
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=2, epsilon=eps, spin=spin)
    got = T.nusht2d2(L, L, _t(ain), _t(loc), spin=spin).cpu().numpy()
    ref = osht.synth_direct(ain, L, spin, theta, phi)
    if spin:
        ref = np.array([ref.real, ref.imag])
    assert _rel2(got, ref) < 10 * max(eps, 1e-13)
    x = rng.standard_normal((nc, npts))
    xin = x[0] if spin == 0 else x
    ad = T.nusht2d1(L, L, _t(xin), _t(loc), spin=spin).cpu().numpy().reshape(nc, -1)
    vals = xin if spin == 0 else (x[0] + 1j * x[1])
    aref = osht.adjoint_direct(vals, L, spin, theta, phi).reshape(nc, -1)
    assert _rel2(ad, aref) < 30 * max(eps, 1e-13)


# ------------------------------------------------------------------ lensing
@pytest.mark.parametrize("spin", [0, 2])
def test_gclm2lenmap_vs_oracle(spin):
    import cunusht_amd
    rng = np.random.default_rng(77 + spin)
    L = 48
    nphi = 2 * L + 1  # 97: prime -> Rader
    ll = np.arange(L + 1)
    dlm = almutil.synalm(np.where(ll >= 2, 1e-4 / (ll + 1.0) ** 2, 0.0), L, rng)
    nc = 1 if spin == 0 else 2
    g = almutil.random_alm(L, spin, nc, rng)
    gin = g[0] if spin == 0 else g
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L, "nphi": nphi}), nuFFTtype=2, epsilon=1e-12,
                                           spin=spin)
    got = T.gclm2lenmap(_t(gin), L, L, dlm_scaled=_t(dlm), spin=spin).cpu().numpy()
    rings = ogeom.gl_geometry(L, nphi=nphi)
    ref = pipeline.gclm2lenmap(gin, dlm, L, spin, rings, 1e-13)
    assert _relmax(got, ref) < 1e-9
    x = rng.standard_normal(got.shape)
    ad = T.lenmap2gclm(_t(x), _t(dlm), None, L, L, spin=spin).cpu().numpy().reshape(nc, -1)
    refad = pipeline.lenmap2gclm(x, dlm, L, spin, nc, rings, 1e-13)
    assert _relmax(ad, refad) < 1e-9


@pytest.mark.parametrize("spin", [1, 2])
@pytest.mark.parametrize("amp", [1e-3, 3e-2])
def test_gclm2lenmap_rotation_by_parallel_transport(spin, amp):
    """The library's spin-s lensed map of a constant-ambient-tensor field (band-limited at l = s) equals its
    exact lensing by parallel transport along each pixel's deflection geodesic (oracle/transport.py): this
    pins the pointing, the angle gamma and the direction of the e^{i s gamma} rotation
    (CPU_nufft_transformer.py:158-160,356-368) without the oracle's formulas."""
    import cunusht_amd
    rng = np.random.default_rng(170 + 10 * spin + int(amp * 1e4))
    L = 24
    rings = ogeom.gl_geometry(L)
    Tn = transport.random_tensor(spin, rng)
    th, ph = rings.pixel_angles()
    w = np.repeat(rings.weight, rings.nph)
    EB = osht.adjoint_direct(transport.ambient_field(spin, Tn, th, ph) * w, L, spin, th, ph)
    ll = np.arange(L + 1)
    dlm = almutil.synalm(np.where(ll >= 2, amp / (ll + 1.0) ** 2, 0.0), L, rng)
    _, _, _, (dt, dp) = pipeline.dlm2pointing(dlm, L, rings)
    exact = transport.lensed_ambient_field(spin, Tn, th, ph, dt, dp)
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=2, epsilon=1e-12, spin=spin)
    print(f"case rotation by transport spin={spin} amp={amp}", flush=True)
    got = T.gclm2lenmap(_t(EB), L, L, dlm_scaled=_t(dlm), spin=spin).cpu().numpy()
    _sync()
    err = np.max(np.abs(got[0] + 1j * got[1] - exact)) / np.max(np.abs(exact))
    print(f"  max rel err vs parallel transport {err:.2e}")
    assert err < 1e-9


# ------------------------------------------------------------------ batched synthesis
@pytest.mark.parametrize("nmaps", [1, 3, 8])
def test_batch_matches_single_calls(nmaps):
    """nusht_nusht2d2_batch (one Legendre recursion for all maps) == independent nusht2d2 calls."""
    import cunusht_amd
    from cunusht_amd.transformer import NuPlan
    rng = np.random.default_rng(300 + nmaps)
    L = 36
    npts = 2000
    loc = np.stack([np.arccos(rng.uniform(-1, 1, npts)), rng.uniform(0, 2 * np.pi, npts)], 1)
    th, ph = _t(loc[:, 0]), _t(loc[:, 1])
    gam = _t(rng.uniform(-0.1, 0.1, npts))
    a = np.stack([almutil.random_alm(L, 2, 2, rng) for _ in range(nmaps)])
    at = _t(a)
    plan = NuPlan(L, 2, 2, 1e-10, npts)
    plan.set_pts(th, ph)
    out = torch.zeros((nmaps, 2, npts), dtype=torch.float64, device="cuda")
    nalm = a.shape[2]
    _lib().call("nusht_nusht2d2_batch", plan._h, nmaps, _p(at), 4 * nalm, _p(out[0, 0]), _p(out[0, 1]), 2 * npts,
                _p(gam), _stream())
    _sync()
    for b in range(nmaps):
        ref = torch.zeros((2, npts), dtype=torch.float64, device="cuda")
        ab = _t(a[b])
        _lib().call("nusht_nusht2d2", plan._h, _p(ab), _p(ref[0]), _p(ref[1]), _p(gam), _stream())
        _sync()
        assert torch.equal(out[b], ref) or _relmax(out[b].cpu().numpy(), ref.cpu().numpy()) < 1e-13


@pytest.mark.parametrize("spin", [0, 2])
def test_batch_every_variant(spin, monkeypatch):
    """Every Legendre synthesis variant, batched (2 maps per launch) and single, at lmax 200 (two groups of
    ring pairs, underflow near the poles), against variant 0's single-map result, within 1e-12 relative."""
    from cunusht_amd.transformer import NuPlan
    rng = np.random.default_rng(500 + spin)
    L, npts, nmaps = 200, 3000, 3
    nc = 1 if spin == 0 else 2
    loc = np.stack([np.arccos(rng.uniform(-1, 1, npts)), rng.uniform(0, 2 * np.pi, npts)], 1)
    th, ph = _t(loc[:, 0]), _t(loc[:, 1])
    gam = _t(rng.uniform(-0.1, 0.1, npts)) if spin else None
    a = np.stack([almutil.random_alm(L, spin, nc, rng) for _ in range(nmaps)])
    at = _t(a)
    nalm = a.shape[2]
    nv = _lib().load().nusht_legendre_variants()
    ref = None
    for v in range(nv):
        monkeypatch.setenv("NUSHT_LEGENDRE_VARIANT", str(v))
        plan = NuPlan(L, spin, nc, 1e-10, npts)
        assert plan.info()["legendre_variant"] == v
        plan.set_pts(th, ph)
        out = torch.zeros((nmaps, 2, npts), dtype=torch.float64, device="cuda")
        _lib().call("nusht_nusht2d2_batch", plan._h, nmaps, _p(at), 2 * nc * nalm, _p(out[0, 0]),
                    _p(out[0, 1]) if spin else None, 2 * npts, _p(gam) if spin else None, _stream())
        single = torch.zeros((nmaps, 2, npts), dtype=torch.float64, device="cuda")
        for b in range(nmaps):
            ab = _t(a[b])
            _lib().call("nusht_nusht2d2", plan._h, _p(ab), _p(single[b, 0]), _p(single[b, 1]) if spin else None,
                        _p(gam) if spin else None, _stream())
        _sync()
        got_b, got_s = out.cpu().numpy(), single.cpu().numpy()
        if ref is None:
            ref = got_s
        for b in range(nmaps):
            assert _relmax(got_s[b], ref[b]) < 1e-12, (v, b)
            assert _relmax(got_b[b], ref[b]) < 1e-12, (v, b)
        del plan


@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("nmaps", [4, 5, 7, 8])
def test_batch_matrix_core_path(spin, nmaps, monkeypatch):
    """Batches of >= 3 maps run the Legendre contraction on the matrix cores (4 maps per launch, 7 or 8 in one
    launch of two column groups; the remainder on the vector kernels).  Compared with the vector-only path
    (NUSHT_LEGENDRE_MFMA=0) at lmax 200 (two groups of ring pairs, underflow near the poles), within 1e-12
    relative."""
    from cunusht_amd.transformer import NuPlan
    rng = np.random.default_rng(700 + 10 * spin + nmaps)
    L, npts = 200, 3000
    nc = 1 if spin == 0 else 2
    loc = np.stack([np.arccos(rng.uniform(-1, 1, npts)), rng.uniform(0, 2 * np.pi, npts)], 1)
    th, ph = _t(loc[:, 0]), _t(loc[:, 1])
    gam = _t(rng.uniform(-0.1, 0.1, npts)) if spin else None
    a = np.stack([almutil.random_alm(L, spin, nc, rng) for _ in range(nmaps)])
    at = _t(a)
    nalm = a.shape[2]
    res = {}
    for mf in ("1", "0"):
        monkeypatch.setenv("NUSHT_LEGENDRE_MFMA", mf)
        plan = NuPlan(L, spin, nc, 1e-10, npts)
        plan.set_pts(th, ph)
        out = torch.zeros((nmaps, 2, npts), dtype=torch.float64, device="cuda")
        _lib().call("nusht_nusht2d2_batch", plan._h, nmaps, _p(at), 2 * nc * nalm, _p(out[0, 0]),
                    _p(out[0, 1]) if spin else None, 2 * npts, _p(gam) if spin else None, _stream())
        _sync()
        res[mf] = out.cpu().numpy()
        del plan
    for b in range(nmaps):
        assert _relmax(res["1"][b], res["0"][b]) < 1e-12, b


# ------------------------------------------------------------------ magnification (inverse lensing)
@pytest.mark.parametrize("curl", [False, True])
@pytest.mark.parametrize("exact", [True, False])
def test_dlm2A_vs_oracle(curl, exact):
    """dlm2A on the GL pointing grid vs the oracle restatement (direct summation), within 1e-12."""
    import cunusht_amd
    from oracle import lensing
    L, nphi = 32, 65
    rng = np.random.default_rng(40 + curl + 2 * exact)
    ll = np.arange(L + 1)
    cl = np.where(ll >= 1, 0.02 / (ll + 1.0) ** 2, 0.0)
    dlm = almutil.synalm(cl, L, rng)
    dclm = almutil.synalm(cl, L, rng) * 0.5 if curl else None
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L, "nphi": nphi}))
    got = cunusht_amd.dlm2A(_t(dlm), L, L, T, dclm=_t(dclm) if curl else None, exact=exact).cpu().numpy()
    _sync()
    rings = ogeom.gl_geometry(L, nphi=nphi)
    th, ph = rings.pixel_angles()
    ref = lensing.magnification(dlm, L, th, ph, dclm, exact=exact)
    assert np.max(np.abs(got - ref)) < 1e-12
    # exact=True is the solid-angle ratio of the deflection map (independent 4th-order finite
    # differences; step 4e-4 keeps the truncation error ~1e-9 at lmax 32)
    if exact:
        sel = rng.choice(th.size, 40, replace=False)
        R = lensing.solid_angle_ratio(dlm, L, th[sel], ph[sel], dclm, h=4e-4)
        assert np.max(np.abs(got[sel] - R)) < 1e-8


def test_dlm2A_zero_deflection_is_one():
    import cunusht_amd
    L = 16
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L, "nphi": 33}))
    A = T.dlm2A(torch.zeros((L + 1) * (L + 2) // 2, dtype=torch.complex128, device="cuda"))
    assert torch.equal(A, torch.ones_like(A))


# ------------------------------------------------------------------ small lmax (fine grid widened to a non-wrapping window)
@pytest.mark.parametrize("L,spin", [(1, 0), (2, 0), (4, 0), (8, 0), (11, 0), (2, 2), (4, 2), (8, 2), (11, 2)])
@pytest.mark.parametrize("eps", [1e-12, 1e-7, 1e-5])
def test_small_lmax_vs_direct(L, spin, eps):
    """nusht2d2 / nusht2d1 at small lmax, with points on both poles, against direct Y_lm summation (forward within
    10 epsilon, 30 epsilon in FP32; the adjoint within 30 epsilon); the fine grid of these plans would be narrower
    than the rows a point at a pole touches without widening."""
    import cunusht_amd
    rng = np.random.default_rng(500 + 10 * L + spin)
    npts = 400
    theta = np.concatenate([[0.0, np.pi, 1e-3, np.pi - 1e-3], np.arccos(rng.uniform(-1, 1, npts - 4))])
    phi = np.concatenate([[0.3, 2 * np.pi - 1e-12, -1.0, 7.0], rng.uniform(0, 2 * np.pi, npts - 4)])
    loc = np.stack([theta, phi], 1)
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    ain = a[0] if spin == 0 else a
    print(f"case L={L} spin={spin} eps={eps}", flush=True)
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=2, epsilon=eps, spin=spin)
    got = T.nusht2d2(L, L, _t(ain), _t(loc), spin=spin).cpu().numpy()
    ref = osht.synth_direct(ain, L, spin, theta, phi)
    if spin:
        ref = np.array([ref.real, ref.imag])
    tol = 10 * max(eps, 1e-13) if eps < 1e-6 else 30 * eps
    assert np.all(np.isfinite(got))
    assert _rel2(got, ref) < tol
    x = rng.uniform(-0.5, 0.5, (nc, npts))
    xin = x[0] if spin == 0 else x
    ad = T.nusht2d1(L, L, _t(xin), _t(loc), spin=spin).cpu().numpy().reshape(nc, -1)
    vals = xin if spin == 0 else (x[0] + 1j * x[1])
    aref = osht.adjoint_direct(vals, L, spin, theta, phi).reshape(nc, -1)
    # the adjoint sums 400 kernel-weighted values into a handful of coefficients: 30 epsilon
    assert _rel2(ad, aref) < 30 * max(eps, 1e-12)


@pytest.mark.parametrize("L", [2, 8])
@pytest.mark.parametrize("spin", [0, 2])
def test_small_lmax_gclm2lenmap_vs_oracle(L, spin):
    import cunusht_amd
    rng = np.random.default_rng(600 + spin + 10 * L)
    ll = np.arange(L + 1)
    dlm = almutil.synalm(np.where(ll >= 2, 1e-3 / (ll + 1.0) ** 2, 0.0), L, rng)
    nc = 1 if spin == 0 else 2
    g = almutil.random_alm(L, spin, nc, rng)
    gin = g[0] if spin == 0 else g
    print(f"case gclm2lenmap L={L} spin={spin}", flush=True)
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L, "nphi": 2 * L + 1}), nuFFTtype=2, epsilon=1e-12,
                                           spin=spin)
    out = T.gclm2lenmap(_t(gin), L, L, dlm_scaled=_t(dlm), spin=spin).cpu().numpy()
    rings = ogeom.gl_geometry(L, nphi=2 * L + 1)
    ref = pipeline.gclm2lenmap(g, dlm, L, spin, rings, 1e-12)
    assert _relmax(out, ref) < 1e-9


def test_lmax_below_spin_is_rejected():
    """A spin-s field has no modes l < s: plans with lmax < spin are refused on the host."""
    from cunusht_amd import _lib
    from cunusht_amd.transformer import NuPlan
    with pytest.raises(_lib.NuShtError, match="lmax must be >= spin"):
        NuPlan(1, 2, 2, 1e-10, 10)


@pytest.mark.parametrize("spin", [0, 2])
def test_adjointness_reference_configuration(spin):
    """The reference's dot-product test at its own settings (test/test_adjointOperator.py:247-277): lmax = mmax
    = 1024, 20 000 points (theta = pi U, phi = 2 pi U), epsilon 1e-8, rng(48), values U(-0.5, 0.5), rtol 1e-9;
    spin 2 in the GRAD_ONLY form (one component)."""
    import cunusht_amd
    lmax, npix, eps = 1024, 20000, 1e-8
    rng = np.random.default_rng(48)
    loc = np.stack([np.pi * rng.uniform(0, 1, npix), 2 * np.pi * rng.uniform(0, 1, npix)], 1)
    a = almutil.random_alm(lmax, spin, 1, rng)
    x = rng.uniform(-0.5, 0.5, (1 if spin == 0 else 2, npix))
    print(f"case reference adjointness spin={spin}", flush=True)
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": lmax}), nuFFTtype=1, epsilon=eps, spin=spin)
    y = T.nusht2d2(lmax, lmax, _t(a[0] if spin == 0 else a), _t(loc), spin=spin).cpu().numpy().reshape(x.shape)
    alm_out = torch.empty((1, almutil.nalm(lmax)), dtype=torch.complex128, device="cuda")
    ad = T.nusht2d1(lmax, lmax, _t(x[0] if spin == 0 else x), _t(loc), alm=alm_out,
                    spin=spin).cpu().numpy().reshape(1, -1)
    v1 = almutil.myalmdot(a[0], ad[0], lmax)
    v2 = float(np.sum(x * y))
    np.testing.assert_allclose(v1, v2, rtol=1e-9)


# ------------------------------------------------------------------ every ES kernel width
# (epsilon, sigma) -> kernel width at lmax 30: FP32 nuFFT for epsilon >= 1e-6 (widths 2..10), FP64 below (7..16
# through sigma 2, 22 through sigma 1.25 = the runtime-width kernels above 16)
_WIDTH_CASES = [(0.1, 0.0, 2), (0.03, 0.0, 3), (1e-2, 0.0, 4), (1e-3, 0.0, 5), (3e-4, 0.0, 6), (1e-4, 0.0, 7),
                (1e-5, 0.0, 8), (3e-6, 0.0, 9), (1e-6, 0.0, 10), (5e-7, 2.0, 7), (1e-7, 2.0, 8), (3e-9, 0.0, 9),
                (1e-9, 0.0, 10), (1e-10, 0.0, 11), (1e-11, 0.0, 12), (1e-12, 0.0, 13), (1e-13, 0.0, 14),
                (1e-14, 0.0, 15), (1e-15, 0.0, 16), (1e-14, 1.25, 22)]


@pytest.mark.parametrize("eps,sigma,width", _WIDTH_CASES)
def test_every_kernel_width(eps, sigma, width):
    """Interpolation and spreading at every unrolled kernel width (and one runtime width), spin 2: nusht2d2 against
    direct Y_lm summation, nusht2d1 against the direct adjoint and the dot-product test between the two (the
    spreading is the exact transpose of the interpolation: 1e-9 in FP64, 1e-5 in FP32)."""
    import cunusht_amd
    rng = np.random.default_rng(7000 + width)
    L, spin, npts = 30, 2, 2000
    theta = np.arccos(rng.uniform(-1, 1, npts))
    phi = rng.uniform(0, 2 * np.pi, npts)
    loc = np.stack([theta, phi], 1)
    a = almutil.random_alm(L, spin, 2, rng)
    print(f"case eps={eps:g} sigma={sigma} width={width}", flush=True)
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=2, epsilon=eps, spin=spin, sigma=sigma)
    y = T.nusht2d2(L, L, _t(a), _t(loc), spin=spin).cpu().numpy()
    widths = {p.info()["kernel_width"] for p in T._plans.values()}
    assert widths == {width}, widths
    single = eps >= 1e-6
    ref = osht.synth_direct(a, L, spin, theta, phi)
    assert _rel2(y, np.array([ref.real, ref.imag])) < (30 if single else 10) * max(eps, 1e-13)
    x = rng.uniform(-0.5, 0.5, (2, npts))
    ad = T.nusht2d1(L, L, _t(x), _t(loc), spin=spin).cpu().numpy().reshape(2, -1)
    aref = osht.adjoint_direct(x[0] + 1j * x[1], L, spin, theta, phi).reshape(2, -1)
    assert _rel2(ad, aref) < (30 if single else 10) * max(eps, 1e-12)
    lhs = float(np.sum(x * y))
    rhs = sum(almutil.myalmdot(a[c], ad[c], L) for c in range(2))
    assert abs(lhs - rhs) < (1e-5 if single else 1e-9) * abs(lhs), (lhs, rhs)
