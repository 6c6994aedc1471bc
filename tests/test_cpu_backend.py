"""The CPU backend, get_transformer('CPU') over libnusht_cpu.so, against the oracle (no GPU needed).

Bars: the Legendre core within 1e-13 (lmax <= 40) / 2e-13 (lmax 200) relative max error of the oracle's C
restatement; the operators within 10 epsilon (epsilon = 1e-12) of direct FP64 summation at the deflected
points (oracle/sht.py) and of the oracle pipeline; adjoint pairs to 1e-12 relative in the dot-product test
(test/test_adjointOperator.py:247-277 of the reference); the ES nuFFT within 10 epsilon of a direct
non-uniform DFT.  Single-precision mode (epsilon > 1e-6) within 30 epsilon, float32 outputs.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import almutil, clib, geometry as ogeom, pipeline, sht as osht

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nusht_cpu.h")
LIB = os.path.join(ROOT, "cunusht_amd", "libnusht_cpu.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libnusht_cpu.so not built")


def _relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _declared():
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(nusht_cpu_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from cunusht_amd import cpu_transformer as ct
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (nusht_cpu_[a-z0-9_]+)", out))
    assert set(_declared()) <= exported
    assert set(_declared()) == set(ct.SYMBOLS)
    assert ct.load().nusht_cpu_version() == 1


def test_get_transformer_cpu():
    """The reference's factory: default / lenspyx and duccnufft solvers (cunusht/cunusht.py:18-27,63-73)."""
    import cunusht_amd
    from cunusht_amd.cpu_transformer import CPU_nusht_lenspyx_transformer, CPU_nusht_transformer
    assert cunusht_amd.get_transformer("CPU") is CPU_nusht_lenspyx_transformer
    assert cunusht_amd.get_transformer("CPU", solver="lenspyx") is CPU_nusht_lenspyx_transformer
    assert cunusht_amd.get_transformer("CPU", solver="duccnufft") is CPU_nusht_transformer
    with pytest.raises(AssertionError):
        cunusht_amd.get_transformer("CPU", solver="finufft")
    with pytest.raises(AssertionError):
        CPU_nusht_transformer(("gl", {"lmax": 8}), nuFFTtype=2, epsilon=1e-10)


@pytest.mark.parametrize("lmax,spin,ncomp", [(20, 0, 1), (20, 1, 1), (21, 2, 2), (33, 3, 2), (200, 2, 2),
                                             (200, 0, 1)])
def test_legendre_core_vs_oracle(lmax, spin, ncomp):
    from cunusht_amd import cpu_transformer as ct
    rng = np.random.default_rng(lmax + 10 * spin)
    a = np.atleast_2d(almutil.random_alm(lmax, spin, ncomp, rng))
    th = np.concatenate([[0.0, np.pi, 1e-9], np.sort(np.arccos(rng.uniform(-1, 1, 40)))])
    G = ct.synth_rings(a, lmax, spin, th, 4)
    tol = 1e-13 if lmax <= 40 else 2e-13
    assert _relmax(G, clib.sht_synth_rings(a, lmax, spin, th)) < tol
    Gp = rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape)
    A = ct.adjoint_rings(Gp, lmax, spin, ncomp, th, 4)
    assert _relmax(A, clib.sht_adjoint_rings(Gp, lmax, spin, ncomp, th)) < 4 * tol


def test_legendre_core_rejects_bad_arguments():
    from cunusht_amd import cpu_transformer as ct
    a = np.zeros((1, almutil.nalm(4)), dtype=np.complex128)
    with pytest.raises(ValueError):
        ct.synth_rings(a, 4, 0, np.array([4.0]))          # theta outside [0, pi]
    with pytest.raises(ValueError):
        ct.synth_rings(np.zeros((2, a.shape[1])), 4, 0, np.array([1.0]))  # two components at spin 0


@pytest.mark.parametrize("spin", [0, 2])
def test_es_nufft_vs_direct_and_adjoint(spin):
    from cunusht_amd.cpu_transformer import EsGrid
    L, eps, n = 12, 1e-12, 300
    g = EsGrid(L, spin, eps)
    rng = np.random.default_rng(5 + spin)
    F = rng.standard_normal((g.ks.size, g.ms.size)) + 1j * rng.standard_normal((g.ks.size, g.ms.size))
    th = np.concatenate([[0.0, np.pi], rng.uniform(0, np.pi, n - 2)])
    ph = rng.uniform(-7, 14, n)
    gam = rng.uniform(-1, 1, n)
    u = g.u2nu(F.copy(), th, ph, 3, gamma=gam, spin_rot=spin)
    E = np.exp(1j * (np.outer(th, g.ks)[:, :, None] + np.outer(ph, g.ms)[:, None, :]))
    rot = np.exp(1j * spin * gam)
    assert _relmax(u, rot * np.einsum("pkm,km->p", E, F)) < 10 * eps
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    Fp = g.nu2u(v, th, ph, 3, gamma=gam, spin_rot=spin)
    assert _relmax(Fp, np.einsum("pkm,p->km", np.conj(E), np.conj(rot) * v)) < 10 * eps
    lhs = np.vdot(v, u)
    assert abs(lhs - np.vdot(Fp, F)) < 1e-12 * abs(lhs)
    deg, err = g.kernel_info()
    assert 2 <= deg <= 24 and err <= g.grid.ktol


def _case(L, spin, rng, scale=1e-5):
    ll = np.arange(L + 1)
    dlm_s = almutil.synalm(np.where(ll >= 2, scale / (ll + 1.0) ** 2, 0.0), L, rng)
    dlm = almutil.almxfl(dlm_s, np.sqrt(ll * (ll + 1.0)), L)
    nc = 1 if spin == 0 else 2
    return dlm_s, dlm, almutil.random_alm(L, spin, nc, rng), nc


def _transformer(L, geom=None, eps=1e-12):
    import cunusht_amd
    return cunusht_amd.get_transformer("CPU", solver="duccnufft")(geom or ("gl", {"lmax": L, "nphi": 2 * L + 1}),
                                                                  epsilon=eps)


@pytest.mark.parametrize("L,spin", [(24, 0), (24, 2), (40, 1), (33, 3), (64, 2)])
def test_gclm2lenmap_vs_direct_and_oracle(L, spin):
    rng = np.random.default_rng(100 + L + spin)
    dlm_s, dlm, g, nc = _case(L, spin, rng)
    T = _transformer(L)
    out = T.gclm2lenmap(g, dlm, L, L, spin, 1e-12, 4, verbose=0)
    assert out.shape == ((T.geom.npix(),) if spin == 0 else (2, T.geom.npix()))
    tp, pp = T.deflectionlib.ptg()
    ex = osht.synth_direct(g, L, spin, tp, pp)
    if spin:
        ex = ex * np.exp(1j * spin * T.deflectionlib.gamma())
        ex = np.array([ex.real, ex.imag])
    assert _relmax(out, ex) < 1e-11
    rings = ogeom.gl_geometry(L, nphi=2 * L + 1)
    assert _relmax(out, pipeline.gclm2lenmap(g, dlm_s, L, spin, rings, 1e-12)) < 1e-11


@pytest.mark.parametrize("L,spin,nc", [(24, 0, 1), (24, 2, 2), (24, 2, 1), (31, 1, 2)])
def test_lenmap2gclm_is_adjoint_and_matches_oracle(L, spin, nc):
    rng = np.random.default_rng(200 + L + spin + nc)
    dlm_s, dlm, g, _ = _case(L, spin, rng)
    g = np.atleast_2d(g)[:nc]
    T = _transformer(L)
    out = T.gclm2lenmap(g, dlm, L, L, spin, 1e-12, 4, verbose=0)
    x = rng.standard_normal(out.shape)
    gout = np.zeros((nc, almutil.nalm(L)), dtype=np.complex128)
    a = T.lenmap2gclm(x, dlm, gout, spin, L, L, 4, epsilon=1e-12, verbose=0)
    assert a.shape == (nc, almutil.nalm(L)) and np.array_equal(a, gout)
    lhs = float(np.sum(out * x))
    rhs = sum(almutil.myalmdot(g[c], a[c], L) for c in range(nc))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    rings = ogeom.gl_geometry(L, nphi=2 * L + 1)
    assert _relmax(a, pipeline.lenmap2gclm(x, dlm_s, L, spin, nc, rings, 1e-12)) < 1e-10


def test_dlm2pointing_vs_oracle():
    L = 30
    rng = np.random.default_rng(8)
    dlm_s, dlm, _, _ = _case(L, 0, rng, scale=1e-3)
    T = _transformer(L)
    ptg = T.dlm2pointing(dlm, L, False, 2)
    rings = ogeom.gl_geometry(L, nphi=2 * L + 1)
    tp, pp, gam, _ = pipeline.dlm2pointing(dlm_s, L, rings)
    assert _relmax(ptg[0], tp) < 1e-13
    assert np.max(np.abs(ptg[1] - pp)) < 1e-12
    assert np.max(np.abs(T.deflectionlib.gamma() - gam)) < 1e-12
    assert T.dlm2pointing(dlm, L, True, 2).dtype == np.float32


def test_healpix_geometry_spin2():
    """Ragged rings (12 nside^2 pixels; rings of 4 .. 4 nside points)."""
    nside, L = 8, 16
    rng = np.random.default_rng(9)
    _, dlm, g, _ = _case(L, 2, rng)
    T = _transformer(L, ("healpix", {"nside": nside, "lmax": L}))
    out = T.gclm2lenmap(g, dlm, L, L, 2, 1e-12, 4, verbose=0)
    tp, pp = T.deflectionlib.ptg()
    ex = osht.synth_direct(g, L, 2, tp, pp) * np.exp(2j * T.deflectionlib.gamma())
    assert _relmax(out, np.array([ex.real, ex.imag])) < 1e-11


def test_single_precision_mode():
    L, eps = 24, 1e-5
    rng = np.random.default_rng(10)
    _, dlm, g, _ = _case(L, 2, rng)
    T = _transformer(L, eps=eps)
    out = T.gclm2lenmap(g, dlm, L, L, 2, eps, 4, verbose=0)
    assert out.dtype == np.float32
    tp, pp = T.deflectionlib.ptg()
    ex = osht.synth_direct(g, L, 2, tp, pp) * np.exp(2j * T.deflectionlib.gamma())
    assert _relmax(out, np.array([ex.real, ex.imag])) < 30 * eps
    a = T.lenmap2gclm(out, dlm, np.zeros((2, almutil.nalm(L)), np.complex64), 2, L, L, 4, epsilon=eps, verbose=0)
    assert a.dtype == np.complex64


def test_synthesis_and_analysis_on_geometry():
    """Map synthesis = direct Y_lm sums; weighted adjoint synthesis on GL inverts it for band-limited alm;
    the unweighted one is its exact transpose."""
    L = 20
    rng = np.random.default_rng(11)
    T = _transformer(L)
    th, ph = T.geom.pixel_angles()
    for spin in (0, 2):
        nc = 1 if spin == 0 else 2
        a = np.atleast_2d(almutil.random_alm(L, spin, nc, rng))
        m = T.synthesis(a, spin=spin)
        ref = osht.synth_direct(a[0] if spin == 0 else a, L, spin, th, ph)
        assert _relmax(m, ref[None] if spin == 0 else np.array([ref.real, ref.imag])) < 1e-12
        back = T.adjoint_synthesis(m, spin=spin)
        assert _relmax(back, a) < 1e-11
        x = rng.standard_normal(m.shape)
        at = T.adjoint_synthesis(x, spin=spin, apply_weights=False)
        lhs = float(np.sum(m * x))
        assert abs(lhs - sum(almutil.myalmdot(a[c], at[c], L) for c in range(nc))) < 1e-12 * abs(lhs)


def test_gclm2lenpixs_matches_lenmap():
    L = 24
    rng = np.random.default_rng(12)
    _, dlm, g, _ = _case(L, 2, rng)
    T = _transformer(L)
    out = T.gclm2lenmap(g, dlm, L, L, 2, 1e-12, 4, verbose=0)
    pixs = rng.choice(T.geom.npix(), 17, replace=False)
    q, u = T.gclm2lenpixs(g, L, 2, pixs)
    assert _relmax(np.array([q, u]), out[:, pixs]) < 1e-11
    g0 = almutil.random_alm(L, 0, 1, rng)
    m0 = T.gclm2lenmap(g0, dlm, L, L, 0, 1e-12, 4, verbose=0)
    assert _relmax(T.gclm2lenpixs(g0, L, 0, pixs), m0[pixs]) < 1e-11


def test_mmax_below_lmax_is_the_padded_layout():
    L, mmax = 20, 12
    rng = np.random.default_rng(13)
    _, dlm, _, _ = _case(L, 0, rng)
    n = almutil.nalm(L, mmax)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a[:L + 1] = a[:L + 1].real
    full = np.zeros(almutil.nalm(L), dtype=np.complex128)
    full[:n] = a
    T = _transformer(L)
    m1 = T.gclm2lenmap(a, dlm, L, mmax, 0, 1e-12, 2, verbose=0)
    m2 = T.gclm2lenmap(full, dlm, L, L, 0, 1e-12, 2, verbose=0)
    assert np.array_equal(m1, m2)
    b = T.lenmap2gclm(m1, dlm, None, 0, L, mmax, 2, epsilon=1e-12, verbose=0)
    assert b.shape == (1, n)


def test_execmodes(tmp_path, monkeypatch):
    monkeypatch.setenv("NUSHT_TIMING_DIR", str(tmp_path))
    L = 16
    rng = np.random.default_rng(14)
    _, dlm, g, _ = _case(L, 2, rng)
    T = _transformer(L)
    ret = T.gclm2lenmap(g, dlm, L, L, 2, 1e-12, 2, verbose=0, execmode="debug")
    for k in ("dlm2pointing", "synthesis", "doubling", "C2C", "nuFFT2d2", "rotate"):
        assert k in ret, k
    nt, nph = T.ntheta_CAR, T.nphi_CAR
    assert ret["synthesis"].shape == (2, nt, nph) and ret["doubling"].shape == (2 * nt - 2, nph)
    d = ret["doubling"]
    assert _relmax(ret["C2C"], np.fft.fft2(d) / d.size) < 1e-14
    cc = np.pi * np.arange(nt) / (nt - 1)
    tt = np.repeat(cc, nph)
    pp = np.tile(2 * np.pi * np.arange(nph) / nph, nt)
    car = osht.synth_direct(g, L, 2, tt, pp).reshape(nt, nph)
    assert _relmax(ret["synthesis"], np.array([car.real, car.imag])) < 1e-12
    # the doubled rows follow D[r >= nt][phi] = (-1)^s M[2 nt - 2 - r][phi + pi] (s = 2: no sign)
    h = nph // 2
    assert np.array_equal(d[nt + 3, :h], d[nt - 5, h:])
    T.gclm2lenmap(g, dlm, L, L, 2, 1e-12, 2, verbose=0, execmode="timing")
    for k in ("synthesis", "doubling", "nuFFT2d2", "pointing", "gclm2lenmap"):
        assert k in T.timer.keys, k
    assert (tmp_path / "CPU_nusht_transformer" / "gclm2lenmap" / "lmax16_epsilon1e-12").exists()


def test_error_paths():
    L = 8
    rng = np.random.default_rng(15)
    _, dlm, g, _ = _case(L, 2, rng)
    T = _transformer(L)
    with pytest.raises(ValueError, match="coefficients"):
        T.gclm2lenmap(g[:, :-1], dlm, L, L, 2, 1e-12, 1, verbose=0)
    ptg = np.stack(T.dlm2pointing(dlm, L, False, 1), axis=1)
    T2 = _transformer(L)
    with pytest.raises(ValueError, match="polrot"):
        T2.gclm2lenmap(g, None, L, L, 2, 1e-12, 1, ptg=ptg, verbose=0)
    # without rotation the given pointing is used as is
    m = T2.gclm2lenmap(g, None, L, L, 2, 1e-12, 1, ptg=ptg, polrot=False, verbose=0)
    tp, pp = ptg[:, 0], ptg[:, 1]
    ex = osht.synth_direct(g, L, 2, tp, pp)
    assert _relmax(m, np.array([ex.real, ex.imag])) < 1e-11
    bad = ptg.copy()
    bad[0, 0] = 4.0
    with pytest.raises(ValueError, match="colatitudes"):
        T2.gclm2lenmap(g, None, L, L, 2, 1e-12, 1, ptg=bad, polrot=False, verbose=0)
    with pytest.raises(ValueError):
        T.gclm2lenmap(g, dlm, L, L, 2, 1.5, 1, verbose=0)


def test_lenspyx_signatures():
    """The default CPU solver's signatures (CPU_nufft_transformer.py:560-762) give the same operators."""
    import cunusht_amd
    L, eps = 20, 1e-12
    rng = np.random.default_rng(16)
    _, dlm, g, _ = _case(L, 2, rng)
    geom = ("gl", {"lmax": L, "nphi": 2 * L + 1})
    T = cunusht_amd.get_transformer("CPU")(geom, dglm=dlm, mmax_dlm=L, nthreads=2, verbose=0, epsilon=eps,
                                           single_prec=False)
    D = _transformer(L)
    ref = D.gclm2lenmap(g, dlm, L, L, 2, eps, 2, verbose=0)
    out = T.gclm2lenmap(g, L, L, 2, 2)                       # deflection from the constructor
    assert np.array_equal(out, ref)
    ptg = T.dlm2pointing(dlm, L, False, 2)
    assert ptg.shape == (T.geom.npix(), 2)
    assert np.array_equal(T.gclm2lenmap(g, L, L, 2, 2, ptg=ptg), ref)   # rotation angle kept for its own ptg
    x = rng.standard_normal(ref.shape)
    a = T.lenmap2gclm(x, dlm, 2, L, L, 2, epsilon=eps, gclm_out=np.zeros((2, almutil.nalm(L)), np.complex128))
    assert np.array_equal(a, D.lenmap2gclm(x, dlm, None, 2, L, L, 2, epsilon=eps, verbose=0))
    # synthesis_general: real (ncomp, npts) maps; its adjoint takes them back
    loc = ptg
    m = T.synthesis_general(L, L, g, loc, 2, eps, 2, "STANDARD", 0)
    assert m.shape == (2, loc.shape[0])
    y = rng.standard_normal(m.shape)
    b = T.adjoint_synthesis_general(L, L, y, loc, 2, eps, "STANDARD", 2, None, 0)
    lhs = float(np.sum(m * y))
    assert abs(lhs - sum(almutil.myalmdot(g[c], b[c], L) for c in range(2))) < 1e-12 * abs(lhs)
    q, u = T.gclm2lenpixs(g, L, 2, np.array([3, 50, 400]))
    assert _relmax(np.array([q, u]), ref[:, [3, 50, 400]]) < 1e-11
    assert T.synthesis(g, 2).shape == (2, T.geom.npix())


@pytest.mark.parametrize("L,spin", [(1, 0), (2, 0), (2, 2), (3, 1)])
def test_small_lmax(L, spin):
    """Tiny lmax: the fine grid is the whole torus (rows wrap) and the GL geometry has a few rings."""
    rng = np.random.default_rng(700 + 10 * L + spin)
    _, dlm, g, nc = _case(L, spin, rng, scale=1e-3)
    T = _transformer(L)
    out = T.gclm2lenmap(g, dlm, L, L, spin, 1e-12, 2, verbose=0)
    tp, pp = T.deflectionlib.ptg()
    ex = osht.synth_direct(g, L, spin, tp, pp)
    if spin:
        ex = ex * np.exp(1j * spin * T.deflectionlib.gamma())
        ex = np.array([ex.real, ex.imag])
    assert _relmax(out, ex) < 1e-11
    x = rng.standard_normal(out.shape)
    a = T.lenmap2gclm(x, dlm, None, spin, L, L, 2, epsilon=1e-12, verbose=0)
    lhs = float(np.sum(out * x))
    rhs = sum(almutil.myalmdot(np.atleast_2d(g)[c], a[c], L) for c in range(nc))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


@pytest.mark.parametrize("spin", [0, 2])
def test_adjointness_reference_configuration(spin):
    """The reference's dot-product test at its own settings (test/test_adjointOperator.py:247-277): lmax = mmax
    = 1024, 20 000 points, epsilon 1e-8, rng(48), rtol 1e-9; spin 2 GRAD_ONLY; through the reference-signature
    synthesis_general / adjoint_synthesis_general."""
    lmax, npix, eps = 1024, 20000, 1e-8
    rng = np.random.default_rng(48)
    loc = np.stack([np.pi * rng.uniform(0, 1, npix), 2 * np.pi * rng.uniform(0, 1, npix)], 1)
    a = almutil.random_alm(lmax, spin, 1, rng)
    x = rng.uniform(-0.5, 0.5, npix) + (1j * rng.uniform(-0.5, 0.5, npix) if spin else 0)
    T = _transformer(lmax, eps=eps)
    mode = "GRAD_ONLY" if spin else "STANDARD"
    y = T.synthesis_general(None, lmax, lmax, a, loc, spin, eps, 8, mode, 0)
    ad = T.adjoint_synthesis_general(lmax, lmax, x, loc, spin, eps, 8, mode, None, 0)
    v1 = almutil.myalmdot(a[0], ad[0], lmax)
    v2 = float(np.real(np.vdot(y, x)))
    np.testing.assert_allclose(v1, v2, rtol=1e-9)


@pytest.mark.parametrize("spin", [0, 2])
def test_lensgclm_identity_at_zero_deflection(spin):
    """lensgclm = weighted adjoint synthesis of gclm2lenmap (CPU_nufft_transformer.py:948-971): with no
    deflection it returns the input alm (GL quadrature is exact for band-limited maps)."""
    L = 16
    rng = np.random.default_rng(800 + spin)
    nc = 1 if spin == 0 else 2
    g = np.atleast_2d(almutil.random_alm(L, spin, nc, rng))
    T = _transformer(L)
    zero = np.zeros(almutil.nalm(L), dtype=np.complex128)
    out = T.lensgclm(g, zero, spin, L, nthreads=2, epsilon=1e-12)
    assert out.shape == g.shape
    assert _relmax(out, g) < 1e-10
    grad = T.lensgclm(g, zero, spin, L, nthreads=2, epsilon=1e-12, out_sht_mode="GRAD_ONLY")
    assert grad.shape == (1, g.shape[1]) and _relmax(grad[0], g[0]) < 1e-10
