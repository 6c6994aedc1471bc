"""Pointing geometries: pixel weights and quadrature exactness (host logic, no GPU).

The weights follow the ducc0 convention that cunusht/geometry.py:228-284 builds on (ring quadrature
weight x 2 pi / nphi).  They are pinned here by what the quadrature rules fix: the total solid angle, exact
integration of polynomials in cos(theta) up to each rule's degree, and analysis(synthesis(alm)) = alm on
grids that resolve the band limit.
"""
import numpy as np
import pytest

from cunusht_amd import geometry as G
from oracle import almutil, sht as osht
from oracle import geometry as ogeom


def _rule_exact(nodes_theta, w, degree):
    x = np.cos(nodes_theta)
    for k in range(degree + 1):
        exact = 2.0 / (k + 1) if k % 2 == 0 else 0.0
        assert abs(np.sum(w * x ** k) - exact) < 1e-13, (k, np.sum(w * x ** k), exact)


@pytest.mark.parametrize("n", [1, 2, 3, 8, 17, 64])
def test_cc_weights_exact(n):
    th = np.linspace(0.0, np.pi, n) if n > 1 else np.array([0.5 * np.pi])
    _rule_exact(th, G.cc_weights(n), n - 1 if n > 1 else 1)


@pytest.mark.parametrize("n", [1, 2, 5, 16, 33])
def test_f1_weights_exact(n):
    th = np.pi * (np.arange(n) + 0.5) / n
    _rule_exact(th, G.f1_weights(n), n - 1 if n > 1 else 1)


def test_cc_weights_known_values():
    # Simpson's rule is Clenshaw-Curtis with three points: (1, 4, 1) / 3
    assert np.allclose(G.cc_weights(3), [1 / 3, 4 / 3, 1 / 3], rtol=0, atol=1e-15)
    # five points: 1/15, 8/15, 12/15, 8/15, 1/15
    assert np.allclose(G.cc_weights(5), np.array([1, 8, 12, 8, 1]) / 15, rtol=0, atol=1e-15)


@pytest.mark.parametrize("geo", [("gl", {"lmax": 31}), ("gl", {"lmax": 20, "nphi": 45}),
                                 ("cc", {"ntheta": 40, "nphi": 64}), ("f1", {"ntheta": 33, "nphi": 70})])
def test_pixel_weights_sum_to_sphere(geo):
    g = G.get_geom(geo)
    w = np.repeat(g.weight, g.nph)
    assert abs(np.sum(w) - 4 * np.pi) < 1e-12


def test_gl_weights_match_oracle():
    g = G.get_geom(("gl", {"lmax": 47}))
    o = ogeom.gl_geometry(47)
    assert np.allclose(g.theta, o.theta, rtol=0, atol=1e-15)
    # nodes agree to the ulp; the weight 2 sin^2(theta) / (n P_{n-1})^2 moves by ~n ulp per ulp of theta
    assert np.allclose(g.weight, o.weight, rtol=1e-13, atol=0)


@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("geo", [("gl", {"lmax": 12}), ("cc", {"ntheta": 28, "nphi": 27}),
                                 ("f1", {"ntheta": 27, "nphi": 26})])
def test_weighted_adjoint_inverts_synthesis(spin, geo):
    """Quadrature analysis (the adjoint applied to weight x map) returns the alm of a band-limited map:
    GL with lmax + 1 rings, CC / F1 with more than 2 lmax + 1 rings (exact to degree n - 1 in cos(theta))."""
    L = 12
    rng = np.random.default_rng(spin)
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    ain = a[0] if spin == 0 else a
    g = G.get_geom(geo)
    th, ph = g.pixel_angles()
    m = osht.synth_direct(ain, L, spin, th, ph)
    w = np.repeat(g.weight, g.nph)
    back = osht.adjoint_direct(w * m, L, spin, th, ph)
    assert np.max(np.abs(np.reshape(back, (nc, -1)) - a)) < 1e-12


# ------------------------------------------------------------------ HEALPix, thinned GL, helpers
def test_healpix_nside1_known_pixels():
    """HEALPix nside 1: 3 rings of 4 pixels at z = 2/3, 0, -2/3; pixel 0 at phi = pi/4, pixel 4 at phi = 0."""
    g = G.get_geom(("healpix", {"nside": 1}))
    assert np.allclose(np.cos(g.theta), [2 / 3, 0, -2 / 3], atol=1e-15)
    assert list(g.nph) == [4, 4, 4] and list(g.ofs) == [0, 4, 8]
    assert np.allclose(g.phi0, [np.pi / 4, 0.0, np.pi / 4], atol=1e-15)


@pytest.mark.parametrize("nside", [2, 8, 32])
def test_healpix_rings(nside):
    g = G.get_geom(("healpix", {"nside": nside}))
    assert g.npix() == 12 * nside ** 2 and g.theta.size == 4 * nside - 1
    assert abs(g.fsky() - 1.0) < 1e-14
    # ring z values of the polar cap, the cap / belt seam (z = 2/3) and the equator; north-south mirror
    z = np.cos(g.theta)
    i = np.arange(1, nside)
    assert np.allclose(z[:nside - 1], 1 - i ** 2 / (3.0 * nside ** 2), rtol=0, atol=1e-14)
    assert abs(z[nside - 1] - 2 / 3) < 1e-14 and abs(z[2 * nside - 1]) < 1e-15
    assert np.allclose(g.theta + g.theta[::-1], np.pi, atol=1e-14)
    assert list(g.nph[:nside]) == list(4 * np.arange(1, nside + 1))
    # the polar rings are exact even at tiny colatitude
    assert abs(g.theta[0] - 2 * np.arcsin(1 / (np.sqrt(6) * nside))) == 0


def test_thingauss_matches_rule_and_thins():
    g = G.get_geom(("tgl", {"lmax": 300, "smax": 2}))
    full = G.get_geom(("gl", {"lmax": 300}))
    assert np.array_equal(g.theta, full.theta)
    assert abs(g.fsky() - 1.0) < 1e-13
    assert g.nph.min() < full.nph.min() and g.nph.max() <= full.nph.max()
    m = np.maximum(G.st2mmax(2, g.theta, 300), G.st2mmax(-2, g.theta, 300))
    assert np.all(g.nph >= np.minimum(2 * m + 1, 2 * 300 + 1) - 1e-9)


def test_restrict_split_thinout():
    g = G.get_geom(("gl", {"lmax": 40}))
    bands = g.split(3)
    assert sum(b.npix() for b in bands) == g.npix()
    pix = np.sort(np.concatenate([G.Geom.rings2pix(b, np.arange(b.theta.size)) for b in bands]))
    assert np.array_equal(pix, np.arange(g.npix()))
    r = g.restrict(0.0, 0.5, False, update_ringstart=True)
    assert np.all(r.theta <= 0.5) and r.ofs[0] == 0
    t = g.thinout(0)
    assert abs(t.fsky() - 1.0) < 1e-13 and t.npix() <= g.npix()
    phs = G.Geom.phis(g, 3)
    assert phs.size == g.nph[3] and np.all((phs >= 0) & (phs < 2 * np.pi))
    assert "healpix" in G.Geom.get_supported_geometries()


def test_pbounds():
    pb = G.pbounds(0.1, 0.4)
    assert pb.contains(np.array([0.0, 0.29, 2 * np.pi - 0.09])).all()
    assert not pb.contains(np.array([0.31, np.pi])).any()
    assert pb.get_range() == pytest.approx(0.4) and pb == G.pbounds(0.1, 0.4)
    assert G.pbdGeometry(G.get_geom(("healpix", {"nside": 2})), pb).npix() == 48


@pytest.mark.parametrize("n", [1, 2, 7, 512, 4096])
def test_gl_rule_is_exact_to_degree_2n_minus_1(n):
    """The n-point rule integrates every Legendre polynomial of degree < 2n exactly (P_0 to 2, the others to 0);
    checked at degrees 0, 1, 2, n, 2n - 2 and 2n - 1 by the three-term recurrence at the nodes."""
    import time
    t0 = time.perf_counter()
    th, w = G.gauss_legendre_thetas(n)
    assert time.perf_counter() - t0 < 5.0  # the Newton loop stops once converged
    x = np.cos(th)
    degs = sorted(d for d in {0, 1, 2, n, 2 * n - 2, 2 * n - 1} if 0 <= d < 2 * n)
    p0, p1 = np.ones_like(x), x.copy()
    vals = {0: p0, 1: p1}
    for k in range(1, max(degs)):
        p0, p1 = p1, ((2 * k + 1) * x * p1 - k * p0) / (k + 1)
        vals[k + 1] = p1
    for d in degs:
        got = float(np.dot(w, vals[d]))
        assert abs(got - (2.0 if d == 0 else 0.0)) < 1e-13, (n, d, got)
