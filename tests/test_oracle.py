"""Pins the oracle (CPU, FP64) against mathematics and the reference's own known answers.

Each check is chosen so that a plausible mistake (dropped term, sign, index, transposed
operand) fails one of them; see the oracle module headers for which check pins which function.
"""
import math

import mpmath as mp
import numpy as np
import pytest

from oracle import almutil, clib, geometry, nufft, pipeline, pointing, sht, transport


def _wigner_mp(j, m1, m2, b):
    """Explicit-sum Wigner small-d d^j_{m1 m2}(b) in mpmath (textbook formula)."""
    s = mp.mpf(0)
    for k in range(0, 2 * j + 2):
        a1, a2, a3, a4 = j + m2 - k, k, m1 - m2 + k, j - m1 - k
        if min(a1, a2, a3, a4) < 0:
            continue
        s += ((-1) ** (m1 - m2 + k) * mp.sqrt(mp.factorial(j + m1) * mp.factorial(j - m1) * mp.factorial(j + m2)
                                               * mp.factorial(j - m2))
              / (mp.factorial(a1) * mp.factorial(a2) * mp.factorial(a3) * mp.factorial(a4))
              * mp.cos(b / 2) ** (2 * j + m2 - m1 - 2 * k) * mp.sin(b / 2) ** (m1 - m2 + 2 * k))
    return s


def _eth_sY(l, m, s, t):
    """sY_lm(t, 0) from its definition by the eth operator on mpmath's Y_lm (s = 0, 1, 2)."""
    mp.mp.dps = 30
    Y = lambda a, b: mp.spherharm(l, m, a, b)
    def eth0(f, a, b):
        return -(mp.diff(lambda x: f(x, b), a) + 1j / mp.sin(a) * mp.diff(lambda y: f(a, y), b))
    def eth1(f, a, b):
        return -(mp.diff(lambda x: f(x, b), a) + 1j / mp.sin(a) * mp.diff(lambda y: f(a, y), b)
                 - mp.cot(a) * f(a, b))
    t = mp.mpf(t)
    if s == 0:
        return Y(t, 0)
    if s == 1:
        return eth0(Y, t, 0) / mp.sqrt(l * (l + 1))
    g = lambda a, b: eth0(Y, a, b)
    return eth1(g, t, 0) / mp.sqrt((l - 1) * l * (l + 1) * (l + 2))


def test_known_harmonics():
    th = np.array([0.0, 0.4, 1.3, np.pi])
    lam = sht.spin_lambda(3, 0, 0, th)
    np.testing.assert_allclose(lam[0], 1 / math.sqrt(4 * math.pi), rtol=1e-15)
    np.testing.assert_allclose(lam[1], math.sqrt(3 / (4 * math.pi)) * np.cos(th), rtol=1e-14, atol=1e-16)
    lam1 = sht.spin_lambda(3, 0, 1, th)
    np.testing.assert_allclose(lam1[1], -math.sqrt(3 / (8 * math.pi)) * np.sin(th), atol=1e-15)


@pytest.mark.parametrize("spin", [0, 1, 2])
def test_spin_lambda_matches_eth_definition(spin):
    """sY_lm from the eth operator applied to mpmath's Y_lm (pins sign and normalisation per spin)."""
    for (l, m) in [(3, 0), (4, 1), (5, 3), (4, -2), (6, -5)]:
        if l < spin:
            continue
        t = 0.83
        ref = complex(_eth_sY(l, m, spin, t))
        got = sht.spin_lambda(l, spin, m, np.array([t]))[l, 0]
        assert abs(ref.imag) < 1e-20
        assert abs(ref.real - got) < 1e-13 * max(1.0, abs(got)), (l, m, spin, ref, got)


def test_spin_lambda_vs_wigner_sum_incl_poles():
    th = np.array([0.0, 0.3, np.pi / 2, 2.9, np.pi])
    L = 8
    mp.mp.dps = 30
    err = 0.0
    for s in (0, 1, 2, -2):
        for m in range(-L, L + 1):
            lam = sht.spin_lambda(L, s, m, th)
            for l in range(max(abs(m), abs(s)), L + 1):
                for i, t in enumerate(th):
                    ref = (-1) ** s * mp.sqrt((2 * l + 1) / (4 * mp.pi)) * _wigner_mp(l, m, -s, mp.mpf(t))
                    err = max(err, abs(float(ref) - lam[l, i]))
    assert err < 1e-13


def test_high_m_underflow_is_handled():
    """At lmax 2000, m near lmax, rings near the pole underflow: values must be 0 or finite, never NaN,
    and the equator value must match the closed-form sectoral magnitude."""
    L = 2000
    th = np.array([1e-3, 0.05, np.pi / 2])
    lam = sht.spin_lambda(L, 0, L, th)
    assert np.all(np.isfinite(lam))
    # |Y_LL(pi/2)| = sqrt((2L+1)!!/(4 pi (2L)!!)) -> log form
    logv = 0.5 * (math.lgamma(2 * L + 2) - math.lgamma(L + 1) - math.lgamma(L + 1.5) - math.log(4 * math.pi)
                  - 2 * L * math.log(2) + 0.5 * math.log(math.pi) - math.log(2) + math.log(2))
    ref = math.sqrt((2 * L + 1) / (4 * math.pi)) * math.exp(0.5 * math.lgamma(2 * L + 1) - math.lgamma(L + 1) - L * math.log(2))
    assert abs(abs(lam[L, 2]) - ref) < 1e-12 * ref
    assert lam[L, 0] == 0.0


@pytest.mark.parametrize("spin", [0, 1, 2])
def test_c_rings_match_numpy(spin):
    rng = np.random.default_rng(10 + spin)
    L = 14
    th = np.concatenate([[0.0], np.sort(rng.uniform(0, np.pi, 7)), [np.pi]])
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    Gc = clib.sht_synth_rings(a, L, spin, th)
    Gn = sht.synth_ring_coeffs(a[0] if spin == 0 else a, L, spin, th)
    if spin == 0:
        Gn = Gn[L:]
    assert np.max(np.abs(Gc - Gn)) < 1e-13


@pytest.mark.parametrize("spin", [0, 2])
def test_c_ring_adjoint_dot(spin):
    rng = np.random.default_rng(20 + spin)
    L = 17
    th = np.sort(rng.uniform(0, np.pi, 11))
    nc = 1 if spin == 0 else 2
    a = almutil.random_alm(L, spin, nc, rng)
    G = clib.sht_synth_rings(a, L, spin, th)
    Gp = rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape)
    if spin == 0:
        lhs = np.sum((np.conj(Gp[0]) * G[0]).real) + 2 * np.sum((np.conj(Gp[1:]) * G[1:]).real)
    else:
        lhs = np.sum((np.conj(Gp) * G).real)
    ad = clib.sht_adjoint_rings(Gp, L, spin, nc, th)
    rhs = sum(almutil.myalmdot(a[c], ad[c], L) for c in range(nc))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_gl_nodes():
    for n in (1, 2, 7, 64, 301):
        th, w = geometry.gl_nodes(n)
        x, wr = np.polynomial.legendre.leggauss(n)
        np.testing.assert_allclose(np.sort(np.cos(th)), np.sort(x), atol=2e-15)
        # numpy's leggauss weights are themselves only ~1e-10 accurate near the poles
        np.testing.assert_allclose(np.sort(w), np.sort(wr), rtol=1e-9)
    # arbitrate with mpmath: Newton on P_n(cos theta) in 40 digits, w = 2 sin^2 / (n P_{n-1})^2
    mp.mp.dps = 40
    n = 301
    th, w = geometry.gl_nodes(n)
    for j in (0, 1, 150):
        t = mp.mpf(th[j])
        for _ in range(8):
            t = t - mp.legendre(n, mp.cos(t)) / mp.diff(lambda s: mp.legendre(n, mp.cos(s)), t)
        wt = 2 * mp.sin(t) ** 2 / (n * mp.legendre(n - 1, mp.cos(t))) ** 2
        assert abs(float(t) - th[j]) < 1e-17 + 1e-15 * th[j]
        assert abs(float(wt) - w[j]) < 1e-13 * w[j]
    th, w = geometry.gl_nodes(1024)
    assert abs(w.sum() - 2) < 1e-13
    x = np.cos(th)
    for deg in (0, 2, 10, 2047):
        exact = 2.0 / (deg + 1) if deg % 2 == 0 else 0.0
        assert abs(np.sum(w * x ** deg) - exact) < 1e-12


def test_good_size_matches_reference_logs():
    # first_steps/quickstart.ipynb:209 -> CC (ntheta, nphi) = (1584, 3168) at lmax 1575
    assert geometry.reference_cc_sizes(1575) == (1584, 3168)
    # quickstart.ipynb:239 -> GL nphi 3200 at lmax 1575
    assert geometry.good_size(2 * 1575 + 1, True) == 3200
    assert geometry.reference_cc_sizes(4095) == (4116, 8192)


def test_nufft_vs_brute_force():
    rng = np.random.default_rng(3)
    F = rng.standard_normal((18, 10)) + 1j * rng.standard_normal((18, 10))
    x1 = rng.uniform(-1, 7, 60)
    x2 = rng.uniform(-2 * np.pi, 4 * np.pi, 60)
    c = rng.standard_normal(60) + 1j * rng.standard_normal(60)
    for eps in (1e-5, 1e-9, 1e-12):
        t2 = nufft.nufft2d2(F, x1, x2, eps)
        b2 = nufft.nudft2d2(F, x1, x2)
        assert np.linalg.norm(t2 - b2) <= eps * np.linalg.norm(b2)
        t1 = nufft.nufft2d1(c, x1, x2, F.shape, eps)
        b1 = nufft.nudft2d1(c, x1, x2, F.shape)
        assert np.linalg.norm(t1 - b1) <= eps * np.linalg.norm(b1)
    # adjointness of the pair
    lhs = np.vdot(c, nufft.nufft2d2(F, x1, x2, 1e-12))
    rhs = np.vdot(nufft.nufft2d1(c, x1, x2, F.shape, 1e-12), F)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_pointing_and_gamma_vs_3d_geometry():
    rng = np.random.default_rng(4)
    t = rng.uniform(0.01, np.pi - 0.01, 2000)
    p = rng.uniform(0, 2 * np.pi, 2000)
    for sig in (1e-4, 3e-3, 0.2):
        dt = rng.normal(0, sig, t.size)
        dp = rng.normal(0, sig, t.size)
        tp, pp = pointing.deflected_angles(t, p, dt, dp)
        g = pointing.gamma(t, dt, dp)
        t3, p3, g3 = pointing.rotation_3d(t, p, dt, dp)
        assert np.max(np.abs(tp - t3)) < 1e-14
        assert np.max(np.abs((pp - p3 + np.pi) % (2 * np.pi) - np.pi)) < 1e-13
        assert np.max(np.abs((g - g3 + np.pi) % (2 * np.pi) - np.pi)) < 1e-13
    # special cases: identity, pure theta deflection
    tp, pp = pointing.deflected_angles(t, p, 0 * t, 0 * t)
    np.testing.assert_allclose(tp, t, atol=1e-15)
    np.testing.assert_allclose(pp, p, atol=1e-15)
    d = 1e-3
    m = (t > 0.1) & (t < np.pi - 0.1)
    tp, pp = pointing.deflected_angles(t[m], p[m], d + 0 * t[m], 0 * t[m])
    np.testing.assert_allclose(tp, t[m] + d, atol=1e-14)


@pytest.mark.parametrize("spin", [0, 2])
def test_pipeline_vs_direct_and_adjoint(spin):
    rng = np.random.default_rng(5 + spin)
    L = 16
    rings = geometry.gl_geometry(L, nphi=2 * L + 1)
    ll = np.arange(L + 1)
    dlm = almutil.synalm(np.where(ll >= 2, 1e-4 / (ll + 1.0) ** 2, 0), L, rng)
    nc = 1 if spin == 0 else 2
    g = almutil.random_alm(L, spin, nc, rng)
    gin = g[0] if spin == 0 else g
    out = pipeline.gclm2lenmap(gin, dlm, L, spin, rings, 1e-12)
    tp, pp, gam, _ = pipeline.dlm2pointing(dlm, L, rings)
    ref = sht.synth_direct(gin, L, spin, tp, pp)
    if spin:
        ref = ref * np.exp(1j * spin * gam)
        ref = np.array([ref.real, ref.imag])
    assert np.max(np.abs(out - ref)) < 1e-11 * np.max(np.abs(ref))
    x = rng.standard_normal(out.shape)
    ad = pipeline.lenmap2gclm(x, dlm, L, spin, nc, rings, 1e-12)
    lhs = np.sum(x * out)
    rhs = sum(almutil.myalmdot(g[c], ad[c], L) for c in range(nc))
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)  # test/test_adjointOperator.py:268 rtol


def _ambient_alm(spin, Tn, L, rings):
    """E/B of a constant-ambient-tensor field by exact quadrature analysis on the GL rings."""
    th, ph = rings.pixel_angles()
    w = np.repeat(rings.weight, rings.nph)
    return sht.adjoint_direct(transport.ambient_field(spin, Tn, th, ph) * w, L, spin, th, ph)


@pytest.mark.parametrize("spin", [1, 2])
def test_ambient_tensor_field_is_spin_s_at_l_s(spin):
    """v . m and m^T T m / 2 with m = e_theta + i e_phi are spin +s fields with power at l = s only (the
    conjugate choice is spin -s and is not reproduced): fixes which field the transport check lenses."""
    rng = np.random.default_rng(40 + spin)
    L = 8
    rings = geometry.gl_geometry(L)
    Tn = transport.random_tensor(spin, rng)
    EB = _ambient_alm(spin, Tn, L, rings)
    ll = np.concatenate([np.arange(m, L + 1) for m in range(L + 1)])
    assert np.max(np.abs(EB[:, ll != spin])) < 1e-14 * np.max(np.abs(EB))
    t = np.arccos(rng.uniform(-1, 1, 40))
    p = rng.uniform(0, 2 * np.pi, 40)
    f = transport.ambient_field(spin, Tn, t, p)
    assert np.max(np.abs(sht.synth_direct(EB, L, spin, t, p) - f)) < 1e-13 * np.max(np.abs(f))
    # the conjugate projection is not a spin +s field
    th, ph = rings.pixel_angles()
    w = np.repeat(rings.weight, rings.nph)
    fbar = np.conj(transport.ambient_field(spin, Tn, th, ph))
    EBbar = sht.adjoint_direct(fbar * w, L, spin, th, ph)
    assert np.max(np.abs(sht.synth_direct(EBbar, L, spin, t, p) - np.conj(f))) > 0.1 * np.max(np.abs(f))


@pytest.mark.parametrize("spin", [1, 2])
@pytest.mark.parametrize("amp", [1e-3, 3e-2])
def test_rotation_direction_by_parallel_transport(spin, amp):
    """pipeline.gclm2lenmap (pointing, gamma and e^{+i s gamma}, CPU_nufft_transformer.py:158-160,356-368)
    equals the exact lensing of a constant-ambient-tensor field by parallel transport along each pixel's
    deflection geodesic; the opposite rotation (e^{-i s gamma}) and no rotation both fail."""
    rng = np.random.default_rng(50 + 10 * spin + int(amp * 1e4))
    L = 12
    rings = geometry.gl_geometry(L)
    Tn = transport.random_tensor(spin, rng)
    EB = _ambient_alm(spin, Tn, L, rings)
    ll = np.arange(L + 1)
    dlm = almutil.synalm(np.where(ll >= 2, amp / (ll + 1.0) ** 2, 0.0), L, rng)
    th, ph = rings.pixel_angles()
    _, _, gam, (dt, dp) = pipeline.dlm2pointing(dlm, L, rings)
    exact = transport.lensed_ambient_field(spin, Tn, th, ph, dt, dp)
    scale = np.max(np.abs(exact))
    QU = pipeline.gclm2lenmap(EB, dlm, L, spin, rings, 1e-12)
    assert np.max(np.abs(QU[0] + 1j * QU[1] - exact)) < 1e-11 * scale
    unrot = pipeline.gclm2lenmap(EB, dlm, L, spin, rings, 1e-12, polrot=False)
    unrot = unrot[0] + 1j * unrot[1]
    flipped = unrot * np.exp(-1j * spin * gam)  # a planted sign mistake in the rotation
    tiny = 1e-3 * np.max(np.abs(gam)) * scale
    assert np.max(np.abs(flipped - exact)) > tiny
    assert np.max(np.abs(unrot - exact)) > tiny


# ------------------------------------------------------------------ magnification (dlm2A)
def _lens_case(amp, curl, seed=3, L=12):
    from oracle import almutil
    rng = np.random.default_rng(seed)
    ll = np.arange(L + 1)
    cl = np.where(ll >= 1, 0.02 / (ll + 1.0) ** 2, 0.0)
    dlm = almutil.synalm(cl, L, rng) * amp
    dclm = almutil.synalm(cl, L, rng) * amp if curl else None
    th = np.arccos(rng.uniform(-0.9, 0.9, 30))
    ph = rng.uniform(0, 2 * np.pi, 30)
    return dlm, dclm, L, th, ph


@pytest.mark.parametrize("curl", [False, True])
@pytest.mark.parametrize("amp", [0.01, 0.3, 1.0])
def test_magnification_is_solid_angle_jacobian(amp, curl):
    """exact=True equals dOmega'/dOmega of the deflection map (finite differences of the deflected
    3-vector; |d| up to 0.26 rad) within the differencing error."""
    from oracle import lensing
    dlm, dclm, L, th, ph = _lens_case(amp, curl)
    A = lensing.magnification(dlm, L, th, ph, dclm, exact=True)
    R = lensing.solid_angle_ratio(dlm, L, th, ph, dclm, h=1e-3)
    assert np.max(np.abs(A - R)) < 2e-9


def test_magnification_reference_sign_is_second_order():
    """The reference's shear sign (exact=False) agrees with the Jacobian through O(|d|^2): its error
    grows as the cube of the deflection amplitude."""
    from oracle import lensing
    e = []
    for amp in (0.01, 0.03):
        dlm, dclm, L, th, ph = _lens_case(amp, True)
        A = lensing.magnification(dlm, L, th, ph, dclm, exact=False)
        R = lensing.solid_angle_ratio(dlm, L, th, ph, dclm, h=1e-3)
        e.append(np.max(np.abs(A - R)))
    assert 1e-9 < e[0] < 1e-7
    assert 24 < e[1] / e[0] < 30


def test_magnification_zero_deflection():
    from oracle import lensing
    z = np.zeros(5)
    assert np.array_equal(lensing.combine(z, z, z, z, z, z), np.ones(5))
