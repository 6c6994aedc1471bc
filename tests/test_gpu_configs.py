"""The five benchmark configurations of this workload as GPU tests, at their full sizes.

Full-size results are checked through properties that do not need a full-size oracle run: the value at a random
sample of points against direct FP64 Y_lm summation there (the oracle's C Legendre restatement evaluated on the
sampled colatitudes, then the sum over m), and the dot-product adjointness of each forward / adjoint pair.

* lmax 255, spin 0, GL 256 x 511, eps 1e-10: the lensed map at 2000 pixels within 10 eps of the max.
* lmax 2047, 8.4M random points, spin 0, eps 1e-5 / 1e-7 / 1e-10: nusht2d2 at 300 points within 10 eps (30 eps
  in FP32) of the max; <x, A y> = <A^H x, y> to 1e-9 (FP64 nuFFT) / 1e-5 (FP32 nuFFT).
* lmax 4095, spin 2, GL 4096 x 8191, eps 1e-6 (FP32 FFT and nuFFT): gclm2lenmap at 200 pixels within 30 eps, the
  gclm2lenmap / lenmap2gclm pair adjoint to 1e-5.
* 64 transforms of lmax {511, ..., 2047} x spin {0, 2} and the mixed serving loop are covered by
  tests/test_gpu_serving.py at reduced sizes; bench.py --uneven-batch 64 runs the full one.
"""
import numpy as np
import pytest
import torch

from oracle import almutil, clib

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _direct(alm, lmax, spin, theta, phi):
    """Direct FP64 synthesis at points: the oracle's Legendre restatement on the points' colatitudes, then the
    sum over m.  spin 0 -> real values, spin s -> Q + iU."""
    G = clib.sht_synth_rings(alm, lmax, spin, theta)
    if spin == 0:
        e = np.exp(1j * np.outer(np.arange(lmax + 1), phi))
        e[1:] *= 2.0
        return np.einsum("mp,mp->p", G, e).real
    e = np.exp(1j * np.outer(np.arange(-lmax, lmax + 1), phi))
    return np.einsum("mp,mp->p", G, e)


def _report(name, got, ref, scale):
    d = np.asarray(got) - np.asarray(ref)
    mx = float(np.max(np.abs(d)) / scale)
    rms = float(np.sqrt(np.mean(np.abs(d) ** 2)) / np.sqrt(np.mean(np.abs(np.asarray(ref)) ** 2)))
    print(f"{name}: max error / max |map| = {mx:.2e}, rms error / rms = {rms:.2e}", flush=True)
    return mx


def _deflection(lmax, rng):
    """dlm_scaled of a deflection with C_l^dd ~ l^-2 (2 <= l <= lmax), 2.5 arcmin rms."""
    ll = np.arange(lmax + 1, dtype=np.float64)
    cl = np.zeros(lmax + 1)
    cl[2:] = ll[2:] ** -2.0
    rms = np.sqrt(np.sum((2 * ll + 1) * cl) / (4 * np.pi))
    cl *= (2.5 / 60.0 * np.pi / 180.0 / rms) ** 2
    dlm = almutil.synalm(cl, lmax, rng)
    with np.errstate(divide="ignore"):
        fl = np.where(ll > 0, 1.0 / np.sqrt(ll * (ll + 1.0)), 0.0)
    return almutil.almxfl(dlm, fl, lmax)


def test_config_plain_spin0_lmax255():
    import cunusht_amd
    L, eps = 255, 1e-10
    rng = np.random.default_rng(2550)
    ll = np.arange(L + 1)
    g = almutil.synalm(np.where(ll >= 2, 1.0 / (ll + 1.0) ** 2, 0.0), L, rng)
    dlm = _deflection(L, rng)
    print("case config lmax 255 spin 0", flush=True)
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L, "nphi": 2 * L + 1}), nuFFTtype=2, epsilon=eps)
    m = T.gclm2lenmap(_t(g), L, L, dlm_scaled=_t(dlm)).cpu().numpy()
    assert m.shape == (256 * 511,)
    ptg = T.ptg.cpu().numpy()
    sel = rng.choice(m.size, 2000, replace=False)
    ref = _direct(g[None, :], L, 0, ptg[0, sel], ptg[1, sel])
    assert _report("lmax 255 spin 0 eps 1e-10", m[sel], ref, np.max(np.abs(m))) < 10 * eps


@pytest.mark.parametrize("eps", [1e-5, 1e-7, 1e-10])
def test_config_adjoint_pair_lmax2047(eps):
    import cunusht_amd
    L = 2047
    npts = 2048 * 4095
    rng = np.random.default_rng(2047)
    theta = np.arccos(rng.uniform(-1, 1, npts))
    phi = rng.uniform(0, 2 * np.pi, npts)
    ll = np.arange(L + 1)
    a = almutil.synalm(1.0 / (ll + 1.0) ** 2, L, rng)
    x = rng.standard_normal(npts)
    print(f"case config lmax 2047 adjoint pair eps={eps}", flush=True)
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=2, epsilon=eps)
    loc = _t(np.stack([theta, phi], 1))
    y = T.nusht2d2(L, L, _t(a), loc).cpu().numpy()
    ad = T.nusht2d1(L, L, _t(x), loc).cpu().numpy()
    lhs = float(np.dot(x, y))
    rhs = almutil.myalmdot(a, ad, L)
    assert abs(lhs - rhs) < (1e-9 if eps < 1e-6 else 1e-5) * abs(lhs)
    sel = rng.choice(npts, 300, replace=False)
    ref = _direct(a[None, :], L, 0, theta[sel], phi[sel])
    tol = 10 * eps if eps < 1e-6 else 30 * eps
    assert _report(f"lmax 2047 nusht2d2 eps {eps:g}", y[sel], ref, np.max(np.abs(y))) < tol
    print(f"lmax 2047 adjointness eps {eps:g}: |lhs - rhs| / |lhs| = {abs(lhs - rhs) / abs(lhs):.2e}", flush=True)


def test_config_full_chain_spin2_lmax4095_fp32():
    import cunusht_amd
    L, eps = 4095, 1e-6
    rng = np.random.default_rng(4095)
    ll = np.arange(L + 1, dtype=np.float64)
    cle = np.zeros(L + 1)
    cle[2:] = np.exp(-(ll[2:] / 3000.0) ** 2) / (ll[2:] * (ll[2:] + 1))
    g = np.array([almutil.synalm(cle, L, rng), almutil.synalm(0.1 * cle, L, rng)])
    dlm = _deflection(L, rng)
    print("case config lmax 4095 spin 2 eps 1e-6", flush=True)
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L, "nphi": 2 * L + 1}), nuFFTtype=2, epsilon=eps,
                                           spin=2)
    gd, dd = _t(g), _t(dlm)
    m = T.gclm2lenmap(gd, L, L, dlm_scaled=dd, spin=2).cpu().numpy()
    assert m.shape == (2, 4096 * 8191)
    ptg = T.ptg.cpu().numpy()
    gam = T._gamma.cpu().numpy()
    sel = rng.choice(m.shape[1], 200, replace=False)
    ref = _direct(g, L, 2, ptg[0, sel], ptg[1, sel]) * np.exp(2j * gam[sel])
    scale = np.max(np.abs(m))
    assert _report("lmax 4095 spin 2 eps 1e-6 (FP32 nuFFT)", m[:, sel], np.array([ref.real, ref.imag]),
                   scale) < 30 * eps
    x = rng.standard_normal(m.shape)
    a = T.lenmap2gclm(_t(x), dd, None, L, L, spin=2).cpu().numpy()
    lhs = float(np.sum(x * m))
    rhs = sum(almutil.myalmdot(g[c], a[c], L) for c in range(2))
    print(f"lmax 4095 spin 2 adjointness: |lhs - rhs| / |lhs| = {abs(lhs - rhs) / abs(lhs):.2e}", flush=True)
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)
