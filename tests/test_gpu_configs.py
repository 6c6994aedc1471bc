"""The five benchmark configurations of this workload as GPU tests, at their full sizes.

Full-size results are checked through properties that do not need a full-size oracle run: the value at a random
sample of points against direct FP64 Y_lm summation there (the oracle's C Legendre restatement evaluated on the
sampled colatitudes, then the sum over m), and the dot-product adjointness of each forward / adjoint pair.

* lmax 255, spin 0, GL 256 x 511, eps 1e-10: the lensed map at 2000 pixels within 10 eps of the max.
* lmax 2047, 8.4M random points, spin 0, eps 1e-5 / 1e-7 / 1e-10: nusht2d2 at 300 points within 10 eps (30 eps
  in FP32) of the max; <x, A y> = <A^H x, y> to 1e-9 (FP64 nuFFT) / 1e-5 (FP32 nuFFT).
* lmax 4095, spin 2, GL 4096 x 8191: eps 1e-6 (FP32 FFT and nuFFT) and the headline eps 1e-7 (FP64):
  gclm2lenmap at 200 pixels against direct summation at the oracle's own pointing of those pixels (the
  deflection summed directly there, then the closed-form pointing and gamma), within 30 eps in FP32 and 1e-7
  (the headline bar) in FP64; the gclm2lenmap / lenmap2gclm pair adjoint to 1e-5 (FP32) / 1e-9 (FP64).
* 64 transforms of lmax {511, 1023, 1535, 2047} x spin {0, 2}, eps 1e-7, each with its own gclm and deflection,
  as one hipGraph-captured batch replay (the Legendre syntheses of each (lmax, spin) fused onto the matrix
  cores): every job equal to its single call (1e-13), 200 pixels of four jobs against direct summation at the
  oracle's pointing.
* The serving loop on a Poisson mix of lmax 255 .. 4095, spin 0 / 2, eps 1e-5 / 1e-7 / 1e-10, forward and
  adjoint: every result equal to the same request run alone (1e-13 forward / 1e-12 adjoint in FP64, eps with
  the FP32 nuFFT), and one forward request per (lmax, spin) against direct summation at 100 pixels.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import almutil, clib, pointing as opointing

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


def _pointing_at(dlm_scaled, lmax, theta, phi):
    """The oracle's pointing and gamma at given pixels: the deflection (spin-1 gradient of
    sum dlm_scaled Y_lm) summed directly there, then the closed-form pointing (pointing.cu:205-223) and gamma
    (CPU_nufft_transformer.py:158-160)."""
    ll = np.arange(lmax + 1, dtype=np.float64)
    dlm = almutil.almxfl(dlm_scaled, np.sqrt(ll * (ll + 1.0)), lmax)
    d = _direct(dlm[None, :], lmax, 1, theta, phi)
    tp, pp = opointing.deflected_angles(theta, phi, d.real, d.imag)
    return tp, pp, opointing.gamma(theta, d.real, d.imag)


def _lensed_at(gclm, dlm_scaled, lmax, spin, theta, phi):
    """Exact lensed values at given pointing pixels: direct summation at the oracle's deflected positions,
    rotated by e^{i s gamma} (spin > 0); spin 0 real, spin s complex Q + iU."""
    tp, pp, gam = _pointing_at(dlm_scaled, lmax, theta, phi)
    v = _direct(np.atleast_2d(gclm), lmax, spin, tp, pp)
    return v if spin == 0 else v * np.exp(1j * spin * gam)


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


@pytest.mark.parametrize("eps", [1e-6, 1e-7])
def test_config_full_chain_spin2_lmax4095(eps):
    import cunusht_amd
    L = 4095
    rng = np.random.default_rng(4095)
    ll = np.arange(L + 1, dtype=np.float64)
    cle = np.zeros(L + 1)
    cle[2:] = np.exp(-(ll[2:] / 3000.0) ** 2) / (ll[2:] * (ll[2:] + 1))
    g = np.array([almutil.synalm(cle, L, rng), almutil.synalm(0.1 * cle, L, rng)])
    dlm = _deflection(L, rng)
    print(f"case config lmax 4095 spin 2 eps {eps:g}", flush=True)
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L, "nphi": 2 * L + 1}), nuFFTtype=2, epsilon=eps,
                                           spin=2)
    gd, dd = _t(g), _t(dlm)
    m = T.gclm2lenmap(gd, L, L, dlm_scaled=dd, spin=2).cpu().numpy()
    assert m.shape == (2, 4096 * 8191)
    sel = np.sort(rng.choice(m.shape[1], 200, replace=False))
    th, ph = T.geom.pixel_angles()
    ref = _lensed_at(g, dlm, L, 2, th[sel], ph[sel])
    scale = np.max(np.abs(m))
    fp32 = eps >= 1e-6
    err = _report(f"lmax 4095 spin 2 eps {eps:g} ({'FP32' if fp32 else 'FP64'} nuFFT)", m[:, sel],
                  np.array([ref.real, ref.imag]), scale)
    assert err < (30 * eps if fp32 else 1e-7)
    x = rng.standard_normal(m.shape)
    a = T.lenmap2gclm(_t(x), dd, None, L, L, spin=2).cpu().numpy()
    lhs = float(np.sum(x * m))
    rhs = sum(almutil.myalmdot(g[c], a[c], L) for c in range(2))
    print(f"lmax 4095 spin 2 eps {eps:g} adjointness: |lhs - rhs| / |lhs| = {abs(lhs - rhs) / abs(lhs):.2e}",
          flush=True)
    assert abs(lhs - rhs) < (1e-5 if fp32 else 1e-9) * abs(lhs)


def _single_call(plan, kind, inp, dlm):
    """The single-call entry point on a plan, on the current stream."""
    from cunusht_amd import _lib
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: ctypes.c_void_p(0 if t is None else t.data_ptr())
    if kind == "gclm2lenmap":
        out = torch.empty((2, plan.npix()), dtype=torch.float64, device="cuda")
        _lib.call("nusht_gclm2lenmap", plan.nuplan._h, plan.ring._h, p(inp), p(dlm), p(out[0]),
                  p(out[1]) if plan.spin else None, 1, s)
        return out[0] if plan.spin == 0 else out
    m = inp.reshape(1 if plan.spin == 0 else 2, plan.npix())
    res = torch.empty((plan.ncomp, plan.nalm()), dtype=torch.complex128, device="cuda")
    _lib.call("nusht_lenmap2gclm", plan.nuplan._h, plan.ring._h, p(m[0]), p(m[1]) if plan.spin else None, p(dlm),
              p(res), 1, s)
    return res


def test_config_uneven_batch_64_jobs_graph():
    """64 independent transforms of uneven size as one captured hipGraph pass (the uneven-batch workload)."""
    from cunusht_amd.batch import BatchJob, CapturedBatch, LensingPlan
    rng = np.random.default_rng(44)
    lmaxs = rng.choice([511, 1023, 1535, 2047], 64)
    spins = rng.choice([0, 2], 64)
    plans, jobs, host = {}, [], []
    for i, (L, s) in enumerate(zip(lmaxs, spins)):
        L, s = int(L), int(s)
        pl = plans.setdefault((L, s), [LensingPlan(L, s, 1e-7, max_maps=4), LensingPlan(L, s, 1e-7, max_maps=4)])
        nc = 1 if s == 0 else 2
        cl = 1.0 / (np.arange(L + 1) + 1.0) ** 2
        g = np.stack([almutil.synalm(cl, L, rng) for _ in range(nc)])
        d = _deflection(L, rng)
        jobs.append(BatchJob("gclm2lenmap", pl[i % 2], _t(d), _t(g)))
        host.append((g, d))
    print(f"case uneven batch: 64 jobs over {len(plans)} (lmax, spin) pairs", flush=True)
    cap = CapturedBatch(jobs)
    for j in jobs:
        j.out.fill_(np.nan)
    cap.replay()
    torch.cuda.synchronize()
    got = [j.out.cpu().numpy() for j in jobs]
    for k, j in enumerate(jobs):
        ref = _single_call(j.plan, "gclm2lenmap", j.inp, j.dlm_scaled).cpu().numpy()
        torch.cuda.synchronize()
        assert np.max(np.abs(got[k] - ref)) <= 1e-13 * np.max(np.abs(ref)), k
    checked = set()
    for k, j in enumerate(jobs):
        key = (j.plan.lmax, j.plan.spin)
        if key in checked or len(checked) >= 4:
            continue
        checked.add(key)
        L, s = key
        g, d = host[k]
        th, ph = j.plan.geom.pixel_angles()
        sel = np.sort(rng.choice(th.size, 200, replace=False))
        ref = _lensed_at(g, d, L, s, th[sel], ph[sel])
        o = got[k]
        gv = o[sel] if s == 0 else o[0, sel] + 1j * o[1, sel]
        err = _report(f"uneven batch job {k} lmax {L} spin {s} eps 1e-7", gv, ref, np.max(np.abs(o)))
        assert err < 1e-7


def test_config_served_mix():
    """The mixed serving configuration (lmax 255 .. 4095, spin 0 / 2, eps 1e-5 / 1e-7 / 1e-10, forward and
    adjoint, Poisson arrivals), every result equal to the request alone on the same plan."""
    from cunusht_amd.serving import LensingServer, Request, poisson_arrivals, serve
    rng = np.random.default_rng(55)
    combos = [(L, s, e, k) for L in (255, 511, 1023, 2047, 4095) for s in (0, 2) for e in (1e-5, 1e-7, 1e-10)
              for k in ("gclm2lenmap", "lenmap2gclm")]
    # every lmax, spin, eps and kind at least once, lmax 4095 in both directions, then random draws
    pick = [(4095, 2, 1e-7, "gclm2lenmap"), (4095, 2, 1e-7, "lenmap2gclm"), (4095, 0, 1e-10, "gclm2lenmap"),
            (4095, 0, 1e-5, "lenmap2gclm"), (255, 2, 1e-5, "gclm2lenmap"), (511, 0, 1e-7, "lenmap2gclm"),
            (1023, 2, 1e-10, "lenmap2gclm"), (2047, 0, 1e-5, "gclm2lenmap")]
    pick += [combos[i] for i in rng.choice(len(combos), 16, replace=False)]
    defl = {L: _deflection(L, rng) for L in (255, 511, 1023, 2047, 4095)}
    ddev = {L: _t(d) for L, d in defl.items()}
    reqs, host = [], []
    for L, s, e, kind in pick:
        nc = 1 if s == 0 else 2
        if kind == "gclm2lenmap":
            cl = 1.0 / (np.arange(L + 1) + 1.0) ** 2
            x = np.stack([almutil.synalm(cl, L, rng) for _ in range(nc)])
        else:
            x = rng.standard_normal((nc, (L + 1) * (2 * L + 2)))
        reqs.append(Request(kind, L, s, e, _t(x if kind == "gclm2lenmap" or s else x[0]), ddev[L]))
        host.append(x)
    print(f"case served mix: {len(reqs)} requests", flush=True)
    server = LensingServer(hbm_budget_bytes=200 << 30, plans_per_key=1, max_batch=8)
    stats = serve(server, reqs, poisson_arrivals(len(reqs), 200.0, seed=5))
    assert stats["requests"] == len(reqs)
    print(f"served mix: p50 {stats['p50_ms']:.1f} ms, p99 {stats['p99_ms']:.1f} ms, {stats['plans']} plans", flush=True)
    direct_done = set()
    for r, x in zip(reqs, host):
        plan = server.slots[r.key][0].plan
        alone = _single_call(plan, r.kind, r.data, r.dlm_scaled).cpu().numpy()
        torch.cuda.synchronize()
        got = r.result.cpu().numpy().reshape(alone.shape)
        fp32 = r.epsilon >= 1e-6
        # FP32 nuFFT: the spreading's global flush adds FP32 values in a varying order, so results move at
        # the FP32 rounding level (1.9e-6 of the max for an lmax-4095 adjoint at eps 1e-5): within eps
        tol = max(1e-6, r.epsilon) if fp32 else (1e-13 if r.kind == "gclm2lenmap" else 1e-12)
        assert np.max(np.abs(got - alone)) <= tol * np.max(np.abs(alone)), (r.kind, r.lmax, r.spin, r.epsilon)
        if r.kind == "gclm2lenmap" and (r.lmax, r.spin) not in direct_done:
            direct_done.add((r.lmax, r.spin))
            th, ph = plan.geom.pixel_angles()
            sel = np.sort(rng.choice(th.size, 100, replace=False))
            ref = _lensed_at(x, defl[r.lmax], r.lmax, r.spin, th[sel], ph[sel])
            gv = got[sel] if r.spin == 0 else got[0, sel] + 1j * got[1, sel]
            err = _report(f"served lmax {r.lmax} spin {r.spin} eps {r.epsilon:g}", gv, ref, np.max(np.abs(got)))
            assert err < (30 * r.epsilon if fp32 else max(10 * r.epsilon, 1e-12))
