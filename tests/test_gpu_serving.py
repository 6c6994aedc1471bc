"""GPU tests of the batch executor (nusht_batch_execute), its hipGraph capture and the serving loop.

Every result of a batch, a graph replay or a served mix is compared with the same transform run alone
through the single-call entry points.  FP64 nuFFT (epsilon < 1e-6): forward (interpolation) results agree to
1e-13 relative (the batched synthesis of maps sharing a deflection runs the matrix-core kernel, a different
summation order), adjoint results, whose spreading accumulates with atomics in a varying order, to 1e-12.
FP32 nuFFT (epsilon >= 1e-6): both to 1e-6, a few FP32 roundings of the varying accumulation order.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import almutil

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _relmax(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _deflection(L, rng, amp=1e-3):
    ll = np.arange(L + 1)
    return almutil.synalm(np.where(ll >= 2, amp / (ll + 1.0) ** 2, 0.0), L, rng)


def _alone(plan, kind, inp, dlm, polrot=True):
    """The single-call entry point on the same plan."""
    from cunusht_amd import _lib
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: ctypes.c_void_p(0 if t is None else t.data_ptr())
    if kind == "gclm2lenmap":
        out = torch.empty((2, plan.npix()), dtype=torch.float64, device="cuda")
        _lib.call("nusht_gclm2lenmap", plan.nuplan._h, plan.ring._h, p(inp), p(dlm), p(out[0]),
                  p(out[1]) if plan.spin else None, int(polrot), s)
        res = out[0] if plan.spin == 0 else out
    else:
        m = inp.reshape(1 if plan.spin == 0 else 2, plan.npix())
        res = torch.empty((plan.ncomp, plan.nalm()), dtype=torch.complex128, device="cuda")
        _lib.call("nusht_lenmap2gclm", plan.nuplan._h, plan.ring._h, p(m[0]), p(m[1]) if plan.spin else None,
                  p(dlm), p(res), int(polrot), s)
    torch.cuda.synchronize()
    return res.cpu().numpy()


def _jobs(rng):
    from cunusht_amd.batch import BatchJob, LensingPlan
    plans = [LensingPlan(24, 0, 1e-10), LensingPlan(32, 2, 1e-7), LensingPlan(40, 2, 1e-10), LensingPlan(16, 0, 1e-5)]
    jobs = []
    for i, pl in enumerate(plans + plans[:2]):  # two plans carry two jobs each (ordered on one stream)
        dlm = _t(_deflection(pl.lmax, rng))
        if i % 2 == 0:
            g = almutil.random_alm(pl.lmax, pl.spin, pl.ncomp, rng)
            jobs.append(BatchJob("gclm2lenmap", pl, dlm, _t(g)))
        else:
            x = rng.standard_normal((1 if pl.spin == 0 else 2, pl.npix()))
            jobs.append(BatchJob("lenmap2gclm", pl, dlm, _t(x)))
    return jobs


def _tol_for(kind, epsilon):
    if epsilon >= 1e-6:
        return 1e-6
    return 1e-13 if kind == "gclm2lenmap" else 1e-12


def _tol(job):
    return _tol_for(job.kind, job.plan.epsilon)


def test_batch_matches_single_calls():
    from cunusht_amd.batch import execute
    rng = np.random.default_rng(1)
    jobs = _jobs(rng)
    execute(jobs)
    torch.cuda.synchronize()
    got = [j.out.cpu().numpy() for j in jobs]
    for j, g in zip(jobs, got):
        ref = _alone(j.plan, j.kind, j.inp, j.dlm_scaled)
        assert _relmax(g.reshape(ref.shape), ref) < _tol(j), (j.kind, j.plan.lmax, j.plan.spin)


def test_batch_rejects_bad_job_before_enqueueing():
    from cunusht_amd import _lib
    from cunusht_amd.batch import execute
    rng = np.random.default_rng(2)
    jobs = _jobs(rng)[:2]
    for j in jobs:
        j.out.fill_(7.0)
    bad = jobs[1].c_job()
    bad.dlm_scaled = None
    arr = (_lib.Job * 2)(jobs[0].c_job(), bad)
    with pytest.raises(_lib.NuShtError, match="batch job 1"):
        _lib.call("nusht_batch_execute", arr, 2, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    assert torch.all(jobs[0].out == 7.0)  # nothing ran


def test_captured_batch_replays_match_eager():
    """A hipGraph of the batch, replayed three times with new inputs written in place and eager work between
    the replays; every replay equals the eager batch on the same inputs."""
    from cunusht_amd.batch import CapturedBatch, execute
    rng = np.random.default_rng(3)
    jobs = _jobs(rng)
    cap = CapturedBatch(jobs)
    for rep in range(3):
        for j in jobs:
            if j.kind == "gclm2lenmap":
                j.inp.copy_(_t(almutil.random_alm(j.plan.lmax, j.plan.spin, j.plan.ncomp, rng)))
            else:
                j.inp.copy_(_t(rng.standard_normal(tuple(j.inp.shape))))
        torch.cuda.synchronize()
        cap.replay()
        torch.cuda.synchronize()
        replayed = [j.out.clone() for j in jobs]
        _ = torch.randn(1 << 20, device="cuda").sum()  # unrelated eager work between replays
        execute(jobs)
        torch.cuda.synchronize()
        for j, r in zip(jobs, replayed):
            assert _relmax(r.cpu().numpy(), j.out.cpu().numpy()) < _tol(j), (rep, j.kind, j.plan.lmax)


def test_server_results_same_alone_and_in_mix():
    """Mixed forward / adjoint requests of several configurations arriving as a Poisson stream; forward
    requests that share a deflection are batched.  Each result equals the request run alone."""
    import cunusht_amd
    from cunusht_amd.serving import LensingServer, Request, poisson_arrivals, serve
    rng = np.random.default_rng(4)
    configs = [(24, 0, 1e-10), (32, 2, 1e-7), (40, 2, 1e-10)]
    shared = {c: _t(_deflection(c[0], rng)) for c in configs}
    reqs = []
    for i in range(18):
        L, spin, eps = configs[i % 3]
        dlm = shared[configs[i % 3]] if i % 2 == 0 else _t(_deflection(L, rng))
        if i % 4 != 3:
            nc = 1 if spin == 0 else 2
            g = almutil.random_alm(L, spin, nc, rng)
            reqs.append(Request("gclm2lenmap", L, spin, eps, _t(g if spin else g[0]), dlm))
        else:
            # lenmaps on the server's default pointing grid (GL with lmax + 1 rings)
            npix = cunusht_amd.get_geom(("gl", {"lmax": L})).npix()
            x = rng.standard_normal(npix if spin == 0 else (2, npix))
            reqs.append(Request("lenmap2gclm", L, spin, eps, _t(x), dlm))
    server = LensingServer(hbm_budget_bytes=64 << 30, plans_per_key=2, max_batch=4)
    stats = serve(server, reqs, poisson_arrivals(len(reqs), 2000.0, seed=5))
    assert stats["requests"] == len(reqs)
    assert stats["p99_ms"] >= stats["p50_ms"] > 0
    for r in reqs:
        T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": r.lmax}), nuFFTtype=2, epsilon=r.epsilon,
                                               spin=r.spin)
        if r.kind == "gclm2lenmap":
            ref = T.gclm2lenmap(r.data, r.lmax, r.lmax, dlm_scaled=r.dlm_scaled, spin=r.spin).cpu().numpy()
        else:
            ref = T.lenmap2gclm(r.data, r.dlm_scaled, None, r.lmax, r.lmax, spin=r.spin).cpu().numpy()
        tol = _tol_for(r.kind, r.epsilon)
        got = r.result.cpu().numpy()
        assert _relmax(got.reshape(ref.shape), ref) < tol, (r.kind, r.lmax, r.spin)


@pytest.mark.parametrize("spin,eps", [(2, 1e-10), (0, 1e-7), (2, 1e-5)])
def test_fused_forward_batch_equals_single_calls(spin, eps):
    """Seven forward jobs of one (lmax, spin), each with its own deflection, on two plans that hold 4 maps:
    the executor runs their Legendre syntheses together (4 + 3 maps on the matrix cores) and each job's pointing
    and interpolation from its slice; every result equals the job's single call.  A captured replay too."""
    from cunusht_amd.batch import BatchJob, CapturedBatch, LensingPlan, execute
    L = 36
    rng = np.random.default_rng(970 + spin)
    plans = [LensingPlan(L, spin, eps, max_maps=4), LensingPlan(L, spin, eps, max_maps=4)]
    nc = 1 if spin == 0 else 2
    jobs = [BatchJob("gclm2lenmap", plans[k % 2], _t(_deflection(L, rng)),
                     _t(almutil.random_alm(L, spin, nc, rng))) for k in range(7)]
    print(f"case fused forward batch spin={spin} eps={eps}", flush=True)
    execute(jobs)
    torch.cuda.synchronize()
    got = [j.out.cpu().numpy() for j in jobs]
    tol = 1e-6 if eps >= 1e-6 else 1e-13
    for k, j in enumerate(jobs):
        ref = _alone(j.plan, "gclm2lenmap", j.inp, j.dlm_scaled)
        assert _relmax(got[k], ref) < tol, k
    cap = CapturedBatch(jobs)
    for j in jobs:
        j.out.fill_(np.nan)
    for _ in range(2):
        cap.replay()
        torch.cuda.synchronize()
        for k, j in enumerate(jobs):
            assert _relmax(j.out.cpu().numpy(), got[k]) < tol, k


def test_server_budget_counts_staging_and_outputs():
    """A slot's bytes include its batch staging and one launch's outputs: with a budget that holds the plans
    of two slots but not their staging, a configuration gets one slot, and every request still completes."""
    import cunusht_amd
    from cunusht_amd.batch import LensingPlan
    from cunusht_amd.serving import LensingServer, Request, serve
    from cunusht_amd.utils import Alm
    L, spin, eps = 40, 2, 1e-10
    plan_bytes = LensingPlan(L, spin, eps, max_maps=4).nbytes()
    nalm, npix = Alm.getsize(L, L), cunusht_amd.get_geom(("gl", {"lmax": L})).npix()
    slot_bytes = plan_bytes + 4 * 2 * nalm * 16 + max(4 * 2 * npix * 8, 2 * nalm * 16)
    rng = np.random.default_rng(61)
    reqs = [Request("gclm2lenmap", L, spin, eps, _t(almutil.random_alm(L, spin, 2, rng)), _t(_deflection(L, rng)))
            for _ in range(5)]
    server = LensingServer(hbm_budget_bytes=2 * plan_bytes + 1024, plans_per_key=2, max_batch=4)
    stats = serve(server, reqs, np.zeros(len(reqs)))
    assert 2 * plan_bytes + 1024 < 2 * slot_bytes
    assert stats["requests"] == 5 and stats["plans"] == 1
    assert server.used_bytes == slot_bytes
    for r in reqs:
        ref = _alone(server.slots[r.key][0].plan, "gclm2lenmap", r.data, r.dlm_scaled)
        assert _relmax(r.result.cpu().numpy(), ref) < 1e-12


def test_server_admission_respects_budget():
    """With a budget that holds one plan, two configurations alternate: idle plans are released."""
    from cunusht_amd.serving import LensingServer, Request, serve
    rng = np.random.default_rng(6)
    reqs = []
    for i in range(6):
        L = 16 if i % 2 == 0 else 20
        reqs.append(Request("gclm2lenmap", L, 0, 1e-10, _t(almutil.random_alm(L, 0, 1, rng)[0]),
                            _t(_deflection(L, rng))))
    server = LensingServer(hbm_budget_bytes=1, plans_per_key=1, max_batch=1)
    stats = serve(server, reqs, np.zeros(len(reqs)))
    assert stats["requests"] == 6 and stats["plans"] == 1


@pytest.mark.parametrize("spin,eps,nmaps", [(2, 1e-10, 5), (0, 1e-10, 3), (2, 1e-5, 2), (2, 1e-10, 1)])
def test_gclm2lenmap_batch_equals_single_calls(spin, eps, nmaps):
    """nusht_gclm2lenmap_batch (pointing once, one synthesis launch for the maps) == nmaps gclm2lenmap calls
    with the same deflection; the call is repeated on the same plan."""
    from cunusht_amd import _lib
    from cunusht_amd.batch import LensingPlan
    L = 40
    rng = np.random.default_rng(900 + spin + nmaps)
    plan = LensingPlan(L, spin, eps, max_maps=8)
    nc = 1 if spin == 0 else 2
    dlm = _t(_deflection(L, rng))
    g = _t(np.stack([almutil.random_alm(L, spin, nc, rng) for _ in range(nmaps)]))
    npix = plan.npix()
    out = torch.full((nmaps, 2, npix), np.nan, dtype=torch.float64, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: ctypes.c_void_p(0 if t is None else t.data_ptr())
    print(f"case gclm2lenmap_batch spin={spin} eps={eps} nmaps={nmaps}", flush=True)
    for _ in range(2):
        _lib.call("nusht_gclm2lenmap_batch", plan.nuplan._h, plan.ring._h, nmaps, p(g[0]), 2 * g[0].numel(), p(dlm),
                  p(out[0, 0]), p(out[0, 1]) if spin else None, 2 * npix, 1, s)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        for k in range(nmaps):
            ref = _alone(plan, "gclm2lenmap", g[k], dlm)
            tol = 1e-6 if eps >= 1e-6 else 1e-12
            assert _relmax(got[k, 0] if spin == 0 else got[k], ref) < tol
