"""Benchmark: spin-2 lensed maps/s at lmax 4095, eps 1e-7 on one MI355X.

Workload (synthetic, seeded): E alm with C_l ~ exp(-(l/3000)^2) / (l(l+1)), B with 0.1 times that
spectrum; deflection with C_l^dd ~ l^-2 (2 <= l <= lmax) normalised to 2.5 arcmin rms; pointing grid
GL (lmax+1) x (2 lmax + 1).  All inputs are resident in HBM before the timed region.

Paths timed:
  plain   : one gclm2lenmap per step, each with its own deflection (spin-1 SHT + pointing,
            point binning, spin-2 nuSHT with the polarisation rotation);
  batched : B maps per step sharing one deflection (pointing once per step, one Legendre
            recursion for all B maps, then the per-map DFS / nuFFT passes);
  batched, own deflection : B maps per step, each with its own deflection, through the batch
            executor (Legendre launches shared, pointing and binning per map);
  served  : requests replayed through the continuous-batching server (Poisson arrivals);
  adjoint : lenmap2gclm one call at a time.
`value` is the batched rate (lensed maps/s); the other rates, p50/p99 latencies, the roofline of the
dominant kernel (live hipEvent timing inside the library) and an oracle CPU baseline (one unscaled map of
the same workload) are reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--lmax L] [--eps E] [--batch B] [--no-cpu]
       python bench.py --uneven-batch 64      (64 independent transforms of uneven lmax / spin)
       python bench.py --serve-mix 48         (served mix of lmax 255..4095, spin 0/2, three eps, both directions)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6


def synth_inputs(lmax: int, rng: np.random.Generator):
    from cunusht_amd.utils import almxfl, synalm
    ll = np.arange(lmax + 1, dtype=np.float64)
    clee = np.zeros(lmax + 1)
    clee[2:] = np.exp(-(ll[2:] / 3000.0) ** 2) / (ll[2:] * (ll[2:] + 1)) * 1e3
    elm = synalm(clee, lmax, rng=rng)
    blm = synalm(0.1 * clee, lmax, rng=rng)
    cldd = np.zeros(lmax + 1)
    cldd[2:] = ll[2:] ** -2.0
    rms = math.sqrt(np.sum((2 * ll + 1) * cldd) / (4 * math.pi))
    cldd *= (2.5 / 60.0 * math.pi / 180.0 / rms) ** 2
    dlm = synalm(cldd, lmax, rng=rng)
    with np.errstate(divide="ignore"):
        fl = np.where(ll > 0, 1.0 / np.sqrt(ll * (ll + 1)), 0.0)
    dlm_scaled = almxfl(dlm, fl, lmax, False)
    return np.stack([elm, blm]), dlm_scaled


def stage_work(stage: str, info: dict, lmax: int, npix: int, maps_per_launch: int = 1, rinfo: dict | None = None):
    """(flops or bytes, unit) of one launch of a stage: the algorithmic work per unit times the units
    one launch processes.

    Legendre synthesis: (l, m, ring pair) units per map at 22 flops for a spin-2 E+B map and 14 for
    the spin-1 gradient pair.  The units counted are the ones a launch processes (the plan's
    legendre_pair_steps): l-ranges in which a whole group of 64 ring pairs is below 2^-300 are
    skipped.  The full grid, N_lm * ceil(n_rings / 2) units, is reported beside it.
    nuFFT interpolation: N_pts (16 + 16) + nrow nf b_grid bytes (FP64 coordinates, complex FP64
    values; the fine rows actually formed)."""
    nlm = (lmax + 1) * (lmax + 2) // 2
    nt = info["ntheta"]
    single = info["single_precision"]
    bc = 8 if single else 16
    if stage == "legendre_synth_cc_s2":
        return info["legendre_pair_steps"] * 22 * maps_per_launch, "flop"
    if stage == "legendre_synth_ring_s1" and rinfo is not None:
        return rinfo["legendre_pair_steps"] * 14, "flop"
    if stage == "nufft_interp":
        return npix * (16 + 16) + info["nrow_fine"] * info["nfine"] * bc, "byte"
    if stage == "theta_pass_t2":
        return (2 * lmax + 1) * nt * 16 + info["nrow_fine"] * (2 * lmax + 1) * bc, "byte"
    if stage == "phi_pass_t2":
        return info["nrow_fine"] * ((2 * lmax + 1) + info["nfine"]) * bc, "byte"
    if stage == "ring_dft_pointing":
        return (lmax + 1) * (2 * lmax + 1) * 16 + npix * 24, "byte"
    return None, None


def executed_flops(stage: str, info: dict, maps_per_launch: int = 1, rinfo: dict | None = None):
    """Flops the Legendre kernels actually run: executed (l, ring pair) steps x (7 flops of
    recursion + 16 of contraction per E + B map, 8 for the E-only gradient of the pointing)."""
    if stage == "legendre_synth_cc_s2":
        return info["legendre_pair_steps"] * (7 + 16 * maps_per_launch)
    if stage == "legendre_synth_ring_s1" and rinfo is not None:
        return rinfo["legendre_pair_steps"] * (7 + 8)
    return None


def profiled_traffic(stage: str, maps_per_launch: int):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (profiles/traffic_spin2_lmax4095.md, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_spin2_lmax4095.md")
    if not os.path.exists(path):
        return None
    key = f"traffic_bytes_per_launch[{stage}, {maps_per_launch} maps] = "
    with open(path) as f:
        for line in f:
            if line.strip().startswith(key):
                return int(line.strip()[len(key):].split()[0])
    return None


def cpu_baseline(lmax: int, eps: float):
    """Oracle (C/OpenMP FP64 restatement of the reference CPU algorithm, oracle/pipeline.py) timed on the
    host's cores: one spin-2 gclm2lenmap of the benchmark workload itself (lmax, eps, GL (lmax+1) x
    (2 lmax + 1), the same input spectra), spin-1 pointing included.  No scaling: the value is 1 / that time."""
    from oracle import geometry as ogeom, pipeline
    rng = np.random.default_rng(2024)
    gclm, dlm_scaled = synth_inputs(lmax, rng)
    rings = ogeom.gl_geometry(lmax, nphi=2 * lmax + 1)
    t0 = time.perf_counter()
    pipeline.gclm2lenmap(gclm, dlm_scaled, lmax, 2, rings, eps)
    t = time.perf_counter() - t0
    return {"value": 1.0 / t, "unit": "lensed maps/s", "cores": _host_threads(), "kind": "port",
            "sample": f"oracle gclm2lenmap spin-2 lmax={lmax} eps={eps:g}, one map on the GL {lmax + 1}x{2 * lmax + 1} "
                      f"grid (the benchmark workload, pointing included), {t:.1f} s"}


def cpu_backend_rate(lmax: int, eps: float):
    """This package's CPU backend (get_transformer('CPU', solver='duccnufft'), libnusht_cpu.so) on the same
    workload as the oracle baseline: one spin-2 gclm2lenmap at the benchmark's lmax and eps, pointing
    included, unscaled.  Informational: the product's host path beside the GPU, not the baseline."""
    import cunusht_amd
    from cunusht_amd.utils import almxfl
    rng = np.random.default_rng(2024)
    gclm, dlm_scaled = synth_inputs(lmax, rng)
    ll = np.arange(lmax + 1, dtype=np.float64)
    dlm = almxfl(dlm_scaled, np.sqrt(ll * (ll + 1)))
    T = cunusht_amd.get_transformer("CPU", solver="duccnufft")(("gl", {"lmax": lmax, "nphi": 2 * lmax + 1}),
                                                                epsilon=eps)
    nth = _host_threads()
    t0 = time.perf_counter()
    T.gclm2lenmap(gclm, dlm, lmax, lmax, 2, eps, nth, verbose=0)
    t = time.perf_counter() - t0
    return {"value": 1.0 / t, "unit": "lensed maps/s", "cores": nth,
            "sample": f"CPU backend gclm2lenmap spin-2 lmax={lmax} eps={eps:g}, one map on the GL "
                      f"{lmax + 1}x{2 * lmax + 1} grid, pointing included, {t:.1f} s"}


def _host_threads() -> int:
    """Host threads the CPU legs use: the OpenMP thread count when set (16 per GPU on the GPU hosts),
    else the cores this process may run on."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(env)) if env and env.isdigit() and int(env) > 0 else n


def rotated_deflections(dlm, L, n):
    """n deflections for the same seeded spectrum: dlm with its m-modes' phases turned by m * 0.3 k rad (k = 0..n-1),
    each a different field of the same statistics."""
    import torch
    m = torch.as_tensor(np.concatenate([np.full(L + 1 - mm, mm) for mm in range(L + 1)]), device="cuda")
    return [dlm * torch.exp(1j * 0.3 * k * m) for k in range(n)]


def own_deflection_batch(args, L, eps, gclms, dlms, geominfo):
    """B lensed maps per step, each with its own deflection, through the batch executor: the Legendre syntheses
    of the B maps share matrix-core launches (4 maps each) while every map computes its own pointing, binning and
    interpolation.  The like-for-like counterpart of the plain rate (nothing of the pointing is shared)."""
    import torch
    from cunusht_amd.batch import BatchJob, LensingPlan, execute
    B = args.batch
    plan = LensingPlan(L, 2, eps, geominfo, max_maps=8)
    defl = rotated_deflections(dlms[0], L, B)
    jobs = [BatchJob("gclm2lenmap", plan, defl[k], gclms[k % len(gclms)]) for k in range(B)]
    for _ in range(args.warmup):
        execute(jobs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        execute(jobs)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    return {"value": args.steps * B / t, "unit": "lensed maps/s", "maps_per_step": B,
            "ms_per_step": t / args.steps * 1e3, "deflections_per_step": B}


def serve_headline(args, L, eps, gclms, dlms, rng):
    """Replay spin-2 gclm2lenmap requests through the continuous-batching server: Poisson arrivals at
    args.serve_rate, four deflections each shared by a quarter of the requests (same-deflection requests
    queued together are batched, up to 8), two resident plans.  Returns throughput and latency."""
    import torch
    from cunusht_amd.serving import LensingServer, Request, poisson_arrivals, serve
    from cunusht_amd.utils import synalm
    geo = lambda lmax: ("gl", {"lmax": lmax, "nphi": 2 * lmax + 1})
    ndefl = 4
    defl = [dlms[i % len(dlms)] for i in range(min(ndefl, len(dlms)))]
    if len(defl) < ndefl:  # further deflections: the first one with its phases rotated by m * 0.3 k rad
        defl += rotated_deflections(dlms[0], L, ndefl)[len(defl):]
    server = LensingServer(hbm_budget_bytes=160 << 30, plans_per_key=2, max_batch=8, geometry=geo)
    key = (L, 2, float(eps), 2)

    def requests(n, seed):
        r = np.random.default_rng(seed)
        return [Request("gclm2lenmap", L, 2, eps, gclms[i % len(gclms)], defl[int(r.integers(ndefl))])
                for i in range(n)]

    server.prewarm(key)
    serve(server, requests(4, 1), np.zeros(4))  # warm-up: second slot, kernels, caches
    server.finished.clear()
    n = args.serve_requests
    st = serve(server, requests(n, 2), poisson_arrivals(n, args.serve_rate, seed=3))
    return {"maps_per_s": st["requests_per_s"], "requests": n, "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
            "arrival_rate_per_s": args.serve_rate, "deflections": ndefl, "max_batch": 8, "plans": st["plans"],
            "plan_bytes": st["plan_bytes"], "wall_s": st["wall_s"]}


def uneven_batch(njobs: int, steps: int, warmup: int):
    """Independent transforms of uneven size: lmax ~ U{511, 1023, 1535, 2047}, spin ~ U{0, 2}, eps 1e-7, each
    with its own seeded gclm (C_l ~ (l+1)^-2) and deflection (2.5 arcmin rms, C_l ~ l^-2).  Jobs of one
    (lmax, spin) share two resident plans that hold 4 maps, so the batch executor runs their Legendre syntheses
    together on the matrix cores.  Timed: the jobs as sequential single calls on one stream, the same jobs
    through nusht_batch_execute, and the batch captured once as a hipGraph and replayed."""
    import torch
    from cunusht_amd.batch import BatchJob, CapturedBatch, LensingPlan, execute
    from cunusht_amd.utils import synalm
    rng = np.random.default_rng(44)
    lmaxs = rng.choice([511, 1023, 1535, 2047], njobs)
    spins = rng.choice([0, 2], njobs)
    plans, jobs = {}, []
    for i, (L, s) in enumerate(zip(lmaxs, spins)):
        L, s = int(L), int(s)
        pl = plans.setdefault((L, s), [LensingPlan(L, s, 1e-7, max_maps=4), LensingPlan(L, s, 1e-7, max_maps=4)])
        ll = np.arange(L + 1, dtype=np.float64)
        cl = 1.0 / (ll + 1.0) ** 2
        nc = 1 if s == 0 else 2
        g = np.stack([synalm(cl, L, rng=rng) for _ in range(nc)])
        cldd = np.zeros(L + 1)
        cldd[2:] = ll[2:] ** -2.0
        cldd *= (2.5 / 60.0 * math.pi / 180.0) ** 2 / (np.sum((2 * ll + 1) * cldd) / (4 * math.pi))
        with np.errstate(divide="ignore"):
            fl = np.where(ll > 0, 1.0 / np.sqrt(ll * (ll + 1)), 0.0)
        from cunusht_amd.utils import almxfl
        d = almxfl(synalm(cldd, L, rng=rng), fl, L, False)
        jobs.append(BatchJob("gclm2lenmap", pl[i % 2], torch.as_tensor(d, device="cuda"),
                             torch.as_tensor(g, device="cuda")))
    from cunusht_amd import _lib
    import ctypes
    sp = lambda: ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def sequential():
        for j in jobs:
            cj = j.c_job()
            _lib.call("nusht_gclm2lenmap", ctypes.c_void_p(cj.plan), ctypes.c_void_p(cj.ring), ctypes.c_void_p(cj.alm_in),
                      ctypes.c_void_p(cj.dlm_scaled), ctypes.c_void_p(cj.map_out0),
                      ctypes.c_void_p(cj.map_out1) if cj.map_out1 else None, cj.polrot, sp())

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    t_seq = timed(sequential)
    t_batch = timed(lambda: execute(jobs))
    cap = CapturedBatch(jobs)
    t_graph = timed(cap.replay)
    return {"metric": f"{njobs} uneven lensing transforms per pass (lmax 511..2047, spin 0/2, eps 1e-7)",
            "unit": "ms per pass", "sequential_ms": t_seq * 1e3, "batch_ms": t_batch * 1e3, "graph_ms": t_graph * 1e3,
            "graph_speedup": t_seq / t_graph, "jobs_per_s_graph": njobs / t_graph,
            "plans": sum(len(v) for v in plans.values()), "steps": steps, "warmup": warmup,
            "mix": {f"{L}/{s}": int(np.sum((lmaxs == L) & (spins == s))) for (L, s) in sorted(plans)}}


def serve_mix(nreq: int, rate: float):
    """The serving configuration: Poisson arrivals at `rate`, each request forward (gclm2lenmap) or adjoint
    (lenmap2gclm), lmax ~ U{255, 511, 1023, 2047, 4095}, spin ~ U{0, 2}, eps ~ U{1e-5, 1e-7, 1e-10}; seeded
    gclm (C_l ~ (l+1)^-2) or N(0, 1) maps, one 2.5 arcmin deflection per lmax (shared, so forward requests of one
    configuration can batch).  Served through LensingServer (one resident plan per configuration, admission
    under a 200 GiB budget, batches of up to 8) twice: cold (plans built on demand) and warm (plans resident);
    beside it the plain path: the same requests one call at a time on resident plans.  Latencies: arrival to
    completion (served), call to completion (plain)."""
    import ctypes

    import torch
    from cunusht_amd.batch import LensingPlan
    from cunusht_amd.serving import LensingServer, Request, poisson_arrivals, serve
    from cunusht_amd.utils import almxfl, synalm
    rng = np.random.default_rng(55)
    lmaxs = [int(x) for x in rng.choice([255, 511, 1023, 2047, 4095], nreq)]
    spins = [int(x) for x in rng.choice([0, 2], nreq)]
    epss = [float(x) for x in rng.choice([1e-5, 1e-7, 1e-10], nreq)]
    kinds = ["gclm2lenmap" if x else "lenmap2gclm" for x in rng.integers(0, 2, nreq)]
    defl = {}
    for L in sorted(set(lmaxs)):
        ll = np.arange(L + 1, dtype=np.float64)
        cldd = np.zeros(L + 1)
        cldd[2:] = ll[2:] ** -2.0
        cldd *= (2.5 / 60.0 * math.pi / 180.0) ** 2 / (np.sum((2 * ll + 1) * cldd) / (4 * math.pi))
        with np.errstate(divide="ignore"):
            fl = np.where(ll > 0, 1.0 / np.sqrt(ll * (ll + 1)), 0.0)
        defl[L] = torch.as_tensor(almxfl(synalm(cldd, L, rng=rng), fl, L, False), device="cuda")
    data = []
    for L, sp, kind in zip(lmaxs, spins, kinds):
        nc = 1 if sp == 0 else 2
        if kind == "gclm2lenmap":
            cl = 1.0 / (np.arange(L + 1) + 1.0) ** 2
            data.append(torch.as_tensor(np.stack([synalm(cl, L, rng=rng) for _ in range(nc)]), device="cuda"))
        else:
            npix = (L + 1) * (2 * L + 2)
            data.append(torch.as_tensor(rng.standard_normal((nc, npix)), device="cuda"))

    def requests():
        return [Request(k, L, sp, e, d, defl[L]) for k, L, sp, e, d in zip(kinds, lmaxs, spins, epss, data)]

    offsets = poisson_arrivals(nreq, rate, seed=5)
    server = LensingServer(hbm_budget_bytes=200 << 30, plans_per_key=1, max_batch=8)
    cold = serve(server, requests(), offsets)
    server.finished = []
    warm = serve(server, requests(), offsets)
    warm["plan_bytes"] = server.used_bytes
    server.slots.clear()  # release the served plans before the plain path builds its own
    del server
    import gc
    gc.collect()
    # plain: one call at a time on resident plans of each configuration
    plans = {}
    s = lambda: ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: ctypes.c_void_p(0 if t is None else t.data_ptr())

    def one(i):
        L, sp, e, kind = lmaxs[i], spins[i], epss[i], kinds[i]
        nc = 1 if sp == 0 else 2
        pl = plans.get((L, sp, e))
        if pl is None:
            pl = plans[(L, sp, e)] = LensingPlan(L, sp, e)
        d = data[i]
        if kind == "gclm2lenmap":
            out = torch.empty((2, pl.npix()), dtype=torch.float64, device="cuda")
            _lib_call("nusht_gclm2lenmap", pl.nuplan._h, pl.ring._h, p(d), p(defl[L]), p(out[0]),
                      p(out[1]) if sp else None, 1, s())
        else:
            res = torch.empty((nc, pl.nalm()), dtype=torch.complex128, device="cuda")
            _lib_call("nusht_lenmap2gclm", pl.nuplan._h, pl.ring._h, p(d[0]), p(d[1]) if sp else None, p(defl[L]),
                      p(res), 1, s())

    from cunusht_amd import _lib
    _lib_call = _lib.call
    for i in range(nreq):  # build the plans and run each request once
        one(i)
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for i in range(nreq):
        ta = time.perf_counter()
        one(i)
        torch.cuda.synchronize()
        lat.append(1e3 * (time.perf_counter() - ta))
    t_plain = time.perf_counter() - t0
    mix = {}
    for L, sp, e, k in zip(lmaxs, spins, epss, kinds):
        key = f"{L}/{sp}"
        mix[key] = mix.get(key, 0) + 1
    return {"metric": f"served lensing requests/s, {nreq} requests of mixed lmax 255..4095, spin 0/2, "
                      f"eps 1e-5/1e-7/1e-10, forward and adjoint; p50/p99 ms",
            "unit": "requests/s", "arrival_rate_per_s": rate, "requests": nreq,
            "served_cold": cold, "served_warm": warm,
            "plain": {"requests_per_s": nreq / t_plain, "p50_ms": float(np.percentile(lat, 50)),
                      "p99_ms": float(np.percentile(lat, 99))},
            "served_warm_over_plain": warm["requests_per_s"] / (nreq / t_plain),
            "forward": kinds.count("gclm2lenmap"), "adjoint": kinds.count("lenmap2gclm"), "mix_lmax_spin": mix}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--lmax", type=int, default=4095)
    ap.add_argument("--eps", type=float, default=1e-7)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--split-batch", action="store_true",
                    help="batched step as pointing, then nusht2d2_batch (no overlap of pointing and Legendre)")
    ap.add_argument("--serve-requests", type=int, default=32,
                    help="requests replayed through the serving loop (0 = skip); they share 4 deflections")
    ap.add_argument("--serve-rate", type=float, default=400.0, help="Poisson arrival rate of served requests (1/s)")
    ap.add_argument("--serve-mix", type=int, default=0,
                    help="run the mixed serving configuration with this many requests instead of the headline")
    ap.add_argument("--uneven-batch", type=int, default=0,
                    help="instead of the headline: N independent gclm2lenmap jobs of lmax in {511, 1023, 1535, 2047}, "
                         "spin in {0, 2}, eps 1e-7, as one hipGraph-captured batch vs N sequential calls")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the timed region, copy a buffer of this many MiB once (a plain copy kernel "
                         "of known traffic, used to calibrate FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only (N=1). Independent lensing requests shard across GPUs "
              "without communication, so no multi-GPU figure is produced here.", file=sys.stderr)
        sys.exit(2)

    import torch
    import cunusht_amd
    from cunusht_amd import _lib

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.uneven_batch > 0:
        print(json.dumps(uneven_batch(args.uneven_batch, args.steps, args.warmup)))
        return
    if args.serve_mix > 0:
        print(json.dumps(serve_mix(args.serve_mix, args.serve_rate)))
        return
    L, eps, B = args.lmax, args.eps, args.batch
    print(f"[bench] lmax={L} eps={eps:g} batch={B} steps={args.steps} warmup={args.warmup}", file=sys.stderr,
          flush=True)
    rng = np.random.default_rng(1234)
    nin = 2
    gclms, dlms = [], []
    for i in range(max(nin, 1)):
        g, d = synth_inputs(L, rng)
        gclms.append(torch.as_tensor(g, device="cuda"))
        dlms.append(torch.as_tensor(d, device="cuda"))
    gb = torch.stack([gclms[i % nin] for i in range(B)]).contiguous()
    print("[bench] inputs ready", file=sys.stderr, flush=True)

    geominfo = ("gl", {"lmax": L, "nphi": 2 * L + 1})
    T = cunusht_amd.get_transformer("GPU")(geominfo, nuFFTtype=2, epsilon=eps, spin=2)
    npix = T.npix()
    out = torch.empty((2, npix), dtype=torch.float64, device="cuda")

    def plain_step(i):
        T.gclm2lenmap(gclms[i % nin], L, L, dlm_scaled=dlms[i % nin], lenmap=out, spin=2)

    # batched: B maps, one shared deflection per step (pointing once, one recursion for all B)
    plan = T._plan_for(2, 2, eps, npix)
    ring = T._ring(1)
    ptg = torch.empty((3, npix), dtype=torch.float64, device="cuda")
    outb = torch.empty((B, 2, npix), dtype=torch.float64, device="cuda")
    stream = lambda: __import__("ctypes").c_void_p(torch.cuda.current_stream().cuda_stream)
    P = lambda t: __import__("ctypes").c_void_p(t.data_ptr())
    nalm2 = gb.shape[1] * gb.shape[2] * 2  # doubles per map (2 components, complex)
    nb_done = 0

    def batched_step(i):
        d = dlms[i % nin]
        plan.pts_gen = 0  # the batch entry sets the plan's points itself
        done = 0
        if not args.split_batch:
            # one call: pointing and binning once, then the matrix-core synthesis and the interpolations of
            # the first (up to 8) maps
            nb = min(8, B)
            _lib.call("nusht_gclm2lenmap_batch", plan._h, ring._h, nb, P(gb[0]), nalm2, P(d), P(outb[0, 0]),
                      P(outb[0, 1]), 2 * npix, 1, stream())
            done = nb
            if done < B:  # later chunks reuse the points; their rotation angle comes from the ring plan
                _lib.call("nusht_ringplan_dlm2pointing", ring._h, P(d), P(ptg[0]), P(ptg[1]), P(ptg[2]), stream())
        else:
            _lib.call("nusht_ringplan_dlm2pointing", ring._h, P(d), P(ptg[0]), P(ptg[1]), P(ptg[2]), stream())
            _lib.call("nusht_plan_set_pts_async", plan._h, P(ptg[0]), P(ptg[1]), npix, stream())
        while done < B:
            nb = min(8, B - done)
            _lib.call("nusht_nusht2d2_batch", plan._h, nb, P(gb[done]), nalm2, P(outb[done, 0]), P(outb[done, 1]),
                      2 * npix, P(ptg[2]), stream())
            done += nb

    # warmup (also builds every plan and table)
    for i in range(args.warmup):
        plain_step(i)
        batched_step(i)
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)
    info = plan.info()
    rinfo = ring.info()
    nlm = (L + 1) * (L + 2) // 2

    # plain: per-call latency with events on the launch stream
    _lib.timing_reset()
    _lib.timing_enable(True)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):  # deflections alternate from the last warm-up one: every call computes its pointing
        evs[i][0].record()
        plain_step(i + args.warmup)
        evs[i][1].record()
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    stages_plain = _lib.timing_read()
    lat = np.array([a.elapsed_time(b) for a, b in evs])
    print(f"[bench] plain: {args.steps / t_plain:.2f} maps/s", file=sys.stderr, flush=True)

    # batched
    _lib.timing_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        batched_step(i + 1)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    stages_batch = _lib.timing_read()
    _lib.timing_enable(False)
    rate_batch = args.steps * B / t_batch
    rate_plain = args.steps / t_plain
    print(f"[bench] batched: {rate_batch:.2f} maps/s", file=sys.stderr, flush=True)

    # the adjoint (lenmap2gclm), one call at a time with its own deflection, like the plain forward path
    _lib.timing_reset()
    _lib.timing_enable(True)
    T.lenmap2gclm(out, dlms[0], None, L, L, spin=2)
    torch.cuda.synchronize()
    _lib.timing_reset()
    evs_a = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        evs_a[i][0].record()
        T.lenmap2gclm(out, dlms[(i + 1) % nin], None, L, L, spin=2)
        evs_a[i][1].record()
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    stages_adj = _lib.timing_read()
    _lib.timing_enable(False)
    lat_a = np.array([a.elapsed_time(b) for a, b in evs_a])
    adjoint = {"value": args.steps / t_adj, "unit": "lenmap2gclm calls/s", "p50_ms": float(np.percentile(lat_a, 50)),
               "p99_ms": float(np.percentile(lat_a, 99)),
               "stages_ms_per_call": {k: v[0] / args.steps for k, v in sorted(stages_adj.items())}}
    print(f"[bench] adjoint: {adjoint['value']:.2f} calls/s", file=sys.stderr, flush=True)

    # accuracy of both paths against exact FP64 synthesis at sampled pixels of the deflected grid (the
    # library's point plan: Legendre recursion on the pixels' own colatitudes and a direct sum over m, no nuFFT)
    batched_step(0)  # deflection dlms[0]; map 0 of the batch is gclms[0]
    plain_step(0)    # the same map and deflection alone; leaves T's pointing and rotation angle in place
    torch.cuda.synchronize()
    pix = np.sort(np.random.default_rng(77).choice(npix, 256, replace=False))
    q, u = T.gclm2lenpixs(gclms[0], L, 2, pix)
    ref = torch.stack([q, u]).cpu().numpy()

    def _err(m):
        d = m[:, pix].cpu().numpy() - ref
        return {"max_rel": float(np.max(np.abs(d)) / np.max(np.abs(ref))),
                "rms_rel": float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(ref * ref)))}

    accuracy = {"reference": "exact FP64 synthesis at 256 random pixels of the deflected grid (point plan)",
                "batched": _err(outb[0]), "plain": _err(out)}
    print(f"[bench] accuracy: {accuracy}", file=sys.stderr, flush=True)

    # roofline of the dominant stage of the headline (batched) path
    dom = max(stages_batch.items(), key=lambda kv: kv[1][0])
    name, (tot_ms, calls) = dom
    launches_cc = stages_batch.get("legendre_synth_cc_s2", (0, 1))[1]
    mpl = B * args.steps / max(launches_cc, 1) if name == "legendre_synth_cc_s2" else 1
    mpl = int(round(mpl))
    work, unit = stage_work(name, info, L, npix, mpl, rinfo)
    roof = None
    if work is not None:
        avg_s = tot_ms / calls / 1e3
        traffic = profiled_traffic(name, mpl)
        if unit == "flop":
            ach = work / avg_s / 1e12
            roof = {"bound": "mfma", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                    "frac": ach / PEAK_FP64_TFLOPS, "traffic": traffic, "kernel": name, "avg_ms": tot_ms / calls,
                    "maps_per_launch": mpl, "algorithmic_flop": work,
                    "full_grid_flop": (L + 1) * (L + 2) // 2 * math.ceil(info["ntheta"] / 2) * 22 * mpl
                    if name == "legendre_synth_cc_s2" else None}
            ex = executed_flops(name, info, mpl, rinfo)
            if ex is not None:
                roof["executed_flop"] = ex
                roof["executed_tflops"] = ex / avg_s / 1e12
                roof["executed_frac"] = ex / avg_s / 1e12 / PEAK_FP64_TFLOPS
        else:
            ach = work / avg_s / 1e9
            roof = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": ach / PEAK_HBM_GBS, "traffic": traffic, "kernel": name, "avg_ms": tot_ms / calls}
    # the nuFFT interpolation against the HBM roofline
    interp = stages_batch.get("nufft_interp")
    nufft_roof = None
    if interp:
        w_i, _ = stage_work("nufft_interp", info, L, npix)
        a = w_i / (interp[0] / interp[1] / 1e3) / 1e9
        nufft_roof = {"achieved_GBs": a, "frac": a / PEAK_HBM_GBS, "avg_ms": interp[0] / interp[1],
                      "traffic": profiled_traffic("nufft_interp", 1)}

    own = own_deflection_batch(args, L, eps, gclms, dlms, geominfo)
    print(f"[bench] batched, own deflection per map: {own['value']:.2f} maps/s", file=sys.stderr, flush=True)

    served = None
    if args.serve_requests > 0:
        served = serve_headline(args, L, eps, gclms, dlms, rng)
        print(f"[bench] served: {served['maps_per_s']:.2f} maps/s, p50 {served['p50_ms']:.1f} ms, "
              f"p99 {served['p99_ms']:.1f} ms", file=sys.stderr, flush=True)

    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cpu = cpu_be = None
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(L, eps)
        print(f"[bench] cpu baseline: {cpu['sample']}", file=sys.stderr, flush=True)
        cpu_be = cpu_backend_rate(L, eps)
        print(f"[bench] cpu backend: {cpu_be['sample']}", file=sys.stderr, flush=True)

    res = {
        "metric": "spin-2 lensed maps/s at lmax 4095, eps 1e-7 (served vs plain); p50/p99 ms",
        "value": rate_batch,
        "unit": "lensed maps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_batch / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if info["single_precision"] else "f64",
        "data": "synthetic (seeded E/B alm and deflection, GL 4096x8191 pointing grid)",
        "config": {"workload": f"spin-2 gclm2lenmap lmax={L} eps={eps:g}, GL {L + 1}x{2 * L + 1}, "
                               f"batched {B} maps per shared deflection",
                   "lmax": L, "epsilon": eps, "batch": B, "npix": npix,
                   "nufft": {"nfine": info["nfine"], "nrow": info["nrow_fine"], "w": info["kernel_width"],
                             "sigma": round(info["sigma"], 4)},
                   "legendre_variant": info["legendre_variant"],
                   "plan_bytes": info["workspace_bytes"],
                   "legendre_pair_steps": {"nusht_cc": info["legendre_pair_steps"],
                                           "full_cc": nlm * math.ceil(info["ntheta"] / 2),
                                           "pointing_gl": rinfo["legendre_pair_steps"],
                                           "full_gl": nlm * math.ceil((L + 1) / 2)}},
        "plain": {"value": rate_plain, "p50_ms": float(np.percentile(lat, 50)),
                  "p99_ms": float(np.percentile(lat, 99)), "latencies_ms": [round(float(x), 3) for x in lat]},
        "served": served,
        "adjoint": adjoint,
        "batched_over_plain": rate_batch / rate_plain,
        "batched_own_deflection": own,
        "batched_own_deflection_over_plain": own["value"] / rate_plain,
        "served_over_plain": (served["maps_per_s"] / rate_plain) if served else None,
        "accuracy": accuracy,
        "roofline": roof,
        "nufft_interp_roofline": nufft_roof,
        "stages_ms_per_step": {k: v[0] / args.steps for k, v in sorted(stages_batch.items())},
        "stages_plain_ms_per_call": {k: v[0] / args.steps for k, v in sorted(stages_plain.items())},
        "cpu_baseline": cpu,
        "cpu_backend": cpu_be,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
