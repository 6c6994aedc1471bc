"""Continuous-batching lensing server on one GPU.

No counterpart in the reference (cunusht runs one host-synchronous call at a time).  Requests
(forward gclm2lenmap or adjoint lenmap2gclm, any lmax / spin / epsilon) are queued on arrival and
admitted, in arrival order, as soon as a resident plan of their configuration is free and the plans fit
the HBM budget:

* a "slot" is one LensingPlan (nuSHT plan + spin-1 ring plan) with its own HIP stream; a configuration
  gets up to `plans_per_key` slots, so requests of one configuration overlap on the device, and requests of
  different configurations always do;
* queued forward requests of one configuration that share a deflection (the same dlm_scaled tensor) are
  admitted together, up to `max_batch`, and run as one pointing + one batched synthesis
  (nusht_gclm2lenmap_batch: the pointing once, one Legendre recursion for all of their maps);
* plans stay resident; when a new plan would exceed the budget, idle plans of other configurations are
  released, least recently used first.  A slot's bytes in the budget are its plans, its batch staging
  buffer (max_batch gclm, allocated with the slot) and the outputs of one launch (max_batch lensed maps,
  or one gclm for an adjoint request), which exist while the launch is in flight and then belong to the
  caller;
* completion is observed through events (no device-wide synchronization), and every request records its
  arrival, admission and completion times, from which p50 / p99 latency follow.
"""
from __future__ import annotations

import ctypes
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .batch import LensingPlan
from .transformer import _dev, _ptr
from .utils import Alm


@dataclass
class Request:
    """kind 'gclm2lenmap' (data = gclm, (ncomp, nalm) complex) or 'lenmap2gclm' (data = lenmap, (npix,) or
    (2, npix) real); dlm_scaled (nalm,) complex."""
    kind: str
    lmax: int
    spin: int
    epsilon: float
    data: torch.Tensor
    dlm_scaled: torch.Tensor
    polrot: bool = True
    arrival: float | None = None
    admitted: float | None = None
    done: float | None = None
    result: torch.Tensor | None = None
    _ready: torch.cuda.Event | None = field(default=None, repr=False)

    @property
    def key(self):
        ncomp = 1 if self.spin == 0 else (self.data.shape[0] if self.kind == "gclm2lenmap" and self.data.dim() == 2
                                          else 2)
        return (self.lmax, self.spin, float(self.epsilon), ncomp)

    @property
    def latency_ms(self) -> float:
        return 1e3 * (self.done - self.arrival)


class _Slot:
    def __init__(self, key, geometry, max_batch):
        lmax, spin, eps, ncomp = key
        self.key = key
        self.plan = LensingPlan(lmax, spin, eps, geometry, ncomp=ncomp, max_maps=max_batch)
        self.stream = torch.cuda.Stream()  # default priority
        self.busy = False
        self.last_used = time.perf_counter()
        nalm, npix = self.plan.nalm(), self.plan.npix()
        # batch staging (forward batches of 2 .. max_batch maps gather their gclm here)
        self.stage = torch.empty((max_batch, ncomp, nalm), dtype=torch.complex128, device="cuda") \
            if max_batch > 1 else None
        out_bytes = max(max_batch * 2 * npix * 8, ncomp * nalm * 16)
        self.nbytes = self.plan.nbytes() + (self.stage.nbytes if self.stage is not None else 0) + out_bytes


class LensingServer:
    def __init__(self, hbm_budget_bytes: int, plans_per_key: int = 2, max_batch: int = 4, geometry=None):
        """geometry: callable lmax -> geometry description or Geom for the pointing grid (default GL
        with lmax + 1 rings)."""
        if not 1 <= max_batch <= 8:
            raise ValueError("max_batch must be in [1, 8]")
        self.budget = int(hbm_budget_bytes)
        self.plans_per_key = plans_per_key
        self.max_batch = max_batch
        self.geometry = geometry or (lambda lmax: ("gl", {"lmax": lmax}))
        self.queue: deque[Request] = deque()
        self.slots: dict = {}
        self.inflight: list = []
        self.finished: list[Request] = []
        self.used_bytes = 0

    # ------------------------------------------------------------------ public
    def submit(self, req: Request) -> Request:
        if req.kind not in ("gclm2lenmap", "lenmap2gclm"):
            raise ValueError(f"unknown request kind {req.kind}")
        if req.arrival is None:
            req.arrival = time.perf_counter()
        req.dlm_scaled = _dev(req.dlm_scaled, torch.complex128, "dlm_scaled").reshape(-1)
        req.data = _dev(req.data, torch.complex128 if req.kind == "gclm2lenmap" else torch.float64, "data")
        if req.dlm_scaled.numel() != Alm.getsize(req.lmax, req.lmax):
            raise ValueError("dlm_scaled must hold nalm(lmax) coefficients")
        req._ready = torch.cuda.Event()
        req._ready.record()  # inputs were produced on the submitter's stream
        self.queue.append(req)
        return req

    def prewarm(self, key) -> None:
        """Create one resident slot for (lmax, spin, epsilon, ncomp) ahead of traffic."""
        if not self.slots.get(key):
            self._new_slot(key)

    def poll(self) -> int:
        """Retire finished work and admit queued requests; returns the number of requests still pending."""
        self._retire()
        self._admit()
        return len(self.queue) + sum(len(b[1]) for b in self.inflight)

    def run_until_idle(self) -> None:
        while self.poll():
            pass

    def stats(self) -> dict:
        lat = np.array([r.latency_ms for r in self.finished])
        return {"requests": len(self.finished),
                "p50_ms": float(np.percentile(lat, 50)) if lat.size else None,
                "p99_ms": float(np.percentile(lat, 99)) if lat.size else None,
                "plans": sum(len(v) for v in self.slots.values()), "plan_bytes": self.used_bytes}

    # ------------------------------------------------------------------ internals
    def _retire(self):
        keep = []
        for slot, reqs, ev in self.inflight:
            if ev.query():
                now = time.perf_counter()
                for r in reqs:
                    r.done = now
                    self.finished.append(r)
                slot.busy = False
                slot.last_used = now
            else:
                keep.append((slot, reqs, ev))
        self.inflight = keep

    def _new_slot(self, key):
        slot = _Slot(key, self.geometry(key[0]), self.max_batch)
        self.slots.setdefault(key, []).append(slot)
        self.used_bytes += slot.nbytes
        self._evict(exclude=key)
        return slot

    def _evict(self, exclude):
        """Release idle slots of other configurations, least recently used first, while over budget."""
        while self.used_bytes > self.budget:
            idle = [(s.last_used, k, s) for k, v in self.slots.items() if k != exclude for s in v if not s.busy]
            if not idle:
                return
            _, k, s = min(idle, key=lambda t: t[0])
            self.slots[k].remove(s)
            self.used_bytes -= s.nbytes
            del s

    def _free_slot(self, key):
        for s in self.slots.get(key, []):
            if not s.busy:
                return s
        n = len(self.slots.get(key, []))
        if n >= self.plans_per_key:
            return None
        est = self.slots[key][0].nbytes if n else 0
        if n and self.used_bytes + est > self.budget:
            self._evict(exclude=key)
            if self.used_bytes + est > self.budget:
                return None
        if not n and self.inflight and self.used_bytes >= self.budget:
            return None  # wait for running work before growing past the budget
        return self._new_slot(key)

    def _admit(self):
        pending = deque()
        while self.queue:
            r = self.queue.popleft()
            slot = self._free_slot(r.key)
            if slot is None:
                pending.append(r)
                continue
            group = [r]
            if r.kind == "gclm2lenmap":
                rest = deque()
                while self.queue and len(group) < self.max_batch:
                    q = self.queue.popleft()
                    if q.kind == "gclm2lenmap" and q.key == r.key and q.polrot == r.polrot \
                            and q.dlm_scaled.data_ptr() == r.dlm_scaled.data_ptr():
                        group.append(q)
                    else:
                        rest.append(q)
                rest.extend(self.queue)
                self.queue = rest
            self._launch(slot, group)
        self.queue = pending

    def _launch(self, slot: _Slot, reqs: list[Request]):
        now = time.perf_counter()
        slot.busy = True
        p = slot.plan
        st = slot.stream
        sp = ctypes.c_void_p(st.cuda_stream)
        with torch.cuda.stream(st):
            for r in reqs:
                st.wait_event(r._ready)
                r.admitted = now
                for t in (r.data, r.dlm_scaled):
                    t.record_stream(st)
            r0 = reqs[0]
            if r0.kind == "lenmap2gclm":
                m = r0.data.reshape(1 if p.spin == 0 else 2, p.npix())
                out = torch.empty((p.ncomp, p.nalm()), dtype=torch.complex128, device="cuda")
                _lib.call("nusht_lenmap2gclm", p.nuplan._h, p.ring._h, _ptr(m[0]), _ptr(m[1] if p.spin else None),
                          _ptr(r0.dlm_scaled), _ptr(out), 1 if r0.polrot else 0, sp)
                r0.result = out if p.ncomp > 1 else out[0]
            elif len(reqs) == 1:
                out = torch.empty((2, p.npix()), dtype=torch.float64, device="cuda")
                _lib.call("nusht_gclm2lenmap", p.nuplan._h, p.ring._h, _ptr(r0.data), _ptr(r0.dlm_scaled),
                          _ptr(out[0]), _ptr(out[1] if p.spin else None), 1 if r0.polrot else 0, sp)
                r0.result = out[0] if p.spin == 0 else out
            else:
                nb, nalm, npix = len(reqs), p.nalm(), p.npix()
                alm = slot.stage[:nb]  # the slot runs one launch at a time: free again once it is idle
                for b, r in enumerate(reqs):
                    alm[b].copy_(r.data.reshape(p.ncomp, nalm))
                out = torch.empty((nb, 2, npix), dtype=torch.float64, device="cuda")
                _lib.call("nusht_gclm2lenmap_batch", p.nuplan._h, p.ring._h, nb, _ptr(alm), 2 * p.ncomp * nalm,
                          _ptr(r0.dlm_scaled), _ptr(out[:, 0]), _ptr(out[:, 1]) if p.spin else None, 2 * npix,
                          1 if r0.polrot else 0, sp)
                for b, r in enumerate(reqs):
                    r.result = out[b, 0] if p.spin == 0 else out[b]
            ev = torch.cuda.Event()
            ev.record(st)
        self.inflight.append((slot, reqs, ev))


def poisson_arrivals(n: int, rate_per_s: float, seed: int = 0) -> np.ndarray:
    """Arrival offsets (s) of a Poisson process of the given rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, n))


def serve(server: LensingServer, requests: list[Request], offsets) -> dict:
    """Replay `requests` arriving at `offsets` (s from now) through the server; returns its statistics
    plus the throughput over the whole run."""
    order = np.argsort(offsets)
    t0 = time.perf_counter()
    i = 0
    while i < len(order) or server.poll():
        now = time.perf_counter() - t0
        while i < len(order) and offsets[order[i]] <= now:
            r = requests[order[i]]
            r.arrival = t0 + offsets[order[i]]
            server.submit(r)
            i += 1
        server.poll()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = server.stats()
    out["wall_s"] = wall
    out["requests_per_s"] = len(requests) / wall
    return out
