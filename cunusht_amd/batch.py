"""Batches of independent transforms of uneven lmax / spin / epsilon (C ABI nusht_batch_execute).

No counterpart in the reference, which runs one transform at a time on the legacy stream
(cunusht/deflection/GPU_nufft_transformer.py:486-673).  A batch forks its jobs onto up to four
internal streams (jobs that share a plan stay ordered on one of them) and joins them back into the
caller's stream; `CapturedBatch` records one batch as a hipGraph (torch.cuda.CUDAGraph) and replays it
on the same buffers.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .geometry import Geom, get_geom
from .transformer import NuPlan, RingSHT, _dev, _ptr, _stream
from .utils import Alm

_KINDS = {"gclm2lenmap": _lib.JOB_GCLM2LENMAP, "lenmap2gclm": _lib.JOB_LENMAP2GCLM,
          "nusht2d2": _lib.JOB_NUSHT2D2, "nusht2d1": _lib.JOB_NUSHT2D1}


class LensingPlan:
    """Resident plans of one lensing configuration: the nuSHT plan (lmax, spin, ncomp, epsilon) and the
    spin-1 ring plan of the pointing geometry.  One LensingPlan runs one job at a time; a batch runs jobs
    on different LensingPlans concurrently."""

    def __init__(self, lmax: int, spin: int, epsilon: float, geometry=None, ncomp: int | None = None,
                 max_maps: int = 1, sigma: float = 0.0):
        """sigma: nuFFT oversampling, 0 automatic, < 0 chosen by measurement at plan creation."""
        self.geom: Geom = geometry if isinstance(geometry, Geom) else get_geom(geometry or ("gl", {"lmax": lmax}))
        self.lmax, self.spin, self.epsilon = lmax, spin, float(epsilon)
        self.ncomp = (1 if spin == 0 else 2) if ncomp is None else ncomp
        self.nuplan = NuPlan(lmax, spin, self.ncomp, self.epsilon, self.geom.npix(), sigma, max_maps)
        self.ring = RingSHT(self.geom, lmax, 1)

    def npix(self) -> int:
        return self.geom.npix()

    def nalm(self) -> int:
        return Alm.getsize(self.lmax, self.lmax)

    def nbytes(self) -> int:
        return self.nuplan.nbytes() + int(self.ring.info()["workspace_bytes"])


@dataclass
class BatchJob:
    """One job: kind in {'gclm2lenmap', 'lenmap2gclm'} on a LensingPlan, with device tensors.
    Forward: alm (ncomp, nalm) complex -> out (npix,) or (2, npix).  Adjoint: maps -> out (ncomp, nalm)."""
    kind: str
    plan: LensingPlan
    dlm_scaled: torch.Tensor
    inp: torch.Tensor
    out: torch.Tensor | None = None
    polrot: bool = True
    _keep: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        if self.kind not in ("gclm2lenmap", "lenmap2gclm"):
            raise ValueError(f"unknown job kind {self.kind}")
        p = self.plan
        self.dlm_scaled = _dev(self.dlm_scaled, torch.complex128, "dlm_scaled").reshape(-1)
        if self.dlm_scaled.numel() != p.nalm():
            raise ValueError("dlm_scaled must hold nalm(lmax) coefficients")
        if self.kind == "gclm2lenmap":
            self.inp = _dev(self.inp, torch.complex128, "gclm").reshape(p.ncomp, p.nalm())
            shape = (p.npix(),) if p.spin == 0 else (2, p.npix())
            if self.out is None:
                self.out = torch.empty(shape, dtype=torch.float64, device="cuda")
            if tuple(self.out.shape) != shape or self.out.dtype != torch.float64 or not self.out.is_contiguous():
                raise ValueError(f"out must be a contiguous float64 tensor of shape {shape}")
        else:
            self.inp = _dev(self.inp, torch.float64, "lenmap")
            self.inp = self.inp.reshape(1 if p.spin == 0 else 2, p.npix())
            if self.out is None:
                self.out = torch.empty((p.ncomp, p.nalm()), dtype=torch.complex128, device="cuda")
            if self.out.numel() != p.ncomp * p.nalm() or self.out.dtype != torch.complex128 \
                    or not self.out.is_contiguous():
                raise ValueError("out must be a contiguous complex128 tensor of ncomp x nalm")

    def c_job(self) -> _lib.Job:
        p = self.plan
        j = _lib.Job()
        j.kind = _KINDS[self.kind]
        j.plan = p.nuplan._h.value
        j.ring = p.ring._h.value
        j.dlm_scaled = self.dlm_scaled.data_ptr()
        j.polrot = 1 if self.polrot else 0
        if self.kind == "gclm2lenmap":
            j.alm_in = self.inp.data_ptr()
            j.map_out0 = (self.out if p.spin == 0 else self.out[0]).data_ptr()
            j.map_out1 = self.out[1].data_ptr() if p.spin else None
        else:
            j.map_in0 = self.inp[0].data_ptr()
            j.map_in1 = self.inp[1].data_ptr() if p.spin else None
            j.alm_out = self.out.data_ptr()
        return j


def execute(jobs: list[BatchJob]) -> None:
    """Enqueue every job on the current stream (forked over internal streams, joined back)."""
    arr = (_lib.Job * len(jobs))(*[j.c_job() for j in jobs])
    _lib.call("nusht_batch_execute", arr, len(jobs), _stream())


class CapturedBatch:
    """One batch recorded as a hipGraph; replay() re-runs it on the same input / output tensors (refill the
    inputs in place between replays)."""

    def __init__(self, jobs: list[BatchJob]):
        self.jobs = list(jobs)
        self.graph = torch.cuda.CUDAGraph()
        execute(self.jobs)  # eager pass first: creates the executor's streams outside the capture
        torch.cuda.synchronize()
        with torch.cuda.graph(self.graph):
            execute(self.jobs)
        torch.cuda.synchronize()

    def replay(self) -> None:
        self.graph.replay()
