"""GPU nuSHT operator: the MI355X counterpart of GPU_cufinufft_transformer.

Mirrors cunusht/deflection/GPU_nufft_transformer.py:107-680 -- constructor, plan, set_pts,
nusht2d2, nusht2d1, gclm2lenmap, lenmap2gclm, dlm2pointing, synthesis, adjoint_synthesis,
npix, get_grid_uniform, hashdict -- on PyTorch-ROCm device tensors instead of CuPy arrays.
Every numerical step runs in libnusht_amd.so (hand-written gfx950 kernels).

Differences from the reference, all additive:
  * `spin` (0, 1, 2) and `polrot` keyword arguments, with the spin-s semantics of the reference
    CPU backend (cunusht/deflection/CPU_nufft_transformer.py:429-529); spin > 0 maps are (Q, U) rows.
  * one instance serves both directions (the reference plans one nuFFT type per instance).
  * `out` / `pointmap` / `alm` / `lenmap` arguments are written to when given.
  * the pointing and point caches compare a caller's array found at the address of the last one used
    with a device copy of it (the reference's full equality compare, GPU_nufft_transformer.py:417-420,
    470-473,550-556), so a buffer refilled in place or a recycled temporary is never mistaken for the old
    points; an array at another address is taken as new without a compare.
Errors: AssertionError for lmax/mmax/epsilon mismatches (as the reference), ValueError for unknown
options, nusht NuShtError (RuntimeError) for failures inside the library.
"""
from __future__ import annotations

import ctypes
import itertools
import warnings

import numpy as np
import torch

from . import _lib, diagnostics
from .geometry import Geom, get_geom
from .utils import Alm


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _dev(x, dtype, name):
    """Device tensor of `dtype` (copies host arrays with a warning, like GPU_nufft_transformer.py:502-513)."""
    if isinstance(x, torch.Tensor):
        if not x.is_cuda:
            warnings.warn(f"{name} is a host tensor and has been transferred to the device", stacklevel=3)
            x = x.cuda()
        if x.dtype != dtype:
            x = x.to(dtype)
        return x.contiguous()
    warnings.warn(f"{name} is not a device tensor and has been transferred to the device", stacklevel=3)
    return torch.as_tensor(np.asarray(x), device="cuda").to(dtype).contiguous()


_GENERATION = itertools.count(1)


class _Held:
    """Device copy of the last array a cache was built from.  `same(t)` holds when `t` sits at the address
    of that array and has the same contents: the reference's full equality compare
    (GPU_nufft_transformer.py:417-420,550-556; one device compare and one stream synchronization) for a
    buffer passed again, which may have been refilled in place (through a raw pointer, which torch's version
    counter does not see) or be a recycled temporary.  An array at another address is taken as new without
    a compare (no synchronization): equal contents there cost one recomputation, never a stale result."""

    def __init__(self):
        self.copy = None
        self.ptr = None
        self.gen = 0

    def same(self, t: torch.Tensor) -> bool:
        c = self.copy
        return (c is not None and t.data_ptr() == self.ptr and c.shape == t.shape and c.dtype == t.dtype
                and bool(torch.equal(c, t)))

    def hold(self, t: torch.Tensor) -> int:
        self.copy = t.detach().clone()
        self.ptr = t.data_ptr()
        self.gen = next(_GENERATION)
        return self.gen


class RingSHT:
    """SHT on an iso-latitude geometry (the pointing grid), one spin."""

    def __init__(self, geom: Geom, lmax: int, spin: int):
        self.geom, self.lmax, self.spin = geom, lmax, spin
        th = np.ascontiguousarray(geom.theta, dtype=np.float64)
        p0 = np.ascontiguousarray(geom.phi0, dtype=np.float64)
        nph = np.ascontiguousarray(geom.nph, dtype=np.int64)
        ofs = np.ascontiguousarray(geom.ofs, dtype=np.int64)
        wt = np.ascontiguousarray(geom.weight, dtype=np.float64)
        d = _lib.RingDesc(lmax, spin, th.size, th.ctypes.data, p0.ctypes.data, nph.ctypes.data, ofs.ctypes.data,
                          wt.ctypes.data)
        h = ctypes.c_void_p()
        _lib.call("nusht_ringplan_create", ctypes.byref(d), _stream(), ctypes.byref(h))
        self._h = h

    def info(self) -> dict:
        i = _lib.RingInfo()
        _lib.call("nusht_ringplan_get_info", self._h, ctypes.byref(i))
        return {f: getattr(i, f) for f, _ in _lib.RingInfo._fields_}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:  # at interpreter exit the module may be gone
            _lib.load().nusht_ringplan_destroy(h)
            self._h = None


class PointPlan:
    """Exact synthesis at a few points (colatitudes fixed at creation, host array)."""

    def __init__(self, lmax: int, spin: int, theta):
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64).reshape(-1))
        h = ctypes.c_void_p()
        _lib.call("nusht_pointplan_create", lmax, spin, th.size, th.ctypes.data, _stream(), ctypes.byref(h))
        self._h = h
        self.lmax, self.spin, self.npts = lmax, spin, th.size

    def synthesis(self, alm: torch.Tensor, ncomp: int, phi: torch.Tensor, gamma=None):
        alm = _dev(alm, torch.complex128, "alm").contiguous()
        phi = _dev(phi, torch.float64, "phi").contiguous()
        if alm.numel() != ncomp * Alm.getsize(self.lmax, self.lmax):
            raise ValueError(f"alm has {alm.numel()} entries, expected {ncomp} x nalm(lmax={self.lmax})")
        if phi.numel() != self.npts or (gamma is not None and gamma.numel() != self.npts):
            raise ValueError(f"phi / gamma must hold one value per point ({self.npts})")
        if gamma is not None:
            gamma = _dev(gamma, torch.float64, "gamma").contiguous()
        out = torch.empty((2, self.npts), dtype=torch.float64, device="cuda")
        _lib.call("nusht_pointplan_synthesis", self._h, _ptr(alm), ncomp, _ptr(phi), _ptr(gamma), _ptr(out[0]),
                  _ptr(out[1] if self.spin else None), _stream())
        return out[0] if self.spin == 0 else out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:  # at interpreter exit the module may be gone
            _lib.load().nusht_pointplan_destroy(h)
            self._h = None


class NuPlan:
    """One resident nuSHT plan (lmax, spin, ncomp, epsilon)."""

    def __init__(self, lmax: int, spin: int, ncomp: int, epsilon: float, npts_max: int, sigma: float = 0.0,
                 max_maps: int = 0):
        d = _lib.PlanDesc(lmax, lmax, spin, ncomp, float(epsilon), int(npts_max), float(sigma), int(max_maps))
        h = ctypes.c_void_p()
        _lib.call("nusht_plan_create", ctypes.byref(d), _stream(), ctypes.byref(h))
        self._h = h
        self.lmax, self.spin, self.ncomp, self.epsilon, self.npts_max = lmax, spin, ncomp, epsilon, npts_max
        self.pts_gen = 0  # generation of the point set last given to set_pts (0: none)

    def info(self) -> dict:
        i = _lib.PlanInfo()
        _lib.call("nusht_plan_get_info", self._h, ctypes.byref(i))
        return {f: getattr(i, f) for f, _ in _lib.PlanInfo._fields_}

    def nbytes(self) -> int:
        """Device bytes held by the plan."""
        return int(_lib.load().nusht_plan_workspace_bytes(self._h))

    def set_pts(self, theta: torch.Tensor, phi: torch.Tensor, checked: bool = True):
        """checked: bounds-check the points (one stream synchronization, raises on a bad point);
        otherwise bad points give NaN in nusht2d2 and are ignored by nusht2d1."""
        fn = "nusht_plan_set_pts" if checked else "nusht_plan_set_pts_async"
        _lib.call(fn, self._h, _ptr(theta), _ptr(phi), theta.numel(), _stream())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:  # at interpreter exit the module may be gone
            _lib.load().nusht_plan_destroy(h)
            self._h = None


class GPU_deflection:
    """`deflectionlib` of the GPU transformer (cunusht/deflection/GPU_nufft_transformer.py:42-104, built at
    :164): the pointing grid's geometry, dlm2pointing, and the SHT on the pointing grid that the reference
    reaches through `__getattr__` on its SHTns transformer (cunusht/sht/GPU_sht_transformer.py:47-121).
    Everything runs on the owning transformer's ring plans.  Maps are ring-major in the geometry's order, as
    every other map of this package (the reference's SHTns grid is theta-contiguous here, a layout its
    lensed maps do not share)."""

    def __init__(self, owner: "GPU_nusht_transformer"):
        self._owner = owner
        self.geominfo = owner.geominfo
        self.geom = owner.geom
        self.execmode = "normal"

    def dlm2pointing(self, dlm_scaled, mmax_dlm=None, verbose=0, pointing_theta=None, pointing_phi=None):
        """GPU_nufft_transformer.py:64-104: deflected (theta, phi), written into pointing_theta / pointing_phi
        when given; returns the (2, npix) pair."""
        ptg = self._owner.dlm2pointing(dlm_scaled, mmax_dlm, verbose)
        for dst, src in ((pointing_theta, ptg[0]), (pointing_phi, ptg[1])):
            if isinstance(dst, torch.Tensor):
                dst.copy_(src.reshape(dst.shape))
        return ptg

    def _lm(self, lmax, mmax):
        L = self._owner.lmax
        return L if lmax is None else lmax, L if mmax is None else mmax

    def synthesis(self, gclm, lmax=None, mmax=None, mode=None, nthreads=None):
        """GPU_sht_transformer.py:47-54: spin-0 map of gclm, flat."""
        lmax, mmax = self._lm(lmax, mmax)
        return self._owner.synthesis(gclm, None, lmax, mmax).reshape(-1)

    def synthesis_cupy(self, gclm, out, lmax=None, mmax=None, mode=None, nthreads=None):
        """GPU_sht_transformer.py:56-62 (the reference ignores `out`; it is written here when it fits)."""
        lmax, mmax = self._lm(lmax, mmax)
        return self._owner.synthesis(gclm, out, lmax, mmax)

    def adjoint_synthesis_cupy(self, synthmap, gclm=None, lmax=None, mmax=None, mode=None, nthreads=None, iw=None):
        """GPU_sht_transformer.py:65-73: SHTns analys (quadrature-weighted adjoint) of synthmap * iw."""
        lmax, mmax = self._lm(lmax, mmax)
        if iw is not None:
            synthmap = _dev(synthmap, torch.float64, "synthmap") * _dev(iw, torch.float64, "iw")
        return self._owner.adjoint_synthesis(synthmap, lmax, mmax, gclm)

    def synthesis_der1(self, gclm, out=None, nthreads=None):
        """GPU_sht_transformer.py:76-85: (d/dtheta, d/dphi / sin theta) of sum gclm Y_lm, shape (2, npix)."""
        return self.synthesis_der1_cupy(gclm)

    def synthesis_der1_cupy(self, gclm, out_theta=None, out_phi=None, nthreads=None):
        """GPU_sht_transformer.py:88-96 (SHTns synth_grad): the gradient pair, (2, npix); written into
        out_theta / out_phi when given."""
        o = self._owner
        alm = _dev(gclm, torch.complex128, "gclm").reshape(-1)
        assert alm.numel() == Alm.getsize(o.lmax, o.lmax), "gclm must have the transformer's lmax"
        res = torch.empty((2, o.npix()), dtype=torch.float64, device="cuda")
        _lib.call("nusht_sht_synth_grad_gl", o._ring(1)._h, _ptr(alm), _ptr(res[0]), _ptr(res[1]), _stream())
        for dst, src in ((out_theta, res[0]), (out_phi, res[1])):
            if isinstance(dst, torch.Tensor):
                dst.copy_(src.reshape(dst.shape))
        return res

    def analysis(self, map, lmax=None, mmax=None, nthreads=None, alm=None, mode=None):
        """GPU_sht_transformer.py:99-105: quadrature analysis of a spin-0 map, shape (1, nalm)."""
        lmax, mmax = self._lm(lmax, mmax)
        return self._owner.adjoint_synthesis(map, lmax, mmax, alm).reshape(1, -1)

    def adjoint_synthesis(self, map, **kwargs):
        """GPU_sht_transformer.py:108-118: analysis of a flat or 2-D spin-0 map."""
        return self.analysis(map, **kwargs)

    def map2alm(self, m, **kwargs):
        return self.analysis(m, **kwargs)

    def alm2map(self, gclm, **kwargs):
        return self.synthesis(gclm, **kwargs)


class GPU_nusht_transformer:
    """nuSHT operators on one MI355X (cunusht/deflection/GPU_nufft_transformer.py:107-680)."""

    def __init__(self, geominfo_deflection, shttransformer_desc="shtns", nuFFTtype=None, epsilon=None,
                 spin: int = 0, sigma: float = 0.0):
        if shttransformer_desc not in ("shtns", "ducc", "cunusht", "amd"):
            raise ValueError('shttransformer_desc must be either "ducc" or "shtns" or "cunusht"')
        if nuFFTtype:
            assert epsilon is not None, "epsilon must be provided if planned is True"
        self.backend = "GPU"
        self.shttransformer_desc = shttransformer_desc
        self.nuFFTtype = nuFFTtype
        self.geominfo = geominfo_deflection
        self.geom = get_geom(geominfo_deflection)
        self.lmax = int(geominfo_deflection[1]["lmax"])
        self.epsilon = epsilon
        self.spin = spin
        self.sigma = sigma
        self.execmode = "normal"
        self.dlm_scaled = None
        self.ptg = None
        self.ret = {}
        self.timer = diagnostics.Timer()
        self._plans = {}
        self._rings = {}
        self._src = _Held()   # the ptg or dlm_scaled the current pointing came from
        self._src_kind = None  # "ptg" or "dlm"
        self._loc = _Held()   # the last loc of nusht2d2 / nusht2d1
        self._cur_pts = None  # (theta, phi, generation, checked)
        self._gamma = None
        self.deflectionlib = GPU_deflection(self)
        if nuFFTtype:
            self.plan(epsilon=epsilon, nuFFTtype=nuFFTtype)

    # ------------------------------------------------------------------ planning
    def _plan_for(self, spin: int, ncomp: int, epsilon: float, npts: int) -> NuPlan:
        key = (spin, ncomp, float(epsilon))
        p = self._plans.get(key)
        if p is None or p.npts_max < npts:
            p = NuPlan(self.lmax, spin, ncomp, epsilon, max(npts, self.npix()), self.sigma)
            self._plans[key] = p
        return p

    def _ring(self, spin: int) -> RingSHT:
        r = self._rings.get(spin)
        if r is None:
            r = RingSHT(self.geom, self.lmax, spin)
            self._rings[spin] = r
        return r

    def plan(self, epsilon, nuFFTtype=1):
        """GPU_nufft_transformer.py:227-246: builds the resident plan (both transform types)."""
        self.epsilon = epsilon
        self.nuFFTtype = nuFFTtype
        self.single_prec = epsilon >= 1e-6
        ncomp = 1 if self.spin == 0 else 2
        return self._plan_for(self.spin, ncomp, epsilon, self.npix())

    def set_pts(self, ptg):
        """GPU_nufft_transformer.py:249-251: ptg (2, npts) rows (theta, phi)."""
        ptg = _dev(ptg, torch.float64, "ptg")
        self.ptg = ptg
        self._gamma = None
        gen = self._src.hold(ptg)
        self._src_kind = "ptg"
        self._cur_pts = (ptg[0].contiguous(), ptg[1].contiguous(), gen, True)

    def _ensure_pts(self, plan: NuPlan):
        theta, phi, gen, checked = self._cur_pts
        if plan.pts_gen != gen:
            # pointing computed here from dlm_scaled lies in [0, pi] by construction: no host check
            plan.set_pts(theta, phi, checked=checked)
            plan.pts_gen = gen

    def _loc_pts(self, plan: NuPlan, loc: torch.Tensor):
        """Points of nusht2d2 / nusht2d1: set on the plan unless it holds these very values."""
        if not self._loc.same(loc):
            self._loc.hold(loc)
        if plan.pts_gen != self._loc.gen:
            plan.set_pts(loc[:, 0].contiguous(), loc[:, 1].contiguous())
            plan.pts_gen = self._loc.gen

    # ------------------------------------------------------------------ geometry helpers
    def npix(self) -> int:
        return self.geom.npix()

    def get_grid_uniform(self):
        """GPU_nufft_transformer.py:326-336: (theta, phi) of every pixel of the pointing grid."""
        th, ph = self.geom.pixel_angles()
        return torch.as_tensor(np.array([th, ph]), device="cuda")

    @staticmethod
    def hashdict():
        """GPU_nufft_transformer.py:676-680 (delensalot compatibility)."""
        return "GPU_nusht_transformer"

    def _check_lm(self, lmax, mmax):
        assert lmax == self.lmax, f"Expected lmax={self.lmax}. lmax must be same as during init, but is {lmax}"
        assert mmax == self.lmax, f"Expected mmax={self.lmax}. mmax must be same as lmax during init, but is {mmax}"

    def _eps(self, epsilon):
        if self.nuFFTtype:
            assert epsilon is None or epsilon == self.epsilon, \
                f"expected epsilon = {self.epsilon} as set during initialisation of nuFFT plan, but is {epsilon}"
            return self.epsilon
        assert epsilon is not None, "epsilon must be provided if not planned"
        self.epsilon = epsilon
        return epsilon

    @staticmethod
    def _ncomp(alm, spin):
        if spin == 0:
            return 1
        return 1 if alm.dim() == 1 or alm.shape[0] == 1 else 2

    # ------------------------------------------------------------------ bare operators
    def nusht2d2(self, lmax, mmax, alm, loc, pointmap=None, epsilon=None, verbose=0, spin=None):
        """alm -> values at loc (npts, 2) = (theta, phi) (GPU_nufft_transformer.py:386-435)."""
        self._check_lm(lmax, mmax)
        spin = self.spin if spin is None else spin
        eps = self._eps(epsilon)
        alm = _dev(alm, torch.complex128, "alm")
        loc = _dev(loc, torch.float64, "loc")
        npts = loc.shape[0]
        ncomp = self._ncomp(alm, spin)
        plan = self._plan_for(spin, ncomp, eps, npts)
        self._loc_pts(plan, loc)
        out = self._out(pointmap, spin, npts)
        _lib.call("nusht_nusht2d2", plan._h, _ptr(alm), _ptr(out[0]), _ptr(out[1] if spin else None), None, _stream())
        return out[0] if spin == 0 else out

    def nusht2d1(self, lmax, mmax, pointmap, loc, epsilon=None, alm=None, verbose=0, spin=None):
        """Exact adjoint of nusht2d2 (GPU_nufft_transformer.py:437-484)."""
        self._check_lm(lmax, mmax)
        spin = self.spin if spin is None else spin
        eps = self._eps(epsilon)
        loc = _dev(loc, torch.float64, "loc")
        npts = loc.shape[0]
        vals = self._values(pointmap, spin)
        ncomp = 1 if spin == 0 else (alm.shape[0] if alm is not None and alm.dim() == 2 else 2)
        plan = self._plan_for(spin, ncomp, eps, npts)
        self._loc_pts(plan, loc)
        res = self._alm_out(alm, ncomp)
        _lib.call("nusht_nusht2d1", plan._h, _ptr(vals[0]), _ptr(vals[1]), None, _ptr(res), _stream())
        return res

    # ------------------------------------------------------------------ lensing
    def dlm2pointing(self, dlm_scaled, mmax_dlm=None, verbose=0, single_prec=False, gamma=False):
        """Deflected (theta, phi) of every pointing pixel (GPU_nufft_transformer.py:341-351, 64-104).
        Returns (2, npix) float64; with gamma=True also the rotation angle (CPU_nufft_transformer.py:158-160)."""
        dlm = _dev(dlm_scaled, torch.complex128, "dlm_scaled").reshape(-1)
        assert dlm.numel() == Alm.getsize(self.lmax, self.lmax), "dlm_scaled must have the plan's lmax"
        n = self.npix()
        out = torch.empty((3 if gamma else 2, n), dtype=torch.float64, device="cuda")
        _lib.call("nusht_ringplan_dlm2pointing", self._ring(1)._h, _ptr(dlm), _ptr(out[0]), _ptr(out[1]),
                  _ptr(out[2] if gamma else None), _stream())
        return out

    def gclm2lenpixs(self, gclm, mmax, spin, pixs, polrot=True, ptg=None, nthreads=None):
        """Lensed values at the pointing pixels `pixs`, by exact (direct) synthesis at their deflected
        positions (CPU_nufft_transformer.py:531-558).  Uses `ptg` if given, else the pointing of the last
        gclm2lenmap / dlm2pointing call.  Returns a 1-D tensor (spin 0) or the tuple (Q, U)."""
        assert spin >= 0, spin
        gclm = _dev(gclm, torch.complex128, "gclm")
        if ptg is None:
            assert self.ptg is not None, "no pointing yet: pass ptg, or run gclm2lenmap / dlm2pointing first"
            ptg = self.ptg
        ptg = _dev(ptg, torch.float64, "ptg")
        pixs = np.asarray(pixs, dtype=np.int64).reshape(-1)
        if pixs.size == 0 or pixs.min() < 0 or pixs.max() >= ptg.shape[1]:
            raise IndexError(f"pixs must be non-empty indices into the {ptg.shape[1]} pointing pixels")
        pix = torch.as_tensor(pixs, device="cuda")
        th = ptg[0].index_select(0, pix).contiguous()
        ph = ptg[1].index_select(0, pix).contiguous()
        gam = None
        if spin and polrot:
            assert self._gamma is not None, "polrot needs the rotation angle: compute the pointing from dlm_scaled"
            gam = self._gamma.index_select(0, pix).contiguous()
        plan = PointPlan(self.lmax, spin, th.cpu().numpy())
        out = plan.synthesis(gclm, self._ncomp(gclm, spin), ph, gam)
        torch.cuda.current_stream().synchronize()  # the plan is freed on return
        return out if spin == 0 else (out[0], out[1])

    def lensgclm(self, gclm, dlm, spin, lmax_out, nthreads=None, mmax=None, mmax_out=None, gclm_out=None,
                 polrot=True, out_sht_mode="STANDARD"):
        """Remapping alm -> lensed map -> adjoint synthesis on the pointing grid
        (CPU_nufft_transformer.py:948-971).  `dlm` is the scaled deflection, as for gclm2lenmap.
        out_sht_mode 'GRAD_ONLY' returns the gradient component only."""
        assert lmax_out == self.lmax and (mmax_out is None or mmax_out == self.lmax), \
            "lensgclm: lmax_out must equal the transformer's lmax"
        if out_sht_mode not in ("STANDARD", "GRAD_ONLY"):
            raise ValueError(f"unknown out_sht_mode {out_sht_mode}")
        m = self.gclm2lenmap(gclm, self.lmax, self.lmax, dlm_scaled=dlm, spin=spin, polrot=polrot)
        nc = 1 if (spin == 0 or out_sht_mode == "GRAD_ONLY") else 2
        return self.adjoint_synthesis(m, lmax=self.lmax, mmax=self.lmax, out=gclm_out, spin=spin, ncomp=nc)

    def dlm2A(self, dlm, dclm=None, exact: bool = True):
        """Determinant of the magnification matrix of the deflection with gradient dlm (and curl dclm) on
        the pointing grid (cunusht/helper_GPU.py:113-161).  Note: dlm, not dlm_scaled.  See helpers.dlm2A."""
        from .helpers import magnification_from_alm
        dlm = _dev(dlm, torch.complex128, "dlm").reshape(-1)
        assert dlm.numel() == Alm.getsize(self.lmax, self.lmax), "dlm must have the plan's lmax"
        dcl = None if dclm is None else _dev(dclm, torch.complex128, "dclm").reshape(-1)
        return magnification_from_alm(self, dlm, dcl, exact)

    def _pointing(self, dlm_scaled, ptg, want_gamma):
        """Current pointing from `ptg` or `dlm_scaled`; recomputed unless the array equals the one the
        current pointing came from (GPU_nufft_transformer.py:546-558)."""
        if ptg is not None:
            ptg = _dev(ptg, torch.float64, "ptg")
            if self._src_kind != "ptg" or not self._src.same(ptg):
                self.set_pts(ptg)
            return
        assert dlm_scaled is not None or self.dlm_scaled is not None, "Need to provide dlm_scaled if ptg is None"
        if dlm_scaled is None:
            dlm_scaled = self.dlm_scaled
        dlm = _dev(dlm_scaled, torch.complex128, "dlm_scaled")
        if self._src_kind != "dlm" or not self._src.same(dlm):
            pg = self.dlm2pointing(dlm, gamma=True)
            self.dlm_scaled = dlm
            self.ptg = pg[:2]
            self._gamma = pg[2].contiguous()
            gen = self._src.hold(dlm)
            self._src_kind = "dlm"
            self._cur_pts = (pg[0].contiguous(), pg[1].contiguous(), gen, False)

    def gclm2lenmap(self, gclm, lmax, mmax, ptg=None, dlm_scaled=None, epsilon=None, lenmap=None, verbose=1,
                    execmode="normal", runid=None, GPUdesc="", spin=None, polrot=True):
        """Lensed map(s) at the deflected pointing grid (GPU_nufft_transformer.py:486-578;
        spin-s with rotation as CPU_nufft_transformer.py:429-484).  execmode 'timing' stores and dumps the
        stage times (self.timer.keys), 'debug' returns the reference's intermediate arrays (self.ret)."""
        assert execmode in ["normal", "debug", "timing"]
        self.execmode = execmode
        run = lambda: self._gclm2lenmap(gclm, lmax, mmax, ptg, dlm_scaled, epsilon, lenmap, spin, polrot)
        if execmode == "timing":
            return diagnostics.timed(self, "gclm2lenmap", lmax, runid, GPUdesc, run)
        res = run()
        if execmode == "debug":
            sp = self.spin if spin is None else spin
            g = _dev(gclm, torch.complex128, "gclm")
            self.ret = diagnostics.debug_forward(self, g, sp, self._ncomp(g, sp), res)
            self.ret["dlm2pointing"] = self.ptg.cpu().numpy()
            print("::debug:: Returned component results")
            return self.ret
        return res

    def _gclm2lenmap(self, gclm, lmax, mmax, ptg, dlm_scaled, epsilon, lenmap, spin, polrot):
        self._check_lm(lmax, mmax)
        spin = self.spin if spin is None else spin
        eps = self._eps(epsilon)
        gclm = _dev(gclm, torch.complex128, "gclm")
        self._pointing(dlm_scaled, ptg, spin > 0)
        npts = self._cur_pts[0].numel()
        ncomp = self._ncomp(gclm, spin)
        plan = self._plan_for(spin, ncomp, eps, npts)
        self._ensure_pts(plan)
        out = self._out(lenmap, spin, npts)
        gam = self._gamma if (spin > 0 and polrot) else None
        if spin > 0 and polrot and gam is None:
            raise ValueError("polrot needs the rotation angle: pass dlm_scaled, not ptg")
        _lib.call("nusht_nusht2d2", plan._h, _ptr(gclm), _ptr(out[0]), _ptr(out[1] if spin else None), _ptr(gam),
                  _stream())
        return out[0] if spin == 0 else out

    def lenmap2gclm(self, lenmap, dlm_scaled, gclm, lmax, mmax, epsilon=None, ptg=None, verbose=1,
                    execmode="normal", runid=None, GPUdesc="", spin=None, polrot=True):
        """Exact adjoint of gclm2lenmap (GPU_nufft_transformer.py:580-673).  execmode 'timing' as for
        gclm2lenmap; 'debug' returns self.ret with the pointing and the result ('adjointsynthesis'): the
        adjoint chain is fused and forms no reference-layout intermediate."""
        assert execmode in ["normal", "debug", "timing"]
        self.execmode = execmode
        run = lambda: self._lenmap2gclm(lenmap, dlm_scaled, gclm, lmax, mmax, epsilon, ptg, spin, polrot)
        if execmode == "timing":
            return diagnostics.timed(self, "lenmap2gclm", lmax, runid, GPUdesc, run)
        res = run()
        if execmode == "debug":
            self.ret = {"dlm2pointing": self.ptg.cpu().numpy(), "adjointsynthesis": res.cpu().numpy()}
            print("::debug:: Returned component results")
            return self.ret
        return res

    def _lenmap2gclm(self, lenmap, dlm_scaled, gclm, lmax, mmax, epsilon, ptg, spin, polrot):
        self._check_lm(lmax, mmax)
        spin = self.spin if spin is None else spin
        eps = self._eps(epsilon)
        self._pointing(dlm_scaled, ptg, spin > 0)
        npts = self._cur_pts[0].numel()
        vals = self._values(lenmap, spin)
        ncomp = 1 if spin == 0 else (gclm.shape[0] if gclm is not None and gclm.dim() == 2 else 2)
        plan = self._plan_for(spin, ncomp, eps, npts)
        self._ensure_pts(plan)
        res = self._alm_out(gclm, ncomp)
        gam = self._gamma if (spin > 0 and polrot) else None
        _lib.call("nusht_nusht2d1", plan._h, _ptr(vals[0]), _ptr(vals[1]), _ptr(gam), _ptr(res), _stream())
        return res

    # ------------------------------------------------------------------ SHT on the pointing grid
    def synthesis(self, alm, out=None, lmax=None, mmax=None, spin=0):
        """alm -> map on the pointing grid (GPU_nufft_transformer.py:353-367)."""
        self._check_lm(self.lmax if lmax is None else lmax, self.lmax if mmax is None else mmax)
        alm = _dev(alm, torch.complex128, "alm")
        outs = self._out(out, spin, self.npix())
        _lib.call("nusht_ringplan_synthesis", self._ring(spin)._h, _ptr(alm), self._ncomp(alm, spin),
                  _ptr(outs[0]), _ptr(outs[1] if spin else None), _stream())
        return outs[0] if spin == 0 else outs

    def adjoint_synthesis(self, synthmap, lmax=None, mmax=None, out=None, spin=0, ncomp=None, apply_weights=True):
        """Adjoint synthesis on the pointing grid with the pixel weights applied, i.e. quadrature analysis
        (GPU_nufft_transformer.py:372-383 -> SHTns analys, GPU_sht_transformer.py:65-73; the CPU backend's
        ducc adjoint_synthesis with apply_weights, CPU_sht_transformer.py:37-48).  apply_weights=False gives
        the exact transpose of synthesis."""
        self._check_lm(self.lmax if lmax is None else lmax, self.lmax if mmax is None else mmax)
        vals = self._values(synthmap, spin)
        nc = 1 if spin == 0 else (2 if ncomp is None else ncomp)
        res = self._alm_out(out, nc)
        _lib.call("nusht_ringplan_analysis" if apply_weights else "nusht_ringplan_adjoint", self._ring(spin)._h,
                  _ptr(vals[0]), _ptr(vals[1]), _ptr(res), nc, _stream())
        return res

    # ------------------------------------------------------------------ buffers
    def _out(self, given, spin, n):
        if spin == 0:
            if isinstance(given, torch.Tensor) and given.is_cuda and given.dtype == torch.float64 \
                    and given.numel() == n and given.is_contiguous():
                return (given.reshape(-1), None)
            return (torch.empty(n, dtype=torch.float64, device="cuda"), None)
        if isinstance(given, torch.Tensor) and given.is_cuda and given.dtype == torch.float64 \
                and tuple(given.shape) == (2, n) and given.is_contiguous():
            return given
        return torch.empty((2, n), dtype=torch.float64, device="cuda")

    def _values(self, m, spin):
        if isinstance(m, torch.Tensor) and m.is_complex():
            m = m.real if spin == 0 else torch.stack([m.real, m.imag]) if m.dim() == 1 else m.real
        m = _dev(m, torch.float64, "map")
        if spin == 0:
            return (m.reshape(-1).contiguous(), None)
        assert m.shape[0] == 2, "spin > 0 maps must be (2, npix) = (Q, U)"
        return (m[0].contiguous(), m[1].contiguous())

    def _alm_out(self, given, ncomp):
        n = Alm.getsize(self.lmax, self.lmax)
        if isinstance(given, torch.Tensor) and given.is_cuda and given.dtype == torch.complex128 \
                and given.numel() == ncomp * n and given.is_contiguous():
            return given.reshape(ncomp, n) if ncomp > 1 else given.reshape(-1) if given.dim() == 1 else given
        shape = (ncomp, n) if ncomp > 1 else (n,)
        return torch.empty(shape, dtype=torch.complex128, device="cuda")
