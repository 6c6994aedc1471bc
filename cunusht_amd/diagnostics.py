"""execmode='timing' and execmode='debug' of the lensing operators.

timing (cunusht/helper_GPU.py:10-51, cunusht/utils.py:201-303): the library's stage timer (hipEvents on the
launch stream, no device-wide synchronization between stages) is read after the call and folded into the
reference's timer keys; the flat {key: seconds} dictionary is kept on `transformer.timer.keys` and written as
JSON to <root>/<class>/<function>/lmax<L>_epsilon<eps>_run<runid:02d>_<GPUdesc>, root = $NUSHT_TIMING_DIR or
./timings (the reference writes below its package's test/benchmark/timings).

debug (cunusht/helper_GPU.py:76-89): the forward operator also returns the reference's intermediate arrays,
recomputed with the stage entry points on the reference's own CAR grid (GPU_nufft_transformer.py:138-141):
'synthesis' (CAR map, theta-contiguous (nphi, ntheta), flattened), 'doubling' (the doubled map (nphi,
2 ntheta - 2), flattened), 'C2C' (its 2D FFT with norm 'forward', (nphi, 2 ntheta - 2)) and 'nuFFT2d2' (the
values at the points).  The fused pipeline never forms these arrays; they are for inspection only.
"""
from __future__ import annotations

import json
import os
import time

from .geometry import Geom, good_size

# library stage -> the reference's timer key (its decorated function names, helper_GPU.py:13)
_KEYS = (("legendre_prep_ring_s1", "spin-1 synth"), ("legendre_synth_ring_s1", "spin-1 synth"),
         ("ring_dft_pointing", "pointing"), ("set_pts", "setpts"),
         ("legendre_prep_cc", "synthesis"), ("legendre_synth_cc", "synthesis"),
         ("theta_pass_t2", "doubling"), ("phi_pass_t2", "C2C"), ("nufft_interp", "nuFFT2d2"),
         ("grid_zero", "nuFFT2d1"), ("nufft_spread", "nuFFT2d1"), ("phi_pass_t1", "iC2C"),
         ("theta_pass_t1", "adjointdoubling"), ("legendre_adjoint_cc", "adjointsynthesis"))


def stage_key(stage: str) -> str:
    for prefix, key in _KEYS:
        if stage.startswith(prefix):
            return key
    return stage


class Timer:
    """Flat {key: seconds} of the last timed call (the reference timer's `keys`)."""

    def __init__(self):
        self.keys: dict = {}

    def dumpjson(self, fn: str) -> None:
        with open(fn, "w") as f:
            json.dump(self.keys, f)

    def __repr__(self):
        return "\n".join(f"{k:>20s}: {v * 1e3:9.3f} ms" for k, v in self.keys.items())


def timing_root() -> str:
    return os.environ.get("NUSHT_TIMING_DIR", os.path.join(os.getcwd(), "timings"))


def dump_timer(transformer, name: str, fname: str) -> str:
    """Writes transformer.timer.keys as JSON to <root>/<class>/<name>/<fname>; returns the path."""
    d = os.path.join(timing_root(), type(transformer).__name__, name)
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, fname)
    transformer.timer.dumpjson(path)
    return path


def timed(transformer, name: str, lmax: int, runid, GPUdesc: str, fn):
    """Run fn() with stage timing and store / dump the timer keys."""
    import torch

    from . import _lib
    _lib.timing_reset()
    _lib.timing_enable(True)
    try:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = fn()
        torch.cuda.synchronize()
        total = time.perf_counter() - t0
        stages = _lib.timing_read()
    finally:
        _lib.timing_enable(False)
    keys = {}
    for st, (ms, _calls) in stages.items():
        k = stage_key(st)
        keys[k] = keys.get(k, 0.0) + ms / 1e3
    keys[name] = total
    transformer.timer.keys = keys
    dump_timer(transformer, name,
               "lmax{}_epsilon{}_run{:02d}_{}".format(lmax, transformer.epsilon, int(runid or 0), GPUdesc))
    print(15 * "- " + "Timing {}: {:.3f} seconds".format(name, total) + 15 * "- ")
    return res


def reference_car_sizes(lmax: int):
    """(ntheta, nphi) of the reference's CAR grid (GPU_nufft_transformer.py:138-140)."""
    return (good_size(lmax + 2) + 3) // 4 * 4, 2 * good_size(lmax + 1)


def debug_forward(transformer, alm, spin: int, ncomp: int, values) -> dict:
    """The reference's forward intermediates on its CAR grid (see module docstring)."""
    import torch

    from . import _lib
    from .transformer import RingSHT, _ptr, _stream
    L = transformer.lmax
    nt, nph = reference_car_sizes(L)
    car_geom = Geom.get_cc_geometry(nt, nph)
    rs = RingSHT(car_geom, L, spin)
    maps = torch.empty((2, nt * nph), dtype=torch.float64, device="cuda")
    _lib.call("nusht_ringplan_synthesis", rs._h, _ptr(alm), ncomp, _ptr(maps[0]), _ptr(maps[1] if spin else None),
              _stream())
    ntd = 2 * nt - 2
    if spin == 0:
        car = maps[0].reshape(nt, nph).t().contiguous()  # theta-contiguous (nphi, ntheta)
        dmap = torch.empty((nph, ntd), dtype=torch.float64, device="cuda")
        _lib.call("nusht_doubling_f64", _ptr(car), nt, nph, 0, _ptr(dmap), _stream())
        cdmap = dmap.to(torch.complex128)
    else:
        car = torch.complex(maps[0], maps[1]).reshape(nt, nph).t().contiguous()
        cdmap = torch.empty((nph, ntd), dtype=torch.complex128, device="cuda")
        _lib.call("nusht_doubling_c128", _ptr(car), nt, nph, spin, _ptr(cdmap), _stream())
        dmap = cdmap
    fc = cdmap.clone()
    _lib.call("nusht_fft2d_c2c_f64", _ptr(fc), nph, ntd, 1, -1, _stream())
    fc /= nph * ntd
    vals = values if spin == 0 else torch.complex(values[0], values[1])
    torch.cuda.synchronize()
    return {"synthesis": car.reshape(-1).cpu().numpy(), "doubling": dmap.reshape(-1).cpu().numpy(),
            "C2C": fc.cpu().numpy(), "nuFFT2d2": vals.cpu().numpy()}
