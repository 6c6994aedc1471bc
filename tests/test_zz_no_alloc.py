"""GPU test of the allocation-free point synthesis (nusht_pointplan_synthesis_noalloc) against the entry point it
is a variant of (nusht_pointplan_synthesis on a point plan), both through the C ABI.

Two calls in a row on one GPU, the second with a smaller input (fewer points, lower lmax), each against the
point plan: FP64 results within 1e-11 of the largest value (the two evaluate the same sums with different
recursion tables and summation orders).  Points include both poles and colatitudes within 1e-3 of them, where
high-m terms underflow.  Sizes beyond the scratch array are refused before any GPU work.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import almutil

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _p(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _s():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _points(rng, n):
    th = np.arccos(rng.uniform(-1, 1, n))
    th[:4] = [0.0, np.pi, 1e-3, np.pi - 5e-4]
    return th, rng.uniform(0, 2 * np.pi, n)


def _both(lmax, spin, ncomp, th, ph, rng, with_gamma):
    from cunusht_amd import _lib
    from cunusht_amd.transformer import PointPlan
    n = th.size
    alm = _t(almutil.random_alm(lmax, spin, ncomp, rng))
    phi = _t(ph)
    gam = _t(rng.uniform(-0.1, 0.1, n)) if with_gamma else None
    theta = _t(th)
    out = torch.full((2, n), np.nan, dtype=torch.float64, device="cuda")
    print(f"case noalloc lmax={lmax} spin={spin} ncomp={ncomp} npts={n}", flush=True)
    _lib.call("nusht_pointplan_synthesis_noalloc", lmax, spin, n, _p(theta), _p(alm), ncomp, _p(phi), _p(gam),
              _p(out[0]), _p(out[1]) if spin else None, _s())
    ref = PointPlan(lmax, spin, th).synthesis(alm, ncomp, phi, gam)
    torch.cuda.synchronize()
    got = out[0] if spin == 0 else out
    r = ref.cpu().numpy()
    err = float(np.max(np.abs(got.cpu().numpy() - r)) / np.max(np.abs(r)))
    print(f"  max |noalloc - point plan| / max = {err:.2e}", flush=True)
    return err


@pytest.mark.parametrize("spin,ncomp", [(0, 1), (2, 2), (1, 1)])
def test_noalloc_matches_point_plan_two_calls(spin, ncomp):
    rng = np.random.default_rng(700 + spin)
    th, ph = _points(rng, 1000)
    assert _both(1023, spin, ncomp, th, ph, rng, with_gamma=spin > 0) < 1e-11
    th2, ph2 = _points(rng, 37)
    assert _both(300, spin, ncomp, th2, ph2, rng, with_gamma=False) < 1e-11


def test_noalloc_lmax4095_spin2():
    rng = np.random.default_rng(710)
    th, ph = _points(rng, 256)
    assert _both(4095, 2, 2, th, ph, rng, with_gamma=True) < 1e-11


def test_noalloc_refuses_sizes_beyond_scratch():
    from cunusht_amd import _lib
    from cunusht_amd._lib import NuShtError
    n = 300
    th, ph = _points(np.random.default_rng(711), n)
    theta, phi = _t(th), _t(ph)
    alm = _t(np.zeros((2, almutil.nalm(4095)), dtype=np.complex128))
    out = torch.zeros((2, n), dtype=torch.float64, device="cuda")
    with pytest.raises(NuShtError):  # 8191 rows x 300 points > 2^21
        _lib.call("nusht_pointplan_synthesis_noalloc", 4095, 2, n, _p(theta), _p(alm), 2, _p(phi), None,
                  _p(out[0]), _p(out[1]), _s())
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0
