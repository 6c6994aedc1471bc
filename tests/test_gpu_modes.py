"""GPU tests of execmode 'timing' (stage times in the reference's JSON layout) and 'debug' (the reference's
intermediate arrays on its own CAR grid), and of the 2D FFT on the reference's non-smooth doubled sizes.

Tolerances: intermediates within 1e-12 relative of the oracle's stages (FP64 throughout); the 2D FFT on
non-smooth sides within 1e-12 of numpy.fft.
"""
import ctypes
import json
import os

import numpy as np
import pytest
import torch

from oracle import almutil, clib, pipeline
from oracle.geometry import reference_cc_sizes

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


def test_timing_mode_json(tmp_path, monkeypatch):
    import cunusht_amd
    monkeypatch.setenv("NUSHT_TIMING_DIR", str(tmp_path))
    L = 32
    rng = np.random.default_rng(1)
    g = almutil.random_alm(L, 2, 2, rng)
    dlm_np = _deflection(L, rng)
    dlm = _t(dlm_np)
    T0 = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=2, epsilon=1e-10, spin=2)
    plain = T0.gclm2lenmap(_t(g), L, L, dlm_scaled=dlm).cpu().numpy()
    # a fresh transformer: the pointing is computed inside the timed call (a transformer that has seen
    # these deflection values already would reuse its pointing, as the reference's equality compare does)
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=2, epsilon=1e-10, spin=2)
    timed = T.gclm2lenmap(_t(g), L, L, dlm_scaled=_t(dlm_np), execmode="timing", runid=3,
                          GPUdesc="mi355x").cpu().numpy()
    assert np.array_equal(plain, timed)
    fn = tmp_path / "GPU_nusht_transformer" / "gclm2lenmap" / "lmax32_epsilon1e-10_run03_mi355x"
    keys = json.loads(fn.read_text())
    for k in ("spin-1 synth", "pointing", "setpts", "synthesis", "doubling", "C2C", "nuFFT2d2", "gclm2lenmap"):
        assert keys[k] > 0, k
    assert keys == T.timer.keys
    x = _t(np.random.default_rng(2).standard_normal((2, T.npix())))
    T.lenmap2gclm(x, dlm, None, L, L, execmode="timing")
    keys = json.loads((tmp_path / "GPU_nusht_transformer" / "lenmap2gclm" / "lmax32_epsilon1e-10_run00_").read_text())
    for k in ("nuFFT2d1", "iC2C", "adjointdoubling", "adjointsynthesis", "lenmap2gclm"):
        assert keys[k] > 0, k


@pytest.mark.parametrize("spin", [0, 2])
def test_debug_mode_intermediates_vs_oracle(spin):
    """The reference's _synthesis / _doubling / _C2C arrays (theta-contiguous, its CAR sizes) vs the oracle's
    stages, and the final values equal to the normal call."""
    import cunusht_amd
    L = 40
    rng = np.random.default_rng(10 + spin)
    nc = 1 if spin == 0 else 2
    g = almutil.random_alm(L, spin, nc, rng)
    gin = g[0] if spin == 0 else g
    dlm = _t(_deflection(L, rng))
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=2, epsilon=1e-10, spin=spin)
    plain = T.gclm2lenmap(_t(gin), L, L, dlm_scaled=dlm).cpu().numpy()
    ret = T.gclm2lenmap(_t(gin), L, L, dlm_scaled=dlm, execmode="debug")
    assert set(ret) >= {"synthesis", "doubling", "C2C", "nuFFT2d2", "dlm2pointing"}
    nt, nph = reference_cc_sizes(L)
    op = pipeline.NuSHT(L)
    G = clib.sht_synth_rings(gin if spin else gin[None, :], L, spin, op.thetas)
    cmap = pipeline.phi_synth(G, L, spin, nph)                      # (nt, nph) ring-major
    assert _relmax(ret["synthesis"].reshape(nph, nt), cmap.T) < 1e-12
    d = op._doubling(cmap, spin)                                      # (2 nt - 2, nph)
    dref = d.real if spin == 0 else d
    assert _relmax(ret["doubling"].reshape(nph, 2 * nt - 2), dref.T) < 1e-12
    F = np.fft.fft2(d) / d.size
    assert _relmax(ret["C2C"], F.T) < 1e-12
    vals = ret["nuFFT2d2"]
    want = plain if spin == 0 else plain[0] + 1j * plain[1]
    assert np.array_equal(vals, want)


@pytest.mark.parametrize("L", [40, 1023, 4095])
def test_fft2d_on_reference_doubled_sizes(L):
    """Sides (nphi, 2 ntheta - 2) of the reference's doubled CAR grid: 2 ntheta - 2 is never smooth there
    (8230 = 2 5 823 at lmax 4095), so the rows take the split / Bluestein DFT."""
    from cunusht_amd import _lib
    nt, nph = reference_cc_sizes(L)
    n0, n1 = (nph, 2 * nt - 2) if L < 4095 else (16, 2 * nt - 2)  # a 16-row slab at lmax 4095
    rng = np.random.default_rng(L)
    x = rng.standard_normal((n0, n1)) + 1j * rng.standard_normal((n0, n1))
    d = _t(x)
    _lib.call("nusht_fft2d_c2c_f64", ctypes.c_void_p(d.data_ptr()), n0, n1, 1, -1,
              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert _relmax(d.cpu().numpy(), np.fft.fft2(x)) < 1e-12
    # and the transposed layout (non-smooth columns)
    y = _t(x.T)
    _lib.call("nusht_fft2d_c2c_f64", ctypes.c_void_p(y.data_ptr()), n1, n0, 1, 1,
              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert _relmax(y.cpu().numpy(), np.fft.ifft2(x.T) * x.size) < 1e-12
