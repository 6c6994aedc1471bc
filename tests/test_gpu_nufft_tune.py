"""GPU tests of the plan-time nuFFT autotuner and the register row-window spreading, through the C ABI.

* Plans created with sigma < 0 time their candidates (fine-grid oversampling x spreading tile height) and keep
  the fastest; whichever they keep, nusht2d2 matches direct Y_lm summation (10 eps FP64, 30 eps FP32) and
  nusht2d1 the direct adjoint.
* Every oversampling the tuner may pick at the headline eps 1e-7 (1.25, 1.5, 2) and every tile height of the
  row-window spreading (1 .. 32 rows) gives the direct adjoint.
* The row-window spreading equals the LDS-atomic spreading (NUSHT_SPREAD_ROWS=0) to 1e-12, also when tiles
  hold more points than one pass of the kernel's LDS records (many points on a small grid).
* Workgroups that walk several tiles down a column strip (NUSHT_SPREAD_SEG, carried halo rows, plain stores
  for cells no other workgroup touches) give the direct adjoint for every segment length and tile height,
  including one segment for the whole strip.
"""
import numpy as np
import pytest
import torch

from oracle import almutil, sht as osht

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _relmax(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _points(rng, n):
    return np.arccos(rng.uniform(-1, 1, n)), rng.uniform(0, 2 * np.pi, n)


def _T(L, eps, spin, sigma=0.0):
    import cunusht_amd
    return cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=2, epsilon=eps, spin=spin, sigma=sigma)


def _pair(T, L, spin, th, ph, rng):
    """nusht2d2 and nusht2d1 at the points, with the direct references."""
    nc = 1 if spin == 0 else 2
    alm = almutil.random_alm(L, spin, nc, rng)
    a = alm[0] if spin == 0 else alm
    loc = _t(np.stack([th, ph], 1))
    v = T.nusht2d2(L, L, _t(a), loc, spin=spin).cpu().numpy()
    x = rng.standard_normal(th.size) + (1j * rng.standard_normal(th.size) if spin else 0)
    xin = _t(x.real if spin == 0 else np.array([x.real, x.imag]))
    ad = T.nusht2d1(L, L, xin, loc, spin=spin).cpu().numpy()
    got = v if spin == 0 else v[0] + 1j * v[1]
    return got, osht.synth_direct(a, L, spin, th, ph), ad, osht.adjoint_direct(x, L, spin, th, ph)


@pytest.mark.parametrize("eps", [1e-5, 1e-7, 1e-10])
@pytest.mark.parametrize("spin", [0, 2])
def test_autotuned_plan_parity(eps, spin):
    L = 48
    rng = np.random.default_rng(800 + spin)
    T = _T(L, eps, spin, sigma=-1.0)
    th, ph = _points(rng, 3000)
    print(f"case autotuned plan eps={eps} spin={spin}", flush=True)
    got, ref, ad, refad = _pair(T, L, spin, th, ph, rng)
    info = next(iter(T._plans.values())).info()
    print(f"  chose sigma {info['sigma']:.4f} w {info['kernel_width']} tile {info['nufft_tile_rows']}x"
          f"{info['nufft_tile_cols']} of {info['nufft_candidates']} candidates", flush=True)
    assert info["nufft_candidates"] >= 1
    tol = 30 * eps if eps >= 1e-6 else 10 * eps
    assert _relmax(got, ref) < tol
    # the adjoint's max error relative to the largest coefficient: 30 eps, as in the small-lmax adjoint tests
    assert _relmax(ad.reshape(refad.shape), refad) < 30 * eps


@pytest.mark.parametrize("sigma", [1.25, 1.5, 2.0])
@pytest.mark.parametrize("spin", [0, 2])
def test_every_tuner_sigma_at_1e7(sigma, spin):
    L = 40
    rng = np.random.default_rng(810 + spin)
    T = _T(L, 1e-7, spin, sigma=sigma)
    th, ph = _points(rng, 2000)
    print(f"case sigma {sigma} spin {spin} eps 1e-7", flush=True)
    got, ref, ad, refad = _pair(T, L, spin, th, ph, rng)
    assert _relmax(got, ref) < 1e-6
    assert _relmax(ad.reshape(refad.shape), refad) < 1e-6


def _rows_kernel_used(info):
    """The plan spreads with the register row-window kernel (tiles 16 (17 - w) columns wide)."""
    used = info["nufft_spread_rows"] == 1
    assert not used or info["nufft_tile_cols"] == 16 * (17 - info["kernel_width"])
    return used


@pytest.mark.parametrize("rows", [1, 4, 8, 16, 32])
@pytest.mark.parametrize("eps", [1e-7, 1e-10])
def test_spread_rows_every_tile_height(rows, eps, monkeypatch):
    monkeypatch.setenv("NUSHT_SPREAD_TR", str(rows))
    L = 56
    rng = np.random.default_rng(820 + rows)
    T = _T(L, eps, 2)
    th, ph = _points(rng, 2500)
    th[:2] = [0.0, np.pi]
    print(f"case spread tile rows {rows}", flush=True)
    got, ref, ad, refad = _pair(T, L, 2, th, ph, rng)
    info = next(iter(T._plans.values())).info()
    assert _rows_kernel_used(info) and info["nufft_tile_rows"] == rows
    assert _relmax(got, ref) < 10 * eps
    assert _relmax(ad.reshape(refad.shape), refad) < 10 * eps


@pytest.mark.parametrize("eps,spin,sigma,L,npts", [(1e-7, 2, 0.0, 24, 400_000), (1e-10, 0, 0.0, 24, 400_000),
                                                   (1e-6, 2, 0.0, 24, 400_000), (1e-10, 0, 128 / 97, 48, 3000),
                                                   (1e-12, 2, 0.0, 40, 5000), (1e-5, 0, 0.0, 64, 20_000)])
def test_spread_rows_equals_lds_atomic_kernel(eps, spin, sigma, L, npts, monkeypatch):
    """Both spreading kernels on the same points: many points on a small grid (tiles spread in several passes),
    the oversampling 128/97 (w 15) the tuner picks at lmax 48 and eps 1e-10, and the widths of eps 1e-12 and
    1e-5."""
    rng = np.random.default_rng(830 + spin + L)
    th, ph = _points(rng, npts)
    x = rng.standard_normal((2, th.size))
    loc = _t(np.stack([th, ph], 1))
    xin = _t(x if spin else x[0])
    outs = []
    for rows in ("1", "0"):
        monkeypatch.setenv("NUSHT_SPREAD_ROWS", rows)
        T = _T(L, eps, spin, sigma=sigma)
        print(f"case spread kernels rows={rows} eps={eps} spin={spin}", flush=True)
        outs.append(T.nusht2d1(L, L, xin, loc, spin=spin).cpu().numpy())
        assert _rows_kernel_used(next(iter(T._plans.values())).info()) == (rows == "1")
        del T
    tol = 1e-12 if eps < 1e-6 else 1e-5  # FP32 grid: the global flush adds FP32 values in a varying order
    assert _relmax(outs[0], outs[1]) < tol


@pytest.mark.parametrize("seg,rows", [(2, 8), (3, 1), (5, 4), (7, 16), (1000, 8), (1000, 32)])
@pytest.mark.parametrize("eps,spin", [(1e-7, 2), (1e-10, 0), (1e-5, 2)])
def test_spread_strip_segments(seg, rows, eps, spin, monkeypatch):
    monkeypatch.setenv("NUSHT_SPREAD_SEG", str(seg))
    monkeypatch.setenv("NUSHT_SPREAD_TR", str(rows))
    L = 64
    rng = np.random.default_rng(840 + seg + rows)
    th, ph = _points(rng, 6000)
    th[:2] = [0.0, np.pi]
    T = _T(L, eps, spin)
    print(f"case spread strip segments seg={seg} rows={rows} eps={eps} spin={spin}", flush=True)
    x = rng.standard_normal(th.size) + (1j * rng.standard_normal(th.size) if spin else 0)
    xin = _t(x.real if spin == 0 else np.array([x.real, x.imag]))
    ad = T.nusht2d1(L, L, xin, _t(np.stack([th, ph], 1)), spin=spin).cpu().numpy()
    info = next(iter(T._plans.values())).info()
    assert _rows_kernel_used(info) and info["nufft_tile_rows"] == rows
    refad = osht.adjoint_direct(x, L, spin, th, ph)
    tol = 30 * eps if eps >= 1e-6 else 10 * eps
    assert _relmax(ad.reshape(refad.shape), refad) < tol
