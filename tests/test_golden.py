"""Checks against the committed golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py from
the oracle's exact routines: direct summation at the target points, closed-form pointing, the doubling
index map, parallel transport).

CPU: the oracle and this package's CPU backend still reproduce the frozen vectors (so the oracle cannot drift
silently); the nuFFT paths at epsilon 1e-12 within 1e-11 of the largest value.
GPU (marked): the HIP library through its C ABI at epsilon 1e-12 within 1e-9 relative max error; the
pointing within 1e-12 rad; the doubling bit-exact.
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import almutil, pipeline, pointing, transport
from oracle.geometry import Rings

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _rings(g):
    return Rings(g["r_theta"], g["r_phi0"], g["r_nph"], g["r_ofs"], g["r_weight"])


def _relmax(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _angle_err(a, b):
    return float(np.max(np.abs((np.asarray(a) - np.asarray(b) + np.pi) % (2 * np.pi) - np.pi)))


def test_golden_files_present():
    names = sorted(f for f in os.listdir(GOLD) if f.endswith(".npz"))
    assert names == sorted(["doubling_7x10.npz", "lensing_spin0_gl16.npz", "lensing_spin2_gl16.npz",
                            "pointing_gl31.npz", "points_spin0_l20.npz", "points_spin2_l20.npz",
                            "transport_spin2_gl12.npz"])


# ------------------------------------------------------------------ oracle and CPU backend (CPU)
@pytest.mark.parametrize("spin", [0, 2])
def test_oracle_reproduces_lensing_vectors(spin):
    g = _load(f"lensing_spin{spin}_gl16")
    L, rings = int(g["lmax"]), _rings(g)
    gclm = g["gclm"][0] if spin == 0 else g["gclm"]
    out = pipeline.gclm2lenmap(gclm, g["dlm_scaled"], L, spin, rings, 1e-12)
    assert _relmax(np.atleast_2d(out), g["lenmap"]) < 1e-11
    tp, pp, gam, _ = pipeline.dlm2pointing(g["dlm_scaled"], L, rings)
    assert np.max(np.abs(tp - g["ptg_theta"])) < 1e-14 and _angle_err(pp, g["ptg_phi"]) < 1e-13
    assert _angle_err(gam, g["gamma"]) < 1e-13
    x = g["adj_in"][0] if spin == 0 else g["adj_in"]
    ad = pipeline.lenmap2gclm(x, g["dlm_scaled"], L, spin, 1 if spin == 0 else 2, rings, 1e-12)
    assert _relmax(np.atleast_2d(ad), g["adj_out"]) < 1e-11


@pytest.mark.parametrize("spin", [0, 2])
def test_oracle_reproduces_point_vectors(spin):
    g = _load(f"points_spin{spin}_l20")
    L = int(g["lmax"])
    op = pipeline.NuSHT(L)
    alm = g["alm"] if spin else g["alm"][0]
    v = op.synthesis_general(alm, spin, g["loc"][:, 0], g["loc"][:, 1], 1e-12)
    assert _relmax(v, g["values"]) < 1e-11
    a = op.adjoint_synthesis_general(g["adj_in"], spin, 1 if spin == 0 else 2, g["loc"][:, 0], g["loc"][:, 1], 1e-12)
    assert _relmax(np.atleast_2d(a), g["adj_out"]) < 1e-11


def test_oracle_reproduces_pointing_vectors():
    g = _load("pointing_gl31")
    th, ph = _rings(g).pixel_angles()
    tp, pp = pointing.deflected_angles(th, ph, g["d_theta"], g["d_phi"])
    assert np.max(np.abs(tp - g["ptg_theta"])) < 1e-15 and _angle_err(pp, g["ptg_phi"]) < 1e-14
    assert _angle_err(pointing.gamma(th, g["d_theta"], g["d_phi"]), g["gamma"]) < 1e-14
    # and the independent 3-vector construction
    t3, p3, g3 = pointing.rotation_3d(th, ph, g["d_theta"], g["d_phi"])
    assert np.max(np.abs(t3 - g["ptg_theta"])) < 1e-13 and _angle_err(g3, g["gamma"]) < 1e-12


def test_doubling_vectors_vs_roll_formula():
    g = _load("doubling_7x10")
    M, Dm = g["map"], g["dmap"]
    nt, nph = M.shape
    for spin in (0, 1, 2):
        sg = -1.0 if spin % 2 else 1.0
        D = np.concatenate([M, sg * np.roll(M[nt - 2:0:-1], -(nph // 2), axis=1)])
        assert np.array_equal(D, g[f"doubled_s{spin}"])
        A = Dm[:nt].copy()
        A[1:nt - 1] += sg * np.roll(Dm[:nt - 1:-1][:nt - 2], -(nph // 2), axis=1)
        assert np.array_equal(A, g[f"adjoint_s{spin}"])


def test_transport_vector_and_oracle_pipeline():
    g = _load("transport_spin2_gl12")
    L, rings = int(g["lmax"]), _rings(g)
    th, ph = rings.pixel_angles()
    _, _, _, (dt, dp) = pipeline.dlm2pointing(g["dlm_scaled"], L, rings)
    f = transport.lensed_ambient_field(2, g["tensor"], th, ph, dt, dp)
    assert _relmax(np.array([f.real, f.imag]), g["lenmap"]) < 1e-13
    out = pipeline.gclm2lenmap(g["gclm"], g["dlm_scaled"], L, 2, rings, 1e-12)
    assert _relmax(out, g["lenmap"]) < 1e-11


@pytest.mark.parametrize("spin", [0, 2])
def test_cpu_backend_reproduces_lensing_vectors(spin):
    import cunusht_amd
    g = _load(f"lensing_spin{spin}_gl16")
    L = int(g["lmax"])
    T = cunusht_amd.get_transformer("CPU", solver="duccnufft")(("gl", {"lmax": L, "nphi": 2 * L + 1}))
    ll = np.arange(L + 1)
    dlm = almutil.almxfl(g["dlm_scaled"], np.sqrt(ll * (ll + 1.0)), L)  # the CPU class takes the unscaled dlm
    out = T.gclm2lenmap(g["gclm"], dlm, L, L, spin, 1e-12, 4)
    assert _relmax(np.atleast_2d(out), g["lenmap"]) < 1e-10
    x = g["adj_in"][0] if spin == 0 else g["adj_in"]
    ad = T.lenmap2gclm(x, dlm, np.zeros_like(g["gclm"]), spin, L, L, 4, epsilon=1e-12)
    assert _relmax(np.atleast_2d(ad), g["adj_out"]) < 1e-10


# ------------------------------------------------------------------ the HIP library (GPU)
def _dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _T(g, spin, nphi=None):
    import cunusht_amd
    L = int(g["lmax"])
    info = {"lmax": L} if nphi is None else {"lmax": L, "nphi": nphi}
    T = cunusht_amd.get_transformer("GPU")(("gl", info), nuFFTtype=2, epsilon=1e-12, spin=spin)
    assert np.max(np.abs(T.geom.theta - g["r_theta"])) < 1e-14
    return T


@pytest.mark.gpu
@pytest.mark.parametrize("spin", [0, 2])
def test_gpu_lensing_vectors(spin):
    import torch
    g = _load(f"lensing_spin{spin}_gl16")
    L = int(g["lmax"])
    T = _T(g, spin, nphi=2 * L + 1)
    gclm = g["gclm"][0] if spin == 0 else g["gclm"]
    dlm = _dev(g["dlm_scaled"])
    print(f"case golden lensing spin={spin}", flush=True)
    out = T.gclm2lenmap(_dev(gclm), L, L, dlm_scaled=dlm, spin=spin)
    ptg = T.dlm2pointing(dlm, gamma=True)
    x = g["adj_in"][0] if spin == 0 else g["adj_in"]
    ad = T.lenmap2gclm(_dev(x), dlm, None, L, L, spin=spin)
    torch.cuda.synchronize()
    assert _relmax(np.atleast_2d(out.cpu().numpy()), g["lenmap"]) < 1e-9
    p = ptg.cpu().numpy()
    assert np.max(np.abs(p[0] - g["ptg_theta"])) < 1e-12 and _angle_err(p[1], g["ptg_phi"]) < 1e-12
    assert _angle_err(p[2], g["gamma"]) < 1e-12
    assert _relmax(np.atleast_2d(ad.cpu().numpy()), g["adj_out"]) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("spin", [0, 2])
def test_gpu_point_vectors(spin):
    import torch
    import cunusht_amd
    g = _load(f"points_spin{spin}_l20")
    L = int(g["lmax"])
    T = cunusht_amd.get_transformer("GPU")(("gl", {"lmax": L}), nuFFTtype=2, epsilon=1e-12, spin=spin)
    loc = _dev(g["loc"])
    alm = g["alm"] if spin else g["alm"][0]
    print(f"case golden points spin={spin}", flush=True)
    v = T.nusht2d2(L, L, _dev(alm), loc, spin=spin)
    x = g["adj_in"]
    xin = _dev(x.real if spin == 0 else np.array([x.real, x.imag]))
    a = T.nusht2d1(L, L, xin, loc, spin=spin)
    torch.cuda.synchronize()
    vv = v.cpu().numpy()
    got = vv if spin == 0 else vv[0] + 1j * vv[1]
    assert _relmax(got, g["values"]) < 1e-9
    assert _relmax(np.atleast_2d(a.cpu().numpy()), g["adj_out"]) < 1e-9


@pytest.mark.gpu
def test_gpu_pointing_vectors():
    import torch
    from cunusht_amd import _lib
    g = _load("pointing_gl31")
    n = g["d_theta"].size
    ring = [_dev(g["r_theta"]), _dev(g["r_phi0"]), _dev(g["r_nph"].astype(np.int64)), _dev(g["r_ofs"].astype(np.int64))]
    d = [_dev(g["d_theta"]), _dev(g["d_phi"])]
    o = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3)]
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    print("case golden pointing", flush=True)
    _lib.call("nusht_pointing_f64", p(ring[0]), p(ring[1]), p(ring[2]), p(ring[3]), int(g["r_theta"].size),
              int(g["r_nph"].max()), n, p(d[0]), p(d[1]), p(o[0]), p(o[1]), p(o[2]), s)
    torch.cuda.synchronize()
    assert np.max(np.abs(o[0].cpu().numpy() - g["ptg_theta"])) < 1e-14
    assert _angle_err(o[1].cpu().numpy(), g["ptg_phi"]) < 1e-12
    assert _angle_err(o[2].cpu().numpy(), g["gamma"]) < 1e-12


@pytest.mark.gpu
def test_gpu_doubling_vectors():
    import torch
    from cunusht_amd import _lib
    g = _load("doubling_7x10")
    M, Dm = g["map"], g["dmap"]
    nt, nph = M.shape
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    mt = _dev(M.T)  # the reference's doubling input is theta-contiguous (nphi, ntheta)
    dmt = _dev(Dm)
    for spin in (0, 1, 2):
        out = torch.empty((nph, 2 * nt - 2), dtype=torch.float64, device="cuda")
        adj = torch.empty((nt, nph), dtype=torch.float64, device="cuda")
        _lib.call("nusht_doubling_f64", p(mt), nt, nph, spin, p(out), s)
        _lib.call("nusht_adjoint_doubling_f64", p(dmt), nt, nph, spin, p(adj), s)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy().T, g[f"doubled_s{spin}"])
        assert np.array_equal(adj.cpu().numpy(), g[f"adjoint_s{spin}"])


@pytest.mark.gpu
def test_gpu_transport_vector():
    import torch
    g = _load("transport_spin2_gl12")
    L = int(g["lmax"])
    T = _T(g, 2)
    print("case golden transport", flush=True)
    out = T.gclm2lenmap(_dev(g["gclm"]), L, L, dlm_scaled=_dev(g["dlm_scaled"]), spin=2)
    torch.cuda.synchronize()
    assert _relmax(out.cpu().numpy(), g["lenmap"]) < 1e-9
