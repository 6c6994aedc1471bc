"""The two backends agree: get_transformer('GPU') (HIP library) against get_transformer('CPU') (host library).

Each is checked against direct FP64 summation at the deflected points (oracle/sht.py) within 5 epsilon
(relative max error; the ES kernel parameters target epsilon and reach 1-2 epsilon on these cases, the
max over points is a small multiple of the norm-wise error), and they agree with each other within 8 epsilon,
epsilon = 1e-10; the adjoints agree within 8 epsilon; the pointing (the same closed form in FP64 on both
sides) within 1e-12.
"""
import numpy as np
import pytest
import torch

from oracle import almutil, sht as osht

pytestmark = pytest.mark.gpu


def _relmax(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(np.asarray(b))))


@pytest.mark.parametrize("geom,spin", [(("gl", {"lmax": 64}), 0), (("gl", {"lmax": 64}), 2),
                                       (("healpix", {"nside": 32, "lmax": 64}), 2)])
def test_gpu_and_cpu_backends_agree(geom, spin):
    import cunusht_amd
    L, eps = 64, 1e-10
    rng = np.random.default_rng(300 + spin)
    ll = np.arange(L + 1)
    dlm_s = almutil.synalm(np.where(ll >= 2, 1e-5 / (ll + 1.0) ** 2, 0.0), L, rng)
    dlm = almutil.almxfl(dlm_s, np.sqrt(ll * (ll + 1.0)), L)
    nc = 1 if spin == 0 else 2
    g = almutil.random_alm(L, spin, nc, rng)
    print(f"case geom={geom[0]} spin={spin}", flush=True)
    G = cunusht_amd.get_transformer("GPU")(geom, nuFFTtype=2, epsilon=eps, spin=spin)
    C = cunusht_amd.get_transformer("CPU", solver="duccnufft")(geom, epsilon=eps)
    og = G.gclm2lenmap(torch.as_tensor(g, device="cuda"), L, L, dlm_scaled=torch.as_tensor(dlm_s, device="cuda"),
                       spin=spin)
    torch.cuda.synchronize()
    oc = C.gclm2lenmap(g, dlm, L, L, spin, eps, 8, verbose=0)
    assert _relmax(G.ptg.cpu().numpy(), C.deflectionlib.ptg()) < 1e-12
    tp, pp = C.deflectionlib.ptg()
    ex = osht.synth_direct(g, L, spin, tp, pp)
    if spin:
        ex = ex * np.exp(1j * spin * C.deflectionlib.gamma())
        ex = np.array([ex.real, ex.imag])
    assert _relmax(oc, ex) < 5 * eps
    assert _relmax(og.cpu().numpy(), ex) < 5 * eps
    assert _relmax(og.cpu().numpy(), oc) < 8 * eps
    x = rng.standard_normal(oc.shape)
    ag = G.lenmap2gclm(torch.as_tensor(x, device="cuda"), torch.as_tensor(dlm_s, device="cuda"), None, L, L,
                       spin=spin)
    torch.cuda.synchronize()
    ac = C.lenmap2gclm(x, dlm, None, spin, L, L, 8, epsilon=eps, verbose=0)
    assert _relmax(np.atleast_2d(ag.cpu().numpy()), ac) < 8 * eps
