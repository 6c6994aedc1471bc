"""Host-side logic without a GPU: timer keys and JSON layout, reference grid sizes, serving helpers."""
import json

import numpy as np

from cunusht_amd import diagnostics
from oracle.geometry import reference_cc_sizes


def test_stage_keys_follow_reference_names():
    assert diagnostics.stage_key("legendre_synth_cc_s2") == "synthesis"
    assert diagnostics.stage_key("legendre_prep_cc_s0") == "synthesis"
    assert diagnostics.stage_key("legendre_synth_ring_s1") == "spin-1 synth"
    assert diagnostics.stage_key("ring_dft_pointing") == "pointing"
    assert diagnostics.stage_key("nufft_interp") == "nuFFT2d2"
    assert diagnostics.stage_key("nufft_spread") == "nuFFT2d1"
    assert diagnostics.stage_key("legendre_adjoint_cc_s2") == "adjointsynthesis"
    assert diagnostics.stage_key("something_else") == "something_else"


def test_timer_dumps_flat_json(tmp_path):
    t = diagnostics.Timer()
    t.keys = {"synthesis": 0.5, "nuFFT2d2": 0.25}
    fn = tmp_path / "lmax31_epsilon1e-07_run00_"
    t.dumpjson(str(fn))
    assert json.loads(fn.read_text()) == {"synthesis": 0.5, "nuFFT2d2": 0.25}


def test_reference_car_sizes_match_oracle():
    for L in (31, 255, 1023, 2047, 4095):
        assert diagnostics.reference_car_sizes(L) == reference_cc_sizes(L)


def test_poisson_arrivals_rate():
    from cunusht_amd.serving import poisson_arrivals
    t = poisson_arrivals(20000, 100.0, seed=1)
    assert np.all(np.diff(t) > 0) and abs(t[-1] / 20000 - 0.01) < 0.0005
