"""The oracle against committed golden vectors of the reference's own C kernels
(tests/golden/ref_c_vectors.npz, written by tests/golden/make_golden.py from
pycbf/cpu/pycbf.c) and against the interpolator-accuracy KAT of
examples/compare_interpolation_methods.ipynb:571.  Needs neither the reference
tree nor a GPU."""
import os

import numpy as np
import pytest

import oracle
from tests import kat_interp

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_c_vectors.npz")
SUMTYPES = ("none", "tx_only", "rx_only", "tx_and_rx")
KINDS = (("cubic", 1), ("nearest", 1), ("nearest", 3), ("nearest", 8))


@pytest.fixture(scope="module")
def g():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_spline_y2_matches_golden(g):
    for y, y2 in zip(g["spline_y"], g["spline_y2"]):
        assert np.array_equal(oracle.spline_y2(y, float(g["spline_dx"])), y2)


def test_cubic_points_match_golden(g):
    got = oracle.cubic_points(g["pts_y"], float(g["pts_x0"]), float(g["pts_dx"]), g["pts_x"])
    assert np.array_equal(got, g["pts_val"])


@pytest.mark.parametrize("sumtype", SUMTYPES)
@pytest.mark.parametrize("kind,usf", KINDS, ids=[f"{k}{u}" for k, u in KINDS])
def test_das_matches_golden(g, kind, usf, sumtype):
    tabs = [g[k] for k in ("das_tautx", "das_apodtx", "das_taurx", "das_apodrx")]
    got = oracle.das_tabbed(g["das_rf"], float(g["das_t0"]), float(g["das_dt"]), *tabs,
                            kind=kind, usf=usf, thresh=float(g["das_thresh"]), sumtype=sumtype)
    assert np.array_equal(got, g[f"das_{kind}{usf}_{sumtype}"])


def test_golden_cases_are_not_trivial(g):
    """The fixtures exercise fills, thresholds and non-zero sums."""
    img = g["das_cubic1_tx_and_rx"]
    assert np.count_nonzero(img) > 0.8 * img.size
    assert np.count_nonzero(g["das_nearest1_none"] == 0) > 0   # thresholded / filled samples
    assert not np.array_equal(g["das_nearest1_tx_and_rx"], g["das_nearest8_tx_and_rx"])


@pytest.mark.slow
def test_interpolation_kat_nearest_vs_cubic():
    """ROI RMSE of the nearest image against the cubic image on the regenerated
    notebook dataset reproduces the reference's reported 35.48316740429086."""
    d = kat_interp.dataset()
    cub = oracle.das_tabbed(d["rf"], d["t0"], d["dt"], *d["tabs"], kind="cubic")
    near = oracle.das_tabbed(d["rf"], d["t0"], d["dt"], *d["tabs"], kind="nearest")
    r = kat_interp.roi_rmse(near, cub, d)
    assert r == pytest.approx(kat_interp.RMSE_NEAREST_VS_CUBIC, rel=1e-12, abs=0)
