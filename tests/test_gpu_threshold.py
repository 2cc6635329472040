"""pycbf's GPU weight rule through the C-ABI: with thresh =
PYCBF_GPU_REFERENCE_THRESH the Tabbed DAS drops exactly the samples with
|apodtx*apodrx| <= 0.01 in double (pycbf/gpu/__engines__.cu:8,762,851).
Weights are drawn so that many products sit on float32(0.01), just below and
just above it; the expected outputs mask the oracle's per-sample values with
that rule evaluated in float64 NumPy."""
import numpy as np
import pytest

import oracle
from pycbf_amd.gpu import TabbedBeamformer
from pycbf_amd.gpu import _engine as E
from tests.helpers import assert_close_to_peak, small_tabbed

pytestmark = pytest.mark.gpu


def _straddling(seed=5):
    w = small_tabbed(seed=seed)
    tautx, apodtx, taurx, apodrx = w["tabs"]
    rng = np.random.default_rng(seed)
    one_hundredth = np.float32(0.01)
    cand_tx = np.float32([1.0, 0.1, -0.1, 0.5, 0.25])
    cand_rx = np.float32([0.01, 0.02, 0.04, 0.1, -0.1, 0.0,
                          np.nextafter(one_hundredth, np.float32(0.0)),
                          np.nextafter(one_hundredth, np.float32(1.0)), 0.0100001, 0.5])
    apodtx = rng.choice(cand_tx, size=apodtx.shape).astype(np.float32)
    apodrx = rng.choice(cand_rx, size=apodrx.shape).astype(np.float32)
    return w, (tautx, apodtx, taurx, apodrx)


@pytest.mark.parametrize("kind", ["linear", "cubic"])
def test_reference_gpu_rule(kind):
    w, tabs = _straddling()
    tautx, apodtx, taurx, apodrx = tabs
    prod = apodtx[:, None, :] * apodrx[None, :, :]                 # float32 products
    gpu_rule = np.abs(prod).astype(np.float64) > 0.01
    cpu_rule = np.abs(prod) >= np.float32(1e-2)
    assert int((gpu_rule != cpu_rule).sum()) > 1000                  # the rules differ here
    per = oracle.das_tabbed(w["rf"], w["t0"], w["dt"], *tabs, kind=kind, thresh=0.0,
                            sumtype="none")
    exp_none = np.where(gpu_rule, per, np.float32(0.0))
    common = dict(tautx=tautx, taurx=taurx, apodtx=apodtx, apodrx=apodrx, t0=w["t0"],
                  dt=w["dt"], nt=w["nt"], interp={"kind": kind},
                  thresh=E.GPU_REFERENCE_THRESH)
    got_none = TabbedBeamformer(**common, sumtype="none")(w["rf"])
    if kind == "linear":
        # one product per output: identical float32 arithmetic
        assert np.array_equal(got_none, exp_none), f"{int((got_none != exp_none).sum())} differ"
    else:
        assert_close_to_peak(got_none, exp_none, 1e-5, "cubic per-sample")
    got = TabbedBeamformer(**common, sumtype="tx_and_rx")(w["rf"])
    exp = exp_none.astype(np.float64).sum(axis=(0, 1))
    assert_close_to_peak(got, exp, 1e-5, f"{kind} image under the GPU rule")
    # and the CPU default keeps the samples at float32(0.01) that the GPU rule drops
    got_cpu = TabbedBeamformer(**{**common, "thresh": 1e-2}, sumtype="tx_and_rx")(w["rf"])
    exp_cpu = np.where(cpu_rule, per, np.float32(0.0)).astype(np.float64).sum(axis=(0, 1))
    assert_close_to_peak(got_cpu, exp_cpu, 1e-5, f"{kind} image under the CPU rule")
    # the two rules give images apart by ten times the parity tolerance
    assert np.abs(exp_cpu - exp).max() > 1e-4 * np.abs(exp).max()
