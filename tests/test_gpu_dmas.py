"""GPU DMAS against the FP64 restatement of the notebook formulas."""
import numpy as np
import pytest

import oracle
from oracle import numerics
from pycbf_amd.gpu import SyntheticDMASBeamformer, TabbedDMASBeamformer
from tests.helpers import assert_close_to_peak, small_synthetic, small_tabbed

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    return small_tabbed(nrx=24)


@pytest.mark.parametrize("dmastype", ["ssr", "power", "pssr"])
@pytest.mark.parametrize("sumlags", [True, False])
@pytest.mark.parametrize("lagaxis", ["rx", "tx"])
def test_tabbed_dmas(small, dmastype, sumlags, lagaxis):
    tautx, apodtx, taurx, apodrx = small["tabs"]
    lags = np.array([1, 2, 3, 5]) if lagaxis == "rx" else np.array([1, 2])
    bf = TabbedDMASBeamformer(tautx=tautx, taurx=taurx, apodtx=apodtx, apodrx=apodrx,
                              t0=small["t0"], dt=small["dt"], nt=small["nt"],
                              interp={"kind": "cubic"}, lags=lags, lagaxis=lagaxis,
                              dmastype=dmastype, sumlags=sumlags)
    got = bf(small["rf"])
    sumtype = "tx_only" if lagaxis == "rx" else "rx_only"
    imsep = oracle.das_tabbed(small["rf"], small["t0"], small["dt"], *small["tabs"],
                              kind="cubic", sumtype=sumtype)
    ref = numerics.dmas(imsep, lags, dmastype, sumlags)
    assert got.shape == ref.shape
    assert_close_to_peak(got, ref, 1e-4, f"{dmastype} sumlags={sumlags} axis={lagaxis}")


def test_dmas_lag_values_are_used(small):
    """Lag 3 alone differs from lag 1 alone (the reference used 0..nlag-1)."""
    tautx, apodtx, taurx, apodrx = small["tabs"]
    outs = []
    for lags in ([1], [3]):
        bf = TabbedDMASBeamformer(tautx=tautx, taurx=taurx, apodtx=apodtx, apodrx=apodrx,
                                  t0=small["t0"], dt=small["dt"], nt=small["nt"],
                                  interp={"kind": "linear"}, lags=np.array(lags),
                                  dmastype="power")
        outs.append(bf(small["rf"]))
    assert not np.allclose(outs[0], outs[1])


def test_synthetic_dmas_frames():
    w = small_synthetic(ntx=4, nrx=16)
    p = w["params"]
    rfs = np.stack([w["rf"], 2.0 * w["rf"]])
    bf = SyntheticDMASBeamformer(**p, t0=w["t0"], dt=w["dt"], nt=w["nt"],
                                 interp={"kind": "cubic"}, lags=np.arange(1, 4),
                                 dmastype="power")
    got = bf(rfs)
    for f in range(2):
        imsep = oracle.das_synthetic(rfs[f], w["t0"], w["dt"], p["ovectx"], p["nvectx"],
                                     p["t0tx"], p["alatx"], p["doftx"], p["ovecrx"], p["nvecrx"],
                                     p["alarx"], p["c0"], p["pnts"], kind="cubic",
                                     sumtype="tx_only")
        ref = numerics.dmas(imsep, np.arange(1, 4), "power", True)
        assert_close_to_peak(got[f], ref, 1e-4, f"frame {f}")
