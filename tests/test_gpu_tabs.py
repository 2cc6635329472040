"""Device delay tabs (pycbf_make_tabs_2d through pycbf_amd.gpu.helpers)
against the oracle's FP64 restatement of make_tabs_and_apods_2D: the float32
tables equal the oracle's float64 tables converted to float32, bit for bit,
and the apodisations agree everywhere."""
import time

import numpy as np
import pytest

from oracle import tabs as otabs

pytestmark = pytest.mark.gpu

C = 1540.0


def _array(n=128, pitch=0.3e-3):
    return pitch * (np.arange(n) - (n - 1) / 2)


CASES = {
    "plane_waves": dict(steers=np.radians(np.linspace(-16, 16, 9)), r0=0.0, ctx=C),
    "plane_waves_ctx1480": dict(steers=np.radians([-10.0, -3.0, 0.0, 7.0]), r0=0.0, ctx=1480.0),
    "divergent": dict(steers=np.radians([-20.0, 0.0, 12.0]), r0=-15e-3, ctx=C),
    "focused": dict(steers=np.radians([-8.0, 0.0, 5.0]), r0=25e-3, ctx=C),
}


@pytest.mark.parametrize("case", list(CASES))
def test_tabs_bitwise_equal_oracle(case):
    import torch
    from pycbf_amd.gpu.helpers import make_tabs_and_apods_2D
    print(f"tabs {case}", flush=True)
    p = CASES[case]
    xele = _array()
    xout = np.linspace(-19e-3, 19e-3, 37)
    zout = np.linspace(1e-3, 45e-3, 53)
    got = make_tabs_and_apods_2D(p["steers"], p["r0"], p["ctx"], C, xele, xout, zout, 1.5)
    ref = otabs.make_tabs_and_apods_2D(p["steers"], p["r0"], p["ctx"], C, xele, xout, zout, 1.5)
    torch.cuda.synchronize()
    for k in ("tautx", "taurx"):
        g = got[k].cpu().numpy()
        r = ref[k].astype(np.float32)
        assert g.shape == r.shape
        assert np.array_equal(g, r), f"{k}: {int((g != r).sum())} values differ"
    for k in ("apodtx", "apodrx"):
        g = got[k].cpu().numpy()
        assert np.array_equal(g, ref[k].astype(np.float32)), f"{k}: {int((g != ref[k]).sum())} differ"
    # the apodisations are not trivial
    assert 0.05 < ref["apodrx"].mean() < 0.95


def test_focused_acceptance_threshold_is_exact():
    """vstar is the first double NumPy's arccos accepts: its predecessor is
    rejected."""
    from pycbf_amd.gpu.helpers import arccos_threshold
    for ala in (0.05, 0.3, 1.0, 1.5):
        v = arccos_threshold(ala)
        assert np.arccos(v) < ala
        assert not np.arccos(np.nextafter(v, 0.0)) < ala


def test_point_sources_need_matching_sound_speeds():
    from pycbf_amd.base import BeamformerException
    from pycbf_amd.gpu.helpers import make_tabs_and_apods_2D
    xele = _array(16)
    for r0 in (-10e-3, 10e-3):
        with pytest.raises(BeamformerException):
            make_tabs_and_apods_2D(np.radians([0.0]), r0, 1500.0, C, xele, [0.0], [0.01], 1.5)


def test_tabbed_beamformer_on_device_tables():
    """A TabbedBeamformer fed the device tables images as one fed the NumPy
    tables (the same float32 values); the single-frame cubic gather adds its
    receive chunks with float atomics, so two runs agree to rounding."""
    import torch
    from pycbf_amd import sim
    from pycbf_amd.gpu import TabbedBeamformer
    from pycbf_amd.gpu.helpers import make_tabs_and_apods_2D
    xele = _array(64)
    steers = np.radians([-6.0, 0.0, 6.0])
    xout, zout = np.linspace(-8e-3, 8e-3, 40), np.linspace(5e-3, 25e-3, 64)
    dev_t = make_tabs_and_apods_2D(steers, 0.0, C, C, xele, xout, zout, 1.5)
    ref = otabs.make_tabs_and_apods_2D(steers, 0.0, C, C, xele, xout, zout, 1.5)
    xs, zs, amp = sim.scatterers(5, 50, (-6e-3, 6e-3), (6e-3, 24e-3), seed=5)
    rf = sim.simulate_rf(steers, xele, 1024, xs, zs, amp, seed=5)
    kw = dict(t0=0.0, dt=1.0 / 40e6, nt=1024, interp={"kind": "cubic"}, sumtype="tx_and_rx")
    a = TabbedBeamformer(tautx=dev_t["tautx"], apodtx=dev_t["apodtx"], taurx=dev_t["taurx"],
                         apodrx=dev_t["apodrx"], **kw)(rf)
    b = TabbedBeamformer(tautx=ref["tautx"].astype(np.float32),
                         apodtx=ref["apodtx"].astype(np.float32),
                         taurx=ref["taurx"].astype(np.float32),
                         apodrx=ref["apodrx"].astype(np.float32), **kw)(rf)
    peak = float(np.abs(b).max())
    assert peak > 0
    assert float(np.abs(a - b).max()) <= 1e-6 * peak


def test_tabs_full_size_timing():
    """75 plane waves x 128 rx over 512 x 512 pixels (the headline grid):
    about 420 MB of tables written on the device."""
    import torch
    from pycbf_amd.gpu.helpers import make_tabs_and_apods_2D
    xele = _array()
    steers = np.radians(np.linspace(-16, 16, 75))
    xout, zout = np.linspace(-19e-3, 19e-3, 512), np.linspace(2e-3, 50e-3, 512)
    make_tabs_and_apods_2D(steers, 0.0, C, C, xele, xout, zout, 1.5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = make_tabs_and_apods_2D(steers, 0.0, C, C, xele, xout, zout, 1.5)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    print(f"tables 75 x 128 x 512^2: {ms:.2f} ms (call, including host constants)", flush=True)
    assert out["tautx"].shape == (75, 512 * 512) and out["taurx"].shape == (128, 512 * 512)
    assert bool(torch.isfinite(out["tautx"]).all())
