"""Pins the oracle restatement against the reference's own C kernels
(pycbf/cpu/pycbf.c compiled from the reference tree into oracle/_ref).

Bitwise: natural-spline second derivatives (tie_knots1D_fixed), cubic values
(cubic1D_fixed_point), and full Tabbed DAS images for cubic and nearest(usf)
driven in the reference CPU class's loop order for every sumtype.  Inputs keep
delays below t0 + dt*(nt-2) wherever the reference would read past the end of
its buffers (pycbf/cpu/pycbf.c:52,103), since those reads are undefined.
"""
import ctypes as ct

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.skipif(oracle.ref() is None,
                                reason="oracle/_ref not built (reference tree absent)")


class _Knots(ct.Structure):
    # struct IntrpData1D_Fixed, pycbf/cpu/pycbf.h:12-21
    _fields_ = [("N", ct.c_int), ("xstart", ct.c_float), ("dx", ct.c_float),
                ("y", ct.POINTER(ct.c_float)), ("y2", ct.POINTER(ct.c_float)),
                ("fill", ct.c_float)]


def _ref_y2(y, dx):
    R = oracle.ref()
    y = np.ascontiguousarray(y, np.float32)
    p = R.tie_knots1D_fixed(y.ctypes.data_as(ct.POINTER(ct.c_float)), y.size, dx, 0.0, 0.0, 1)
    try:
        k = ct.cast(p, ct.POINTER(_Knots)).contents
        return np.ctypeslib.as_array(k.y2, shape=(y.size,)).copy()
    finally:
        R.free_IntrpData1D_Fixed(p)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 2048])
@pytest.mark.parametrize("dx", [1.0, 2.5e-8, 1.0 / 30e6])
def test_spline_y2_bitwise(n, dx):
    rng = np.random.default_rng(n)
    y = rng.standard_normal(n).astype(np.float32)
    assert np.array_equal(oracle.spline_y2(y, dx), _ref_y2(y, dx))


def test_cubic_points_bitwise():
    rng = np.random.default_rng(1)
    n, t0, dt = 512, 1e-6, np.float32(2.5e-8)
    y = rng.standard_normal(n).astype(np.float32)
    tmax = np.float32(t0) + dt * np.float32(n - 1)
    x = np.concatenate([rng.uniform(t0 - 5 * dt, float(tmax) - 2 * float(dt), 4000),
                        [t0, t0 + float(dt), float(tmax) - 2 * float(dt)],
                        [t0 - 1e-9, float(tmax) + 1e-9]]).astype(np.float32)
    got = oracle.cubic_points(y, t0, float(dt), x)
    ref = oracle.ref_cubic_points(y, t0, float(dt), x)
    assert np.array_equal(got, ref)


def _case(seed, ntx=3, nrx=5, nt=300, np_=700):
    rng = np.random.default_rng(seed)
    t0, dt = 2e-7, np.float32(2.5e-8)
    rf = rng.standard_normal((ntx, nrx, nt)).astype(np.float32)
    span = float(dt) * (nt - 3)
    # delays from before t0 to just under the last safe knot; weights around thresh
    tautx = rng.uniform(t0 - 0.1 * span, t0 + 0.45 * span, (ntx, np_)).astype(np.float32)
    taurx = rng.uniform(0.0, 0.5 * span, (nrx, np_)).astype(np.float32)
    apodtx = rng.choice([0.0, 0.005, 0.01, 0.2, 1.0, -0.7], (ntx, np_)).astype(np.float32)
    apodrx = rng.uniform(-1, 1, (nrx, np_)).astype(np.float32)
    return rf, t0, float(dt), (tautx, apodtx, taurx, apodrx)


@pytest.mark.parametrize("sumtype", ["none", "tx_only", "rx_only", "tx_and_rx"])
@pytest.mark.parametrize("kind,usf", [("cubic", 1), ("nearest", 1), ("nearest", 3),
                                      ("nearest", 8)])
def test_das_bitwise_vs_reference(kind, usf, sumtype):
    rf, t0, dt, tabs = _case(7)
    got = oracle.das_tabbed(rf, t0, dt, *tabs, kind=kind, usf=usf, sumtype=sumtype)
    ref = oracle.ref_das_tabbed(rf, t0, dt, *tabs, kind=kind, usf=usf, sumtype=sumtype)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)


def test_workload_bitwise_vs_reference():
    """The 5 PW x 128 rx x 2048 sample plane-wave workload, cubic and nearest."""
    from pycbf_amd import sim
    w = sim.workload_tabbed_small()
    for kind, usf in (("cubic", 1), ("nearest", 1), ("nearest", 4)):
        got = oracle.das_tabbed(w["rf"], w["t0"], w["dt"], *w["tabs"], kind=kind, usf=usf)
        ref = oracle.ref_das_tabbed(w["rf"], w["t0"], w["dt"], *w["tabs"], kind=kind, usf=usf)
        assert np.array_equal(got, ref), kind
