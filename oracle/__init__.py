"""CPU oracle for pycbf's beamforming hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or as the timed CPU baseline).  ``pycbf_amd`` never imports it.

Contents
--------
* ``liboracle.so`` (built from ``pycbf_oracle.c`` by ``oracle/Makefile``): a C
  restatement of ``pycbf/cpu/pycbf.c`` (natural spline, cubic / nearest(+usf)
  kernels), the project-defined ``linear`` kind, ``korder_cubic``
  (``pycbf/gpu/__engines__.cu:292-355``), the reference CPU driver's loop order
  and output offsets (``pycbf/cpu/__core_bf_classes__.py:50-61,150-192,220-250``)
  and the synthetic-point geometry (``pycbf/gpu/__engines__.cu:382-470``).
* ``_ref/libpycbf_ref.so``: the reference's own ``pycbf.c`` compiled from the
  read-only reference tree (git-ignored, rebuilt only where the tree exists).
* ``numerics``: FP64 NumPy restatements of DMAS / SLSC and the k-order ``S``
  operator.
* ``diff64``: FP64 torch-autograd restatement of the Tabbed DAS and the
  straight-ray delays.
* ``tabs``: FP64 NumPy restatement of ``make_tabs_and_apods_2D``.
* ``demod``: FP64 NumPy restatement of the notebooks' ``fft_tukey_demod``.

Pinning status of each function is listed in ``oracle/README`` style in the
module docstrings of the tests that pin them (tests/test_oracle_*.py).
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

KINDS = {"nearest": 0, "linear": 1, "cubic": 2, "korder_cubic": 3}
SUMTYPES = {"none": 0, "tx_only": 1, "rx_only": 2, "tx_and_rx": 3}

_f = ct.c_float
_i = ct.c_int
_i64 = ct.c_int64
_pf = ct.POINTER(ct.c_float)


def _ptr(a):
    return a.ctypes.data_as(_pf)


def lib():
    """Load the oracle restatement (``oracle/liboracle.so``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ct.CDLL(path)
        L.orc_spline_y2.argtypes = [_pf, _i, _f, _pf, _pf]
        L.orc_spline_forward_coef.argtypes = [_i, _pf]
        L.orc_index_floor.argtypes = [_f, _f, _f, _i]
        L.orc_index_floor.restype = _i
        L.orc_index_nearest.argtypes = [_f, _f, _f, _i]
        L.orc_index_nearest.restype = _i
        L.orc_cubic_point.argtypes = [_pf, _pf, _i, _f, _f, _f]
        L.orc_cubic_point.restype = _f
        L.orc_cubic_many.argtypes = [_pf, _pf, _i, _f, _f, _pf, _i, _pf]
        L.orc_linear_point.argtypes = [_pf, _i, _f, _f, _f]
        L.orc_linear_point.restype = _f
        L.orc_korder_point.argtypes = [_pf, _i, _f, _f, _pf, _i, _f]
        L.orc_korder_point.restype = _f
        chan = [_f, _f, _i, _pf, _i, _f, _pf, _pf, _pf, _pf, _pf]
        L.orc_bf_cubic.argtypes = chan
        L.orc_bf_linear.argtypes = chan
        L.orc_bf_nearest.argtypes = chan + [_i]
        L.orc_bf_korder.argtypes = chan + [_pf, _i]
        L.orc_upsample.argtypes = [_f, _f, _i, _pf, _i, _pf]
        L.orc_das_tabbed.argtypes = [_pf, _i, _i, _i, _i, _f, _f, _pf, _pf, _pf, _pf, _i64,
                                     _i, _i, _pf, _i, _f, _i, _pf]
        L.orc_das_tabbed_mt.argtypes = [_pf, _i, _i, _i, _f, _f, _pf, _pf, _pf, _pf, _i64,
                                        _i, _i, _pf, _i, _f, _pf, _i]
        L.orc_das_tabbed_mt.restype = _i
        L.orc_geom_tx.argtypes = [_i, _pf, _pf, _f, _f, _f, _f, _pf, _pf, _pf]
        L.orc_geom_rx.argtypes = [_i, _pf, _pf, _f, _f, _pf, _pf, _pf]
        L.orc_indices.argtypes = [_pf, _pf, _i, _i, _i64, _f, _f, _i, _i,
                                  ct.POINTER(ct.c_int32)]
        L.orc_synthetic_tables.argtypes = [_i, _i, _pf, _pf, _pf, _pf, _pf, _i, _pf, _pf, _pf,
                                           _f, _pf, _i64, _pf, _pf, _pf, _pf]
        _LIB = L
    return _LIB


def ref():
    """Load the compiled reference ``pycbf.c`` if it has been built, else None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libpycbf_ref.so")
        if not os.path.exists(path):
            return None
        R = ct.CDLL(path)
        chan = [_f, _f, _i, _pf, _i, _f, _pf, _pf, _pf, _pf, _pf]
        R.beamform_cubic.argtypes = chan
        R.beamform_cubic.restype = None
        R.beamform_nearest.argtypes = chan + [_i]
        R.beamform_nearest.restype = None
        R.tie_knots1D_fixed.argtypes = [_pf, _i, _f, _f, _f, _i]
        R.tie_knots1D_fixed.restype = ct.c_void_p
        R.cubic1D_fixed_point.argtypes = [_f, ct.c_void_p]
        R.cubic1D_fixed_point.restype = _f
        R.cubic1D_fixed.argtypes = [_pf, _i, ct.c_void_p]
        R.cubic1D_fixed.restype = ct.c_void_p
        R.free_IntrpData1D_Fixed.argtypes = [ct.c_void_p]
        R.free_IntrpData1D_Fixed.restype = None
        _REF = R
    return _REF


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# --------------------------------------------------------------------------
# scalar / channel helpers
# --------------------------------------------------------------------------
def spline_y2(y, dx):
    y = _f32(y)
    n = y.size
    y2 = np.empty(n, np.float32)
    u = np.empty(max(n, 1), np.float32)
    lib().orc_spline_y2(_ptr(y), n, dx, _ptr(y2), _ptr(u))
    return y2


def spline_forward_coef(n):
    out = np.empty(n, np.float32)
    lib().orc_spline_forward_coef(n, _ptr(out))
    return out


def cubic_points(y, x0, dx, x):
    y = _f32(y)
    y2 = spline_y2(y, dx)
    x = _f32(x)
    out = np.empty(x.size, np.float32)
    lib().orc_cubic_many(_ptr(y), _ptr(y2), y.size, x0, dx, _ptr(x), x.size, _ptr(out))
    return out


def linear_points(y, x0, dx, x):
    y = _f32(y)
    L = lib()
    return np.array([L.orc_linear_point(_ptr(y), y.size, x0, dx, float(v)) for v in _f32(x)],
                    np.float32)


def korder_points(y, x0, dx, S, x):
    y = _f32(y)
    S = _f32(S)
    k = S.shape[0]
    L = lib()
    return np.array([L.orc_korder_point(_ptr(y), y.size, x0, dx, _ptr(S), k, float(v))
                     for v in _f32(x)], np.float32)


def index_floor(x, x0, dx, n):
    L = lib()
    return np.array([L.orc_index_floor(float(v), x0, dx, n) for v in _f32(x)], np.int64)


def index_nearest(x, x0, dx, n):
    L = lib()
    return np.array([L.orc_index_nearest(float(v), x0, dx, n) for v in _f32(x)], np.int64)


def indices(tautx, taurx, t0, dt, nt, kind="cubic", usf=1):
    """Sample index of every (tx, rx, px) as the reference selects it
    (pycbf/cpu/pycbf.c:81 knot index, :264 nearest on the dt/usf grid);
    int32 (ntx, nrx, np), -1 outside the trace."""
    tautx, taurx = _f32(tautx), _f32(taurx)
    ntx, np_ = tautx.shape
    nrx = taurx.shape[0]
    out = np.empty((ntx, nrx, np_), np.int32)
    nearest = kind == "nearest"
    n = usf * (nt - 1) + 1 if nearest else nt
    d = np.float32(dt) / np.float32(usf) if nearest else dt
    lib().orc_indices(_ptr(tautx), _ptr(taurx), ntx, nrx, np_, t0, float(d), n, int(nearest),
                      out.ctypes.data_as(ct.POINTER(ct.c_int32)))
    return out


def upsample(y, t0, dt, usf):
    y = _f32(y)
    out = np.empty(usf * (y.size - 1) + 1, np.float32)
    lib().orc_upsample(t0, dt, y.size, _ptr(y), usf, _ptr(out))
    return out


# --------------------------------------------------------------------------
# full drivers
# --------------------------------------------------------------------------
def das_tabbed(rf, t0, dt, tautx, apodtx, taurx, apodrx, kind="cubic", usf=1, thresh=1e-2,
               sumtype="tx_and_rx", S=None):
    """Tabbed DAS with the reference CPU driver's semantics.

    rf is (ntx, nrx, nt) or (nfr, ntx, nrx, nt); returns float32 with shape
    (nfr?, <sumtype axes>, np), i.e. the GPU class's output layout.
    """
    rf = _f32(rf)
    squeeze = rf.ndim == 3
    if squeeze:
        rf = rf[None]
    nfr, ntx, nrx, nt = rf.shape
    tautx, apodtx, taurx, apodrx = (_f32(a) for a in (tautx, apodtx, taurx, apodrx))
    np_ = tautx.shape[1]
    st = SUMTYPES[sumtype]
    shape = {0: (ntx, nrx, np_), 1: (nrx, np_), 2: (ntx, np_), 3: (np_,)}[st]
    out = np.empty((nfr,) + shape, np.float32)
    k = 0
    Sp = None
    if kind == "korder_cubic":
        S = _f32(S)
        k = S.shape[0]
        Sp = _ptr(S)
    lib().orc_das_tabbed(_ptr(rf), nfr, ntx, nrx, nt, t0, dt, _ptr(tautx), _ptr(apodtx),
                         _ptr(taurx), _ptr(apodrx), np_, KINDS[kind], int(usf), Sp, k,
                         thresh, st, _ptr(out))
    return out[0] if squeeze else out


def das_tabbed_mt(rf, t0, dt, tautx, apodtx, taurx, apodrx, kind="cubic", usf=1, thresh=1e-2,
                  nthreads=0, S=None):
    """Multi-threaded tx_and_rx DAS (channels split over OpenMP threads with
    private accumulators, summed in double); returns (image, threads_used)."""
    rf = _f32(rf)
    ntx, nrx, nt = rf.shape
    tautx, apodtx, taurx, apodrx = (_f32(a) for a in (tautx, apodtx, taurx, apodrx))
    np_ = tautx.shape[1]
    out = np.empty(np_, np.float32)
    k, Sp = 0, None
    if kind == "korder_cubic":
        S = _f32(S)
        k, Sp = S.shape[0], _ptr(S)
    used = lib().orc_das_tabbed_mt(_ptr(rf), ntx, nrx, nt, t0, dt, _ptr(tautx), _ptr(apodtx),
                                   _ptr(taurx), _ptr(apodrx), np_, KINDS[kind], int(usf), Sp,
                                   k, thresh, _ptr(out), int(nthreads))
    return out, used


def synthetic_tables(ovectx, nvectx, t0tx, alatx, doftx, ovecrx, nvecrx, alarx, c0, pnts):
    """Synthetic-point parameters -> (tautx, apodtx, taurx, apodrx) float32 tables."""
    ovectx, nvectx, t0tx, alatx, doftx, ovecrx, nvecrx, alarx, pnts = (
        _f32(a) for a in (ovectx, nvectx, t0tx, alatx, doftx, ovecrx, nvecrx, alarx, pnts))
    ntx, ndim = ovectx.shape
    nrx = ovecrx.shape[0]
    np_ = pnts.shape[0]
    tautx = np.empty((ntx, np_), np.float32)
    apodtx = np.empty((ntx, np_), np.float32)
    taurx = np.empty((nrx, np_), np.float32)
    apodrx = np.empty((nrx, np_), np.float32)
    lib().orc_synthetic_tables(ndim, ntx, _ptr(ovectx), _ptr(nvectx), _ptr(t0tx), _ptr(alatx),
                               _ptr(doftx), nrx, _ptr(ovecrx), _ptr(nvecrx), _ptr(alarx),
                               float(c0), _ptr(pnts), np_, _ptr(tautx), _ptr(apodtx),
                               _ptr(taurx), _ptr(apodrx))
    return tautx, apodtx, taurx, apodrx


def das_synthetic(rf, t0, dt, ovectx, nvectx, t0tx, alatx, doftx, ovecrx, nvecrx, alarx, c0,
                  pnts, kind="cubic", usf=1, thresh=1e-2, sumtype="tx_and_rx", S=None):
    """Synthetic DAS = tabbed DAS over tables from the same float32 geometry."""
    tabs = synthetic_tables(ovectx, nvectx, t0tx, alatx, doftx, ovecrx, nvecrx, alarx, c0, pnts)
    tautx, apodtx, taurx, apodrx = tabs
    return das_tabbed(rf, t0, dt, tautx, apodtx, taurx, apodrx, kind=kind, usf=usf,
                      thresh=thresh, sumtype=sumtype, S=S)


# --------------------------------------------------------------------------
# reference C (compiled from the reference tree) driven with the reference's
# own loop order and offsets
# --------------------------------------------------------------------------
def ref_das_tabbed(rf, t0, dt, tautx, apodtx, taurx, apodrx, kind="cubic", usf=1, thresh=1e-2,
                   sumtype="tx_and_rx"):
    """Run the reference beamform_cubic / beamform_nearest channel by channel in
    the order of pycbf/cpu/__core_bf_classes__.py:240-245 (single worker)."""
    R = ref()
    if R is None:
        raise RuntimeError("oracle/_ref/libpycbf_ref.so is not built")
    rf = _f32(rf)
    ntx, nrx, nt = rf.shape
    tautx, apodtx, taurx, apodrx = (_f32(a) for a in (tautx, apodtx, taurx, apodrx))
    np_ = tautx.shape[1]
    st = SUMTYPES[sumtype]
    shape = {0: (ntx, nrx, np_), 1: (nrx, np_), 2: (ntx, np_), 3: (np_,)}[st]
    out = np.zeros(shape, np.float32)
    flat = out.reshape(-1)
    for itx in range(ntx):
        for irx in range(nrx):
            sig = np.ascontiguousarray(rf[itx, irx])
            if st == 0:
                off = np_ * (irx + itx * nrx)
            elif st == 1:
                off = np_ * irx
            elif st == 2:
                off = np_ * itx
            else:
                off = 0
            o = flat[off:off + np_]
            args = (t0, dt, nt, _ptr(sig), np_, thresh, _ptr(tautx[itx]), _ptr(apodtx[itx]),
                    _ptr(taurx[irx]), _ptr(apodrx[irx]), _ptr(o))
            if kind == "cubic":
                R.beamform_cubic(*args)
            elif kind == "nearest":
                R.beamform_nearest(*args, int(usf))
            else:
                raise ValueError(f"reference C has no kind {kind!r}")
    return out


def ref_cubic_points(y, x0, dx, x):
    """Reference tie_knots1D_fixed + cubic1D_fixed_point at each x."""
    R = ref()
    if R is None:
        raise RuntimeError("oracle/_ref/libpycbf_ref.so is not built")
    y = _f32(y)
    knots = R.tie_knots1D_fixed(_ptr(y), y.size, dx, x0, 0.0, 1)
    try:
        return np.array([R.cubic1D_fixed_point(float(v), knots) for v in _f32(x)], np.float32)
    finally:
        R.free_IntrpData1D_Fixed(knots)


from . import numerics  # noqa: E402,F401
