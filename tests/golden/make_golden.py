#!/usr/bin/env python3
"""Writes tests/golden/ref_c_vectors.npz: inputs and outputs of the reference's
own C kernels (pycbf/cpu/pycbf.c, built by `make -C oracle` into oracle/_ref
from the reference sources) on small seeded cases.

Vectors:
  spline_y / spline_dx / spline_y2    tie_knots1D_fixed (pycbf.c:114-166), 4 channels
  pts_y / pts_x0 / pts_dx / pts_x / pts_val
                                      cubic1D_fixed_point (pycbf.c:68-105)
  das_*                               beamform_cubic / beamform_nearest (pycbf.c:179-271)
                                      driven channel by channel in the order and with the
                                      output offsets of pycbf/cpu/__core_bf_classes__.py:50-61,240-245,
                                      for kinds cubic, nearest usf 1/3/8 and every sumtype

Delays stay below t0 + dt*(nt-2): past that point the reference reads beyond
its buffers (pycbf.c:52,103) and its output is undefined.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_c_vectors.npz")
SUMTYPES = ("none", "tx_only", "rx_only", "tx_and_rx")
KINDS = (("cubic", 1), ("nearest", 1), ("nearest", 3), ("nearest", 8))


def ref_y2(y, dx):
    import ctypes as ct

    class Knots(ct.Structure):
        # struct IntrpData1D_Fixed, pycbf/cpu/pycbf.h:12-21
        _fields_ = [("N", ct.c_int), ("xstart", ct.c_float), ("dx", ct.c_float),
                    ("y", ct.POINTER(ct.c_float)), ("y2", ct.POINTER(ct.c_float)),
                    ("fill", ct.c_float)]
    R = oracle.ref()
    y = np.ascontiguousarray(y, np.float32)
    p = R.tie_knots1D_fixed(y.ctypes.data_as(ct.POINTER(ct.c_float)), y.size, dx, 0.0, 0.0, 1)
    try:
        k = ct.cast(p, ct.POINTER(Knots)).contents
        return np.ctypeslib.as_array(k.y2, shape=(y.size,)).copy()
    finally:
        R.free_IntrpData1D_Fixed(p)


def main():
    if oracle.ref() is None:
        sys.exit("oracle/_ref/libpycbf_ref.so is not built: run `make -C oracle` "
                 "where the reference tree is present")
    rng = np.random.default_rng(20240611)
    v = {}

    dx = np.float32(1.0 / 30e6)
    y = rng.standard_normal((4, 257)).astype(np.float32)
    y[3] = np.cumsum(y[3]).astype(np.float32)          # smooth-ish channel
    v["spline_y"] = y
    v["spline_dx"] = np.float32(dx)
    v["spline_y2"] = np.stack([ref_y2(c, float(dx)) for c in y])

    n, x0, dt = 200, np.float32(1e-6), np.float32(2.5e-8)
    yp = rng.standard_normal(n).astype(np.float32)
    tmax = float(x0) + float(dt) * (n - 1)
    xp = np.concatenate([rng.uniform(float(x0) - 4 * float(dt), tmax - 2 * float(dt), 600),
                         float(x0) + float(dt) * np.arange(0, n - 2, 17),
                         [float(x0) - 1e-9, tmax + 1e-9]]).astype(np.float32)
    v["pts_y"], v["pts_x0"], v["pts_dx"], v["pts_x"] = yp, x0, dt, xp
    v["pts_val"] = oracle.ref_cubic_points(yp, float(x0), float(dt), xp)

    ntx, nrx, nt, npx = 3, 4, 160, 257
    t0, dt = np.float32(2e-7), np.float32(2.5e-8)
    rf = rng.standard_normal((ntx, nrx, nt)).astype(np.float32)
    span = float(dt) * (nt - 3)
    tautx = rng.uniform(float(t0) - 0.1 * span, float(t0) + 0.45 * span, (ntx, npx)).astype(np.float32)
    taurx = rng.uniform(0.0, 0.5 * span, (nrx, npx)).astype(np.float32)
    apodtx = rng.choice([0.0, 0.005, 0.01, 0.2, 1.0, -0.7], (ntx, npx)).astype(np.float32)
    apodrx = rng.uniform(-1, 1, (nrx, npx)).astype(np.float32)
    v.update(das_rf=rf, das_t0=t0, das_dt=dt, das_tautx=tautx, das_apodtx=apodtx,
             das_taurx=taurx, das_apodrx=apodrx, das_thresh=np.float32(1e-2))
    for kind, usf in KINDS:
        for st in SUMTYPES:
            v[f"das_{kind}{usf}_{st}"] = oracle.ref_das_tabbed(
                rf, float(t0), float(dt), tautx, apodtx, taurx, apodrx, kind=kind, usf=usf,
                thresh=1e-2, sumtype=st)
    np.savez_compressed(OUT, **v)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {len(v)} arrays)")


if __name__ == "__main__":
    main()
