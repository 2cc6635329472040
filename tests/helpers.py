"""Shared workload builders and comparison helpers for the tests."""
from __future__ import annotations

import numpy as np

from pycbf_amd import sim


def assert_close_to_peak(got, ref, rel=1e-5, what=""):
    """max |got - ref| <= rel * max |ref| (the image-peak tolerance)."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    peak = float(np.max(np.abs(ref))) if ref.size else 0.0
    err = float(np.max(np.abs(got - ref))) if ref.size else 0.0
    assert err <= rel * max(peak, 1e-30), f"{what}: max err {err:.3e} > {rel:g} * peak {peak:.3e}"
    return err, peak


def small_tabbed(seed=0, ntx=5, nrx=32, nt=1024, nx=32, nz=64, n_points=12):
    """Reduced plane-wave tabbed workload (same model as the 5x128x2048 case)."""
    xele = sim.linear_array(nrx)
    steers = np.radians(np.linspace(-8, 8, ntx))
    xlim = (float(xele[0]) * 1.2, float(xele[-1]) * 1.2)
    _, _, pnts = sim.grid(nx, nz, xlim, (4e-3, 18e-3))
    xs, zs, amp = sim.scatterers(n_points, 50, (xlim[0] * 0.8, xlim[1] * 0.8), (5e-3, 17e-3), seed)
    rf = sim.simulate_rf(steers, xele, nt, xs, zs, amp, seed=seed)
    tabs = sim.pw_tabs(steers, xele, pnts)
    return dict(rf=rf, steers=steers, xele=xele, pnts=pnts, t0=0.0, dt=1.0 / sim.FS, nt=nt,
                tabs=tabs)


def small_synthetic(seed=0, ntx=8, nrx=32, nt=1024, nx=32, nz=48, mode="pw"):
    """Reduced synthetic-point workload: plane waves as points at -10 m (mode
    'pw'), diverging ('div', r0 < 0), focused ('foc', r0 > 0) or focused with a
    depth of field ('dof')."""
    xele = sim.linear_array(nrx)
    steers = np.radians(np.linspace(-16, 16, ntx))
    xlim = (float(xele[0]) * 1.3, float(xele[-1]) * 1.3)
    _, _, pnts = sim.grid(nx, nz, xlim, (2e-3, 18e-3))
    params = sim.pw_synthetic(steers, xele, pnts)
    if mode in ("div", "foc", "dof"):
        r0 = -6e-3 if mode == "div" else 10e-3
        nvec = np.stack([np.sin(steers), np.cos(steers)], axis=1)
        ovec = r0 * nvec if mode == "div" else r0 * nvec
        params["ovectx"] = ovec.astype(np.float32)
        params["nvectx"] = nvec.astype(np.float32)
        params["t0tx"] = (np.abs(r0) / sim.C0 * (1 if mode != "div" else -1)
                          * np.ones(ntx)).astype(np.float32)
        params["alatx"] = np.full(ntx, 0.35, np.float32)
        params["doftx"] = np.full(ntx, 6e-3 if mode == "dof" else 0.0, np.float32)
    xs, zs, amp = sim.scatterers(10, 40, (xlim[0] * 0.8, xlim[1] * 0.8), (3e-3, 17e-3), seed)
    rf = sim.simulate_rf(steers, xele, nt, xs, zs, amp, seed=seed)
    return dict(rf=rf, params=params, t0=0.0, dt=1.0 / sim.FS, nt=nt)


def float32_inputs(t0, dt, geo=None):
    """The time axis (and slowness-map geometry) as the kernels receive them:
    rounded to float32 (pycbf's RFInfo carries t0 and dt as float32).  FP64
    references of the kernels' functions take these values, so that both
    sides evaluate the same function."""
    import dataclasses
    t32, dt32 = float(np.float32(t0)), float(np.float32(dt))
    if geo is None:
        return t32, dt32
    g32 = dataclasses.replace(geo, **{k: float(np.float32(getattr(geo, k)))
                                      for k in ("xm0", "zm0", "dxm", "dzm")})
    return t32, dt32, g32

