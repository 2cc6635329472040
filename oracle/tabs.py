"""FP64 NumPy restatement of pycbf's delay-tab and apodisation generator
(test infrastructure only: the checker of the GPU generator).

``make_tabs_and_apods_2D`` follows pycbf/helpers/__full_aperture_2D__.py:
* receive tabs and f-number apodisation, :3-35 (``calc_rx_tabs_and_apods``);
* divergent-wave transmit tabs, :37-84 (``r0 < 0``);
* focused transmit tabs through the synthetic point, :86-127 and :270-314
  (``r0 > 0``);
* plane-wave transmit tabs and the aperture-projection apodisation, :129-168
  (``r0 == 0``);
* the dispatch on r0, :170-203.
Pixels are ``meshgrid(xout, zout, indexing='ij')`` flattened (x-major, z
fastest).  Every per-pixel quantity is computed with the reference's float64
operations in the reference's order; ``pixel_loop`` evaluates the same
formulas one pixel at a time (plain Python), the brute-force check of the
vectorised broadcasting.
"""
from __future__ import annotations

import numpy as np


def _grid(xout, zout):
    X, Z = np.meshgrid(np.asarray(xout, np.float64), np.asarray(zout, np.float64), indexing="ij")
    return np.array([X.flatten(), Z.flatten()]).T


def rx_tabs(cm, xele, xout, zout, fnum):
    """pycbf/helpers/__full_aperture_2D__.py:3-35."""
    xele = np.asarray(xele, np.float64)
    P = _grid(xout, zout)
    E = np.array([xele, np.zeros_like(xele)]).T
    d = P[None, :, :] - E[:, None, :]
    tau = np.sqrt(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) / cm
    apod = np.abs(d[..., 0]) - d[..., 1] / (2 * fnum) < 0
    return tau, apod


def tx_tabs_pw(steers, ctx, cm, xele, xout, zout):
    """pycbf/helpers/__full_aperture_2D__.py:129-168."""
    xele = np.asarray(xele, np.float64)
    steers = np.asarray(steers, np.float64)
    P = _grid(xout, zout)
    O = np.array([[xele[-1] if s <= 0 else xele[0] for s in steers], np.zeros(steers.shape)]).T
    st = np.arcsin(cm * np.sin(steers) / ctx)
    N = np.array([np.sin(st), np.cos(st)]).T
    d = P[None, :, :] - O[:, None, :]
    tau = (d[..., 0] * N[:, None, 0] + d[..., 1] * N[:, None, 1]) / cm
    Np = np.array([-np.cos(st), np.sin(st)]).T
    dn = P - np.array([xele[0], 0.0])
    dp = P - np.array([xele[-1], 0.0])
    neg = dn[None, :, 0] * Np[:, None, 0] + dn[None, :, 1] * Np[:, None, 1]
    pos = dp[None, :, 0] * Np[:, None, 0] + dp[None, :, 1] * Np[:, None, 1]
    return tau, (neg <= 0) & (pos >= 0)


def tx_tabs_divergent(steers, r0, ctx, cm, xele, xout, zout):
    """pycbf/helpers/__full_aperture_2D__.py:37-84."""
    if ctx != cm:
        raise ValueError("global SOS correction is not implemented for divergent waves")
    xele = np.asarray(xele, np.float64)
    steers = np.asarray(steers, np.float64)
    r0 = -np.abs(r0)
    O = np.array([r0 * np.sin(steers), r0 * np.cos(steers)]).T
    E = np.array([xele, np.zeros_like(xele)]).T
    doe = E[None, :, :] - O[:, None, :]
    tau0 = np.min(np.sqrt(doe[..., 0] * doe[..., 0] + doe[..., 1] * doe[..., 1]) / ctx, axis=1,
                  keepdims=True)
    P = _grid(xout, zout)
    d = P[None, :, :] - O[:, None, :]
    tau = np.sqrt(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) / cm - tau0
    return tau, np.ones(tau.shape, dtype=bool)


def synthetic_focused(steers, ctx, cm, xele, rpw):
    """Synthetic-point parameters of a focused transmit,
    pycbf/helpers/__full_aperture_2D__.py:270-314."""
    if cm != ctx:
        raise ValueError("speed of sound correction is not implemented for focused waves")
    xele = np.asarray(xele, np.float64)
    steers = np.asarray(steers, np.float64)
    ovec = rpw * np.array([np.sin(steers), np.cos(steers)]).T
    dxmax = np.max(np.hypot(xele[None, :] - ovec[:, 0, None], ovec[:, 1, None]), axis=1)
    t0 = dxmax / cm
    alneg = np.arctan2(ovec[:, 0] - xele[0], ovec[:, 1])
    alpos = np.arctan2(ovec[:, 0] - xele[-1], ovec[:, 1])
    ala = -(alpos - alneg) / 2
    alst = (alpos + alneg) / 2
    nvec = np.array([np.sin(alst), np.cos(alst)]).T
    return ovec, nvec, t0, ala


def tx_tabs_focused(steers, r0, ctx, cm, xele, xout, zout):
    """pycbf/helpers/__full_aperture_2D__.py:86-127."""
    ovec, nvec, t0, ala = synthetic_focused(steers, ctx, cm, xele, r0)
    P = _grid(xout, zout)
    d = P[None, :, :] - ovec[:, None, :]
    proj = d[..., 0] * nvec[:, None, 0] + d[..., 1] * nvec[:, None, 1]
    mag = np.sqrt(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1])
    tau = np.sign(proj) * mag / cm + t0[:, None]
    mag = mag.copy()
    mag[mag < 1e-12] = 1e-12
    with np.errstate(invalid="ignore"):
        apod = np.arccos(np.abs(proj) / mag) < ala[:, None]
    return tau, apod


def make_tabs_and_apods_2D(steers, r0, ctx, cm, xele, xout, zout, fnum):
    """pycbf/helpers/__full_aperture_2D__.py:170-203; float64 arrays and bool
    apodisations, as the reference returns them."""
    if r0 > 0:
        tautx, apodtx = tx_tabs_focused(steers, r0, ctx, cm, xele, xout, zout)
    elif r0 == 0:
        tautx, apodtx = tx_tabs_pw(steers, ctx, cm, xele, xout, zout)
    else:
        tautx, apodtx = tx_tabs_divergent(steers, r0, ctx, cm, xele, xout, zout)
    taurx, apodrx = rx_tabs(cm, xele, xout, zout, fnum)
    return {"tautx": tautx, "taurx": taurx, "apodtx": apodtx, "apodrx": apodrx}


def pixel_loop(steers, r0, ctx, cm, xele, xout, zout, fnum):
    """The same tables one (element, pixel) at a time in plain Python floats
    (IEEE double, same operation order): a check of the vectorised
    broadcasting and pixel order on small inputs."""
    import math
    steers = [float(s) for s in steers]
    xele = [float(x) for x in xele]
    pts = [(float(x), float(z)) for x in xout for z in zout]
    ntx, nrx, npx = len(steers), len(xele), len(pts)
    tt = np.zeros((ntx, npx)), np.zeros((ntx, npx), bool)
    tr = np.zeros((nrx, npx)), np.zeros((nrx, npx), bool)
    for e, xe in enumerate(xele):
        for p, (x, z) in enumerate(pts):
            dx, dz = x - xe, z - 0.0
            tr[0][e, p] = math.sqrt(dx * dx + dz * dz) / cm
            tr[1][e, p] = abs(dx) - dz / (2 * fnum) < 0
    if r0 == 0:
        for t, s in enumerate(steers):
            ox = xele[-1] if s <= 0 else xele[0]
            st = float(np.arcsin(cm * np.sin(s) / ctx))
            nx, nz = float(np.sin(st)), float(np.cos(st))
            px, pz = -float(np.cos(st)), float(np.sin(st))
            for p, (x, z) in enumerate(pts):
                tt[0][t, p] = ((x - ox) * nx + (z - 0.0) * nz) / cm
                neg = (x - xele[0]) * px + (z - 0.0) * pz
                pos = (x - xele[-1]) * px + (z - 0.0) * pz
                tt[1][t, p] = neg <= 0 and pos >= 0
    elif r0 < 0:
        rr = -abs(r0)
        for t, s in enumerate(steers):
            ox, oz = rr * float(np.sin(s)), rr * float(np.cos(s))
            tau0 = min(math.sqrt((xe - ox) ** 2 + (0.0 - oz) ** 2) / ctx for xe in xele)
            for p, (x, z) in enumerate(pts):
                dx, dz = x - ox, z - oz
                tt[0][t, p] = math.sqrt(dx * dx + dz * dz) / cm - tau0
                tt[1][t, p] = True
    else:
        ovec, nvec, t0, ala = synthetic_focused(steers, ctx, cm, xele, r0)
        for t in range(ntx):
            for p, (x, z) in enumerate(pts):
                dx, dz = x - float(ovec[t, 0]), z - float(ovec[t, 1])
                proj = dx * float(nvec[t, 0]) + dz * float(nvec[t, 1])
                mag = math.sqrt(dx * dx + dz * dz)
                sg = float((proj > 0) - (proj < 0))
                tt[0][t, p] = sg * mag / cm + float(t0[t])
                m = mag if mag >= 1e-12 else 1e-12
                v = abs(proj) / m
                tt[1][t, p] = (not math.isnan(v)) and v <= 1.0 and math.acos(v) < float(ala[t])
    return {"tautx": tt[0], "apodtx": tt[1], "taurx": tr[0], "apodrx": tr[1]}
