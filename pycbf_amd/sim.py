"""Seeded synthetic workloads for the beamforming hot path.

Geometry follows the conventions of pycbf/helpers/__full_aperture_2D__.py:
pixels x-major / z-fastest (:387-388), plane-wave origin ``xele[-1]`` for
steer <= 0 else ``xele[0]`` (:153,250), synthetic plane waves as a point at
``rpw = -10 m`` along the wave normal (:226-268), receive acceptance
``atan2(1/2, fnum)`` (:222).  Channel data are Gaussian-enveloped pulses
(scipy.signal.gausspulse's formula, -6 dB reference) at the two-way delay of
each scatterer plus white noise at a given SNR.
"""
from __future__ import annotations

import math

import numpy as np

C0 = 1540.0
FS = 40e6
FC = 5e6


def linear_array(nele: int = 128, pitch: float = 0.3e-3) -> np.ndarray:
    return pitch * (np.arange(nele) - (nele - 1) / 2)


def grid(nx: int, nz: int, xlim=(-19e-3, 19e-3), zlim=(5e-3, 40e-3)):
    """Regular grid; returns (xout, zout, pnts) with pnts (nx*nz, 2) x-major."""
    xout = np.linspace(xlim[0], xlim[1], nx)
    zout = np.linspace(zlim[0], zlim[1], nz)
    X, Z = np.meshgrid(xout, zout, indexing="ij")
    return xout, zout, np.ascontiguousarray(np.stack([X.ravel(), Z.ravel()], axis=1))


def pw_origin(steers, xele):
    return np.array([xele[-1] if s <= 0 else xele[0] for s in np.atleast_1d(steers)])


def pw_tabs(steers, xele, pnts, c=C0, fnum=1.5):
    """Plane-wave tabbed geometry in float32: tautx (ntx, np), apodtx = 1,
    taurx (nrx, np), Hann receive apodisation at f-number ``fnum``."""
    f = np.float32
    steers = np.asarray(steers, np.float64)
    x = pnts[:, 0].astype(f)[None, :]
    z = pnts[:, 1].astype(f)[None, :]
    xr = pw_origin(steers, xele).astype(f)[:, None]
    s = np.sin(steers).astype(f)[:, None]
    co = np.cos(steers).astype(f)[:, None]
    c = f(c)
    tautx = (((x - xr) * s + z * co) / c).astype(f)
    apodtx = np.ones_like(tautx)
    xe = np.asarray(xele).astype(f)[:, None]
    dx = x - xe
    taurx = (np.sqrt(dx * dx + z * z) / c).astype(f)
    half = z / f(2 * fnum)
    ad = np.abs(dx)
    apodrx = np.where(ad <= half, f(0.5) * (f(1) + np.cos(f(np.pi) * ad / half)), f(0)).astype(f)
    return tautx, apodtx, taurx, apodrx


def pw_synthetic(steers, xele, pnts, c=C0, fnum=1.5, rpw=-10.0):
    """Synthetic-point parameters for plane waves (make_synthetic_points_2D
    with r0 = 0); returns a dict of the Synthetic fields (float32)."""
    steers = np.asarray(steers, np.float64)
    nrx = len(xele)
    orig = np.stack([pw_origin(steers, xele), np.zeros(steers.shape)], axis=1)
    nvectx = np.stack([np.sin(steers), np.cos(steers)], axis=1)
    ovectx = rpw * nvectx
    t0tx = np.sum(nvectx * (ovectx - orig), axis=-1) / c
    dxo = orig - ovectx
    alatx = np.arccos(np.abs(np.sum(dxo * nvectx, axis=-1)) / np.linalg.norm(dxo, axis=-1))
    f = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    return dict(ovectx=f(ovectx), nvectx=f(nvectx), t0tx=f(t0tx), doftx=f(np.zeros(len(steers))),
                alatx=f(alatx), ovecrx=f(np.stack([xele, np.zeros(nrx)], axis=1)),
                nvecrx=f(np.stack([np.zeros(nrx), np.ones(nrx)], axis=1)),
                alarx=f(np.full(nrx, np.arctan2(0.5, fnum))), c0=float(c), pnts=f(pnts))


def _pulse_a(fc=FC, bw=0.6, bwr=-6.0):
    ref = 10.0 ** (bwr / 20.0)
    return -((math.pi * fc * bw) ** 2) / (4.0 * math.log(ref))


def scatterers(n_points, n_speckle, xlim, zlim, seed):
    rng = np.random.default_rng(seed)
    xs = rng.uniform(xlim[0], xlim[1], n_points + n_speckle)
    zs = rng.uniform(zlim[0], zlim[1], n_points + n_speckle)
    amp = np.concatenate([np.ones(n_points), 0.05 * rng.standard_normal(n_speckle)])
    return xs, zs, amp


def cyst_speckle(n_speckle, xlim, zlim, center, radius, seed):
    """Diffuse speckle (normal amplitudes) with an anechoic disc removed."""
    rng = np.random.default_rng(seed)
    xs = rng.uniform(xlim[0], xlim[1], n_speckle)
    zs = rng.uniform(zlim[0], zlim[1], n_speckle)
    amp = rng.standard_normal(n_speckle)
    keep = (xs - center[0]) ** 2 + (zs - center[1]) ** 2 > radius ** 2
    return xs[keep], zs[keep], amp[keep]


def two_way_delays(xs, zs, steers, xele, c=C0):
    """(ntx, nrx, ns) plane-wave transmit + receive delays."""
    steers = np.asarray(steers, np.float64)
    xr = pw_origin(steers, xele)
    ttx = ((xs[None, :] - xr[:, None]) * np.sin(steers)[:, None]
           + zs[None, :] * np.cos(steers)[:, None]) / c
    trx = np.sqrt((xs[None, :] - np.asarray(xele)[:, None]) ** 2 + zs[None, :] ** 2) / c
    return ttx[:, None, :] + trx[None, :, :]


def simulate_rf(steers, xele, nt, xs, zs, amp, fs=FS, fc=FC, bw=0.6, snr_db=30.0, seed=0,
                c=C0):
    """NumPy channel data (ntx, nrx, nt) float32 (for small workloads)."""
    a = _pulse_a(fc, bw)
    tau = two_way_delays(xs, zs, steers, xele, c)  # (ntx, nrx, ns)
    ntx, nrx, ns = tau.shape
    half = int(math.ceil(math.sqrt(16.0 / a) * fs)) + 1
    rf = np.zeros((ntx, nrx, nt))
    k = np.arange(-half, half + 1)
    for s in range(ns):
        i0 = np.round(tau[:, :, s] * fs).astype(np.int64)
        idx = i0[..., None] + k
        t = idx / fs - tau[:, :, s, None]
        val = amp[s] * np.exp(-a * t * t) * np.cos(2 * math.pi * fc * t)
        ok = (idx >= 0) & (idx < nt)
        ti, ri, _ = np.nonzero(ok)
        np.add.at(rf, (ti, ri, idx[ok]), val[ok])
    rng = np.random.default_rng(seed + 7919)
    rms = np.sqrt(np.mean(rf * rf)) if np.any(rf) else 1.0
    rf += rng.standard_normal(rf.shape) * rms * 10 ** (-snr_db / 20)
    return rf.astype(np.float32)


def simulate_rf_delays_torch(tautx, taurx, amp, nt, device, fs=FS, fc=FC, bw=0.6, snr_db=30.0,
                             seed=0, chunk=64):
    """torch channel data (ntx, nrx, nt) float32 from given one-way delays
    tautx (ntx, S) and taurx (nrx, S) of S point scatterers (any medium)."""
    import torch
    a = _pulse_a(fc, bw)
    tautx = torch.as_tensor(tautx, device=device).to(torch.float64)
    taurx = torch.as_tensor(taurx, device=device).to(torch.float64)
    amp_t = torch.as_tensor(amp, device=device).to(torch.float64)
    ntx, nrx = tautx.shape[0], taurx.shape[0]
    half = int(math.ceil(math.sqrt(16.0 / a) * fs)) + 1
    k = torch.arange(-half, half + 1, device=device, dtype=torch.float64)
    rf = torch.zeros(ntx * nrx * nt, dtype=torch.float32, device=device)
    base = (torch.arange(ntx * nrx, device=device) * nt).view(ntx, nrx, 1, 1)
    for s0 in range(0, amp_t.numel(), chunk):
        tau = tautx[:, None, s0:s0 + chunk] + taurx[None, :, s0:s0 + chunk]
        am = amp_t[s0:s0 + chunk]
        i0 = torch.round(tau * fs)
        idx = i0[..., None] + k
        t = idx / fs - tau[..., None]
        val = am[None, None, :, None] * torch.exp(-a * t * t) * torch.cos(2 * math.pi * fc * t)
        ok = (idx >= 0) & (idx < nt)
        flat = (base + idx.clamp(0, nt - 1).long()).reshape(-1)
        rf.index_add_(0, flat, torch.where(ok, val, torch.zeros_like(val)).float().reshape(-1))
    rf = rf.view(ntx, nrx, nt)
    g = torch.Generator(device=device)
    g.manual_seed(int(seed) + 7919)
    rms = torch.sqrt(torch.mean(rf * rf)).clamp_min(1e-30)
    rf += torch.randn(rf.shape, generator=g, device=device) * rms * 10 ** (-snr_db / 20)
    return rf


def simulate_rf_torch(steers, xele, nt, xs, zs, amp, device, fs=FS, fc=FC, bw=0.6,
                      snr_db=30.0, seed=0, c=C0, chunk=64):
    """torch channel data (ntx, nrx, nt) float32 on ``device`` (large workloads)."""
    import torch
    a = _pulse_a(fc, bw)
    steers = torch.as_tensor(np.asarray(steers, np.float64), device=device)
    xe = torch.as_tensor(np.asarray(xele, np.float64), device=device)
    xr = torch.as_tensor(pw_origin(steers.cpu().numpy(), xele), device=device)
    ntx, nrx = steers.numel(), xe.numel()
    half = int(math.ceil(math.sqrt(16.0 / a) * fs)) + 1
    k = torch.arange(-half, half + 1, device=device, dtype=torch.float64)
    rf = torch.zeros(ntx * nrx * nt, dtype=torch.float32, device=device)
    xs_t = torch.as_tensor(xs, device=device)
    zs_t = torch.as_tensor(zs, device=device)
    amp_t = torch.as_tensor(amp, device=device)
    base = (torch.arange(ntx * nrx, device=device) * nt).view(ntx, nrx, 1, 1)
    for s0 in range(0, xs_t.numel(), chunk):
        x = xs_t[s0:s0 + chunk]
        z = zs_t[s0:s0 + chunk]
        am = amp_t[s0:s0 + chunk]
        ttx = ((x[None] - xr[:, None]) * torch.sin(steers)[:, None]
               + z[None] * torch.cos(steers)[:, None]) / c
        trx = torch.sqrt((x[None] - xe[:, None]) ** 2 + z[None] ** 2) / c
        tau = ttx[:, None, :] + trx[None, :, :]  # (ntx, nrx, S)
        i0 = torch.round(tau * fs)
        idx = i0[..., None] + k  # (ntx, nrx, S, W)
        t = idx / fs - tau[..., None]
        val = am[None, None, :, None] * torch.exp(-a * t * t) * torch.cos(2 * math.pi * fc * t)
        ok = (idx >= 0) & (idx < nt)
        flat = (base + idx.clamp(0, nt - 1).long()).reshape(-1)
        rf.index_add_(0, flat, torch.where(ok, val, torch.zeros_like(val)).float().reshape(-1))
    rf = rf.view(ntx, nrx, nt)
    g = torch.Generator(device=device)
    g.manual_seed(int(seed) + 7919)
    rms = torch.sqrt(torch.mean(rf * rf)).clamp_min(1e-30)
    rf += torch.randn(rf.shape, generator=g, device=device) * rms * 10 ** (-snr_db / 20)
    return rf


# --------------------------------------------------------------------------
# named workloads
# --------------------------------------------------------------------------
def workload_tabbed_small(seed=0, ntx=5, nrx=128, nt=2048, nx=128, nz=256, n_points=20):
    """Plain Tabbed DAS workload: 5 plane waves at -8..8 deg, 128 elements,
    2048 samples at 40 MHz, 20 scatterers, 256(z) x 128(x) grid over z 5-40 mm,
    x +-19 mm, Hann f/1.5 receive apodisation, float32 geometric tabs."""
    xele = linear_array(nrx)
    steers = np.radians(np.linspace(-8, 8, ntx))
    xout, zout, pnts = grid(nx, nz, (-19e-3, 19e-3), (5e-3, 40e-3))
    xs, zs, amp = scatterers(n_points, 0, (-15e-3, 15e-3), (5e-3, 40e-3), seed)
    rf = simulate_rf(steers, xele, nt, xs, zs, amp, seed=seed)
    tabs = pw_tabs(steers, xele, pnts)
    return dict(rf=rf, steers=steers, xele=xele, pnts=pnts, xout=xout, zout=zout, t0=0.0,
                dt=1.0 / FS, nt=nt, tabs=tabs)


def workload_slsc(nrx=128, ntx=75, nx=256, nz=256, sub=8, xlim=(-19e-3, 19e-3),
                  zlim=(2e-3, 42e-3)):
    """Coherence workload geometry: 75 plane waves over +-16 deg, 128 elements,
    an output grid nx x nz over a 40 mm depth whose axial pixels each span
    ``sub`` fine pixels of c/(2 fs) (one RF sample of two-way time), so an
    8-sample axial kernel is one wavelength at 5 MHz and 40 MHz sampling.
    Returns steers, xele, the synthetic parameters on the fine grid
    (nx * nz*sub pixels, x-major) and the fine grid shape."""
    xele = linear_array(nrx)
    steers = np.radians(np.linspace(-16, 16, ntx))
    nzf = nz * sub
    _, _, pnts = grid(nx, nzf, xlim, zlim)
    return steers, xele, pw_synthetic(steers, xele, pnts), (nx, nzf)


def sos_true_map(nzm=64, nxm=64, xlim=(-19e-3, 19e-3), zlim=(5e-3, 45e-3), seed=0,
                 c_top=1450.0, c_bottom=1540.0, z_layer=10e-3, blobs=6, amp=20.0):
    """Two-layer medium (c_top above z_layer, c_bottom below) with +-amp m/s
    Gaussian blobs at seeded positions, on the (nzm, nxm) node grid covering
    the image region; returns float32 (nzm, nxm) in m/s."""
    rng = np.random.default_rng(seed)
    x = xlim[0] + (np.arange(nxm) + 0.5) * (xlim[1] - xlim[0]) / nxm
    z = zlim[0] + (np.arange(nzm) + 0.5) * (zlim[1] - zlim[0]) / nzm
    Z, X = np.meshgrid(z, x, indexing="ij")
    c = np.where(Z < z_layer, c_top, c_bottom).astype(np.float64)
    for _ in range(blobs):
        cx, cz = rng.uniform(*xlim), rng.uniform(zlim[0] + 5e-3, zlim[1])
        r = rng.uniform(2e-3, 5e-3)
        c += rng.choice([-1.0, 1.0]) * amp * np.exp(-((X - cx) ** 2 + (Z - cz) ** 2) / (2 * r * r))
    return c.astype(np.float32)


def workload_sos_geometry(ntx=31, nrx=128, nx=256, nz=256, nt=4096, xlim=(-19e-3, 19e-3),
                          zlim=(5e-3, 45e-3), nzm=64, nxm=64):
    """Differentiable-DAS workload geometry: 31 plane waves over +-15 deg, 128
    elements, an nx x nz pixel grid over the image region and a 64 x 64
    slowness map covering it."""
    xele = linear_array(nrx)
    steers = np.radians(np.linspace(-15, 15, ntx))
    _, _, pnts = grid(nx, nz, xlim, zlim)
    return dict(steers=steers, xorig=pw_origin(steers, xele), xele=xele, pnts=pnts,
                gridshape=(nx, nz), nt=nt, xlim=xlim, zlim=zlim, nzm=nzm, nxm=nxm)
