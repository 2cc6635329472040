"""The interpolator-accuracy dataset of the reference's
examples/compare_interpolation_methods.ipynb (cells at :49-123 and :217-218),
regenerated offline, and the ROI RMSE it reports (:285-297, :571-576).

64 elements at one-wavelength pitch, 5 MHz Gaussian pulses (bw 0.75) sampled
at 30 MHz, 300 point scatterers from default_rng(0), full synthetic aperture
(64 x 64 x 1078), a 128 x 504 grid with f/1 rectangular apodisation.

The notebook sums scipy.signal.gausspulse over the scatterers in float64; the
same closed form (exp(-a t^2) cos(2 pi fc t), a = -(pi fc bw)^2 / (4 ln 10^(-6/20)))
is evaluated here with torch on the host threads so the dataset builds in
seconds.  The reference's reported values, from the notebook's output cell:
"""
from __future__ import annotations

import math

import numpy as np

# examples/compare_interpolation_methods.ipynb:571-576 (output of the RMSE cell)
RMSE_NEAREST_VS_CUBIC = 35.48316740429086        # CPU nearest and GPU nearest
RMSE_KORDER126_VS_CUBIC = 0.0014545855716605876  # GPU korder_cubic, k = 126

FC, C, NELE = 5e6, 1540.0, 64


def dataset():
    """-> dict(rf (64,64,1078) float64, tabs (tautx, apodtx, taurx, apodrx),
    t0, dt, nt, shape (nx, 1, nz), roi slices (z, x))."""
    import torch
    fs = 6 * FC
    lam = C / FC
    dele = lam
    xele = dele * (np.arange(NELE) - (NELE - 1) / 2)
    ex, ey, ez = np.meshgrid(xele, 0, 0, indexing="ij")
    eles = np.array([ex.flatten(), ey.flatten(), ez.flatten()]).T
    xgrid = dele * (np.arange(0, NELE, 1 / 2) - (NELE - 1) / 2)
    zgrid = np.arange(1 * lam, NELE * dele, lam / 8)
    gx, gy, gz = np.meshgrid(xgrid, 0, zgrid, indexing="ij")
    grid = np.array([gx.flatten(), gy.flatten(), gz.flatten()]).T

    rng = np.random.default_rng(0)
    npnt = 300
    points = 1e-3 * np.array([rng.uniform(-2.5, 2.5, npnt), np.zeros(npnt),
                              rng.uniform(7.5, 12.5, npnt)]).T
    tmax = 2 * np.sqrt((dele * (NELE - 1)) ** 2 + (NELE * dele) ** 2) / C
    t = np.arange(0, tmax, 1 / fs)
    a = -(math.pi * FC * 0.75) ** 2 / (4.0 * math.log(pow(10.0, -6 / 20.0)))
    tt = torch.from_numpy(t)
    rf = torch.zeros((NELE, NELE, t.size), dtype=torch.float64)
    for point in points:
        tau = torch.from_numpy(np.linalg.norm(eles - point, axis=-1) / C)
        T = tt[None, None, :] - (tau[:, None, None] + tau[None, :, None])
        rf += torch.exp(-a * T * T) * torch.cos(2 * math.pi * FC * T)

    dx = grid[None, :, :] - eles[:, None, :]
    tau = np.linalg.norm(dx, axis=-1) / C
    apod = np.abs(dx[:, :, 0]) - dx[:, :, 2] / 2 < 0
    iz0 = int(np.argmin(np.abs(zgrid - 7.5e-3)))
    iz1 = int(np.argmin(np.abs(zgrid - 12.5e-3)))
    ix0 = int(np.argmin(np.abs(xgrid + 2.5e-3)))
    ix1 = int(np.argmin(np.abs(xgrid - 2.5e-3)))
    return dict(rf=rf.numpy(), tabs=(tau, apod, tau, apod), t0=float(t[0]),
                dt=float(t[1] - t[0]), nt=int(t.size), shape=gx.shape,
                roi=(slice(iz0, iz1), slice(ix0, ix1)))


def roi_rmse(im, im_ref, d):
    """RMSE over the scatterer ROI of two flat (np,) images, computed as the
    notebook does on float64 images reshaped to (nz, nx)."""
    a = np.asarray(im, np.float64).reshape(d["shape"]).T.squeeze()
    b = np.asarray(im_ref, np.float64).reshape(d["shape"]).T.squeeze()
    z, x = d["roi"]
    return float(np.sqrt(np.mean(np.abs(a[z, x] - b[z, x]) ** 2)))
