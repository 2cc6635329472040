"""FP64 NumPy restatements used as checkers (test infrastructure only).

* ``make_S`` restates ``__make_S_by_k__`` (pycbf/gpu/__core_bf_classes__.py:366-394):
  the k x k operator mapping k samples to the not-a-knot cubic spline's first
  derivatives at those samples (unit spacing).
* ``dmas`` restates the notebook DMAS formulas
  (examples/advanced_beamforming.ipynb:296-301 "SSR" and
  examples/develope_gpu_slsc_and_dmas_kernels.ipynb:292-303 "power" and
  "P-DMAS-SSR"), with lags taken from the ``lags`` array (the reference CUDA
  kernel ignores it, pycbf/gpu/__advanced_engines__.cu:44-56).
* ``slsc`` restates ``calc_lagn_coherence`` and the pixel loop of
  examples/advanced_beamforming.ipynb:65-72,205-223: per pixel, select the
  elements whose mask is set, need at least max(lags)+1 of them, per lag
  mean-centre each channel over the kernel window, population std (std <
  1e-10 -> NaN), pair-normalised cross correlation averaged over valid pairs
  of *selected* elements (nanmean), then averaged over the lags.
* ``slsc_image`` restates the notebook's pixel loop itself
  (examples/advanced_beamforming.ipynb:205-223) over a per-receive-channel
  image (nrx, nx, nz): output pixel (ix, iz) = (ixo*dsfx, izo*dsfz), axial
  window iz+kofs .. iz+kofs+kwin-1 clipped to the image (the notebook's
  iz-nkern .. iz+nkern is kwin = 2*nkern+1, kofs = -nkern), elements selected
  where the apodisation is non-zero over the whole window.
"""
from __future__ import annotations

import numpy as np


def make_S(k: int) -> np.ndarray:
    """k-order derivative operator S = C^-1 P (float64)."""
    diag = np.full(k, 4.0)
    diag[0] = diag[-1] = 1.0
    up = np.ones(k - 1)
    up[0] = 2.0
    lo = up[::-1].copy()
    C = np.diag(diag) + np.diag(up, 1) + np.diag(lo, -1)
    pu = np.full(k - 1, 3.0)
    pu[0] = 2.0
    pl = -pu[::-1]
    P = np.diag(pu, 1) + np.diag(pl, -1)
    P[0, 0] = -2.5
    P[-1, -1] = 2.5
    P[0, 2] = 0.5
    P[-1, -3] = -0.5
    return np.linalg.solve(C, P)


def dmas(im, lags, dmastype="ssr", sumlags=True):
    """DMAS over the separated axis (axis 0) of ``im`` (nsep, np) in float64."""
    im = np.asarray(im, np.float64)
    nsep = im.shape[0]
    lags = [int(m) for m in np.asarray(lags).ravel()]
    per = []
    for m in lags:
        if m <= 0 or m >= nsep:
            per.append(np.zeros(im.shape[1]))
            continue
        q = im[:-m] * im[m:]
        if dmastype == "ssr":
            per.append(np.sum(np.sign(q) * np.sqrt(np.abs(q)), axis=0))
        else:
            per.append(np.sum(q, axis=0))
    per = np.array(per)
    if dmastype == "pssr":
        if sumlags:
            s = per.sum(axis=0)
            return np.sign(s) * np.sqrt(np.abs(s))
        return np.sign(per) * np.sqrt(np.abs(per))
    return per.sum(axis=0) if sumlags else per


def lag_coherence(sub, lag):
    """calc_lagn_coherence (examples/advanced_beamforming.ipynb:65-72), FP64.

    sub is (window, n_selected)."""
    sub = np.asarray(sub, np.float64)
    n = sub.shape[1]
    centred = sub - sub.mean(axis=0, keepdims=True)
    sig = sub.std(axis=0)
    sig[sig < 1e-10] = np.nan
    cross = np.mean(centred[:, : n - lag] * centred[:, lag:], axis=0)
    r = cross / (sig[: n - lag] * sig[lag:])
    if r.size == 0 or np.all(np.isnan(r)):
        return np.nan
    return np.nanmean(r)


def slsc(X, mask, lags):
    """Short-lag spatial coherence per pixel.

    X: (K, nrx, np) kernel samples per element and pixel; mask: (nrx, np)
    element selection.  Returns (np,) float64 (0 where too few elements are
    selected, NaN where a lag has no valid pair)."""
    X = np.asarray(X, np.float64)
    mask = np.asarray(mask).astype(bool)
    lags = [int(m) for m in np.asarray(lags).ravel()]
    K, nrx, npix = X.shape
    out = np.zeros(npix)
    need = max(lags) + 1
    for p in range(npix):
        sel = np.nonzero(mask[:, p])[0]
        if sel.size < need:
            continue
        sub = X[:, sel, p]
        acc = 0.0
        for m in lags:
            acc += lag_coherence(sub, m) / len(lags)
        out[p] = acc
    return out


def slsc_image(im, apod, lags, kwin, kofs, dsfx=1, dsfz=1):
    """SLSC image of one frame, FP64 (pure-Python pixel loop: small cases).

    im: (nrx, nx, nz); apod: (nrx, nx, nz) or None; returns
    (ceil(nx/dsfx), ceil(nz/dsfz)) float64."""
    im = np.asarray(im, np.float64)
    nrx, nx, nz = im.shape
    lags = [int(m) for m in np.asarray(lags).ravel()]
    need = max(lags) + 1
    out = np.zeros(((nx + dsfx - 1) // dsfx, (nz + dsfz - 1) // dsfz))
    for ix in range(0, nx, dsfx):
        for iz in range(0, nz, dsfz):
            z0 = max(0, iz + kofs)
            z1 = min(nz, iz + kofs + kwin)
            if apod is None:
                sel = np.ones(nrx, bool)
            else:
                sel = np.all(np.asarray(apod)[:, ix, z0:z1] != 0, axis=-1)
            if np.sum(sel) < need:
                continue
            channels = im[sel, ix, z0:z1]
            acc = 0.0
            for m in lags:
                acc += lag_coherence(channels.T, m) / len(lags)
            out[ix // dsfx, iz // dsfz] = acc
    return out
