"""Delay tabs and apodisations of a 2-D linear array, computed on the GPU.

``make_tabs_and_apods_2D`` takes the arguments of pycbf's helper of the same
name (pycbf/helpers/__full_aperture_2D__.py:170-203) and returns the same
dict keys, as float32 CUDA tensors ready for ``TabbedBeamformer`` (the
apodisations as 1.0 / 0.0; the reference returns float64 tabs and bool
apodisations that the beamformer converts to float32).  The per-pixel work
runs in ``pycbf_make_tabs_2d``, in float64 with the reference's operation
order, so the float32 tables equal the reference's tables converted to
float32.  The per-transmit constants below use the reference's own NumPy
calls:
* plane waves, :150-160: the origin (last element when steering <= 0, else
  the first), the steering corrected for ctx, its normal and the normal
  turned by +90 degrees;
* divergent waves, :62-76: the virtual source at -|r0| along the steering
  and tau0, its earliest arrival at an element;
* focused waves, :270-314 (through :110): the synthetic point, its normal,
  t0 and the half-aperture angle; the acceptance test arccos(v) < ala
  (:125) becomes vstar <= v <= 1, with vstar the smallest double that NumPy's
  arccos accepts (bisection over doubles).
"""
from __future__ import annotations

import numpy as np

from ..base import BeamformerException
from . import _engine as E


def _torch():
    import torch
    return torch


def arccos_threshold(ala: float) -> float:
    """Smallest double v in [0, 1] with np.arccos(v) < ala (2.0 when no v
    qualifies).  np.arccos is monotone decreasing, so for v >= 0
    np.arccos(v) < ala holds exactly when vstar <= v <= 1."""
    ala = float(ala)
    acc = lambda v: bool(np.arccos(np.float64(v)) < ala)  # noqa: E731
    if not acc(1.0):
        return 2.0
    if acc(0.0):
        return 0.0
    lo = int(np.float64(0.0).view(np.int64))
    hi = int(np.float64(1.0).view(np.int64))
    while hi - lo > 1:  # acc(lo) is false, acc(hi) true; non-negative doubles order like their bits
        mid = (lo + hi) // 2
        if acc(np.int64(mid).view(np.float64)):
            hi = mid
        else:
            lo = mid
    return float(np.int64(hi).view(np.float64))


def transmit_constants(steers, r0, ctx, cm, xele):
    """(mode, txc): the per-transmit rows of pycbf_make_tabs_2d (ntx x 8
    float64) for a synthetic source radius r0 (0: plane wave, < 0: divergent,
    > 0: focused)."""
    steers = np.atleast_1d(np.asarray(steers, np.float64))
    xele = np.asarray(xele, np.float64)
    ntx = len(steers)
    txc = np.zeros((ntx, 8), np.float64)
    if r0 == 0:
        O = np.array([[xele[-1] if s <= 0 else xele[0] for s in steers], np.zeros(steers.shape)]).T
        st = np.arcsin(cm * np.sin(steers) / ctx)
        txc[:, 0:2] = O
        txc[:, 2] = np.sin(st)
        txc[:, 3] = np.cos(st)
        txc[:, 4] = -np.cos(st)
        txc[:, 5] = np.sin(st)
        txc[:, 6] = xele[0]
        txc[:, 7] = xele[-1]
        return E.TABS["plane"], txc
    if ctx != cm:
        raise BeamformerException("Global SOS correction has not been implemented for "
                                  "divergent or focused wave imaging")
    if r0 < 0:
        r = -np.abs(r0)
        O = np.array([r * np.sin(steers), r * np.cos(steers)]).T
        E_ = np.array([xele, np.zeros_like(xele)]).T
        dX_OE = E_[None, :, :] - O[:, None, :]
        tau0 = np.min(np.linalg.norm(dX_OE, axis=-1) / ctx, axis=1)
        txc[:, 0:2] = O
        txc[:, 2] = tau0
        return E.TABS["divergent"], txc
    ovec = r0 * np.array([np.sin(steers), np.cos(steers)]).T
    dxmax = np.max(np.hypot(xele[None, :] - ovec[:, 0, None], ovec[:, 1, None]), axis=1)
    t0 = dxmax / cm
    alneg = np.arctan2(ovec[:, 0] - xele[0], ovec[:, 1])
    alpos = np.arctan2(ovec[:, 0] - xele[-1], ovec[:, 1])
    ala = -(alpos - alneg) / 2
    alst = (alpos + alneg) / 2
    txc[:, 0:2] = ovec
    txc[:, 2] = np.sin(alst)
    txc[:, 3] = np.cos(alst)
    txc[:, 4] = t0
    txc[:, 5] = [arccos_threshold(a) for a in ala]
    return E.TABS["focused"], txc


def make_tabs_and_apods_2D(steers, r0, ctx, cm, xele, xout, zout, fnum, device=None):
    """GPU make_tabs_and_apods_2D: dict of float32 tensors tautx / apodtx
    (ntx, nx*nz) and taurx / apodrx (nrx, nx*nz) on ``device`` (default
    cuda:0), pixels x-major as meshgrid(xout, zout, indexing='ij')."""
    torch = _torch()
    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    mode, txc = transmit_constants(steers, r0, ctx, cm, xele)
    xele = np.ascontiguousarray(np.asarray(xele, np.float64).ravel())
    xout = np.ascontiguousarray(np.asarray(xout, np.float64).ravel())
    zout = np.ascontiguousarray(np.asarray(zout, np.float64).ravel())
    ntx, nrx, nx, nz = len(txc), len(xele), len(xout), len(zout)
    d = lambda a: torch.as_tensor(a, device=dev)  # noqa: E731
    txc_d, xele_d, xout_d, zout_d = d(np.ascontiguousarray(txc)), d(xele), d(xout), d(zout)
    npx = nx * nz
    out = {k: torch.empty((n, npx), dtype=torch.float32, device=dev)
           for k, n in (("tautx", ntx), ("apodtx", ntx), ("taurx", nrx), ("apodrx", nrx))}
    rc = E.lib.pycbf_make_tabs_2d(mode, ntx, E.ptr(txc_d), nrx, E.ptr(xele_d), E.ptr(xout_d), nx,
                                  E.ptr(zout_d), nz, float(cm), float(fnum), E.ptr(out["tautx"]),
                                  E.ptr(out["apodtx"]), E.ptr(out["taurx"]),
                                  E.ptr(out["apodrx"]), E.stream_of(dev))
    E.check(rc, "pycbf_make_tabs_2d")
    return out
