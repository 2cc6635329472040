"""GPU DMAS and SLSC beamformers (drop-in for pycbf/gpu/__advanced_bf_classes__.py,
plus SLSC, which the reference only has as a notebook loop).

The base DAS runs with sumtype 'tx_only' (lagaxis 'rx') or 'rx_only' (lagaxis
'tx'), then pycbf_dmas correlates the separated images at the requested lags.
Unlike the reference (pycbf/gpu/__advanced_bf_classes__.py:53,62 and
__advanced_engines__.cu:44-56) the lag values are taken from ``lags``, the
separated axis follows ``lagaxis``, and every frame of a batch is processed.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ..base import BeamformerException, DMASBeamformer, SLSCBeamformer
from . import _engine as E
from .core import SyntheticBeamformer, TabbedBeamformer, _torch


class _GPUDMAS:
    def _dmas_setup(self):
        self._lags = self._to_dev(self.lags, dtype=_torch().int32)

    def _dmas_buffer(self, nframes, buffer):
        torch = _torch()
        shape = (self.nop,) if self.sumlags else (self.nlags, self.nop)
        if nframes is not None:
            shape = (nframes,) + shape
        if buffer is None:
            return torch.empty(shape, dtype=torch.float32, device=self._dev)
        if not (isinstance(buffer, torch.Tensor) and tuple(buffer.shape) == shape
                and buffer.dtype == torch.float32 and buffer.device == self._dev
                and buffer.is_contiguous()):
            raise BeamformerException(
                f"buffer must be a contiguous float32 tensor of shape {shape} on {self._dev}")
        return buffer

    def _dmas_call(self, base_call, txrxt, out_as_numpy, buffer):
        rf, nframes = self._coerce_rf(txrxt)
        imsep = base_call(self, rf, out_as_numpy=False)
        nf = 1 if nframes is None else nframes
        nsep = self.nrx if self.lagaxis == "rx" else self.ntx
        imout = self._dmas_buffer(nframes, buffer)
        rc = E.lib.pycbf_dmas(E.ptr(imsep), nf, nsep, self.nop, E.ptr(self._lags), self.nlags,
                              E.DMASTYPE[self.dmastype], int(bool(self.sumlags)), E.ptr(imout),
                              E.stream_of(self._dev))
        E.check(rc, "pycbf_dmas")
        return imout.cpu().numpy() if out_as_numpy else imout


@dataclass(kw_only=True)
class TabbedDMASBeamformer(TabbedBeamformer, DMASBeamformer, _GPUDMAS):
    """Delay multiply and sum over precomputed delay tabs."""

    def __post_init__(self):
        DMASBeamformer.__post_init__(self)
        TabbedBeamformer.__post_init__(self)
        self._dmas_setup()

    def __call__(self, txrxt, out_as_numpy: bool = True, buffer=None):
        return self._dmas_call(TabbedBeamformer.__call__, txrxt, out_as_numpy, buffer)


@dataclass(kw_only=True)
class SyntheticDMASBeamformer(SyntheticBeamformer, DMASBeamformer, _GPUDMAS):
    """Delay multiply and sum over synthetic-point geometry."""

    def __post_init__(self):
        DMASBeamformer.__post_init__(self)
        SyntheticBeamformer.__post_init__(self)
        self._dmas_setup()

    def __call__(self, txrxt, out_as_numpy: bool = True, buffer=None):
        return self._dmas_call(SyntheticBeamformer.__call__, txrxt, out_as_numpy, buffer)


def usable_mask(apod, device):
    """uint8 mask (non-zero apodisation) on ``device``; uint8 tensors pass through."""
    torch = _torch()
    if isinstance(apod, torch.Tensor) and apod.dtype == torch.uint8:
        return apod.to(device).contiguous()
    a = apod if isinstance(apod, torch.Tensor) else torch.as_tensor(np.asarray(apod))
    return (a.to(device) != 0).to(torch.uint8).contiguous()


def slsc_selection(apod, kwin, kofs=None, dsfx=1, dsfz=1, device=None):
    """Usable-element bits per SLSC output pixel (pycbf_slsc_selection): an
    (nxo, nzo, 4) int64 CUDA tensor from the (nrx, nx, nz) receive
    apodisation (non-zero = usable)."""
    torch = _torch()
    if device is None:
        device = apod.device if isinstance(apod, torch.Tensor) and apod.is_cuda else "cuda"
    mask = usable_mask(apod, device)
    if mask.dim() != 3:
        raise BeamformerException("apod must be (nrx, nx, nz)")
    nrx, nx, nz = mask.shape
    kofs = -(int(kwin) // 2) if kofs is None else int(kofs)
    sel = torch.empty(((nx + dsfx - 1) // dsfx, (nz + dsfz - 1) // dsfz, 4), dtype=torch.int64,
                      device=mask.device)
    rc = E.lib.pycbf_slsc_selection(E.ptr(mask), nrx, nx, nz, int(kwin), kofs, int(dsfx),
                                    int(dsfz), E.ptr(sel), E.stream_of(mask.device))
    E.check(rc, "pycbf_slsc_selection")
    return sel


def slsc(im, lags, kwin, kofs=None, apod=None, dsfx=1, dsfz=1, precision="fp16", out=None,
         sel=None):
    """SLSC of per-receive-channel images on the GPU (pycbf_slsc).

    im: float32 CUDA tensor (nrx, nx, nz) or (nfr, nrx, nx, nz).  Elements are
    selected by ``sel`` (from slsc_selection, reusable across frames) or, when
    it is None, by ``apod`` ((nrx, nx, nz), non-zero = usable; None = all).
    Returns (nfr?, ceil(nx/dsfx), ceil(nz/dsfz)) float32."""
    torch = _torch()
    if not (isinstance(im, torch.Tensor) and im.is_cuda and im.dtype == torch.float32):
        raise BeamformerException("im must be a float32 CUDA tensor")
    squeeze = im.dim() == 3
    im4 = (im[None] if squeeze else im).contiguous()
    if im4.dim() != 4:
        raise BeamformerException("im must be (nrx, nx, nz) or (nfr, nrx, nx, nz)")
    nfr, nrx, nx, nz = im4.shape
    kofs = -(int(kwin) // 2) if kofs is None else int(kofs)
    lags_h = np.ascontiguousarray(np.asarray(lags).ravel(), np.int32)
    oshape = ((nx + dsfx - 1) // dsfx, (nz + dsfz - 1) // dsfz)
    if sel is None and apod is not None:
        if tuple(np.shape(apod)) != (nrx, nx, nz):
            raise BeamformerException(f"apod must be {(nrx, nx, nz)}")
        sel = slsc_selection(apod, kwin, kofs, dsfx, dsfz, device=im.device)
    if sel is not None and (tuple(sel.shape) != oshape + (4,) or sel.dtype != torch.int64
                            or not sel.is_contiguous()):
        raise BeamformerException(f"sel must be a contiguous int64 tensor of shape {oshape + (4,)}")
    shape = (nfr,) + oshape
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=im.device)
    elif tuple(out.shape) != shape or out.dtype != torch.float32 or not out.is_contiguous():
        raise BeamformerException(f"out must be a contiguous float32 tensor of shape {shape}")
    rc = E.lib.pycbf_slsc(E.ptr(im4), nfr, nrx, nx, nz, E.ptr(sel), int(kwin), kofs, int(dsfx),
                          int(dsfz), lags_h.ctypes.data, lags_h.size,
                          E.SLSC_PRECISION[precision], E.ptr(out), E.stream_of(im.device))
    E.check(rc, "pycbf_slsc")
    return out[0] if squeeze else out


class _GPUSLSC:
    def _slsc_setup(self, apodrx):
        nx, nz = self.gridshape
        if nx * nz != self.nop:
            raise BeamformerException(
                f"gridshape {self.gridshape} holds {nx * nz} pixels but there are {self.nop}")
        self._slsc_sel = slsc_selection(apodrx.reshape(self.nrx, nx, nz), self.kwin, self.kofs,
                                        self.dsfx, self.dsfz, device=self._dev)

    def _slsc_call(self, base_call, txrxt, out_as_numpy, buffer):
        torch = _torch()
        rf, nframes = self._coerce_rf(txrxt)
        imrx = base_call(self, rf, out_as_numpy=False)
        nf = 1 if nframes is None else nframes
        nx, nz = self.gridshape
        shape = self._slsc_outshape() if nframes is None else (nframes,) + self._slsc_outshape()
        if buffer is None:
            out = torch.empty(shape, dtype=torch.float32, device=self._dev)
        else:
            out = buffer
            if not (isinstance(out, torch.Tensor) and tuple(out.shape) == shape
                    and out.dtype == torch.float32 and out.device == self._dev
                    and out.is_contiguous()):
                raise BeamformerException(
                    f"buffer must be a contiguous float32 tensor of shape {shape} on {self._dev}")
        slsc(imrx.reshape(nf, self.nrx, nx, nz), self.lags, self.kwin, self.kofs, None,
             self.dsfx, self.dsfz, self.precision, out=out.reshape((nf,) + self._slsc_outshape()),
             sel=self._slsc_sel)
        return out.cpu().numpy() if out_as_numpy else out


@dataclass(kw_only=True)
class TabbedSLSCBeamformer(TabbedBeamformer, SLSCBeamformer, _GPUSLSC):
    """Short-lag spatial coherence over precomputed delay tabs; elements are
    selected by ``apodrx``."""

    def __post_init__(self):
        SLSCBeamformer.__post_init__(self)
        TabbedBeamformer.__post_init__(self)
        self._slsc_setup(self._apodrx)

    def __call__(self, txrxt, out_as_numpy: bool = True, buffer=None):
        return self._slsc_call(TabbedBeamformer.__call__, txrxt, out_as_numpy, buffer)


@dataclass(kw_only=True)
class SyntheticSLSCBeamformer(SyntheticBeamformer, SLSCBeamformer, _GPUSLSC):
    """Short-lag spatial coherence over synthetic-point geometry; elements are
    selected by the receive acceptance apodisation of the geometry."""

    def __post_init__(self):
        SLSCBeamformer.__post_init__(self)
        SyntheticBeamformer.__post_init__(self)
        self._slsc_setup(self.tables()[3])

    def __call__(self, txrxt, out_as_numpy: bool = True, buffer=None):
        return self._slsc_call(SyntheticBeamformer.__call__, txrxt, out_as_numpy, buffer)
