"""GPU DMAS beamformers (drop-in for pycbf/gpu/__advanced_bf_classes__.py).

The base DAS runs with sumtype 'tx_only' (lagaxis 'rx') or 'rx_only' (lagaxis
'tx'), then pycbf_dmas correlates the separated images at the requested lags.
Unlike the reference (pycbf/gpu/__advanced_bf_classes__.py:53,62 and
__advanced_engines__.cu:44-56) the lag values are taken from ``lags``, the
separated axis follows ``lagaxis``, and every frame of a batch is processed.
"""
from __future__ import annotations

from dataclasses import dataclass

from ..base import BeamformerException, DMASBeamformer
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
