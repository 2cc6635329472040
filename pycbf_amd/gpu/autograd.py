"""Differentiable Tabbed DAS: a torch.autograd.Function whose forward is
pycbf_das_tabbed and whose backward is pycbf_das_tabbed_bwd (hand-written
adjoint kernels: transposed gather into LDS bins for the channel data, a
per-(element, pixel) pass for the delays and apodisations, and the adjoint of
the natural-spline prefilter, of the k-order Hermite coefficients or of the
Akima / makima coefficients).

The forward is pycbf's Tabbed DAS (pycbf/cpu/pycbf.c:179-271 with the GPU
layouts of pycbf/gpu/__engines__.cu:813-864); pycbf itself has no gradients.
torch provides the device memory, the stream and the autograd graph only.
"""
from __future__ import annotations

import ctypes

from ..base import BeamformerException
from . import _engine as E

DIFF_KINDS = ("nearest", "linear", "cubic", "korder_cubic", "akima", "makima")


def _torch():
    import torch
    return torch


def _check(t, name, shape=None):
    torch = _torch()
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32):
        raise BeamformerException(f"{name} must be a float32 CUDA tensor")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise BeamformerException(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    return t.contiguous()


def _out_shape(sumtype, nfr, ntx, nrx, np_):
    per = {"none": (ntx, nrx, np_), "tx_only": (nrx, np_), "rx_only": (ntx, np_),
           "tx_and_rx": (np_,)}[sumtype]
    return (nfr,) + per


class _TabbedDAS:
    """Container of the non-tensor arguments (hashable-free, kept on ctx)."""

    def __init__(self, t0, dt, kind, sumtype, thresh, k=None, usf=1):
        if kind not in DIFF_KINDS:
            raise BeamformerException(f"differentiable kinds are {DIFF_KINDS}, got {kind!r}")
        if sumtype not in E.SUMTYPE:
            raise BeamformerException(f"unknown sumtype {sumtype!r}")
        self.t0, self.dt, self.kind, self.sumtype = float(t0), float(dt), kind, sumtype
        self.thresh = float(thresh)
        self.param, self.S = 1, None
        if kind == "nearest":
            if int(usf) != usf or usf < 1:
                raise BeamformerException("usf must be an integer >= 1")
            self.param = int(usf)
        elif usf != 1:
            raise BeamformerException("usf applies to the nearest kind only")
        if kind == "korder_cubic":
            from .core import make_S_by_k
            k = 8 if k is None else int(k)
            if k < 4 or k % 2:
                raise BeamformerException("'k' for korder_cubic must be at least four and even")
            self.param = k
            self.S_host = make_S_by_k(k).astype("float32")

    def S_on(self, device):
        """The (k, k) operator on ``device`` (korder_cubic), else None."""
        if self.S is None and self.kind == "korder_cubic":
            self.S = _torch().as_tensor(self.S_host, device=device)
        return self.S


def _das_function():
    torch = _torch()

    class DASFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, rf, tautx, apodtx, taurx, apodrx, cfg, tautx_lo=None, taurx_lo=None):
            nfr, ntx, nrx, nt = rf.shape
            np_ = tautx.shape[1]
            info = E.rfinfo(ntx, nrx, 0, cfg.t0, cfg.dt, nt)
            kind = E.INTERP[cfg.kind]
            need = E.lib.pycbf_das_workspace_bytes(ctypes.byref(info), nfr, np_, kind, cfg.param, 0)
            ws = torch.empty(max(int(need), 1), dtype=torch.uint8, device=rf.device)
            out = torch.empty(_out_shape(cfg.sumtype, nfr, ntx, nrx, np_), dtype=torch.float32,
                              device=rf.device)
            S = cfg.S_on(rf.device)
            if tautx_lo is None and taurx_lo is None:
                rc = E.lib.pycbf_das_tabbed(ctypes.byref(info), E.ptr(rf), nfr, E.ptr(tautx),
                                            E.ptr(apodtx), E.ptr(taurx), E.ptr(apodrx), None, np_,
                                            kind, cfg.param, E.ptr(S), cfg.thresh,
                                            E.SUMTYPE[cfg.sumtype], E.ptr(out), E.ptr(ws),
                                            ws.numel(), E.stream_of(rf.device))
                E.check(rc, "pycbf_das_tabbed")
            else:
                rc = E.lib.pycbf_das_tabbed_residual(
                    ctypes.byref(info), E.ptr(rf), nfr, E.ptr(tautx), E.ptr(apodtx),
                    E.ptr(taurx), E.ptr(apodrx), E.ptr(tautx_lo), E.ptr(taurx_lo), np_, kind,
                    cfg.param, E.ptr(S), cfg.thresh, E.SUMTYPE[cfg.sumtype], E.ptr(out),
                    E.ptr(ws), ws.numel(), E.stream_of(rf.device))
                E.check(rc, "pycbf_das_tabbed_residual")
            ctx.save_for_backward(rf, tautx, apodtx, taurx, apodrx)
            ctx.cfg = cfg
            ctx.lo = (tautx_lo, taurx_lo)
            return out

        @staticmethod
        def backward(ctx, gout):
            rf, tautx, apodtx, taurx, apodrx = ctx.saved_tensors
            cfg = ctx.cfg
            need_rf, need_ttx, need_atx, need_trx, need_arx = ctx.needs_input_grad[:5]
            gout = gout.to(torch.float32).contiguous()
            nfr, ntx, nrx, nt = rf.shape
            np_ = tautx.shape[1]
            mk = lambda ok, like: torch.empty_like(like) if ok else None  # noqa: E731
            grf, gttx, gatx = mk(need_rf, rf), mk(need_ttx, tautx), mk(need_atx, apodtx)
            gtrx, garx = mk(need_trx, taurx), mk(need_arx, apodrx)
            info = E.rfinfo(ntx, nrx, 0, cfg.t0, cfg.dt, nt)
            kind = E.INTERP[cfg.kind]
            need = E.lib.pycbf_das_bwd_workspace_bytes(ctypes.byref(info), nfr, np_, kind,
                                                       cfg.param)
            ws = torch.empty(max(int(need), 1), dtype=torch.uint8, device=rf.device)
            ttx_lo, trx_lo = ctx.lo
            rc = E.lib.pycbf_das_tabbed_bwd(
                ctypes.byref(info), E.ptr(rf), nfr, E.ptr(tautx), E.ptr(apodtx), E.ptr(taurx),
                E.ptr(apodrx), E.ptr(ttx_lo), E.ptr(trx_lo), np_, kind, cfg.param,
                E.ptr(cfg.S_on(rf.device)), cfg.thresh,
                E.SUMTYPE[cfg.sumtype], E.ptr(gout),
                E.ptr(grf), E.ptr(gttx), E.ptr(gatx), E.ptr(gtrx), E.ptr(garx), E.ptr(ws),
                ws.numel(), E.stream_of(rf.device))
            E.check(rc, "pycbf_das_tabbed_bwd")
            return grf, gttx, gatx, gtrx, garx, None, None, None

    return DASFunction


_FN = None


def das_tabbed(rf, tautx, apodtx, taurx, apodrx, t0, dt, interp="cubic", sumtype="tx_and_rx",
               thresh=0.0, k=None, usf=1, tautx_lo=None, taurx_lo=None):
    """Differentiable Tabbed DAS on the GPU.

    rf (ntx, nrx, nt) or (nfr, ntx, nrx, nt); tautx / apodtx (ntx, np);
    taurx / apodrx (nrx, np); all float32 CUDA tensors, any of which may
    require grad.  ``interp`` is 'nearest', 'linear', 'cubic' (natural
    spline), 'korder_cubic' (``k``, default 8), 'akima' or 'makima';
    'nearest' takes ``usf`` (the spline upsampling factor, default 1).  The
    output has the layout of TabbedBeamformer for ``sumtype``.
    ``thresh`` (default 0: every sample) drops samples with
    |apodtx * apodrx| < thresh in the forward and the adjoint alike.
    ``tautx_lo`` / ``taurx_lo`` (optional, constant) are float32 residuals of
    the delays (``ray_delays(..., residuals=True)``): the forward
    (pycbf_das_tabbed_residual) and the delay gradients then place each
    sample at the delays summed in double with their residuals."""
    global _FN
    if _FN is None:
        _FN = _das_function()
    cfg = _TabbedDAS(t0, dt, interp, sumtype, thresh, k, usf)
    squeeze = rf.dim() == 3
    rf4 = rf[None] if squeeze else rf
    if rf4.dim() != 4:
        raise BeamformerException("rf must be (ntx, nrx, nt) or (nfr, ntx, nrx, nt)")
    nfr, ntx, nrx, nt = rf4.shape
    rf4 = _check(rf4, "rf")
    np_ = tautx.shape[1] if tautx.dim() == 2 else -1
    tautx = _check(tautx, "tautx", (ntx, np_))
    apodtx = _check(apodtx, "apodtx", (ntx, np_))
    taurx = _check(taurx, "taurx", (nrx, np_))
    apodrx = _check(apodrx, "apodrx", (nrx, np_))
    if tautx_lo is not None:
        tautx_lo = _check(tautx_lo.detach(), "tautx_lo", (ntx, np_))
    if taurx_lo is not None:
        taurx_lo = _check(taurx_lo.detach(), "taurx_lo", (nrx, np_))
    out = _FN.apply(rf4, tautx, apodtx, taurx, apodrx, cfg, tautx_lo, taurx_lo)
    return out[0] if squeeze else out


def das_beamformer(rf, bf, tables=None):
    """Differentiable forward of a TabbedBeamformer or SyntheticBeamformer
    ``bf`` with respect to the channel data (and, for tables passed in, the
    tables): its interpolation, sumtype and thresh, through das_tabbed.  A
    SyntheticBeamformer's geometry is evaluated once into device tables
    (``bf.tables()``: pycbf_synthetic_tables, the float32 operations its
    gather evaluates per sample), so the geometry itself is a constant.
    ``tables`` (tautx, apodtx, taurx, apodrx) overrides the beamformer's."""
    torch = _torch()
    if tables is None:
        if hasattr(bf, "tables"):
            tables = bf.tables()
        else:
            tables = tuple(torch.as_tensor(getattr(bf, n), dtype=torch.float32,
                                           device=rf.device).contiguous()
                           for n in ("tautx", "apodtx", "taurx", "apodrx"))
    kind = bf.interp["kind"]
    return das_tabbed(rf, *tables, bf.t0, bf.dt, kind, bf.sumtype, bf.thresh,
                      k=bf.interp.get("k") if kind == "korder_cubic" else None,
                      usf=int(bf.interp.get("usf", 1)) if kind == "nearest" else 1)
