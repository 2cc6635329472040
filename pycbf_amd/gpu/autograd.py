"""Differentiable Tabbed DAS: a torch.autograd.Function whose forward is
pycbf_das_tabbed and whose backward is pycbf_das_tabbed_bwd (hand-written
adjoint kernels: transposed gather into LDS bins for the channel data, a
per-(element, pixel) pass for the delays and apodisations, and the adjoint of
the natural-spline prefilter, of the k-order Hermite coefficients or of the
Akima / makima coefficients).

The forward is pycbf's Tabbed DAS (pycbf/cpu/pycbf.c:179-271 with the GPU
layouts of pycbf/gpu/__engines__.cu:813-864); pycbf itself has no gradients.
das_synthetic does the same for the Synthetic DAS (pycbf/gpu/__engines__.cu:
382-470, 620-694) through pycbf_das_synthetic / pycbf_das_synthetic_bwd,
with gradients for the channel data and the speed of sound and no delay table.
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


def _synthetic_function():
    torch = _torch()

    class SyntheticDASFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, rf, c0, bf):
            nfr = rf.shape[0]
            info = bf._rfinfo()
            kind = bf.interp["kind"]
            need = E.lib.pycbf_das_workspace_bytes(ctypes.byref(info), nfr, bf.nop,
                                                   E.INTERP[kind], bf._interp_param(), 1)
            ws = torch.empty(max(int(need), 1), dtype=torch.uint8, device=rf.device)
            out = torch.empty(bf._out_shape(nfr), dtype=torch.float32, device=rf.device)
            c0f = float(c0)
            rc = E.lib.pycbf_das_synthetic(
                ctypes.byref(info), E.ptr(rf), nfr, E.ptr(bf._txpack), E.ptr(bf._rxpack), c0f,
                E.ptr(bf._pnts), E.ptr(bf._perm), bf.nop, E.INTERP[kind], bf._interp_param(),
                E.ptr(bf._S_dev()), float(bf.thresh), E.SUMTYPE[bf.sumtype], E.ptr(out),
                E.ptr(ws), ws.numel(), E.stream_of(rf.device))
            E.check(rc, "pycbf_das_synthetic")
            ctx.save_for_backward(rf)
            ctx.bf = bf
            ctx.c0 = c0f
            ctx.c0_shape = c0.shape if torch.is_tensor(c0) else None
            return out

        @staticmethod
        def backward(ctx, gout):
            (rf,) = ctx.saved_tensors
            bf = ctx.bf
            need_rf, need_c0 = ctx.needs_input_grad[:2]
            gout = gout.to(torch.float32).contiguous()
            nfr = rf.shape[0]
            grf = torch.empty_like(rf) if need_rf else None
            gc0 = torch.empty(1, dtype=torch.float32, device=rf.device) if need_c0 else None
            if grf is None and gc0 is None:
                return None, None, None
            info = bf._rfinfo()
            kind = bf.interp["kind"]
            need = E.lib.pycbf_das_bwd_workspace_bytes(ctypes.byref(info), nfr, bf.nop,
                                                       E.INTERP[kind], bf._interp_param())
            ws = torch.empty(max(int(need), 1), dtype=torch.uint8, device=rf.device)
            rc = E.lib.pycbf_das_synthetic_bwd(
                ctypes.byref(info), E.ptr(rf), nfr, E.ptr(bf._txpack), E.ptr(bf._rxpack),
                ctx.c0, E.ptr(bf._pnts), bf.nop, E.INTERP[kind], bf._interp_param(),
                E.ptr(bf._S_dev()), float(bf.thresh), E.SUMTYPE[bf.sumtype], E.ptr(gout),
                E.ptr(grf), None, None, None, None, E.ptr(gc0), E.ptr(ws), ws.numel(),
                E.stream_of(rf.device))
            E.check(rc, "pycbf_das_synthetic_bwd")
            if gc0 is not None:
                gc0 = gc0.reshape(ctx.c0_shape)
            return grf, gc0, None

    return SyntheticDASFunction


_SFN = None


def das_synthetic(rf, bf, c0=None):
    """Differentiable forward of a SyntheticBeamformer ``bf`` (its geometry,
    interpolation, sumtype and thresh) with respect to the channel data and
    the speed of sound, with no delay or apodisation table anywhere: the
    forward is pycbf_das_synthetic, the backward pycbf_das_synthetic_bwd,
    whose kernels re-evaluate the geometry per tile in registers.  ``c0`` is
    a 0-dim float tensor (set requires_grad for d loss / d c0) or None for
    ``bf.c0``.  The gradient takes every delay as a path length over c0, the
    transmit time intercepts t0tx included, as pycbf's helpers build them
    (t0tx = distance / c): the derivative of the image of the geometry
    rebuilt at another sound speed.  rf is
    (ntx, nrx, nt) or (nfr, ntx, nrx, nt) float32 on the beamformer's device;
    the output has the beamformer's frame layout."""
    global _SFN
    torch = _torch()
    if _SFN is None:
        _SFN = _synthetic_function()
    if not getattr(bf, "_synthetic", False):
        raise BeamformerException("das_synthetic needs a SyntheticBeamformer")
    squeeze = rf.dim() == 3
    rf4 = rf[None] if squeeze else rf
    rf4 = _check(rf4, "rf", (rf4.shape[0], bf.ntx, bf.nrx, bf.nt))
    if c0 is None:
        c0 = torch.tensor(float(bf.c0), dtype=torch.float32)
    out = _SFN.apply(rf4, c0, bf)
    return out[0] if squeeze else out


def das_beamformer(rf, bf, tables=None):
    """Differentiable forward of a TabbedBeamformer or SyntheticBeamformer
    ``bf`` with respect to the channel data (and, for tables passed in, the
    tables): its interpolation, sumtype and thresh.  A SyntheticBeamformer
    without ``tables`` goes through das_synthetic (geometry re-evaluated in
    registers, forward and backward, no table stored); a TabbedBeamformer, or
    ``tables`` (tautx, apodtx, taurx, apodrx) given, through das_tabbed."""
    torch = _torch()
    if tables is None:
        if getattr(bf, "_synthetic", False):
            return das_synthetic(rf, bf)
        else:
            tables = tuple(torch.as_tensor(getattr(bf, n), dtype=torch.float32,
                                           device=rf.device).contiguous()
                           for n in ("tautx", "apodtx", "taurx", "apodrx"))
    kind = bf.interp["kind"]
    return das_tabbed(rf, *tables, bf.t0, bf.dt, kind, bf.sumtype, bf.thresh,
                      k=bf.interp.get("k") if kind == "korder_cubic" else None,
                      usf=int(bf.interp.get("usf", 1)) if kind == "nearest" else 1)
