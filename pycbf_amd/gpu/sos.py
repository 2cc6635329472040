"""Speed-of-sound estimation through the differentiable DAS.

The chain slowness map -> straight-ray delays (pycbf_ray_delays, adjoint
pycbf_ray_delays_bwd) -> per-receive-channel or per-transmit DAS images
(pycbf_das_tabbed_residual, adjoint pycbf_das_tabbed_bwd) -> a scalar loss
is differentiated end to end: each stage is a torch.autograd.Function over
hand-written kernels, and only the scalar loss on the images is written with
torch operations.  Two losses: the cross-angle consistency of the
per-transmit images (SoSProblem.angle_loss) and an image-focus loss (lag-one
coherence and energy of the per-receive images, SoSProblem.loss).

The reference corrects delays with an external eikonal solver
(examples/rough_sos_correction.ipynb); this module states its own
straight-ray model (see pycbf_ray_delays in include/pycbf_hip.h).
"""
from __future__ import annotations

import ctypes  # noqa: F401
from dataclasses import dataclass

import numpy as np

from ..base import BeamformerException
from . import _engine as E
from .autograd import das_tabbed


def _torch():
    import torch
    return torch


@dataclass
class SlownessGrid:
    """Node (i, k) of an (nzm, nxm) map sits at (xm0 + (k + 1/2) dxm,
    zm0 + (i + 1/2) dzm); the map is bilinear and edge-extended."""
    nzm: int
    nxm: int
    xm0: float
    zm0: float
    dxm: float
    dzm: float

    @staticmethod
    def covering(xlim, zlim, nxm, nzm):
        return SlownessGrid(nzm, nxm, float(xlim[0]), float(zlim[0]),
                            (xlim[1] - xlim[0]) / nxm, (zlim[1] - zlim[0]) / nzm)

    def centres(self):
        x = self.xm0 + (np.arange(self.nxm) + 0.5) * self.dxm
        z = self.zm0 + (np.arange(self.nzm) + 0.5) * self.dzm
        return x, z


def _ray_function():
    torch = _torch()

    class RayDelays(torch.autograd.Function):
        @staticmethod
        def forward(ctx, slow, geo, steer, xorig, xele, pvec, c0, nsteps):
            ntx, nrx, np_ = steer.numel(), xele.numel(), pvec.shape[0]
            mk = lambda n: torch.empty((n, np_), dtype=torch.float32, device=slow.device)  # noqa: E731
            tautx, taurx, tautx_lo, taurx_lo = mk(ntx), mk(nrx), mk(ntx), mk(nrx)
            rc = E.lib.pycbf_ray_delays(E.ptr(slow), geo.nzm, geo.nxm, geo.xm0, geo.zm0, geo.dxm,
                                        geo.dzm, E.ptr(steer), E.ptr(xorig), ntx, float(c0),
                                        E.ptr(xele), nrx, E.ptr(pvec), np_, int(nsteps),
                                        E.ptr(tautx), E.ptr(taurx), E.ptr(tautx_lo),
                                        E.ptr(taurx_lo), E.stream_of(slow.device))
            E.check(rc, "pycbf_ray_delays")
            ctx.save_for_backward(slow, steer, xorig, xele, pvec)
            ctx.geo, ctx.c0, ctx.nsteps = geo, float(c0), int(nsteps)
            ctx.mark_non_differentiable(tautx_lo, taurx_lo)
            return tautx, taurx, tautx_lo, taurx_lo

        @staticmethod
        def backward(ctx, gtautx, gtaurx, _glo_tx=None, _glo_rx=None):
            slow, steer, xorig, xele, pvec = ctx.saved_tensors
            geo = ctx.geo
            gslow = torch.empty_like(slow)
            gt = None if gtautx is None else gtautx.to(torch.float32).contiguous()
            gr = None if gtaurx is None else gtaurx.to(torch.float32).contiguous()
            need = E.lib.pycbf_ray_delays_bwd_workspace_bytes(geo.nzm, geo.nxm)
            ws = torch.empty(max(int(need), 1), dtype=torch.uint8, device=slow.device)
            rc = E.lib.pycbf_ray_delays_bwd(E.ptr(slow), geo.nzm, geo.nxm, geo.xm0, geo.zm0,
                                            geo.dxm, geo.dzm, E.ptr(steer), E.ptr(xorig),
                                            steer.numel(), ctx.c0, E.ptr(xele), xele.numel(),
                                            E.ptr(pvec), pvec.shape[0], ctx.nsteps, E.ptr(gt),
                                            E.ptr(gr), E.ptr(gslow), E.ptr(ws), ws.numel(),
                                            E.stream_of(slow.device))
            E.check(rc, "pycbf_ray_delays_bwd")
            return gslow, None, None, None, None, None, None, None

    return RayDelays


_RAY = None


def ray_delays(slow, geo: SlownessGrid, steer, xorig, xele, pvec, c0=1540.0, nsteps=64,
               residuals=False):
    """Straight-ray (tautx (ntx, np), taurx (nrx, np)) through the slowness
    map ``slow`` (nzm, nxm) float32 CUDA tensor (differentiable); steer /
    xorig (ntx), xele (nrx), pvec (np, 2) float32 CUDA tensors.  With
    ``residuals`` also the float32 residuals (tautx_lo, taurx_lo) of the
    delays, evaluated in double (constants for das_tabbed's adjoint)."""
    global _RAY
    torch = _torch()
    if _RAY is None:
        _RAY = _ray_function()
    for name, t in (("slow", slow), ("steer", steer), ("xorig", xorig), ("xele", xele),
                    ("pvec", pvec)):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32):
            raise BeamformerException(f"{name} must be a float32 CUDA tensor")
    if tuple(slow.shape) != (geo.nzm, geo.nxm):
        raise BeamformerException(f"slow must be {(geo.nzm, geo.nxm)}")
    out = _RAY.apply(slow.contiguous(), geo, steer.contiguous(), xorig.contiguous(),
                     xele.contiguous(), pvec.contiguous(), c0, nsteps)
    return out if residuals else out[:2]


def lag1_coherence(imrx, kwin=5):
    """Mean (over element pairs) lag-1 correlation of per-receive-channel images
    imrx (nrx, nx, nz) over a sliding axial window of kwin samples; returns
    (nx, nz).  Differentiable torch operations (the scalar loss only)."""
    torch = _torch()
    F = torch.nn.functional
    x = imrx.permute(1, 0, 2).reshape(-1, 1, imrx.shape[2])  # (nx*nrx, 1, nz)

    def box(v):
        return F.avg_pool1d(v, kwin, stride=1, padding=kwin // 2, count_include_pad=False)

    nrx = imrx.shape[0]
    m = box(x).reshape(imrx.shape[1], nrx, -1)
    m2 = box(x * x).reshape(imrx.shape[1], nrx, -1)
    xs = imrx.permute(1, 0, 2)                                 # (nx, nrx, nz)
    prod = (xs[:, :-1] * xs[:, 1:]).reshape(-1, 1, imrx.shape[2])
    mp = box(prod).reshape(imrx.shape[1], nrx - 1, -1)
    cov = mp - m[:, :-1] * m[:, 1:]
    var = (m2 - m * m).clamp_min(0.0)
    rho = cov / torch.sqrt(var[:, :-1] * var[:, 1:] + 1e-30)
    return rho.mean(dim=1)


def focus_loss(imrx, energy0, kwin=5, weight=0.1):
    """-mean lag-1 coherence - weight * (DAS image energy / energy0)."""
    img = imrx.sum(dim=0)
    return -lag1_coherence(imrx, kwin).mean() - weight * (img * img).mean() / energy0


@dataclass
class SoSProblem:
    """Channel data and geometry of a speed-of-sound estimate: plane waves
    (steer, xorig), receive elements xele, pixels pvec (nx*nz, 2) x-major on an
    (nx, nz) grid, rf (ntx, nrx, nt), and the slowness-map grid."""
    rf: object
    steer: object
    xorig: object
    xele: object
    pvec: object
    gridshape: tuple
    geo: SlownessGrid
    t0: float
    dt: float
    c0: float = 1540.0
    nsteps: int = 64
    kwin: int = 5
    interp: str = "cubic"
    energy_weight: float = 0.1
    smooth_weight: float = 0.0

    def images(self, cmap):
        """Per-receive-channel images (nrx, nx, nz) for the sound-speed map
        cmap (nzm, nxm) in m/s (differentiable)."""
        slow = 1.0 / cmap
        tautx, taurx, tlo, rlo = ray_delays(slow, self.geo, self.steer, self.xorig, self.xele,
                                            self.pvec, self.c0, self.nsteps, residuals=True)
        ones_tx = self._ones(tautx)
        ones_rx = self._ones(taurx)
        imrx = das_tabbed(self.rf, tautx, ones_tx, taurx, ones_rx, self.t0, self.dt, self.interp,
                          "tx_only", 0.0, tautx_lo=tlo, taurx_lo=rlo)
        nx, nz = self.gridshape
        return imrx.reshape(-1, nx, nz)

    def tx_images(self, cmap):
        """Per-transmit images (ntx, nx, nz) for the sound-speed map cmap
        (nzm, nxm) in m/s (differentiable): each plane wave's image summed
        over the receive elements."""
        slow = 1.0 / cmap
        tautx, taurx, tlo, rlo = ray_delays(slow, self.geo, self.steer, self.xorig, self.xele,
                                            self.pvec, self.c0, self.nsteps, residuals=True)
        im = das_tabbed(self.rf, tautx, self._ones(tautx), taurx, self._ones(taurx), self.t0,
                        self.dt, self.interp, "rx_only", 0.0, tautx_lo=tlo, taurx_lo=rlo)
        nx, nz = self.gridshape
        return im.reshape(-1, nx, nz)

    def angle_loss(self, cmap, lag=3):
        """Cross-angle consistency of the map: the energy of the differences
        between the images of plane waves `lag` steering steps apart, relative
        to the images' energy, plus smooth_weight times the mean squared
        neighbour difference of the map in m/s.  With the right delays every
        plane wave images the scatterers at the same place, so the
        differences shrink (the principle of computed ultrasound tomography
        in echo mode).  Kernels in float32, the scalar loss in float64."""
        im = self.tx_images(cmap).double()
        d = im[lag:] - im[:-lag]
        loss = (d * d).sum() / (im * im).sum()
        if self.smooth_weight > 0.0:
            c = cmap.double()
            rough = ((c[1:] - c[:-1]) ** 2).mean() + ((c[:, 1:] - c[:, :-1]) ** 2).mean()
            loss = loss + self.smooth_weight * rough
        return loss

    def _ones(self, like):
        torch = _torch()
        key = ("_ones", tuple(like.shape))
        if not hasattr(self, "_cache"):
            self._cache = {}
        if key not in self._cache:
            self._cache[key] = torch.ones(like.shape, dtype=torch.float32, device=like.device)
        return self._cache[key]

    def loss(self, cmap, energy0):
        """Focus loss of the map (the kernels run in float32; the scalar loss
        on their images is evaluated in float64) plus smooth_weight times the
        mean squared neighbour difference of the map in m/s."""
        loss = focus_loss(self.images(cmap).double(), energy0, self.kwin, self.energy_weight)
        if self.smooth_weight > 0.0:
            c = cmap.double()
            rough = ((c[1:] - c[:-1]) ** 2).mean() + ((c[:, 1:] - c[:, :-1]) ** 2).mean()
            loss = loss + self.smooth_weight * rough
        return loss
