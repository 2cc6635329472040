"""Ragged batches: many Synthetic cubic DAS / SLSC requests (own grids,
transmit and receive counts) beamformed by one persistent gather launch, with
the tile shape picked by an autotuner and the whole run capturable in a HIP
graph.  No reference counterpart: pycbf beamforms one configuration per call
(pycbf/gpu/__core_bf_classes__.py:68-118).

Usage::

    batch = RaggedBatch(requests, nt=4096, dt=1 / 40e6)
    batch.rf_views[r].copy_(channel_data_of_request_r)   # (ntx, nrx, nt) each
    batch.autotune()          # optional: pick tile / receive chunk
    batch.capture()           # optional: record the run in a HIP graph
    batch.run()               # enqueue (graph replay when captured)
    batch.outputs[r]          # DAS (nx, nz) or SLSC (nx/dsfx, nz/dsfz) image
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass, field

import numpy as np

from ..base import BeamformerException
from . import _engine as E
from .advanced import slsc_selection


def _torch():
    import torch
    return torch


@dataclass
class BatchRequest:
    """One request: Synthetic-point geometry fields (pycbf_amd.sim.pw_synthetic
    style dict: ovectx, nvectx, t0tx, alatx, doftx, ovecrx, nvecrx, alarx, c0,
    pnts with 2-D points on an x-major nx * nz grid), the grid shape, and the
    kind: 'das' (image summed over transmits and receives) or 'slsc'
    (per-receive images, then short-lag spatial coherence)."""
    params: dict
    gridshape: tuple
    kind: str = "das"
    lags: tuple = tuple(range(1, 21))
    kwin: int = 8
    kofs: int | None = None
    dsfx: int = 1
    dsfz: int = 1
    precision: str = "fp16"
    thresh: float = 1e-2
    extra: dict = field(default_factory=dict)


class RaggedBatch:
    """Device state, plan and workspace of a ragged batch (see module doc)."""

    def __init__(self, requests, nt, dt, t0=0.0, device=None):
        torch = _torch()
        if not requests:
            raise BeamformerException("empty batch")
        self.device = torch.device(device if device is not None else "cuda")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        dev = self.device
        self.requests = list(requests)
        self.nt, self.dt, self.t0 = int(nt), float(dt), float(t0)
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
        sizes, self._keep = [], []
        for q in self.requests:
            p = q.params
            ntx, nrx = f32(p["ovectx"]).shape[0], f32(p["ovecrx"]).shape[0]
            nx, nz = q.gridshape
            if f32(p["pnts"]).shape != (nx * nz, 2):
                raise BeamformerException("pnts must be (nx*nz, 2) for the request's grid")
            if q.kind not in ("das", "slsc"):
                raise BeamformerException(f"kind must be 'das' or 'slsc', got {q.kind!r}")
            sizes.append((ntx, nrx, nx, nz))
        self.sizes = sizes
        nch = sum(a * b for a, b, _, _ in sizes)
        self.rf_all = torch.zeros(nch * self.nt, dtype=torch.float32, device=dev)
        self.rf_views, off = [], 0
        for ntx, nrx, _, _ in sizes:
            n = ntx * nrx * self.nt
            self.rf_views.append(self.rf_all[off:off + n].view(ntx, nrx, self.nt))
            off += n
        das_px = sum(nx * nz for (_, _, nx, nz), q in zip(sizes, self.requests) if q.kind == "das")
        self._das_out = torch.zeros(max(das_px, 1), dtype=torch.float32, device=dev)
        self.outputs, reqs = [], (E.BatchReq * len(self.requests))()
        doff = 0
        for r, (q, (ntx, nrx, nx, nz)) in enumerate(zip(self.requests, sizes)):
            p = q.params
            txpack = np.zeros((ntx, E.TXPACK_FLOATS), np.float32)
            rxpack = np.zeros((nrx, E.RXPACK_FLOATS), np.float32)
            arrs = [f32(p[k]) for k in ("ovectx", "nvectx", "t0tx", "alatx", "doftx", "ovecrx",
                                        "nvecrx", "alarx")]
            rc = E.lib.pycbf_synthetic_pack(2, ntx, *(E.ptr(a) for a in arrs[:5]), nrx,
                                            *(E.ptr(a) for a in arrs[5:]), E.ptr(txpack),
                                            E.ptr(rxpack))
            E.check(rc, "pycbf_synthetic_pack")
            dtx = torch.as_tensor(txpack, device=dev)
            drx = torch.as_tensor(rxpack, device=dev)
            pv = torch.as_tensor(f32(p["pnts"]), device=dev)
            R = reqs[r]
            R.kind = 0 if q.kind == "das" else 1
            R.ntx, R.nrx, R.nx, R.nz, R.nt = ntx, nrx, nx, nz, self.nt
            R.t0, R.dt, R.c0, R.thresh = self.t0, self.dt, float(p["c0"]), float(q.thresh)
            R.txpack, R.rxpack, R.pvec = E.ptr(dtx), E.ptr(drx), E.ptr(pv)
            keep = [dtx, drx, pv]
            if q.kind == "das":
                out = self._das_out[doff:doff + nx * nz]
                doff += nx * nz
                R.out = E.ptr(out)
                self.outputs.append(out.view(nx, nz))
            else:
                kofs = -(int(q.kwin) // 2) if q.kofs is None else int(q.kofs)
                imrx = torch.empty((nrx, nx * nz), dtype=torch.float32, device=dev)
                oshape = ((nx + q.dsfx - 1) // q.dsfx, (nz + q.dsfz - 1) // q.dsfz)
                out = torch.empty(oshape, dtype=torch.float32, device=dev)
                apodrx = self._apodrx(dtx, drx, pv, float(p["c0"]), ntx, nrx, nx * nz)
                sel = slsc_selection(apodrx.view(nrx, nx, nz), q.kwin, kofs, q.dsfx, q.dsfz,
                                     device=dev)
                lags = np.asarray(q.lags, np.int32).ravel()
                if not 1 <= lags.size <= 64:
                    raise BeamformerException("1..64 lags")
                R.out, R.imrx, R.sel = E.ptr(out), E.ptr(imrx), E.ptr(sel)
                R.kwin, R.kofs, R.dsfx, R.dsfz = int(q.kwin), kofs, int(q.dsfx), int(q.dsfz)
                R.nlag, R.precision = int(lags.size), E.SLSC_PRECISION[q.precision]
                for i, m in enumerate(lags):
                    R.lags[i] = int(m)
                keep += [imrx, out, sel]
                self.outputs.append(out)
            self._keep.append(keep)
        self._reqs = reqs
        handle = ct.c_void_p()
        E.check(E.lib.pycbf_batch_create(ct.cast(reqs, ct.c_void_p), len(self.requests),
                                         ct.byref(handle)), "pycbf_batch_create")
        self._h = handle
        need = E.lib.pycbf_batch_workspace_bytes(self._h)
        self._ws = torch.empty(int(need), dtype=torch.uint8, device=dev)
        self.tile, self.rchunk = 16, 32
        self.configure(self.tile, self.rchunk)
        self._graph = None

    def _apodrx(self, dtx, drx, pv, c0, ntx, nrx, np_):
        torch = _torch()
        info = E.rfinfo(ntx, nrx, 2, self.t0, self.dt, self.nt)
        t = [torch.empty((n, np_), dtype=torch.float32, device=self.device)
             for n in (ntx, ntx, nrx, nrx)]
        rc = E.lib.pycbf_synthetic_tables(ct.byref(info), E.ptr(dtx), E.ptr(drx), c0, E.ptr(pv),
                                          np_, *(E.ptr(x) for x in t),
                                          E.stream_of(self.device))
        E.check(rc, "pycbf_synthetic_tables")
        return t[3]

    @property
    def samples(self):
        """tx * rx * pixel samples gathered per run (SLSC requests: their
        per-receive images)."""
        return sum(ntx * nrx * nx * nz for ntx, nrx, nx, nz in self.sizes)

    def configure(self, tile, rchunk):
        rc = E.lib.pycbf_batch_configure(self._h, int(tile), int(rchunk), E.ptr(self._ws),
                                         self._ws.numel(), E.stream_of(self.device))
        E.check(rc, "pycbf_batch_configure")
        self.tile, self.rchunk = int(tile), int(rchunk)
        self._graph = None

    def autotune(self, reps=3):
        """Times every (tile, receive chunk) pair on this batch and keeps the
        fastest; returns (tile, rchunk, {(tile, rchunk): ms})."""
        bt, bc = ct.c_int(), ct.c_int()
        times = (ct.c_float * 12)()
        rc = E.lib.pycbf_batch_autotune(self._h, E.ptr(self.rf_all), int(reps), E.ptr(self._ws),
                                        self._ws.numel(), E.stream_of(self.device), ct.byref(bt),
                                        ct.byref(bc), ct.cast(times, ct.c_void_p))
        E.check(rc, "pycbf_batch_autotune")
        self.tile, self.rchunk = bt.value, bc.value
        self._graph = None
        grid = {(t, c): times[i * 4 + j] for i, t in enumerate((8, 16, 32))
                for j, c in enumerate((16, 32, 64, 128))}
        return self.tile, self.rchunk, grid

    def _launch(self):
        rc = E.lib.pycbf_batch_run(self._h, E.ptr(self.rf_all), E.ptr(self._ws),
                                   self._ws.numel(), E.stream_of(self.device))
        E.check(rc, "pycbf_batch_run")

    def capture(self):
        """Records one run in a HIP graph (replayed by run())."""
        torch = _torch()
        self._launch()          # warm: everything the run touches exists
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launch()
        self._graph = g

    def run(self):
        if self._graph is not None:
            self._graph.replay()
        else:
            self._launch()

    def latencies_ms(self):
        """Completion time of each request in the last run (ms after its start)."""
        out = (ct.c_double * len(self.requests))()
        rc = E.lib.pycbf_batch_latencies(self._h, E.ptr(self._ws), self._ws.numel(),
                                         ct.cast(out, ct.c_void_p))
        E.check(rc, "pycbf_batch_latencies")
        return np.array(out[:], dtype=np.float64)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            E.lib.pycbf_batch_destroy(h)
            self._h = None
