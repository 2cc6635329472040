"""Beamformer base classes: the API contract of pycbf's
``pycbf/__bf_base_classes__.py`` and ``pycbf/__advanced_bf_classes__.py``.

Field names, defaults, shape rules and exception type are kept so that code
written against ``pycbf.gpu`` runs unchanged.  Departures (each fixes a
reference defect):

* ``korder_cubic`` without ``k`` defaults to k = 8 (the reference crashes,
  ``pycbf/__bf_base_classes__.py:85-88``; 8 is its GPU default).
* ``Tabbed`` also checks ``apodrx.shape[1]`` (the reference checks
  ``apodtx`` twice, ``:126``).
* ``thresh`` (default 1e-2, CPU semantics ``|apodtx*apodrx| >= thresh``,
  ``pycbf/cpu/pycbf.c:201,267``) is a field of every beamformer.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import ClassVar, Literal

import numpy as np

__PYCBF_DATATYPE__ = np.float32

INTERP_KINDS = ("nearest", "linear", "akima", "makima", "korder_cubic", "cubic")
SUMTYPES = ("none", "tx_only", "rx_only", "tx_and_rx")


class BeamformerException(Exception):
    """Validation / engine error (pycbf/__bf_base_classes__.py:16)."""


def _shape(a):
    return tuple(np.shape(a))


def _ndim(a):
    return len(_shape(a))


@dataclass(kw_only=True)
class Beamformer:
    """Base of all beamformers (pycbf/__bf_base_classes__.py:18-95).

    ``sumtype`` picks the axes summed inside the engine: ``'tx_and_rx'`` (DAS
    image), ``'tx_only'`` (one image per receive channel), ``'rx_only'`` (one
    per transmit) or ``'none'`` (every (tx, rx) pair).  ``interp`` is a dict
    with ``kind`` in nearest / linear / akima / makima / korder_cubic / cubic,
    plus ``k`` (korder_cubic, even, >= 4) or ``usf`` (nearest upsampling).
    """
    nbf: ClassVar[int] = 0
    id: int = field(init=False)
    interp: dict = field(default_factory=lambda: {"kind": "cubic"})
    sumtype: Literal["none", "tx_only", "rx_only", "tx_and_rx"] = "tx_and_rx"
    thresh: float = 1e-2

    ntx: int = field(init=False)
    nrx: int = field(init=False)
    nop: int = field(init=False)

    def __post_init__(self):
        self.id = Beamformer.nbf
        Beamformer.nbf += 1
        self.interp = dict(self.interp)
        self.__check_interp__()
        self.__check_sumtype__()

    def __check_sumtype__(self):
        if self.sumtype not in SUMTYPES:
            raise BeamformerException(
                f"sumtype was {self.sumtype} but must be one of the following: "
                + ", ".join(SUMTYPES))

    def __check_interp__(self):
        kind = self.interp.get("kind", None)
        if kind is None or kind not in INTERP_KINDS:
            raise BeamformerException(
                f"interp['kind'] was {kind} but must be one of the following: "
                + ", ".join(INTERP_KINDS))
        if kind == "korder_cubic":
            k = self.interp.get("k", None)
            if k is None:
                k = 8
            k = int(k)
            if (k < 4) or (k % 2 != 0):
                raise BeamformerException(
                    "'k' for interp kind 'korder_cubic' must be at least four and even")
            self.interp["k"] = k
        elif kind == "cubic":
            pass
        else:
            usf = self.interp.get("usf", None)
            if usf is None:
                self.interp["usf"] = 1
            elif int(usf) < 1:
                raise BeamformerException(
                    f"'usf' for interp kind '{kind}' must be an interger greater than 1")
            else:
                self.interp["usf"] = int(usf)
            if kind != "nearest" and self.interp["usf"] != 1:
                raise BeamformerException(f"upsampling is only defined for kind 'nearest'")


@dataclass(kw_only=True)
class Tabbed:
    """Precomputed delay tabs: ``tautx``/``apodtx`` (Ntx, Np) and
    ``taurx``/``apodrx`` (Nrx, Np) (pycbf/__bf_base_classes__.py:97-139)."""
    tautx: np.ndarray = field(init=True)
    taurx: np.ndarray = field(init=True)
    apodtx: np.ndarray = field(init=True)
    apodrx: np.ndarray = field(init=True)
    ntx: int = field(init=False)
    nrx: int = field(init=False)
    nop: int = field(init=False)

    def __post_init__(self):
        all2d = all(_ndim(a) == 2 for a in (self.tautx, self.taurx, self.apodtx, self.apodrx))
        if not all2d:
            raise BeamformerException("all delay tabs and apodizations must be 2D matrices")
        ntx = _shape(self.tautx)[0]
        nrx = _shape(self.taurx)[0]
        nop = _shape(self.tautx)[1]
        if _shape(self.apodtx)[0] != ntx:
            raise BeamformerException(
                f"tautx and apodtx must have the same shape in dim 0 but were "
                f"{_shape(self.tautx)[0]} and {_shape(self.apodtx)[0]}")
        if _shape(self.apodrx)[0] != nrx:
            raise BeamformerException(
                f"taurx and apodrx must have the same shape in dim 0 but were "
                f"{_shape(self.taurx)[0]} and {_shape(self.apodrx)[0]}")
        if not all(_shape(a)[1] == nop for a in (self.taurx, self.apodtx, self.apodrx)):
            raise BeamformerException(
                "tautx, taurx, apodtx, and apodrx must have the same shape in dim 1 but were "
                f"{_shape(self.tautx)[1]}, {_shape(self.taurx)[1]}, {_shape(self.apodtx)[1]}, "
                f"and {_shape(self.apodrx)[1]}")
        self.ntx, self.nrx, self.nop = ntx, nrx, nop


@dataclass(kw_only=True)
class Synthetic:
    """Synthetic-point geometry (pycbf/__bf_base_classes__.py:141-219):
    ``ovectx``/``nvectx`` (Ntx, ndim), ``doftx``/``alatx``/``t0tx`` (Ntx),
    ``ovecrx``/``nvecrx`` (Nrx, ndim), ``alarx`` (Nrx), ``c0`` and ``pnts``
    (Np, ndim) with ndim 2 or 3."""
    ovectx: np.ndarray = field(init=True)
    nvectx: np.ndarray = field(init=True)
    doftx: np.ndarray = field(init=True)
    alatx: np.ndarray = field(init=True)
    t0tx: np.ndarray = field(init=True)
    ovecrx: np.ndarray = field(init=True)
    nvecrx: np.ndarray = field(init=True)
    alarx: np.ndarray = field(init=True)
    c0: float = field(init=True)
    pnts: np.ndarray = field(init=True)
    ntx: int = field(init=False)
    nrx: int = field(init=False)
    nop: int = field(init=False)
    ndimp: int = field(init=False)

    def __post_init__(self):
        all1d = all(_ndim(a) == 1 for a in (self.doftx, self.alatx, self.t0tx, self.alarx))
        all2d = all(_ndim(a) == 2 for a in (self.ovectx, self.nvectx, self.ovecrx, self.nvecrx))
        if not all1d:
            raise BeamformerException(
                "all acceptance angle, depth of field, and t0 vectors must be 1D")
        if not all2d:
            raise BeamformerException("all origin and normal vector matrices must be 2D")
        if _ndim(self.pnts) != 2:
            raise BeamformerException("pnts must be an Np by ndim matrix")
        ntx = _shape(self.ovectx)[0]
        nrx = _shape(self.ovecrx)[0]
        nop, ndimp = _shape(self.pnts)
        if not (_shape(self.nvectx)[0] == ntx and _shape(self.doftx)[0] == ntx
                and _shape(self.alatx)[0] == ntx and _shape(self.t0tx)[0] == ntx):
            raise BeamformerException(
                "All tx variables must have the same shape in the first dimension")
        if not (_shape(self.nvecrx)[0] == nrx and _shape(self.alarx)[0] == nrx):
            raise BeamformerException("All rx variables must have the same shame in the first variable")
        if not all(_shape(a)[1] == ndimp for a in (self.ovectx, self.nvectx, self.ovecrx,
                                                     self.nvecrx)):
            raise BeamformerException("All matrix variables must have the same number of dimensions")
        if ndimp not in (2, 3):
            raise BeamformerException(f"points must be 2D or 3D, got ndim={ndimp}")
        self.ntx, self.nrx, self.nop, self.ndimp = ntx, nrx, nop, ndimp


@dataclass(kw_only=True)
class DMASBeamformer:
    """Delay-multiply-and-sum fields (pycbf/__advanced_bf_classes__.py:7-25).

    ``lags`` are the element offsets correlated along ``lagaxis`` ('rx'
    correlates receive channels after summing transmits, 'tx' the reverse);
    ``dmastype`` is 'ssr' (signed square root of each product), 'power' (plain
    products) or 'pssr' (signed square root of the summed power, the notebooks'
    P-DMAS-SSR); ``sumlags`` sums over lags."""
    lags: np.ndarray = field(init=True)
    lagaxis: Literal["tx", "rx"] = "rx"
    nlags: int = field(init=False)
    sumlags: bool = True
    dmastype: Literal["ssr", "power", "pssr"] = "ssr"

    def __post_init__(self):
        if self.lagaxis == "rx":
            self.sumtype = "tx_only"
        elif self.lagaxis == "tx":
            self.sumtype = "rx_only"
        else:
            raise BeamformerException("lagaxis keyword can only be 'tx' or 'rx'")
        if self.dmastype not in ("ssr", "power", "pssr"):
            raise BeamformerException(
                "dmastype keyword must be 'ssr' (signed-square-root), 'power' or 'pssr'")
        lags = np.asarray(self.lags).ravel()
        if lags.size == 0:
            raise BeamformerException("lags must not be empty")
        if not np.all(np.equal(np.mod(lags, 1), 0)):
            raise BeamformerException("lags must be integers")
        self.lags = lags.astype(np.int32)
        self.nlags = int(lags.size)


@dataclass(kw_only=True)
class SLSCBeamformer:
    """Short-lag spatial coherence fields (no class in the reference; the
    algorithm is examples/advanced_beamforming.ipynb:65-72,205-223).

    The base DAS sums transmits only (per-receive-channel images); ``gridshape``
    = (nx, nz) gives the x-major layout of the Np pixels so the axial window
    can be taken along z.  Output pixel (ixo, izo) is pixel (ixo*dsfx,
    izo*dsfz); its window is z + kofs .. z + kofs + kwin - 1 clipped to the
    grid (the notebook's z - nkern .. z + nkern is kwin = 2*nkern+1,
    kofs = -nkern, the default when kofs is None).  ``precision`` 'fp32'
    computes the lag sums on the vector units, 'fp16' / 'bf16' on MFMA from
    normalised channels in that format (FP32 accumulation)."""
    lags: np.ndarray = field(init=True)
    gridshape: tuple = field(init=True)
    kwin: int = 8
    kofs: int | None = None
    dsfx: int = 1
    dsfz: int = 1
    precision: Literal["fp32", "fp16", "bf16"] = "fp16"
    nlags: int = field(init=False)

    def __post_init__(self):
        self.sumtype = "tx_only"
        lags = np.asarray(self.lags).ravel()
        if lags.size == 0 or lags.size > 64:
            raise BeamformerException("lags must hold 1 to 64 values")
        if not np.all(np.equal(np.mod(lags, 1), 0)) or lags.min() < 1 or lags.max() > 63:
            raise BeamformerException("lags must be integers in 1..63")
        self.lags = lags.astype(np.int32)
        self.nlags = int(lags.size)
        if len(tuple(self.gridshape)) != 2:
            raise BeamformerException("gridshape must be (nx, nz)")
        self.gridshape = tuple(int(g) for g in self.gridshape)
        self.kwin = int(self.kwin)
        if not 1 <= self.kwin <= 32:
            raise BeamformerException("kwin must be in 1..32")
        self.kofs = -(self.kwin // 2) if self.kofs is None else int(self.kofs)
        if not -(self.kwin - 1) <= self.kofs <= 0:
            raise BeamformerException("kofs must be in -(kwin-1)..0 so the window holds its pixel")
        if int(self.dsfx) < 1 or int(self.dsfz) < 1:
            raise BeamformerException("dsfx and dsfz must be >= 1")
        self.dsfx, self.dsfz = int(self.dsfx), int(self.dsfz)
        if self.precision not in ("fp32", "fp16", "bf16"):
            raise BeamformerException("precision must be 'fp32', 'fp16' or 'bf16'")

    def _slsc_outshape(self):
        nx, nz = self.gridshape
        return ((nx + self.dsfx - 1) // self.dsfx, (nz + self.dsfz - 1) // self.dsfz)
