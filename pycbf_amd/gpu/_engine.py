"""ctypes binding of libpycbf_hip.so -- the replacement of pycbf's CuPy engine
module (pycbf/gpu/__engines__.py:1-26).

The reference JIT-compiles CUDA source with CuPy at import and hands out
RawKernel handles plus numpy struct dtypes ``xInfo`` / ``RFInfo``.  Here the
kernels are prebuilt for gfx950 in ``pycbf_amd/libpycbf_hip.so`` and reached
through the C-ABI declared in ``include/pycbf_hip.h``; ``RFInfo`` is the same
32-byte struct as a ``ctypes.Structure``.  There is no fallback: if the
library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as ct
import os

from ..base import BeamformerException

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBPATH = os.path.join(_PKG, "libpycbf_hip.so")

if not os.path.exists(LIBPATH):
    raise BeamformerException(
        f"{LIBPATH} is not built; run `python -m pycbf_amd.build` (hipcc, gfx950)")

# torch-ROCm ships its own libamdhip64 (same soname).  Loading torch first makes
# the engine bind to the HIP runtime torch already uses, so device memory and
# streams handed over from torch are valid in the library; loading the system
# runtime first would put two HIP runtimes into the process.
import torch  # noqa: E402,F401

lib = ct.CDLL(LIBPATH)

INTERP = {"nearest": 0, "linear": 1, "akima": 2, "makima": 3, "korder_cubic": 4, "cubic": 5}
SUMTYPE = {"none": 0, "tx_only": 1, "rx_only": 2, "tx_and_rx": 3}
DMASTYPE = {"ssr": 0, "power": 1, "pssr": 2}
TABS = {"plane": 0, "divergent": 1, "focused": 2}
SLSC_PRECISION = {"fp32": 0, "fp16": 1, "bf16": 2}
# thresh reproducing pycbf's GPU rule |w| > 0.01 (a float against the double
# 0.01, pycbf/gpu/__engines__.cu:8,564) under |w| >= thresh: the smallest
# float above 0.01 (PYCBF_GPU_REFERENCE_THRESH)
GPU_REFERENCE_THRESH = 0.010000000707805157
TXPACK_FLOATS = 16
RXPACK_FLOATS = 12
ABI_VERSION = 3


class xInfo(ct.Structure):
    """struct xInfo, pycbf/gpu/__engines__.cu:360-364."""
    _fields_ = [("x0", ct.c_float), ("dx", ct.c_float), ("nx", ct.c_int32)]


class RFInfo(ct.Structure):
    """struct RFInfo, pycbf/gpu/__engines__.cu:369-374 (32 bytes)."""
    _fields_ = [("ntx", ct.c_int64), ("nrx", ct.c_int64), ("ndim", ct.c_int32),
                ("tInfo", xInfo)]


assert ct.sizeof(RFInfo) == 32


class BatchReq(ct.Structure):
    """struct pycbf_batch_req (include/pycbf_hip.h)."""
    _fields_ = [("kind", ct.c_int32), ("ntx", ct.c_int32), ("nrx", ct.c_int32),
                ("nx", ct.c_int32), ("nz", ct.c_int32), ("nt", ct.c_int32),
                ("t0", ct.c_float), ("dt", ct.c_float), ("c0", ct.c_float),
                ("thresh", ct.c_float), ("txpack", ct.c_void_p), ("rxpack", ct.c_void_p),
                ("pvec", ct.c_void_p), ("out", ct.c_void_p), ("imrx", ct.c_void_p),
                ("sel", ct.c_void_p), ("kwin", ct.c_int32), ("kofs", ct.c_int32),
                ("dsfx", ct.c_int32), ("dsfz", ct.c_int32), ("nlag", ct.c_int32),
                ("precision", ct.c_int32), ("lags", ct.c_int32 * 64)]

_vp = ct.c_void_p
_i32 = ct.c_int32
_i64 = ct.c_int64
_f32 = ct.c_float
_sz = ct.c_size_t
_prf = ct.POINTER(RFInfo)

_SIGS = {
    "pycbf_abi_version": ([], ct.c_int),
    "pycbf_last_error": ([], ct.c_char_p),
    "pycbf_das_workspace_bytes": ([_prf, _i64, _i64, ct.c_int, ct.c_int, ct.c_int], _sz),
    "pycbf_das_tabbed": ([_prf, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _i64, ct.c_int, ct.c_int,
                          _vp, _f32, ct.c_int, _vp, _vp, _sz, _vp], ct.c_int),
    "pycbf_das_tabbed_residual": ([_prf, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i64,
                                   ct.c_int, ct.c_int, _vp, _f32, ct.c_int, _vp, _vp, _sz, _vp],
                                  ct.c_int),
    "pycbf_synthetic_pack": ([ct.c_int, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                              _vp, _vp], ct.c_int),
    "pycbf_das_synthetic": ([_prf, _vp, _i64, _vp, _vp, _f32, _vp, _vp, _i64, ct.c_int,
                             ct.c_int, _vp, _f32, ct.c_int, _vp, _vp, _sz, _vp], ct.c_int),
    "pycbf_das_synthetic_noalloc": ([_prf, _vp, _i64, _vp, _vp, _f32, _vp, _vp, _i64,
                                     ct.c_int, ct.c_int, _vp, _f32, ct.c_int, _vp, _vp],
                                    ct.c_int),
    "pycbf_synthetic_tables": ([_prf, _vp, _vp, _f32, _vp, _i64, _vp, _vp, _vp, _vp, _vp],
                               ct.c_int),
    "pycbf_tile_order": ([_vp, _i64, ct.c_int, _vp], ct.c_int),
    "pycbf_make_tabs_2d": ([ct.c_int, _i64, _vp, _i64, _vp, _vp, _i64, _vp, _i64, ct.c_double,
                            ct.c_double, _vp, _vp, _vp, _vp, _vp], ct.c_int),
    "pycbf_demod_apply": ([_vp, _i64, _i64, _vp, _i64, _vp, _vp, _sz, _vp], ct.c_int),
    "pycbf_demod_workspace_bytes": ([_i64, _i64, _i64], _sz),
    "pycbf_demod_fft": ([_vp, _i64, _i64, _i64, _i64, _i64, _i64, _vp, _vp, _vp], ct.c_int),
    "pycbf_spline_coefs": ([_vp, _i64, _i32, _f32, _vp, _vp, _sz, _vp], ct.c_int),
    "pycbf_spline_y2": ([_vp, _i64, _i32, _f32, _vp, _vp, _sz, _vp], ct.c_int),
    "pycbf_spline_workspace_bytes": ([_i64, _i32], _sz),
    "pycbf_debug_indices": ([_prf, _vp, _vp, _i64, ct.c_int, ct.c_int, _vp, _vp], ct.c_int),
    "pycbf_dmas": ([_vp, _i64, _i64, _i64, _vp, ct.c_int, ct.c_int, ct.c_int, _vp, _vp],
                   ct.c_int),
    "pycbf_slsc": ([_vp, _i64, _i64, _i64, _i64, _vp, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                    _vp, ct.c_int, ct.c_int, _vp, _vp], ct.c_int),
    "pycbf_slsc_selection": ([_vp, _i64, _i64, _i64, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                              _vp, _vp], ct.c_int),
    "pycbf_das_bwd_workspace_bytes": ([_prf, _i64, _i64, ct.c_int, ct.c_int], _sz),
    "pycbf_das_tabbed_bwd": ([_prf, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i64, ct.c_int,
                              ct.c_int, _vp, _f32, ct.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                              _sz, _vp], ct.c_int),
    "pycbf_das_synthetic_bwd": ([_prf, _vp, _i64, _vp, _vp, _f32, _vp, _i64, ct.c_int, ct.c_int,
                                 _vp, _f32, ct.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                 _sz, _vp], ct.c_int),
    "pycbf_ray_delays": ([_vp, ct.c_int, ct.c_int, _f32, _f32, _f32, _f32, _vp, _vp, ct.c_int,
                          _f32, _vp, ct.c_int, _vp, _i64, ct.c_int, _vp, _vp, _vp, _vp, _vp],
                         ct.c_int),
    "pycbf_ray_delays_bwd_workspace_bytes": ([ct.c_int, ct.c_int], _sz),
    "pycbf_ray_delays_bwd": ([_vp, ct.c_int, ct.c_int, _f32, _f32, _f32, _f32, _vp, _vp,
                              ct.c_int, _f32, _vp, ct.c_int, _vp, _i64, ct.c_int, _vp, _vp, _vp,
                              _vp, _sz, _vp], ct.c_int),
    "pycbf_batch_create": ([_vp, ct.c_int, _vp], ct.c_int),
    "pycbf_batch_destroy": ([_vp], None),
    "pycbf_batch_workspace_bytes": ([_vp], _sz),
    "pycbf_batch_configure": ([_vp, ct.c_int, ct.c_int, _vp, _sz, _vp], ct.c_int),
    "pycbf_batch_run": ([_vp, _vp, _vp, _sz, _vp], ct.c_int),
    "pycbf_batch_autotune": ([_vp, _vp, ct.c_int, _vp, _sz, _vp, _vp, _vp, _vp], ct.c_int),
    "pycbf_batch_latencies": ([_vp, _vp, _sz, _vp], ct.c_int),
    "pycbf_measure_bandwidth": ([ct.c_int, _vp, _sz, ct.c_int, _vp, _sz,
                                 ct.POINTER(ct.c_double), _vp], ct.c_int),
    "pycbf_profile_enable": ([ct.c_int], None),
    "pycbf_profile_reset": ([], None),
    "pycbf_profile_read": ([ct.c_char_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_int64)],
                           ct.c_int),
}

for _name, (_args, _res) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.argtypes = _args
    _fn.restype = _res

if lib.pycbf_abi_version() != ABI_VERSION:
    raise BeamformerException(
        f"libpycbf_hip.so ABI {lib.pycbf_abi_version()} != expected {ABI_VERSION}; rebuild it")


def last_error() -> str:
    msg = lib.pycbf_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> None:
    """Raise BeamformerException for a nonzero status of a C-ABI call."""
    if rc != 0:
        raise BeamformerException(f"{what} failed ({rc}): {last_error()}")


def rfinfo(ntx: int, nrx: int, ndim: int, t0: float, dt: float, nt: int) -> RFInfo:
    info = RFInfo()
    info.ntx = int(ntx)
    info.nrx = int(nrx)
    info.ndim = int(ndim)
    info.tInfo.x0 = float(t0)
    info.tInfo.dx = float(dt)
    info.tInfo.nx = int(nt)
    return info


def ptr(t) -> int:
    """Device (or host numpy) address of a contiguous array, 0 for None."""
    if t is None:
        return 0
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def stream_of(device) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream


def profile_read(kernel: str):
    """(total_ms, launches) recorded for `kernel` since the last reset."""
    ms = ct.c_double(0.0)
    n = ct.c_int64(0)
    lib.pycbf_profile_read(kernel.encode(), ct.byref(ms), ct.byref(n))
    return ms.value, n.value


def measure_bandwidth(kind: str, device, nbytes: int = 0, reps: int = 5) -> float:
    """Measured read bandwidth (GB/s) of 'hbm' (a streaming read of nbytes,
    default 4 GiB), 'l2' (workgroups re-reading a 2 MiB region) or 'lds'."""
    import torch
    code = {"hbm": 0, "l2": 1, "lds": 2}[kind]
    if not nbytes:
        nbytes = {"hbm": 4 << 30, "l2": 2 << 20, "lds": 16}[kind]
    buf = torch.ones(nbytes // 4, dtype=torch.float32, device=device)
    sink = torch.empty(8192 * 256, dtype=torch.int32, device=device)
    out = ct.c_double(0.0)
    check(lib.pycbf_measure_bandwidth(code, ptr(buf), buf.numel() * 4, reps, ptr(sink),
                                      sink.numel() * 4, ct.byref(out), stream_of(device)),
          "pycbf_measure_bandwidth")
    return out.value

