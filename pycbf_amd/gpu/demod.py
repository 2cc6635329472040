"""FFT-window demodulation of beamformed RF to IQ on the GPU.

``fft_tukey_demod`` takes the arguments of the notebooks' function of the
same name (examples/SyntheticVsTabbed.ipynb, cell 1): rfft along ``axis``,
keep the bins from the one nearest ``kmin`` to the one nearest ``kmax``,
weight them with a Tukey window, ifftshift them to base band (``shiftit``)
and ifft.

For n a power of two (8..4096) it runs that chain itself, one workgroup per
line in LDS (``pycbf_demod_fft``: real FFT, band, window, shift and a
Bluestein inverse DFT of the L kept bins), reading each line once and writing
its IQ once; the host builds the window / chirp / twiddle tables once in
float64 (cached per device).  For other lengths the chain, linear in the
samples of each line, is one fixed operator M (n x L complex) per (n, band,
window):

    M[k, m] = 1/L * sum_b w_b exp(-2 pi i k (ikmin + b) / n) exp(2 pi i s(b) m / L)

with s(b) the position of bin b after ifftshift ((b - L//2) mod L, or b
without the shift).  The host builds M once in float64 from that closed form
(cached per device), and ``pycbf_demod_apply`` computes iq = rf @ M for all
lines at once, float32-accurate on the BF16 matrix units (a three-way BF16
split of both operands).  The notebook returns complex128 from
NumPy or CuPy input; here the result is complex64 (a CUDA tensor for tensor
input, a NumPy array for NumPy input).
"""
from __future__ import annotations

import functools

import numpy as np

from . import _engine as E


def _torch():
    import torch
    return torch


def band(n: int, kmin: float, kmax: float, ks: float):
    """(ikmin, ikmax): the kept rfft bins [ikmin, ikmax) of an n-sample axis
    sampled at ks (the nearest bins to kmin and kmax, both kept)."""
    freqs = np.fft.rfftfreq(int(n), 1 / ks)
    ikmin = int(np.argmin(np.abs(freqs - kmin)))
    ikmax = int(np.argmin(np.abs(freqs - kmax))) + 1
    return ikmin, ikmax


def demod_operator(n: int, kmin: float, kmax: float, ks: float, alpha: float = 0.25,
                   shiftit: bool = True) -> np.ndarray:
    """The demodulation as a complex128 (n, L) matrix: iq = rf @ M."""
    from scipy.signal.windows import tukey
    ikmin, ikmax = band(n, kmin, kmax, ks)
    L = ikmax - ikmin
    if L < 1:
        raise ValueError(f"empty band: bins [{ikmin}, {ikmax}) for kmin={kmin}, kmax={kmax}")
    w = tukey(L, alpha=alpha)
    b = np.arange(L)
    pos = (b - L // 2) % L if shiftit else b
    k = np.arange(n)
    # forward DFT of each sample onto the kept bins (exponent reduced mod n exactly)
    fwd = np.exp(-2j * np.pi * ((k[:, None] * (ikmin + b)[None, :]) % n) / n) * w[None, :]
    m = np.arange(L)
    inv = np.exp(2j * np.pi * ((pos[:, None] * m[None, :]) % L) / L) / L
    return fwd @ inv


@functools.lru_cache(maxsize=16)
def _device_operator(n, kmin, kmax, ks, alpha, shiftit, dev_index):
    """The operator as the device takes it: transposed, (2L, n) float32 with
    row 2m = Re and row 2m+1 = Im of output sample m."""
    torch = _torch()
    M = demod_operator(n, kmin, kmax, ks, alpha, shiftit)
    opT = np.empty((2 * M.shape[1], n), np.float32)
    opT[0::2] = M.real.T
    opT[1::2] = M.imag.T
    return torch.as_tensor(opT, device=torch.device("cuda", dev_index))


FFT_MAX_N = 4096
FFT_MAX_CONV = 2048


def fft_tables(n: int, kmin: float, kmax: float, ks: float, alpha: float = 0.25,
               shiftit: bool = True):
    """(ikmin, nbin, shift, nconv, tables) of pycbf_demod_fft for an n-sample
    axis (complex128 tables: pre, post, chirp transform, split twiddles,
    FFT twiddles), or None when n or the band is outside its range."""
    from scipy.signal.windows import tukey
    n = int(n)
    if n < 8 or n > FFT_MAX_N or n & (n - 1):
        return None
    ikmin, ikmax = band(n, kmin, kmax, ks)
    m = ikmax - ikmin
    if m < 1:
        raise ValueError(f"empty band: bins [{ikmin}, {ikmax}) for kmin={kmin}, kmax={kmax}")
    L = 1
    while L < 2 * m - 1:
        L *= 2
    if L > FFT_MAX_CONV:
        return None
    w = tukey(m, alpha=alpha)
    shift = m // 2 if shiftit else 0
    k = np.arange(m)
    chirp = np.exp(1j * np.pi * ((k * k) % (2 * m)) / m)  # exponent reduced mod 2m exactly
    pre = w[(k + shift) % m] * chirp
    post = chirp / (m * L)
    b = np.zeros(L, np.complex128)
    b[:m] = np.conj(chirp)
    if m > 1:
        b[L - m + 1:] = np.conj(chirp[1:])[::-1]
    bhat = np.fft.fft(b)
    split = np.exp(-2j * np.pi * (ikmin + (k + shift) % m) / n)
    T = max(n // 2, L)
    tw = np.exp(-2j * np.pi * np.arange(T) / T)
    return ikmin, m, shift, L, np.concatenate([pre, post, bhat, split, tw])


@functools.lru_cache(maxsize=16)
def _device_fft_tables(n, kmin, kmax, ks, alpha, shiftit, dev_index):
    torch = _torch()
    t = fft_tables(n, kmin, kmax, ks, alpha, shiftit)
    if t is None:
        return None
    ikmin, m, shift, L, tab = t
    f = np.empty((tab.size, 2), np.float32)
    f[:, 0] = tab.real
    f[:, 1] = tab.imag
    return ikmin, m, shift, L, torch.as_tensor(f, device=torch.device("cuda", dev_index))


def fft_tukey_demod(rfdata, kmin, kmax, ks, alpha=0.25, axis=-1, shiftit=True, method=None):
    """IQ of real ``rfdata`` along ``axis``: complex64, the axis shortened to
    the L kept bins.  ``method``: None (the FFT kernel where n allows, else
    the operator product), "fft" or "operator"."""
    torch = _torch()
    as_numpy = not isinstance(rfdata, torch.Tensor)
    if as_numpy:
        arr = np.asarray(rfdata)
        if np.iscomplexobj(arr):
            raise TypeError("fft_tukey_demod needs real RF data")
        x = torch.as_tensor(np.ascontiguousarray(arr, np.float32), device="cuda")
    else:
        if rfdata.is_complex():
            raise TypeError("fft_tukey_demod needs real RF data")
        if rfdata.device.type != "cuda":
            raise ValueError("fft_tukey_demod needs a CUDA tensor or a NumPy array")
        x = rfdata.to(torch.float32)
    if x.dim() < 1:
        raise ValueError("fft_tukey_demod needs at least one axis")
    dev = x.device
    with torch.cuda.device(dev):
        x = torch.movedim(x, axis, -1).contiguous()
        n = x.shape[-1]
        lead = tuple(x.shape[:-1])
        rows = int(np.prod(lead)) if lead else 1
        dix = dev.index if dev.index is not None else 0
        ft = None
        if method in (None, "fft"):
            ft = _device_fft_tables(int(n), float(kmin), float(kmax), float(ks), float(alpha),
                                    bool(shiftit), dix)
            if ft is None and method == "fft":
                raise ValueError(f"the FFT demodulation needs n a power of two <= {FFT_MAX_N} "
                                 f"(got {n})")
        elif method != "operator":
            raise ValueError("method must be None, 'fft' or 'operator'")
        if ft is not None:
            ikmin, m, shift, L, tab = ft
            out = torch.empty(lead + (2 * m,), dtype=torch.float32, device=dev)
            rc = E.lib.pycbf_demod_fft(E.ptr(x), rows, n, ikmin, m, shift, L, E.ptr(tab),
                                       E.ptr(out), E.stream_of(dev))
            E.check(rc, "pycbf_demod_fft")
            iq = torch.movedim(torch.view_as_complex(out.view(lead + (m, 2))), -1, axis)
            return iq.cpu().numpy() if as_numpy else iq
        op = _device_operator(int(n), float(kmin), float(kmax), float(ks), float(alpha),
                              bool(shiftit), dev.index if dev.index is not None else 0)
        ncol = op.shape[0]
        out = torch.empty(lead + (ncol,), dtype=torch.float32, device=dev)
        need = int(E.lib.pycbf_demod_workspace_bytes(rows, n, ncol))
        ws = torch.empty(max(need // 4, 1), dtype=torch.float32, device=dev) if need else None
        rc = E.lib.pycbf_demod_apply(E.ptr(x), rows, n, E.ptr(op), ncol, E.ptr(out), E.ptr(ws),
                                     need, E.stream_of(dev))
        E.check(rc, "pycbf_demod_apply")
        iq = torch.view_as_complex(out.view(lead + (ncol // 2, 2)))
        iq = torch.movedim(iq, -1, axis)
    if as_numpy:
        return iq.cpu().numpy()
    return iq


def zout_fft_tukey_demod(zout, kmin, kmax):
    """Axial coordinates of the samples ``fft_tukey_demod`` returns along a
    z axis sampled at ``zout`` (same cell of the notebook)."""
    zout = np.asarray(zout, np.float64)
    dz = zout[1] - zout[0]
    ks = np.fft.rfftfreq(len(zout), dz)
    dk = ks[1] - ks[0]
    ikmin = np.argmin(np.abs(ks - kmin))
    ikmax = np.argmin(np.abs(ks - kmax))
    return zout[0] + np.arange(ikmax - ikmin + 1) / (dk * (ikmax - ikmin))
