"""FP64 NumPy restatement of the notebooks' FFT demodulation (test
infrastructure only: the checker of the GPU demodulation).

``fft_tukey_demod`` follows examples/SyntheticVsTabbed.ipynb, cell 1
(``fft_tukey_demod``): rfft along the axis, keep the bins from the one
nearest kmin to the one nearest kmax, weight them with a Tukey window
(scipy.signal.windows.tukey), optionally ifftshift them to base band, and
ifft.  ``zout_fft_tukey_demod`` (same cell) gives the axial coordinates of
the demodulated samples.  NumPy only (the notebook also accepts CuPy arrays).
"""
from __future__ import annotations

import numpy as np


def band(n: int, kmin: float, kmax: float, ks: float):
    """(ikmin, ikmax) of the kept rfft bins of an n-sample axis."""
    freqs = np.fft.rfftfreq(n, 1 / ks)
    ikmin = int(np.argmin(np.abs(freqs - kmin)))
    ikmax = int(np.argmin(np.abs(freqs - kmax))) + 1
    return ikmin, ikmax


def fft_tukey_demod(rfdata, kmin, kmax, ks, alpha=0.25, axis=-1, shiftit=True):
    from scipy.signal.windows import tukey
    rfdata = np.asarray(rfdata)
    ikmin, ikmax = band(rfdata.shape[axis], kmin, kmax, ks)
    winshape = [1] * rfdata.ndim
    winshape[axis] = ikmax - ikmin
    win = tukey(ikmax - ikmin, alpha=alpha).reshape(winshape)
    sl = [slice(None)] * rfdata.ndim
    sl[axis] = slice(ikmin, ikmax)
    spec = win * np.fft.rfft(rfdata, axis=axis)[tuple(sl)]
    if shiftit:
        spec = np.fft.ifftshift(spec, axes=axis)
    return np.fft.ifft(spec, axis=axis)


def zout_fft_tukey_demod(zout, kmin, kmax):
    zout = np.asarray(zout)
    dz = zout[1] - zout[0]
    ks = np.fft.rfftfreq(len(zout), dz)
    dk = ks[1] - ks[0]
    ikmin = np.argmin(np.abs(ks - kmin))
    ikmax = np.argmin(np.abs(ks - kmax))
    ks_eff = dk * (ikmax - ikmin)
    return zout[0] + np.arange(ikmax - ikmin + 1) / ks_eff
