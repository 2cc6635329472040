"""Host tables of the FFT demodulation (pycbf_amd.gpu.demod.fft_tables): the
chain the kernel runs on them, done here with NumPy FFTs in float64 (real
FFT by an n/2-point complex FFT and the split, kept bins times pre, circular
convolution with the chirp transform, inverse by conj(FFT(conj(.))), times
post), reproduces the notebook's chain (oracle.demod) to rounding; and the
band edges that select the fallback."""
import numpy as np
import pytest

from oracle import demod as od
from pycbf_amd.gpu.demod import fft_tables


def _chain(x, tabs):
    ikmin, m, shift, L, tab = tabs
    pre, post = tab[:m], tab[m:2 * m]
    bhat, split = tab[2 * m:2 * m + L], tab[2 * m + L:3 * m + L]
    n = x.shape[-1]
    half = n // 2
    Z = np.fft.fft(x[..., 0::2] + 1j * x[..., 1::2])
    k = np.arange(m)
    kb = ikmin + (k + shift) % m
    zk = Z[..., np.where(kb == half, 0, kb)]
    zc = np.conj(Z[..., np.where(kb == 0, 0, half - kb)])
    X = 0.5 * (zk + zc) + split * (-0.5j * (zk - zc))
    C = np.zeros(x.shape[:-1] + (L,), complex)
    C[..., :m] = X * pre
    E = np.fft.fft(np.conj(np.fft.fft(C) * bhat))
    return np.conj(E[..., :m]) * post


@pytest.mark.parametrize("n,dz,shiftit", [(2048, 19.55e-6, True), (1024, 20e-6, False),
                                          (4096, 5e-6, True), (64, 20e-6, True)])
def test_tables_reproduce_the_notebook_chain(n, dz, shiftit):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((4, n))
    kmin, kmax, ks = 2 * 5e6 / 1540, 2 * 12e6 / 1540, 1 / dz
    tabs = fft_tables(n, kmin, kmax, ks, 0.25, shiftit)
    assert tabs is not None
    ikmin, m, shift, L, _ = tabs
    assert L >= 2 * m - 1 and L & (L - 1) == 0
    ref = od.fft_tukey_demod(x, kmin, kmax, ks, shiftit=shiftit)
    got = _chain(x, tabs)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_lengths_outside_the_fft_path():
    kmin, kmax, ks = 2 * 5e6 / 1540, 2 * 12e6 / 1540, 1 / 20e-6
    assert fft_tables(777, kmin, kmax, ks) is None      # not a power of two
    assert fft_tables(8192, kmin, kmax, ks) is None     # beyond 4096 samples
    assert fft_tables(4, kmin, kmax, ks) is None        # below 8 samples
