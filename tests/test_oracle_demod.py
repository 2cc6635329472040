"""The oracle's FFT demodulation (fft_tukey_demod of
examples/SyntheticVsTabbed.ipynb, cell 1) pinned by closed forms, and the
host-side operator of pycbf_amd.gpu.demod checked against it."""
import numpy as np
import pytest

from oracle import demod as od


def test_band_bins_by_hand():
    # n = 64 at ks = 64: bin f sits at f cycles per unit; 10.4 -> 10, 20.6 -> 21
    assert od.band(64, 10.4, 20.6, 64.0) == (10, 22)
    assert od.band(64, 0.0, 32.0, 64.0) == (0, 33)


@pytest.mark.parametrize("shiftit", [True, False])
def test_tone_in_flat_band_is_a_pure_exponential(shiftit):
    """cos(2 pi f0 k / n) has rfft n/2 at bin f0; inside the flat part of the
    window the output is n/(2L) exp(2 pi i j m / L), j = f0's position after
    the optional ifftshift."""
    n, ks = 256, 256.0
    kmin, kmax = 40.0, 100.0
    ikmin, ikmax = od.band(n, kmin, kmax, ks)
    L = ikmax - ikmin
    f0 = 70
    x = np.cos(2 * np.pi * f0 * np.arange(n) / n)
    iq = od.fft_tukey_demod(x, kmin, kmax, ks, shiftit=shiftit)
    b = f0 - ikmin
    j = (b - L // 2) % L if shiftit else b
    want = n / (2 * L) * np.exp(2j * np.pi * j * np.arange(L) / L)
    assert iq.shape == (L,)
    np.testing.assert_allclose(iq, want, atol=1e-12)


def test_tone_outside_band_vanishes_and_edges_are_windowed():
    n, ks = 256, 256.0
    iq = od.fft_tukey_demod(np.cos(2 * np.pi * 20 * np.arange(n) / n), 40.0, 100.0, ks)
    assert np.abs(iq).max() < 1e-12
    # the window's first sample is zero: a tone on the first kept bin vanishes
    iq = od.fft_tukey_demod(np.cos(2 * np.pi * 40 * np.arange(n) / n), 40.0, 100.0, ks)
    assert np.abs(iq).max() < 1e-12


def test_axis_argument():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((37, 5))
    a = od.fft_tukey_demod(x, 0.1, 0.3, 1.0, axis=0)
    b = od.fft_tukey_demod(x.T, 0.1, 0.3, 1.0, axis=-1).T
    np.testing.assert_allclose(a, b, atol=1e-13)


def test_zout_matches_band_length():
    zout = np.linspace(1e-3, 40e-3, 777)
    dz = zout[1] - zout[0]
    kmin, kmax = 2 * 5e6 / 1540, 2 * 12e6 / 1540
    ikmin, ikmax = od.band(len(zout), kmin, kmax, 1 / dz)
    z = od.zout_fft_tukey_demod(zout, kmin, kmax)
    assert len(z) == ikmax - ikmin
    assert z[0] == zout[0]
    # spacing 1 / (dk (L - 1)) with dk = 1 / (n dz): L - 1 steps span n dz
    np.testing.assert_allclose((z[1] - z[0]) * (len(z) - 1), dz * len(zout), rtol=1e-12)


@pytest.mark.parametrize("n,shiftit", [(64, True), (255, True), (512, False), (1000, True)])
def test_host_operator_equals_fft_pipeline(n, shiftit):
    """pycbf_amd.gpu.demod.demod_operator (closed-form DFT matrices) applied to
    random lines equals the oracle's FFT chain."""
    from pycbf_amd.gpu.demod import demod_operator
    rng = np.random.default_rng(n)
    x = rng.standard_normal((6, n))
    kmin, kmax, ks = 0.12, 0.31, 1.0
    M = demod_operator(n, kmin, kmax, ks, 0.25, shiftit)
    want = od.fft_tukey_demod(x, kmin, kmax, ks, shiftit=shiftit)
    assert M.shape == (n, want.shape[-1])
    np.testing.assert_allclose(x @ M, want, atol=1e-11 * np.abs(want).max())


def test_host_operator_rejects_empty_band():
    from pycbf_amd.gpu.demod import demod_operator
    with pytest.raises(ValueError):
        demod_operator(64, 0.4, 0.1, 1.0)


def test_host_zout_equals_oracle():
    from pycbf_amd.gpu.demod import zout_fft_tukey_demod
    zout = np.linspace(2e-3, 30e-3, 300)
    np.testing.assert_array_equal(zout_fft_tukey_demod(zout, 6e3, 15e3),
                                  od.zout_fft_tukey_demod(zout, 6e3, 15e3))
