"""Device demodulation (pycbf_demod_apply through pycbf_amd.gpu.demod)
against the oracle's FFT chain, and the split-BF16 product against float64
on ragged shapes."""
import ctypes
import time

import numpy as np
import pytest

from oracle import demod as od

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24


@pytest.mark.parametrize("rows,k,ncol", [(1, 1, 1), (63, 17, 130), (65, 100, 3), (130, 15, 64),
                                         (200, 513, 66), (128, 128, 128), (257, 33, 129),
                                         # split over k (few lines, long lines): partial
                                         # products and an ordered reduction
                                         (64, 4096, 130), (257, 1000, 129), (5, 777, 31)])
def test_product_within_fp32_bound(rows, k, ncol):
    """iq = rf @ opT^T from the three-way BF16 split: six exact partial
    products per term (6k summands in float32) plus the dropped cross terms
    (below 2^-23 |a||b| each) stay within (6k + 4) u (|rf| @ |opT|^T); a
    transposed operand or a wrong lane layout misses it by orders of
    magnitude.  Operands span 2^-20 .. 2^20 so the split sees every binade."""
    import torch
    from pycbf_amd.gpu import _engine as E
    print(f"demod product rows={rows} k={k} ncol={ncol}", flush=True)
    rng = np.random.default_rng(rows * 1000 + k)
    a = (rng.standard_normal((rows, k)) * 2.0 ** rng.integers(-20, 21, (rows, k))).astype(np.float32)
    bt = (rng.standard_normal((ncol, k)) * 2.0 ** rng.integers(-20, 21, (ncol, k))).astype(np.float32)
    dev = torch.device("cuda", 0)
    ad, bd = torch.as_tensor(a, device=dev), torch.as_tensor(bt, device=dev)
    out = torch.full((rows, ncol), float("nan"), dtype=torch.float32, device=dev)
    need = int(E.lib.pycbf_demod_workspace_bytes(rows, k, ncol))
    ws = torch.full((max(need // 4, 1),), float("nan"), dtype=torch.float32, device=dev)
    rc = E.lib.pycbf_demod_apply(E.ptr(ad), rows, k, E.ptr(bd), ncol, E.ptr(out), E.ptr(ws),
                                 need, E.stream_of(dev))
    E.check(rc, "pycbf_demod_apply")
    got = out.cpu().numpy().astype(np.float64)
    a64, b64 = a.astype(np.float64), bt.astype(np.float64)
    want = a64 @ b64.T
    bound = (6 * k + 4) * U32 * (np.abs(a64) @ np.abs(b64).T)
    assert np.all(np.isfinite(got))
    assert np.all(np.abs(got - want) <= bound + 1e-30)
    # typical error is that of a float32 product, far inside the bound
    assert np.median(np.abs(got - want) / (np.abs(a64) @ np.abs(b64).T + 1e-300)) < 4 * U32 * np.sqrt(k)


def _lines(nlines, n, dz, seed):
    """Band-limited RF lines (5-12 MHz echoes at 1540 m/s) plus noise."""
    rng = np.random.default_rng(seed)
    z = np.arange(n) * dz
    k0 = 2 * 7.5e6 / 1540.0
    x = np.zeros((nlines, n))
    for i in range(nlines):
        for zc, amp in zip(rng.uniform(z[0], z[-1], 12), rng.standard_normal(12)):
            x[i] += amp * np.exp(-((z - zc) / (3 / k0)) ** 2) * np.cos(2 * np.pi * k0 * (z - zc))
    return x + 0.01 * rng.standard_normal(x.shape)


@pytest.mark.parametrize("n,shiftit,method,dz", [
    (1024, True, None, 20e-6), (777, True, None, 20e-6), (512, False, None, 20e-6),
    (1024, True, "operator", 20e-6), (2048, True, "fft", 9.625e-6),
    (4096, False, "fft", 5e-6), (4096, True, "fft", 20e-6), (8192, True, None, 20e-6), (64, True, "fft", 20e-6)])
def test_demod_matches_oracle(n, shiftit, method, dz):
    """Both methods (the FFT kernel for power-of-two lengths, the operator
    product otherwise or on request) against the FP64 NumPy chain."""
    import torch
    from pycbf_amd.gpu.demod import fft_tukey_demod
    print(f"demod lines n={n} shiftit={shiftit} method={method}", flush=True)
    kmin, kmax, ks = 2 * 5e6 / 1540, 2 * 12e6 / 1540, 1 / dz
    x = _lines(96, n, dz, n)
    want = od.fft_tukey_demod(x, kmin, kmax, ks, shiftit=shiftit)
    got = fft_tukey_demod(torch.as_tensor(x.astype(np.float32), device="cuda"), kmin, kmax, ks,
                          shiftit=shiftit, method=method)
    assert got.dtype == torch.complex64 and tuple(got.shape) == want.shape
    g = got.cpu().numpy()
    peak = np.abs(want).max()
    assert peak > 0
    # float32 input rounding plus float32 accumulation over n samples (or
    # over the log2 n stages of the transforms)
    print(f"  max error {np.abs(g - want).max() / peak:.2e} of peak", flush=True)
    assert np.abs(g - want).max() <= 2e-5 * peak


def test_demod_axis_and_numpy_input():
    """Axis 0 of a 3-D NumPy array (beamformed (nz, nx, frames)): NumPy
    complex64 out, equal to the oracle along that axis."""
    from pycbf_amd.gpu.demod import fft_tukey_demod
    dz = 25e-6
    kmin, kmax, ks = 2 * 5e6 / 1540, 2 * 12e6 / 1540, 1 / dz
    x = _lines(24, 600, dz, 9).reshape(4, 6, 600).transpose(2, 1, 0).copy()  # (600, 6, 4)
    got = fft_tukey_demod(x, kmin, kmax, ks, axis=0)
    want = od.fft_tukey_demod(x, kmin, kmax, ks, axis=0)
    assert isinstance(got, np.ndarray) and got.dtype == np.complex64
    assert got.shape == want.shape
    assert np.abs(got - want).max() <= 2e-5 * np.abs(want).max()


def test_demod_no_lines():
    import torch
    from pycbf_amd.gpu.demod import fft_tukey_demod
    x = torch.zeros((0, 256), dtype=torch.float32, device="cuda")
    iq = fft_tukey_demod(x, 0.1, 0.3, 1.0)
    ikmin, ikmax = od.band(256, 0.1, 0.3, 1.0)
    assert tuple(iq.shape) == (0, ikmax - ikmin)


def test_demod_full_image_timing():
    """512 lines x 4096 axial samples (a 512-line image at lambda/8), band
    5-12 MHz: one product after the operator is cached."""
    import torch
    from pycbf_amd.gpu.demod import fft_tukey_demod
    dz = 1540 / (2 * 40e6) / 2
    kmin, kmax, ks = 2 * 5e6 / 1540, 2 * 12e6 / 1540, 1 / dz
    x = torch.randn((512, 4096), dtype=torch.float32, device="cuda")
    fft_tukey_demod(x, kmin, kmax, ks)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        iq = fft_tukey_demod(x, kmin, kmax, ks)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e2
    print(f"demod 512 x 4096 -> {tuple(iq.shape)}: {ms:.3f} ms per call", flush=True)
    assert bool(torch.isfinite(torch.view_as_real(iq)).all())
