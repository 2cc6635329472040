"""GPU SLSC (pycbf_slsc, FP32 and MFMA FP16/BF16) against the FP64 oracle
restatement of examples/advanced_beamforming.ipynb:65-72,205-223.

Bars (coherence values lie in [-1, 1]): FP32 within 1e-4 absolute, MFMA FP16
and BF16 within 1e-3 absolute; NaN exactly where the oracle has NaN; zeros
exactly where too few elements are selected."""
import zlib

import numpy as np
import pytest

from oracle import numerics
from tests.helpers import small_synthetic, small_tabbed

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-4, "fp16": 1e-3, "bf16": 1e-3}


def _channels(rng, nrx, nx, nz, snr_db=3.0):
    """Per-rx images with a common component (coherence well inside (0, 1))."""
    common = rng.standard_normal((1, nx, nz))
    gain = rng.uniform(0.5, 1.5, (nrx, 1, 1))
    noise = rng.standard_normal((nrx, nx, nz)) * 10 ** (-snr_db / 20)
    return (gain * common + noise).astype(np.float32)


def _compare(got, ref, tol, what):
    got = np.asarray(got, np.float64)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    assert np.array_equal(np.isnan(got), np.isnan(ref)), f"{what}: NaN pattern differs"
    fin = ~np.isnan(ref)
    assert np.array_equal(got[fin] == 0.0, ref[fin] == 0.0), f"{what}: zero pattern differs"
    err = float(np.max(np.abs(got[fin] - ref[fin]))) if fin.any() else 0.0
    print(f"{what}: max |err| vs FP64 = {err:.3e}", flush=True)
    assert err <= tol, f"{what}: max err {err:.3e} > {tol:g}"
    return err


CASES = {
    # nrx, nx, nz, kwin, kofs, dsfx, dsfz, lags, apod
    "128rx_k8_lags1-25": (128, 4, 40, 8, -4, 1, 1, list(range(1, 26)), False),
    "200rx_k5_lags1-40_apod_ds": (200, 6, 31, 5, -2, 2, 3, list(range(1, 41)), True),
    "256rx_k17_sparse_lags": (256, 2, 24, 17, -8, 1, 2, [1, 3, 7, 31, 32, 33, 63], False),
    "40rx_k3_few_elements": (40, 3, 20, 3, -1, 1, 1, [1, 2, 5], True),
    "96rx_k16_lags1-20": (96, 3, 48, 16, -8, 1, 4, list(range(1, 21)), False),
    "160rx_k32_apod": (160, 2, 70, 32, -16, 1, 8, [1, 2, 4, 8, 16], True),
    # windows that tile the axis (dsfz == kwin): the channel-row kernel, with
    # windows clipped at both ends, partial runs of pixels, unaligned offsets
    "128rx_k8_tiles_lags1-25": (128, 3, 64, 8, -4, 1, 8, list(range(1, 26)), False),
    "200rx_k8_tiles_apod": (200, 4, 50, 8, -3, 2, 8, list(range(1, 31)), True),
    "100rx_k16_tiles_apod": (100, 2, 70, 16, -8, 1, 16, list(range(1, 13)), True),
    "64rx_k32_tiles": (64, 2, 100, 32, -16, 1, 32, [1, 2, 4, 8], False),
    "256rx_k8_tiles_sparse_lags": (256, 2, 40, 8, 0, 1, 8, [1, 3, 7, 31, 32, 33, 63], False),
}


@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("case", list(CASES))
def test_slsc_matches_fp64_oracle(case, precision):
    import torch
    from pycbf_amd.gpu import slsc
    nrx, nx, nz, kwin, kofs, dsfx, dsfz, lags, use_apod = CASES[case]
    print(f"slsc {case} {precision}", flush=True)
    rng = np.random.default_rng(zlib.crc32(case.encode()))
    im = _channels(rng, nrx, nx, nz)
    apod = None
    if use_apod:
        # an aperture that grows with depth and shifts with x, plus scattered holes
        e = np.arange(nrx)[:, None, None]
        half = 0.25 * nrx + 0.6 * nrx * np.arange(nz)[None, None, :] / nz
        ctr = nrx / 2 + (np.arange(nx)[None, :, None] - nx / 2) * 3
        apod = (np.abs(e - ctr) <= half).astype(np.float32)
        apod *= (rng.uniform(size=apod.shape) > 0.03)
    if case.startswith("256"):
        im[17] = 0.25          # constant channels: zero deviation, pairs dropped
        im[100, :, :10] = 0.0
    ref = numerics.slsc_image(im, apod, lags, kwin, kofs, dsfx, dsfz)
    dev = torch.device("cuda", 0)
    got = slsc(torch.as_tensor(im, device=dev), lags, kwin, kofs,
               None if apod is None else torch.as_tensor(apod, device=dev), dsfx, dsfz,
               precision)
    _compare(got.cpu().numpy(), ref, TOL[precision], f"{case} {precision}")


def test_slsc_frames_and_nan_lags():
    """Frames are independent; a lag whose pairs all include a flat channel is
    NaN and makes the pixel NaN (numpy semantics)."""
    import torch
    from pycbf_amd.gpu import slsc
    rng = np.random.default_rng(5)
    nrx, nx, nz = 64, 2, 16
    ims = np.stack([_channels(rng, nrx, nx, nz), _channels(rng, nrx, nx, nz)])
    ims[1, ::2, 1, :] = 3.0                    # frame 1, column 1: lag-1 pairs all flat
    got = slsc(torch.as_tensor(ims, device="cuda"), [1, 2], 6, -3, precision="fp32").cpu().numpy()
    for f in range(2):
        ref = numerics.slsc_image(ims[f], None, [1, 2], 6, -3)
        _compare(got[f], ref, TOL["fp32"], f"frame {f}")
    assert np.all(np.isnan(got[1, 1])) and not np.any(np.isnan(got[0]))


def test_slsc_precisions_agree_on_dense_windows():
    """All three precisions agree on a 128-channel 64x256 image (dense windows)."""
    import torch
    from pycbf_amd.gpu import slsc
    rng = np.random.default_rng(9)
    im = torch.as_tensor(_channels(rng, 128, 64, 256), device="cuda")
    outs = {p: slsc(im, list(range(1, 26)), 8, -4, precision=p).cpu().numpy()
            for p in ("fp32", "fp16", "bf16")}
    for p in ("fp16", "bf16"):
        assert np.max(np.abs(outs[p] - outs["fp32"])) < 1e-3


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_tabbed_slsc_beamformer(precision):
    import oracle
    from pycbf_amd.gpu import TabbedSLSCBeamformer
    w = small_tabbed(nx=16, nz=48)
    tautx, apodtx, taurx, apodrx = w["tabs"]
    nx, nz = 16, 48
    bf = TabbedSLSCBeamformer(tautx=tautx, taurx=taurx, apodtx=apodtx, apodrx=apodrx,
                              t0=w["t0"], dt=w["dt"], nt=w["nt"], interp={"kind": "cubic"},
                              lags=np.arange(1, 9), gridshape=(nx, nz), kwin=5, dsfz=2,
                              precision=precision)
    got = bf(w["rf"])
    imrx = oracle.das_tabbed(w["rf"], w["t0"], w["dt"], *w["tabs"], kind="cubic",
                             sumtype="tx_only")
    ref = numerics.slsc_image(imrx.reshape(-1, nx, nz), apodrx.reshape(-1, nx, nz),
                              np.arange(1, 9), 5, -2, 1, 2)
    _compare(got, ref, TOL[precision], f"tabbed SLSC {precision}")
    assert np.count_nonzero(got) > 0.5 * got.size


def test_synthetic_slsc_beamformer_frames():
    import oracle
    from pycbf_amd.gpu import SyntheticSLSCBeamformer
    w = small_synthetic(nx=12, nz=32)
    p = w["params"]
    nx, nz = 12, 32
    rf2 = np.stack([w["rf"], w["rf"][::-1].copy()])
    bf = SyntheticSLSCBeamformer(**p, t0=w["t0"], dt=w["dt"], nt=w["nt"],
                                 interp={"kind": "cubic"}, lags=np.arange(1, 6),
                                 gridshape=(nx, nz), kwin=3, precision="fp32")
    got = bf(rf2)
    assert got.shape == (2, nx, nz)
    tabs = oracle.synthetic_tables(p["ovectx"], p["nvectx"], p["t0tx"], p["alatx"], p["doftx"],
                                   p["ovecrx"], p["nvecrx"], p["alarx"], p["c0"], p["pnts"])
    for f in range(2):
        imrx = oracle.das_tabbed(rf2[f], w["t0"], w["dt"], *tabs, kind="cubic",
                                 sumtype="tx_only")
        ref = numerics.slsc_image(imrx.reshape(-1, nx, nz), tabs[3].reshape(-1, nx, nz),
                                  np.arange(1, 6), 3, -1)
        _compare(got[f], ref, TOL["fp32"], f"synthetic SLSC frame {f}")


def test_slsc_selection_bits_match_window_rule():
    """pycbf_slsc_selection: bit r set iff the mask is non-zero over the whole
    clipped window of the output pixel."""
    import torch
    from pycbf_amd.gpu import slsc_selection
    rng = np.random.default_rng(21)
    nrx, nx, nz, kwin, kofs, dsfx, dsfz = 200, 5, 37, 6, -3, 2, 4
    mask = (rng.uniform(size=(nrx, nx, nz)) > 0.02).astype(np.uint8)
    sel = slsc_selection(torch.as_tensor(mask, device="cuda"), kwin, kofs, dsfx, dsfz)
    bits = sel.cpu().numpy().view(np.uint64)
    for ixo in range(bits.shape[0]):
        for izo in range(bits.shape[1]):
            ix, iz = ixo * dsfx, izo * dsfz
            z0, z1 = max(0, iz + kofs), min(nz, iz + kofs + kwin)
            ok = np.all(mask[:, ix, z0:z1] != 0, axis=1)
            exp = np.zeros(4, np.uint64)
            for r in np.nonzero(ok)[0]:
                exp[r // 64] |= np.uint64(1) << np.uint64(r % 64)
            assert np.array_equal(bits[ixo, izo], exp), (ixo, izo)
