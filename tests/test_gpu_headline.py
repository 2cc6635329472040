"""GPU parity of the headline path against the oracle, at the headline sizes.

Workload: Synthetic-point cubic DAS of 75 plane waves over +-16 deg (encoded
as synthetic points at -10 m) x 128 receive elements (0.3 mm pitch, f/1.5
acceptance) x 4096 float32 samples at 40 MHz onto the 512 x 512 grid over
z 2-50 mm, x +-19 mm, sumtype tx_and_rx, through ``SyntheticBeamformer``
(``pycbf_das_synthetic``).  The GPU beamforms every pixel with the launch the
library picks for these sizes: 16 frames per workgroup, the transmit axis
split into 75 partial images and their reduction.  20 frames add a second
frame pass with a 4-frame tail.

The oracle (its float32 restatement of pycbf/cpu/pycbf.c:179-207 driven over
the synthetic tables of pycbf/gpu/__engines__.cu:382-470, both pinned bit for
bit to the reference C by the CPU tests) evaluates every 7th pixel: 37 450
pixels that fall in every 256-pixel tile of the processing order.  Bars:
images within 1e-5 of the oracle image's peak, frame by frame; sample indices
and per-sample linear outputs of the in-register geometry bit for bit.
"""
import os
import warnings

import numpy as np
import pytest

import oracle
from oracle import numerics
from pycbf_amd import sim
from pycbf_amd.gpu import SyntheticBeamformer
from tests.helpers import assert_close_to_peak

pytestmark = pytest.mark.gpu

NTX, NRX, NT, NX, NZ = 75, 128, 4096, 512, 512
DT = 1.0 / sim.FS
STRIDE = 7


@pytest.fixture(scope="module")
def geo():
    xele = sim.linear_array(NRX)
    steers = np.radians(np.linspace(-16, 16, NTX))
    _, _, pnts = sim.grid(NX, NZ, (-19e-3, 19e-3), (2e-3, 50e-3))
    params = sim.pw_synthetic(steers, xele, pnts)
    sub = np.arange(0, NX * NZ, STRIDE)
    p = params
    tabs = oracle.synthetic_tables(p["ovectx"], p["nvectx"], p["t0tx"], p["alatx"], p["doftx"],
                                   p["ovecrx"], p["nvecrx"], p["alarx"], p["c0"], pnts[sub])
    return dict(xele=xele, steers=steers, params=params, sub=sub, tabs=tabs)


@pytest.fixture(scope="module")
def rf20(geo):
    """20 seeded frames (50 points + 500 speckle scatterers each, 30 dB SNR)."""
    import torch
    dev = torch.device("cuda", 0)
    rf = torch.empty((20, NTX, NRX, NT), dtype=torch.float32, device=dev)
    for f in range(20):
        xs, zs, amp = sim.scatterers(50, 500, (-19e-3, 19e-3), (2e-3, 50e-3), 900 + f)
        rf[f] = sim.simulate_rf_torch(geo["steers"], geo["xele"], NT, xs, zs, amp, dev,
                                      seed=900 + f)
    torch.cuda.synchronize()
    return rf


def _bf(geo, interp, sumtype="tx_and_rx"):
    return SyntheticBeamformer(**geo["params"], t0=0.0, dt=DT, nt=NT, interp=interp,
                               sumtype=sumtype, device="cuda:0")


def _oracle_frame(geo, rf_frame, kind="cubic", S=None):
    tautx, apodtx, taurx, apodrx = geo["tabs"]
    img, _ = oracle.das_tabbed_mt(rf_frame, 0.0, DT, tautx, apodtx, taurx, apodrx, kind=kind,
                                  S=S)
    return img


@pytest.mark.parametrize("nfr", [16, 20])
def test_headline_cubic_frames_match_oracle(geo, rf20, nfr):
    """Every frame of a 16- and a 20-frame batch, cubic, tx_and_rx."""
    import torch
    print(f"headline cubic {NTX}x{NRX}x{NT} {NX}x{NZ} px, {nfr} frames", flush=True)
    bf = _bf(geo, {"kind": "cubic"})
    out = bf(rf20[:nfr], out_as_numpy=False)
    torch.cuda.synchronize()
    assert tuple(out.shape) == (nfr, NX * NZ)
    got = out[:, torch.as_tensor(geo["sub"], device=out.device)].cpu().numpy()
    assert np.all(np.isfinite(got))
    for f in range(nfr):
        ref = _oracle_frame(geo, rf20[f].cpu().numpy())
        err, peak = assert_close_to_peak(got[f], ref, 1e-5, f"frame {f} of {nfr}")
        assert peak > 0.0
        if f in (0, nfr - 1):
            print(f"  frame {f}: max err {err:.3e} (peak {peak:.3e})", flush=True)


def test_headline_indices_bitwise(geo):
    """The in-register geometry of the staged gather selects the reference's
    sample for every (tx, rx, px): nearest, sumtype none, channel data
    rf[t] = t + 1 and weights 0 / 1, so each output is the selected index + 1
    (or 0 where the pixel is outside the element's cone or the trace)."""
    import torch
    print("headline nearest indices, sumtype none", flush=True)
    dev = torch.device("cuda", 0)
    ramp = torch.arange(1, NT + 1, dtype=torch.float32, device=dev)
    rf = ramp.expand(NTX, NRX, NT).contiguous()
    bf = _bf(geo, {"kind": "nearest"}, "none")
    out = bf(rf, out_as_numpy=False)
    sub = torch.as_tensor(geo["sub"], device=dev)
    got = out[:, :, sub].cpu().numpy()
    del out
    torch.cuda.empty_cache()
    tautx, apodtx, taurx, apodrx = geo["tabs"]
    idx = oracle.indices(tautx, taurx, 0.0, DT, NT, kind="nearest")
    w = apodtx[:, None, :] * apodrx[None, :, :]
    assert np.all((w == 0.0) | (w == 1.0))
    exp = np.where((idx >= 0) & (w >= np.float32(1e-2)), idx + 1, 0).astype(np.float32)
    bad = int((got != exp).sum())
    assert bad == 0, f"{bad} of {exp.size} selected samples differ"
    assert (exp > 0).mean() > 0.2


def test_headline_linear_samples_bitwise(geo, rf20):
    """Per-sample outputs of the linear kind (knot index and both weights from
    the in-register delays) are bit for bit the oracle's."""
    import torch
    print("headline linear per-sample outputs, sumtype none", flush=True)
    dev = torch.device("cuda", 0)
    bf = _bf(geo, {"kind": "linear"}, "none")
    out = bf(rf20[3], out_as_numpy=False)
    sub = torch.as_tensor(geo["sub"], device=dev)
    got = out[:, :, sub].cpu().numpy()
    del out
    torch.cuda.empty_cache()
    ref = oracle.das_tabbed(rf20[3].cpu().numpy(), 0.0, DT, *geo["tabs"], kind="linear",
                            sumtype="none")
    bad = int((got != ref).sum())
    assert bad == 0, f"{bad} of {ref.size} samples differ"


def _max_window(geo, perm, es_pair):
    """Largest sample window (table entries) a 256-pixel tile of the
    processing order spans in one channel: the staged gather's slot holds
    80 entries for the Hermite kinds at 16 frames."""
    p = geo["params"]
    tabs = oracle.synthetic_tables(p["ovectx"], p["nvectx"], p["t0tx"], p["alatx"], p["doftx"],
                                   p["ovecrx"], p["nvecrx"], p["alarx"], p["c0"], p["pnts"])
    tautx, apodtx, taurx, apodrx = tabs
    ntile = (NX * NZ) // 256
    order = perm.reshape(ntile, 256)

    def rng(tau, apod):
        t = np.where(apod != 0, tau, np.nan)[:, order]     # (ne, ntile, 256)
        return np.nanmin(t, axis=2), np.nanmax(t, axis=2)

    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)   # tiles no element reaches
        txlo, txhi = rng(tautx, apodtx)
        rxlo, rxhi = rng(taurx, apodrx)
    lo = txlo[:, None, :] + rxlo[None, :, :]
    hi = txhi[:, None, :] + rxhi[None, :, :]
    jlo = np.floor(np.clip(lo, 0, None) / np.float32(DT))
    jhi = np.floor(np.clip(hi, 0, None) / np.float32(DT))
    return float(np.nanmax(jhi - jlo + 1 + es_pair))


def test_headline_korder_window_fallback_and_direct_gather(geo, rf20):
    """korder_cubic at 9 frames takes the 16-frame gather whose LDS slot (80
    table entries) is smaller than some tiles' windows, so those channels use
    the direct global gather; and the direct-gather kernel (PYCBF_DAS_LDS=0)
    on the cubic headline.  Both against the oracle."""
    import torch
    print("headline korder_cubic 9 frames (window fallback)", flush=True)
    bf = _bf(geo, {"kind": "korder_cubic", "k": 8})
    perm = bf._perm.cpu().numpy()
    assert _max_window(geo, perm, 0) > 80
    out = bf(rf20[:9], out_as_numpy=False)
    sub = torch.as_tensor(geo["sub"], device=out.device)
    got = out[:, sub].cpu().numpy()
    S = numerics.make_S(8).astype(np.float32)
    for f in (0, 8):
        ref = _oracle_frame(geo, rf20[f].cpu().numpy(), "korder_cubic", S)
        assert_close_to_peak(got[f], ref, 1e-5, f"korder frame {f}")
    print("headline cubic, direct gather (PYCBF_DAS_LDS=0), 2 frames", flush=True)
    os.environ["PYCBF_DAS_LDS"] = "0"
    try:
        bfd = _bf(geo, {"kind": "cubic"})
        outd = bfd(rf20[:2], out_as_numpy=False)
        torch.cuda.synchronize()
    finally:
        del os.environ["PYCBF_DAS_LDS"]
    gotd = outd[:, sub].cpu().numpy()
    for f in range(2):
        ref = _oracle_frame(geo, rf20[f].cpu().numpy())
        assert_close_to_peak(gotd[f], ref, 1e-5, f"direct gather frame {f}")


def test_headline_plain_comparator(geo, rf20):
    """The thread-per-sample comparator (PYCBF_DAS_PLAIN=1: pycbf's GPU launch
    scheme with one float atomic per sample) on the cubic headline, 2 frames."""
    import torch
    print("headline cubic, thread-per-sample comparator (PYCBF_DAS_PLAIN=1), 2 frames",
          flush=True)
    os.environ["PYCBF_DAS_PLAIN"] = "1"
    try:
        bf = _bf(geo, {"kind": "cubic"})
        out = bf(rf20[:2], out_as_numpy=False)
        torch.cuda.synchronize()
    finally:
        del os.environ["PYCBF_DAS_PLAIN"]
    got = out[:, torch.as_tensor(geo["sub"], device=out.device)].cpu().numpy()
    for f in range(2):
        ref = _oracle_frame(geo, rf20[f].cpu().numpy())
        assert_close_to_peak(got[f], ref, 1e-5, f"plain comparator frame {f}")


def test_headline_staged_register_gather(geo, rf20):
    """The register-staged gather (PYCBF_DAS_DMA=0, k_das_lds) on the cubic
    headline, 16 frames, and with round-robin workgroup placement
    (PYCBF_XCD_REMAP=0) for 2 frames: both against the oracle."""
    import torch
    print("headline cubic, register-staged gather (PYCBF_DAS_DMA=0), 16 frames", flush=True)
    sub = torch.as_tensor(geo["sub"], device="cuda")
    os.environ["PYCBF_DAS_DMA"] = "0"
    try:
        out = _bf(geo, {"kind": "cubic"})(rf20[:16], out_as_numpy=False)
        torch.cuda.synchronize()
    finally:
        del os.environ["PYCBF_DAS_DMA"]
    got = out[:, sub].cpu().numpy()
    for f in (0, 7, 15):
        ref = _oracle_frame(geo, rf20[f].cpu().numpy())
        assert_close_to_peak(got[f], ref, 1e-5, f"register-staged frame {f}")
    os.environ["PYCBF_XCD_REMAP"] = "0"
    try:
        out = _bf(geo, {"kind": "cubic"})(rf20[:2], out_as_numpy=False)
        torch.cuda.synchronize()
    finally:
        del os.environ["PYCBF_XCD_REMAP"]
    got = out[:, sub].cpu().numpy()
    for f in range(2):
        ref = _oracle_frame(geo, rf20[f].cpu().numpy())
        assert_close_to_peak(got[f], ref, 1e-5, f"round-robin placement frame {f}")


def _windows(tabs, perm, dt, es_pair):
    """Sample window (table entries) of every (tile, tx, rx) of the processing
    order, NaN where no pixel of the tile takes a sample of the channel."""
    tautx, apodtx, taurx, apodrx = tabs
    order = perm.reshape(-1, 256)

    def rng(tau, apod):
        t = np.where(apod != 0, tau, np.nan)[:, order]
        return np.nanmin(t, axis=2), np.nanmax(t, axis=2)

    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        txlo, txhi = rng(tautx, apodtx)
        rxlo, rxhi = rng(taurx, apodrx)
    lo = txlo[:, None, :] + rxlo[None, :, :]
    hi = txhi[:, None, :] + rxhi[None, :, :]
    jlo = np.floor(np.clip(lo, 0, None) / np.float32(dt))
    jhi = np.floor(np.clip(hi, 0, None) / np.float32(dt))
    return jhi - jlo + 1 + es_pair


def test_dma_gather_wide_and_narrow_windows():
    """Pixels in a fine patch (0.6 x 0.45 mm at 20 mm) and on a coarse grid over
    4-12 mm, so some tiles' channel windows fit the 16-frame staging slot (95
    entries) and others take the DMA kernel's direct-gather loop; 9 transmits x
    31 receives x 1025 table entries (an odd count per frame: consecutive
    frames' rows start on alternating parities), 10 frames, cubic tx_and_rx,
    every pixel against the oracle."""
    import torch
    ntx, nrx, nt, nfr = 9, 31, 1024, 10
    dt = 1.0 / sim.FS
    xele = sim.linear_array(nrx)
    steers = np.radians(np.linspace(-12, 12, ntx))
    fx, fz = np.meshgrid(np.linspace(-0.3e-3, 0.3e-3, 32), np.linspace(20e-3, 20.45e-3, 16))
    cx, cz = np.meshgrid(np.linspace(-5e-3, 5e-3, 32), np.linspace(4e-3, 12e-3, 16))
    pnts = np.ascontiguousarray(np.stack([np.concatenate([fx.ravel(), cx.ravel()]),
                                          np.concatenate([fz.ravel(), cz.ravel()])], axis=1))
    params = sim.pw_synthetic(steers, xele, pnts)
    p = params
    tabs = oracle.synthetic_tables(p["ovectx"], p["nvectx"], p["t0tx"], p["alatx"], p["doftx"],
                                   p["ovecrx"], p["nvecrx"], p["alarx"], p["c0"], pnts)
    bf = SyntheticBeamformer(**params, t0=0.0, dt=dt, nt=nt, interp={"kind": "cubic"},
                             sumtype="tx_and_rx", device="cuda:0")
    win = _windows(tabs, bf._perm.cpu().numpy(), dt, 1)
    print(f"wide/narrow windows: {int(np.sum(win > 95))} wide, {int(np.sum(win <= 95))} staged",
          flush=True)
    assert np.sum(win > 95) > 0 and np.sum(win <= 95) > 0
    rng = np.random.default_rng(17)
    rf = rng.standard_normal((nfr, ntx, nrx, nt)).astype(np.float32)
    out = bf(torch.as_tensor(rf, device="cuda"), out_as_numpy=False)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for f in (0, 4, 9):
        ref, _ = oracle.das_tabbed_mt(rf[f], 0.0, dt, *tabs, kind="cubic")
        assert_close_to_peak(got[f], ref, 1e-5, f"wide/narrow frame {f}")
