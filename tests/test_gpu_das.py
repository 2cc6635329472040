"""GPU parity of the DAS hot path against the CPU oracle (through the C-ABI).

Bars: identical sample-index selection (bitwise), per-sample outputs of
nearest / linear bitwise equal to the oracle, images within 1e-5 of the
oracle's image peak (float32 sums in a different order)."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from oracle import numerics
from pycbf_amd import sim
from pycbf_amd.gpu import SyntheticBeamformer, TabbedBeamformer, make_S_by_k
from tests.helpers import assert_close_to_peak, small_synthetic, small_tabbed

pytestmark = pytest.mark.gpu

KINDS = [({"kind": "nearest"}, "nearest", 1), ({"kind": "nearest", "usf": 4}, "nearest", 4),
         ({"kind": "linear"}, "linear", 1), ({"kind": "cubic"}, "cubic", 1),
         ({"kind": "korder_cubic", "k": 8}, "korder_cubic", 1)]


@pytest.fixture(scope="module")
def c0():
    """The plain Tabbed case: 5 PW x 128 rx x 2048 samples, 256x128 grid."""
    return sim.workload_tabbed_small()


@pytest.fixture(scope="module")
def small():
    return small_tabbed()


def _tabbed(w, interp, sumtype="tx_and_rx", **kw):
    tautx, apodtx, taurx, apodrx = w["tabs"]
    return TabbedBeamformer(tautx=tautx, taurx=taurx, apodtx=apodtx, apodrx=apodrx, t0=w["t0"],
                            dt=w["dt"], nt=w["nt"], interp=interp, sumtype=sumtype, **kw)


def _oracle(w, kind, usf, sumtype="tx_and_rx", rf=None):
    S = numerics.make_S(8).astype(np.float32) if kind == "korder_cubic" else None
    return oracle.das_tabbed(w["rf"] if rf is None else rf, w["t0"], w["dt"], *w["tabs"],
                             kind=kind, usf=usf, sumtype=sumtype, S=S)


@pytest.mark.parametrize("interp,kind,usf", KINDS, ids=[k[1] + str(k[2]) for k in KINDS])
def test_tabbed_full_case_matches_oracle(c0, interp, kind, usf):
    print(f"tabbed 5x128x2048 {kind} usf={usf}", flush=True)
    got = _tabbed(c0, interp)(c0["rf"])
    ref = _oracle(c0, kind, usf)
    assert_close_to_peak(got, ref, 1e-5, f"{kind} usf={usf}")


@pytest.mark.parametrize("kind,usf", [("nearest", 1), ("nearest", 4), ("cubic", 1)])
def test_index_selection_bitwise(c0, kind, usf):
    """Every (tx, rx, px) selects the same sample index as the reference rule."""
    import torch
    from pycbf_amd.gpu import _engine as E
    tautx, _, taurx, _ = c0["tabs"]
    ntx, np_ = tautx.shape
    nrx = taurx.shape[0]
    dev = torch.device("cuda")
    ttx = torch.as_tensor(tautx, device=dev)
    trx = torch.as_tensor(taurx, device=dev)
    idx = torch.empty((ntx, nrx, np_), dtype=torch.int32, device=dev)
    info = E.rfinfo(ntx, nrx, 0, c0["t0"], c0["dt"], c0["nt"])
    rc = E.lib.pycbf_debug_indices(ctypes.byref(info), E.ptr(ttx), E.ptr(trx), np_,
                                   E.INTERP[kind], usf, E.ptr(idx), E.stream_of(dev))
    E.check(rc, "pycbf_debug_indices")
    ref = oracle.indices(tautx, taurx, c0["t0"], c0["dt"], c0["nt"], kind=kind, usf=usf)
    got = idx.cpu().numpy()
    assert np.array_equal(got, ref), f"{int((got != ref).sum())} indices differ"
    assert (ref >= 0).mean() > 0.5


@pytest.mark.parametrize("data", ["normal", "sparse", "tail", "bursts"])
@pytest.mark.parametrize("nt", [1, 2, 3, 31, 32, 33, 64, 96, 100, 2047, 2048, 2049, 4095, 4096,
                                4097, 6000])
def test_spline_prefilter_bitwise(nt, data):
    """Second derivatives (pycbf_spline_y2) and the gather table
    (pycbf_spline_coefs) are bit-identical to tie_knots1D_fixed's, for dense
    data, for sparse spikes between exact zeros (whose recurrence tails decay
    through the subnormal range), for echoes followed by a long run of zeros,
    and for short bursts between zero runs of every length.  Up to 4096
    samples one wave sweeps each channel in segments whose guessed carries
    are checked and recomputed (the zero runs force recomputation); longer
    channels take the one-lane-per-channel kernel."""
    import torch
    from pycbf_amd.gpu import _engine as E
    print(f"spline nt={nt} data={data}", flush=True)
    rng = np.random.default_rng(nt)
    nch, dt = 70, np.float32(2.5e-8)
    rf = rng.standard_normal((nch, nt)).astype(np.float32)
    if data == "sparse":
        rf *= (rng.random((nch, nt)) < 0.01).astype(np.float32)
    elif data == "tail":
        # channel c keeps its first 1 + (c * 37) % nt samples
        keep = 1 + (np.arange(nch) * 37) % max(nt, 1)
        rf[np.arange(nt)[None, :] >= keep[:, None]] = 0.0
    elif data == "bursts":
        on = np.zeros((nch, nt), dtype=bool)
        for c in range(nch):
            k = 0
            while k < nt:
                n_on = int(rng.integers(1, 8))
                on[c, k:k + n_on] = True
                k += n_on + int(rng.integers(1, 200))
        rf *= on.astype(np.float32)
    dev = torch.device("cuda")
    x = torch.as_tensor(rf, device=dev)
    wsb = E.lib.pycbf_spline_workspace_bytes(nch, nt)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    y2 = torch.empty((nch, nt), dtype=torch.float32, device=dev)
    tab = torch.empty((nch, nt + 1, 2), dtype=torch.float32, device=dev)
    s = E.stream_of(dev)
    E.check(E.lib.pycbf_spline_y2(E.ptr(x), nch, nt, float(dt), E.ptr(y2), E.ptr(ws), wsb, s), "y2")
    E.check(E.lib.pycbf_spline_coefs(E.ptr(x), nch, nt, float(dt), E.ptr(tab), E.ptr(ws), wsb, s),
            "coefs")
    ref = np.stack([oracle.spline_y2(r, float(dt)) for r in rf])
    got = y2.cpu().numpy()
    assert np.array_equal(got, ref), f"{int((got != ref).sum())} y2 values differ"
    if 64 < nt <= 4096:
        # the one-lane-per-channel sweep (PYCBF_SPLINE_SEG=0) gives the same bits
        os.environ["PYCBF_SPLINE_SEG"] = "0"
        try:
            E.check(E.lib.pycbf_spline_y2(E.ptr(x), nch, nt, float(dt), E.ptr(y2), E.ptr(ws), wsb,
                                          s), "y2")
            assert np.array_equal(y2.cpu().numpy(), ref)
        finally:
            del os.environ["PYCBF_SPLINE_SEG"]
    h = np.float32(np.float32(1.0) / np.float32(6.0)) * (dt * dt)
    t = tab.cpu().numpy()
    assert np.array_equal(t[:, :nt, 0], rf)
    assert np.array_equal(t[:, :nt, 1], (h * ref).astype(np.float32))
    assert np.all(t[:, nt] == 0)


@pytest.mark.parametrize("sumtype", ["none", "tx_only", "rx_only", "tx_and_rx"])
@pytest.mark.parametrize("interp,kind,usf", KINDS, ids=[k[1] + str(k[2]) for k in KINDS])
def test_sumtypes_match_oracle(small, sumtype, interp, kind, usf):
    print(f"sumtype {sumtype} {kind} usf={usf}", flush=True)
    got = _tabbed(small, interp, sumtype)(small["rf"])
    ref = _oracle(small, kind, usf, sumtype)
    assert got.shape == ref.shape
    if sumtype == "none" and kind in ("nearest", "linear"):
        # one product per output: identical float32 arithmetic
        assert np.array_equal(got, ref), f"{int((got != ref).sum())} samples differ"
    else:
        assert_close_to_peak(got, ref, 1e-5, f"{sumtype} {kind}")


def test_frames_batch_equals_single_frames(small):
    """A 4-D batch beamforms every frame (the reference computed frame 0 only)."""
    rfs = np.stack([small["rf"], 0.5 * small["rf"][:, :, ::-1].copy(), -small["rf"]])
    bf = _tabbed(small, {"kind": "cubic"})
    batch = bf(rfs)
    assert batch.shape == (3, small["tabs"][0].shape[1])
    for f in range(3):
        single = bf(rfs[f])
        assert_close_to_peak(batch[f], single, 1e-6, f"frame {f}")
        ref = _oracle(small, "cubic", 1, rf=rfs[f])
        assert_close_to_peak(batch[f], ref, 1e-5, f"frame {f} vs oracle")


def test_output_buffer_and_torch_input(small):
    import torch
    bf = _tabbed(small, {"kind": "linear"})
    rf = torch.as_tensor(small["rf"], device="cuda")
    buf = torch.full((small["tabs"][0].shape[1],), 123.0, device="cuda")
    out = bf(rf, out_as_numpy=False, buffer=buf)
    assert out.data_ptr() == buf.data_ptr()
    ref = _oracle(small, "linear", 1)
    assert_close_to_peak(out.cpu().numpy(), ref, 1e-5)


def test_threshold_semantics(small):
    """|apodtx*apodrx| >= thresh, CPU rule, with a non-default threshold."""
    for thresh in (0.0, 0.3, 0.75):
        got = _tabbed(small, {"kind": "linear"}, "none", thresh=thresh)(small["rf"])
        ref = oracle.das_tabbed(small["rf"], small["t0"], small["dt"], *small["tabs"],
                                kind="linear", thresh=thresh, sumtype="none")
        assert np.array_equal(got, ref)


def test_akima_makima_linear_data_exact(small):
    """Akima / makima reproduce affine channel data: y(q) = 0.01 q - 2 at the
    fractional sample q = (tautx + taurx - t0)/dt, times the weight."""
    nt, t0, dt = small["nt"], small["t0"], small["dt"]
    t = np.arange(nt, dtype=np.float64)
    rf = np.broadcast_to((0.01 * t - 2.0).astype(np.float32), small["rf"].shape).copy()
    tautx, apodtx, taurx, apodrx = small["tabs"]
    x = (tautx[:, None, :] + taurx[None, :, :]).astype(np.float32)
    w = (apodtx[:, None, :] * apodrx[None, :, :]).astype(np.float32)
    tmax = np.float32(t0) + np.float32(dt) * np.float32(nt - 1)
    q = (x.astype(np.float64) - t0) / np.float64(np.float32(dt))
    live = (x >= np.float32(t0)) & (x <= tmax) & (np.abs(w) >= np.float32(1e-2))
    exp = np.where(live, w * (0.01 * q - 2.0), 0.0)
    for kind in ("akima", "makima"):
        got = _tabbed(small, {"kind": kind}, "none")(rf)
        assert_close_to_peak(got, exp, 2e-6, kind)


# ------------------------------------------------------------------ synthetic
SYN_MODES = ["pw", "div", "foc", "dof"]


@pytest.mark.parametrize("mode", SYN_MODES)
def test_synthetic_tables_bitwise(mode):
    w = small_synthetic(mode=mode)
    p = w["params"]
    bf = SyntheticBeamformer(**p, t0=w["t0"], dt=w["dt"], nt=w["nt"], interp={"kind": "cubic"})
    got = [t.cpu().numpy() for t in bf.tables()]
    ref = oracle.synthetic_tables(p["ovectx"], p["nvectx"], p["t0tx"], p["alatx"], p["doftx"],
                                  p["ovecrx"], p["nvecrx"], p["alarx"], p["c0"], p["pnts"])
    names = ["tautx", "apodtx", "taurx", "apodrx"]
    for n, g, r in zip(names, got, ref):
        live = (ref[1] != 0) if n == "tautx" else np.ones_like(r, bool)
        assert np.array_equal(g[live], r[live]), f"{mode} {n}: {int((g[live] != r[live]).sum())} differ"
    # every branch of the geometry is exercised somewhere
    assert ref[1].min() == 0.0 or mode == "pw"
    assert ref[3].min() == 0.0


def test_receive_boundary_and_tiny_distances_bitwise():
    """The receive weight rule and delay at their edge cases, bit for bit
    against the oracle: for four receive elements (acceptance angles 0.32,
    0.5, 0.1 and 1.2 rad, tilted normals), points on arcs through both cone
    edges in steps of 3e-8 rad at 5, 17 and 33 mm, so that the float32
    cosine proj / mag lands on the element's threshold T and on every float
    a few ulps either side of it; and points at the element itself and
    1e-17 .. 4e-12 m away, whose squared distance lies below 2^-96 (the
    square root's rescaled branch) or just above it."""
    alas = np.float32([0.32175055, 0.5, 0.1, 1.2])
    phis = np.array([0.0, 0.2, -0.15, 0.05])
    xo = np.float32([-3e-3, -1e-3, 1e-3, 3e-3])
    ovecrx = np.stack([xo, np.zeros(4, np.float32)], 1).astype(np.float32)
    nvecrx = np.stack([np.sin(phis), np.cos(phis)], 1).astype(np.float32)
    pts = []
    for e in range(4):
        for r in (5e-3, 17e-3, 33e-3):
            for sgn in (1.0, -1.0):
                th = phis[e] + sgn * (float(alas[e]) + np.arange(-150, 151) * 3e-8)
                pts.append(np.stack([xo[e] + r * np.sin(th), r * np.cos(th)], 1))
        o = ovecrx[e].astype(np.float64)
        for off in (0.0, 1e-17, 3e-16, 2e-15, 1e-14, 4e-12):
            pts.append(np.array([[o[0] + off, o[1] + off], [o[0] - off, o[1] + 0.5 * off]]))
    p = dict(ovectx=np.zeros((1, 2), np.float32), nvectx=np.float32([[0.0, 1.0]]),
             t0tx=np.zeros(1, np.float32), alatx=np.zeros(1, np.float32),
             doftx=np.zeros(1, np.float32), ovecrx=ovecrx, nvecrx=nvecrx, alarx=alas, c0=1540.0,
             pnts=np.concatenate(pts).astype(np.float32))
    ref = oracle.synthetic_tables(p["ovectx"], p["nvectx"], p["t0tx"], p["alatx"], p["doftx"],
                                  p["ovecrx"], p["nvecrx"], p["alarx"], p["c0"], p["pnts"])
    # the cases are there: both weights for every element, and squared
    # distances below 2^-96
    assert all(0.0 in ref[3][e] and 1.0 in ref[3][e] for e in range(4))
    d = p["pnts"][None] - p["ovecrx"][:, None]
    assert int(((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) < 2.0 ** -96).sum()) >= 16
    bf = SyntheticBeamformer(**p, t0=0.0, dt=1.0 / sim.FS, nt=1024, interp={"kind": "cubic"})
    got = [t.cpu().numpy() for t in bf.tables()]
    for n, g, r in zip(["tautx", "apodtx", "taurx", "apodrx"], got, ref):
        bad = g.view(np.uint32) != r.view(np.uint32)
        assert not bad.any(), f"{n}: {int(bad.sum())} differ"


@pytest.mark.parametrize("mode", SYN_MODES)
@pytest.mark.parametrize("interp,kind,usf", KINDS, ids=[k[1] + str(k[2]) for k in KINDS])
def test_synthetic_das_matches_oracle(mode, interp, kind, usf):
    w = small_synthetic(mode=mode)
    p = w["params"]
    bf = SyntheticBeamformer(**p, t0=w["t0"], dt=w["dt"], nt=w["nt"], interp=interp)
    got = bf(w["rf"])
    S = numerics.make_S(8).astype(np.float32) if kind == "korder_cubic" else None
    ref = oracle.das_synthetic(w["rf"], w["t0"], w["dt"], p["ovectx"], p["nvectx"], p["t0tx"],
                               p["alatx"], p["doftx"], p["ovecrx"], p["nvecrx"], p["alarx"],
                               p["c0"], p["pnts"], kind=kind, usf=usf, S=S)
    assert_close_to_peak(got, ref, 1e-5, f"{mode} {kind}")


def test_synthetic_sumtypes_and_frames():
    w = small_synthetic(mode="pw", ntx=4, nrx=16)
    p = w["params"]
    rfs = np.stack([w["rf"], -0.25 * w["rf"]])
    for sumtype in ("none", "tx_only", "rx_only", "tx_and_rx"):
        bf = SyntheticBeamformer(**p, t0=w["t0"], dt=w["dt"], nt=w["nt"],
                                 interp={"kind": "cubic"}, sumtype=sumtype)
        got = bf(rfs)
        for f in range(2):
            ref = oracle.das_synthetic(rfs[f], w["t0"], w["dt"], p["ovectx"], p["nvectx"],
                                       p["t0tx"], p["alatx"], p["doftx"], p["ovecrx"],
                                       p["nvecrx"], p["alarx"], p["c0"], p["pnts"],
                                       kind="cubic", sumtype=sumtype)
            assert_close_to_peak(got[f], ref, 1e-5, f"{sumtype} frame {f}")


@pytest.mark.parametrize("sumtype", ["tx_and_rx", "none"])
def test_tabbed_processing_order_keeps_every_pixel(sumtype):
    """TabbedBeamformer groups pixels into tiles by a Morton code of two of
    their receive delays and keeps its tables in that order; its images (every
    output at its own pixel) equal the engine's on the tables in the pixels'
    own order (no processing order).  Sums agree to float rounding: a pixel's
    window may be staged in one grouping and gathered directly in the other."""
    import torch
    from pycbf_amd.gpu import _engine as E
    print(f"tabbed processing order, {sumtype}", flush=True)
    w = small_tabbed(seed=4, ntx=7, nrx=48, nt=1536, nx=48, nz=96)
    rf = np.stack([w["rf"], 0.7 * w["rf"]]).astype(np.float32)
    bf = _tabbed(w, {"kind": "cubic"}, sumtype)
    assert not torch.equal(bf._order.cpu(), torch.arange(bf.nop))
    got = np.asarray(bf(rf))
    dev = torch.device("cuda", 0)
    # (tautx, apodtx, taurx, apodrx) in the pixels' own order
    tabs = [torch.as_tensor(np.asarray(t, np.float32), device=dev).contiguous()
            for t in w["tabs"]]
    rfd = torch.as_tensor(rf, device=dev)
    info = bf._rfinfo()
    kind = E.INTERP["cubic"]
    need = E.lib.pycbf_das_workspace_bytes(ctypes.byref(info), 2, bf.nop, kind, 1, 0)
    ws = torch.empty(int(need), dtype=torch.uint8, device=dev)
    ref = torch.empty(bf._out_shape(2), dtype=torch.float32, device=dev)
    rc = E.lib.pycbf_das_tabbed(ctypes.byref(info), E.ptr(rfd), 2, *(E.ptr(t) for t in tabs),
                                None, bf.nop, kind, 1, None, float(bf.thresh),
                                E.SUMTYPE[sumtype], E.ptr(ref), E.ptr(ws), ws.numel(),
                                E.stream_of(dev))
    E.check(rc, "pycbf_das_tabbed")
    ref = ref.cpu().numpy().reshape(got.shape)
    assert np.isfinite(got).all()
    assert_close_to_peak(got, ref, 1e-6, f"processing order {sumtype}")
