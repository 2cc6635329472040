"""GPU Synthetic adjoint (pycbf_das_synthetic_bwd), whose kernels re-evaluate
the delays and weights of every (element, pixel) pair per tile in registers.

- Against pycbf_das_tabbed_bwd on the tables of the same geometry
  (pycbf_synthetic_tables): the channel-data, delay and apodisation gradients
  are bit for bit equal (the same float32 delays and weights, the same
  fixed-point sums), in every transmit mode, for every kind the adjoint
  supports and every sumtype.
- d loss / d c0 against FP64 central differences (h = 1e-3 m/s) of the same
  pipeline.  Every delay of the geometry is a path length over c (the
  transmit time intercepts too, which pycbf's helpers compute as distances
  over c), so the FP64 side rescales the kernels' float32 delays around c0:
  tau(c) = tau c0 / c.
- Headline transducer (75 plane waves x 128 receive elements x 4096 samples,
  512 x 512 grid): transpose identity <A rf, G> = <rf, A^T G> of the
  channel-data gradient through das_synthetic, and the c0 gradient against
  FP64 central differences on a 48 x 48 sub-grid.
"""
import ctypes

import numpy as np
import pytest

from pycbf_amd import sim
from tests.helpers import float32_inputs, small_synthetic

pytestmark = pytest.mark.gpu


def _bf(w, interp, sumtype="tx_and_rx", thresh=None):
    from pycbf_amd.gpu import SyntheticBeamformer
    kw = {} if thresh is None else {"thresh": thresh}
    return SyntheticBeamformer(**w["params"], t0=w["t0"], dt=w["dt"], nt=w["nt"], interp=interp,
                               sumtype=sumtype, **kw)


def _grads(bf, rf, gout, synthetic):
    """(grf, gtautx, gapodtx, gtaurx, gapodrx[, gc0]) of the synthetic adjoint,
    or of the tabbed adjoint on bf.tables()."""
    import torch
    from pycbf_amd.gpu import _engine as E
    nfr = rf.shape[0]
    info = bf._rfinfo()
    kind = E.INTERP[bf.interp["kind"]]
    need = E.lib.pycbf_das_bwd_workspace_bytes(ctypes.byref(info), nfr, bf.nop, kind,
                                               bf._interp_param())
    ws = torch.empty(int(need), dtype=torch.uint8, device=rf.device)
    grf = torch.empty_like(rf)
    g = [torch.empty((n, bf.nop), dtype=torch.float32, device=rf.device)
         for n in (bf.ntx, bf.ntx, bf.nrx, bf.nrx)]
    st = E.stream_of(rf.device)
    if synthetic:
        gc0 = torch.empty(1, dtype=torch.float32, device=rf.device)
        rc = E.lib.pycbf_das_synthetic_bwd(
            ctypes.byref(info), E.ptr(rf), nfr, E.ptr(bf._txpack), E.ptr(bf._rxpack),
            float(bf.c0), E.ptr(bf._pnts), bf.nop, kind, bf._interp_param(), E.ptr(bf._S_dev()),
            float(bf.thresh), E.SUMTYPE[bf.sumtype], E.ptr(gout), E.ptr(grf),
            *(E.ptr(t) for t in g), E.ptr(gc0), E.ptr(ws), ws.numel(), st)
        E.check(rc, "pycbf_das_synthetic_bwd")
        return (grf, *g, gc0)
    tabs = bf.tables()
    rc = E.lib.pycbf_das_tabbed_bwd(
        ctypes.byref(info), E.ptr(rf), nfr, *(E.ptr(t) for t in tabs), None, None, bf.nop, kind,
        bf._interp_param(), E.ptr(bf._S_dev()), float(bf.thresh), E.SUMTYPE[bf.sumtype],
        E.ptr(gout), E.ptr(grf), *(E.ptr(t) for t in g), E.ptr(ws), ws.numel(), st)
    E.check(rc, "pycbf_das_tabbed_bwd")
    return (grf, *g)


CASES = [("pw", {"kind": "cubic"}, "tx_and_rx"), ("pw", {"kind": "cubic"}, "none"),
         ("pw", {"kind": "cubic"}, "tx_only"), ("pw", {"kind": "cubic"}, "rx_only"),
         ("pw", {"kind": "linear"}, "tx_and_rx"), ("pw", {"kind": "akima"}, "tx_and_rx"),
         ("pw", {"kind": "makima"}, "tx_only"),
         ("pw", {"kind": "korder_cubic", "k": 8}, "tx_and_rx"),
         ("pw", {"kind": "nearest", "usf": 2}, "tx_and_rx"),
         ("pw", {"kind": "nearest", "usf": 1}, "rx_only"),
         ("foc", {"kind": "cubic"}, "tx_and_rx"), ("div", {"kind": "linear"}, "tx_and_rx"),
         ("dof", {"kind": "cubic"}, "none")]


@pytest.mark.parametrize("mode,interp,sumtype", CASES)
def test_synthetic_adjoint_equals_tabbed_adjoint_on_its_tables(mode, interp, sumtype):
    import torch
    print(f"synthetic adjoint vs tabbed: {mode} {interp} {sumtype}", flush=True)
    w = small_synthetic(seed=11, mode=mode)
    bf = _bf(w, interp, sumtype)
    rf = torch.as_tensor(np.stack([w["rf"], 0.5 * w["rf"][::-1].copy()]).astype(np.float32),
                         device="cuda")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(21)
    gout = torch.randn(bf._out_shape(2), generator=gen, device="cuda")
    syn = _grads(bf, rf, gout, True)
    tab = _grads(bf, rf, gout, False)
    for name, a, b in zip(("grf", "gtautx", "gapodtx", "gtaurx", "gapodrx"), syn, tab):
        if not (interp["kind"] == "nearest" and name in ("gtautx", "gtaurx")):
            assert float(b.abs().max()) > 0, name  # nearest: zero delay gradient
        assert torch.equal(a, b), (name, float((a - b).abs().max()), float(b.abs().max()))


def _c0_fd(bf, rf, G, t0, dt, h):
    """FP64 central difference of sum(G * image) in c0 around the kernels'
    float32 delays (cubic, sumtype of bf)."""
    from oracle import diff64
    ttx, atx, trx, arx = (t.double() for t in bf.tables())
    c0 = float(np.float32(bf.c0))
    t32, dt32 = float32_inputs(t0, dt)
    rf64 = rf.double()
    y2 = diff64.natural_y2(rf64, dt32)
    G64 = G.double()

    def L(c):
        s = c0 / c
        im = diff64.das_tabbed(rf64, ttx * s, atx, trx * s, arx, t32, dt32, "cubic",
                               bf.sumtype, float(bf.thresh), x32=False, y2=y2)
        return float((im * G64).sum())

    return (L(c0 + h) - L(c0 - h)) / (2 * h)


@pytest.mark.parametrize("mode", ["pw", "foc", "div", "dof"])
def test_c0_gradient_matches_fp64_central_differences(mode):
    import torch
    from pycbf_amd.gpu.autograd import das_synthetic
    print(f"c0 gradient vs FP64 central differences: {mode}", flush=True)
    w = small_synthetic(seed=5, mode=mode)
    bf = _bf(w, {"kind": "cubic"}, "tx_and_rx")
    rf = torch.as_tensor(w["rf"][None].astype(np.float32), device="cuda")
    c0 = torch.tensor(float(bf.c0), dtype=torch.float32, requires_grad=True)
    out = das_synthetic(rf, bf, c0)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(6)
    G = torch.randn(out.shape, generator=gen, device="cuda")
    (gc,) = torch.autograd.grad((out * G).sum(), [c0])
    fd = _c0_fd(bf, rf, G, w["t0"], w["dt"], 1e-3)
    print(f"  d loss / d c0: adjoint {float(gc):.6e}, FP64 central difference {fd:.6e}",
          flush=True)
    assert fd != 0.0
    assert abs(float(gc) - fd) <= 1e-3 * abs(fd), (float(gc), fd)


NTX, NRX, NT = 75, 128, 4096


def _headline_bf(nx, nz, interp=None):
    from pycbf_amd.gpu import SyntheticBeamformer
    xele = sim.linear_array(NRX)
    steers = np.radians(np.linspace(-16, 16, NTX))
    _, _, pnts = sim.grid(nx, nz, (-19e-3, 19e-3), (2e-3, 50e-3))
    params = sim.pw_synthetic(steers, xele, pnts)
    bf = SyntheticBeamformer(**params, t0=0.0, dt=1.0 / sim.FS, nt=NT,
                             interp=interp or {"kind": "cubic"}, sumtype="tx_and_rx")
    return bf, steers, xele


def _headline_rf(steers, xele, seed):
    import torch
    xs, zs, amp = sim.scatterers(50, 2000, (-19e-3, 19e-3), (2e-3, 50e-3), seed)
    return sim.simulate_rf_torch(steers, xele, NT, xs, zs, amp, torch.device("cuda", 0),
                                 seed=seed)[None].contiguous()


def test_headline_channel_gradient_transpose_identity():
    """512 x 512, one frame: <A rf, G> = <rf, A^T G> within 1e-5, and the
    synthetic adjoint's channel-data gradient equal to the tabbed adjoint's."""
    import torch
    from pycbf_amd.gpu.autograd import das_synthetic
    print("headline synthetic adjoint: transpose identity, 512x512", flush=True)
    bf, steers, xele = _headline_bf(512, 512)
    rf = _headline_rf(steers, xele, 77).requires_grad_()
    out = das_synthetic(rf, bf)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(78)
    G = torch.randn(out.shape, generator=gen, device="cuda")
    (grf,) = torch.autograd.grad((out * G).sum(), [rf])
    lhs = float((out.detach().double() * G.double()).sum())
    rhs = float((rf.detach().double() * grf.double()).sum())
    print(f"  <A rf, G> {lhs:.9e}  <rf, A^T G> {rhs:.9e}", flush=True)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs), (lhs, rhs)
    tab = _grads(bf, rf.detach(), G, False)[0]
    assert torch.equal(grf, tab), float((grf - tab).abs().max())


def test_headline_c0_gradient_matches_fp64_central_differences():
    import torch
    from pycbf_amd.gpu.autograd import das_synthetic
    print("headline synthetic adjoint: c0 gradient, 48x48 sub-grid", flush=True)
    bf, steers, xele = _headline_bf(48, 48)
    rf = _headline_rf(steers, xele, 79)
    c0 = torch.tensor(float(bf.c0), dtype=torch.float32, requires_grad=True)
    out = das_synthetic(rf, bf, c0)
    # the loss of a focus search: mean image energy
    loss = (out.double() ** 2).mean()
    (gc,) = torch.autograd.grad(loss, [c0])
    G = (2.0 * out.detach().double() / out.numel()).float()
    fd = _c0_fd(bf, rf, G, 0.0, 1.0 / sim.FS, 1e-3)
    print(f"  d energy / d c0: adjoint {float(gc):.6e}, FP64 central difference of the "
          f"linearised loss {fd:.6e}", flush=True)
    assert abs(float(gc) - fd) <= 1e-3 * abs(fd), (float(gc), fd)
