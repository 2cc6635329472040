"""GPU adjoint of the Tabbed DAS (pycbf_das_tabbed_bwd through the autograd
Function) against the FP64 differentiable restatement (oracle/diff64.py).

Bars: every gradient within 1e-4 of the reference gradient's peak, and the
forward values within 1e-4 of their peak (float32 kernels, which keep the
reference C's separately rounded interpolation weights a and b, against
float64 autograd with exact weights on the same float32 delay sums); delay
gradients within 1e-3 relative of FP64 central differences; the channel-data
adjoint satisfies <A x, y> = <x, A^T y> to 1e-5 relative at the
differentiable workload's full size (31 tx x 128 rx x 4096 samples, 256 x 256
pixels)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SUMTYPES = ["none", "tx_only", "rx_only", "tx_and_rx"]


def _case(seed, ntx=3, nrx=8, nt=160, np_=300, nfr=1, discrete=False):
    rng = np.random.default_rng(seed)
    t0, dt = 1e-6, 2.5e-8
    span = dt * (nt - 1)
    rf = rng.standard_normal((nfr, ntx, nrx, nt)).astype(np.float32)
    tautx = rng.uniform(t0 - 0.05 * span, t0 + 0.5 * span, (ntx, np_)).astype(np.float32)
    taurx = rng.uniform(0.0, 0.53 * span, (nrx, np_)).astype(np.float32)
    if discrete:
        apodtx = rng.choice([0.0, 0.005, 0.2, 1.0, -0.7], (ntx, np_)).astype(np.float32)
        apodrx = rng.choice([0.003, 0.5, 1.0, -0.25], (nrx, np_)).astype(np.float32)
    else:
        apodtx = rng.uniform(-1, 1, (ntx, np_)).astype(np.float32)
        apodrx = rng.uniform(-1, 1, (nrx, np_)).astype(np.float32)
    return rf, tautx, apodtx, taurx, apodrx, t0, dt


KORDER_K = 8


def _S64():
    """The k-order operator as the GPU uses it (float32), in float64."""
    from oracle import numerics
    return numerics.make_S(KORDER_K).astype(np.float32).astype(np.float64)


def _grads_gpu(arrs, t0, dt, kind, sumtype, thresh, gout):
    import torch
    from pycbf_amd.gpu.autograd import das_tabbed
    t = [torch.as_tensor(a, device="cuda").requires_grad_() for a in arrs]
    out = das_tabbed(*t, t0, dt, kind, sumtype, thresh,
                     k=KORDER_K if kind == "korder_cubic" else None)
    g = torch.as_tensor(gout, device="cuda", dtype=torch.float32)
    (out * g).sum().backward()
    return out.detach().cpu().numpy(), [x.grad.cpu().numpy() for x in t]


def _grads_ref(arrs, t0, dt, kind, sumtype, thresh, gout, x32=True):
    import torch
    from oracle import diff64
    t = [torch.as_tensor(a, dtype=torch.float64).requires_grad_() for a in arrs]
    out = diff64.das_tabbed(*t, t0, dt, kind, sumtype, thresh, x32=x32,
                            S=_S64() if kind == "korder_cubic" else None)
    (out * torch.as_tensor(gout, dtype=torch.float64)).sum().backward()
    return out.detach().numpy(), [np.zeros(x.shape) if x.grad is None else x.grad.numpy()
                                  for x in t]


NAMES = ["rf", "tautx", "apodtx", "taurx", "apodrx"]


@pytest.mark.parametrize("sumtype", SUMTYPES)
@pytest.mark.parametrize("kind", ["nearest", "linear", "cubic", "korder_cubic", "akima", "makima"])
def test_adjoint_matches_fp64_autograd(kind, sumtype):
    print(f"adjoint {kind} {sumtype}", flush=True)
    rf, tautx, apodtx, taurx, apodrx, t0, dt = _case(11, nfr=2)
    arrs = [rf, tautx, apodtx, taurx, apodrx]
    ref_out = _grads_ref(arrs, t0, dt, kind, sumtype, 0.0,
                         np.zeros(1))[0]  # shape probe
    gout = np.random.default_rng(12).standard_normal(ref_out.shape).astype(np.float32)
    out, g = _grads_gpu(arrs, t0, dt, kind, sumtype, 0.0, gout)
    ref_out, gr = _grads_ref(arrs, t0, dt, kind, sumtype, 0.0, gout)
    assert np.max(np.abs(out - ref_out)) <= 1e-4 * np.max(np.abs(ref_out))
    for name, a, b in zip(NAMES, g, gr):
        if kind == "nearest" and name.startswith("tau"):
            assert np.all(a == 0.0)
            continue
        err = np.max(np.abs(a - b))
        assert err <= 1e-4 * np.max(np.abs(b)), f"{kind} {sumtype} d{name}: {err:.3e}"


@pytest.mark.parametrize("kind", ["linear", "cubic", "korder_cubic", "makima"])
def test_adjoint_threshold_matches_forward_gate(kind):
    rf, tautx, apodtx, taurx, apodrx, t0, dt = _case(13, discrete=True)
    arrs = [rf, tautx, apodtx, taurx, apodrx]
    gout = np.random.default_rng(14).standard_normal((1, 300)).astype(np.float32)
    out, g = _grads_gpu(arrs, t0, dt, kind, "tx_and_rx", 1e-2, gout)
    ref_out, gr = _grads_ref(arrs, t0, dt, kind, "tx_and_rx", 1e-2, gout)
    assert np.max(np.abs(out - ref_out)) <= 1e-4 * np.max(np.abs(ref_out))
    for name, a, b in zip(NAMES, g, gr):
        assert np.max(np.abs(a - b)) <= 1e-4 * np.max(np.abs(b)), name


def test_delay_gradient_matches_fp64_central_differences():
    """d loss / d tau from the GPU adjoint against central differences of the
    all-float64 forward (h = 1e-13 s, cubic, 16 seeded entries per table)."""
    import torch
    from oracle import diff64
    rf, tautx, apodtx, taurx, apodrx, t0, dt = _case(15)
    arrs = [rf, tautx, apodtx, taurx, apodrx]
    gout = np.random.default_rng(16).standard_normal((1, 300)).astype(np.float32)
    _, g = _grads_gpu(arrs, t0, dt, "cubic", "tx_and_rx", 0.0, gout)
    base = [torch.as_tensor(a, dtype=torch.float64) for a in arrs]
    G = torch.as_tensor(gout, dtype=torch.float64)

    def L(vals):
        return float((diff64.das_tabbed(*vals, t0, dt, "cubic", x32=False) * G).sum())

    rng = np.random.default_rng(17)
    h = 1e-13
    for ai in (1, 3):
        peak = np.max(np.abs(g[ai]))
        for _ in range(16):
            idx = tuple(int(rng.integers(0, s)) for s in arrs[ai].shape)
            plus = [b.clone() for b in base]
            minus = [b.clone() for b in base]
            plus[ai][idx] += h
            minus[ai][idx] -= h
            fd = (L(plus) - L(minus)) / (2 * h)
            assert abs(g[ai][idx] - fd) <= 1e-3 * max(abs(fd), 1e-2 * peak), (ai, idx, g[ai][idx], fd)


@pytest.mark.parametrize("nt", [1, 2, 3, 31, 33, 64, 100, 517, 2047, 2049, 4096, 4100])
def test_cubic_rf_adjoint_trace_lengths(nt):
    """The transposed spline solve for trace lengths around the segment and
    tile edges, over 70 channels: up to 4096 samples one wave per channel
    (lane segments from checked carries; the sparse cotangent bins of 120
    pixels leave long zero runs that force recomputed carries), beyond that
    one lane per channel.  Channel-data gradients against FP64 autograd."""
    print(f"cubic rf adjoint nt={nt}", flush=True)
    arrs = list(_case(40 + nt, ntx=5, nrx=14, nt=nt, np_=120)[:5])
    t0, dt = 1e-6, 2.5e-8
    rng = np.random.default_rng(nt)
    gout = rng.standard_normal(120).astype(np.float32)
    _, g = _grads_gpu(arrs, t0, dt, "cubic", "tx_and_rx", 0.0, gout)
    _, r = _grads_ref(arrs, t0, dt, "cubic", "tx_and_rx", 0.0, gout)
    peak = max(float(np.abs(r[0]).max()), 1e-30)
    # float32 sample positions x <= t0 + nt*dt carry a rounding of 2^-23 x,
    # a relative error of 2^-23 x / dt in the interpolation weights that the
    # FP64 reference does not have (it reaches 1e-4 near 850 samples here)
    tol = max(1e-4, 2.0 ** -23 * (t0 / dt + nt))
    assert float(np.abs(g[0] - r[0]).max()) <= tol * peak
    if nt > 64:
        # the segment sweeps agree with the one-lane-per-channel sweeps; the
        # cotangent bins come from float atomics in varying order, so two runs
        # differ in the last bits whichever solve follows
        os.environ["PYCBF_SPLINE_SEG"] = "0"
        try:
            _, g1 = _grads_gpu(arrs, t0, dt, "cubic", "tx_and_rx", 0.0, gout)
        finally:
            del os.environ["PYCBF_SPLINE_SEG"]
        assert float(np.abs(g[0] - g1[0]).max()) <= 1e-6 * peak


@pytest.mark.parametrize("kind", ["nearest", "linear", "cubic", "korder_cubic", "akima", "makima"])
def test_rf_adjoint_transpose_identity_full_size(kind):
    """<A rf, G> = <rf, A^T G> for the 31 x 128 x 4096, 256 x 256 workload."""
    import torch
    from pycbf_amd import sim
    from pycbf_amd.gpu.autograd import das_tabbed
    print(f"transpose identity {kind}", flush=True)
    xele = sim.linear_array(128)
    steers = np.radians(np.linspace(-15, 15, 31))
    _, _, pnts = sim.grid(256, 256, (-19e-3, 19e-3), (5e-3, 45e-3))
    tabs = [torch.as_tensor(a, device="cuda") for a in sim.pw_tabs(steers, xele, pnts)]
    g = torch.Generator(device="cuda")
    g.manual_seed(18)
    rf = torch.randn((31, 128, 4096), generator=g, device="cuda", requires_grad=True)
    G = torch.randn((256 * 256,), generator=g, device="cuda")
    out = das_tabbed(rf, *tabs, 0.0, 1.0 / 40e6, kind, "tx_and_rx", 0.0,
                     k=KORDER_K if kind == "korder_cubic" else None)
    (grf,) = torch.autograd.grad((out * G).sum(), [rf])
    lhs = float((out.double() * G.double()).sum())
    rhs = float((rf.detach().double() * grf.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + 1e-30), (lhs, rhs)
    assert float(grf.abs().max()) > 0


@pytest.mark.parametrize("kind,nt,k", [("korder_cubic", 8, 8), ("korder_cubic", 9, 8),
                                        ("korder_cubic", 160, 4), ("korder_cubic", 300, 16),
                                        ("korder_cubic", 4097, 8), ("korder_cubic", 4100, 8),
                                        ("akima", 5, None), ("makima", 6, None),
                                        ("makima", 4097, None), ("akima", 4100, None)])
def test_hermite_adjoint_window_edges_and_long_traces(kind, nt, k):
    """korder_cubic / akima / makima adjoints against FP64 autograd for
    windows and end slopes clamped at both ends (short traces) and for traces
    past one workgroup's LDS bins (the per-tile path with global atomics
    beyond 4097 samples)."""
    import torch
    from oracle import diff64, numerics
    from pycbf_amd.gpu.autograd import das_tabbed
    print(f"{kind} adjoint nt={nt} k={k}", flush=True)
    rf, tautx, apodtx, taurx, apodrx, t0, dt = _case(31 + nt, ntx=2, nrx=5, nt=nt, np_=260)
    arrs = [rf, tautx, apodtx, taurx, apodrx]
    gout = np.random.default_rng(32).standard_normal((1, 260)).astype(np.float32)
    t = [torch.as_tensor(a, device="cuda").requires_grad_() for a in arrs]
    out = das_tabbed(*t, t0, dt, kind, "tx_and_rx", 0.0, k=k)
    (out * torch.as_tensor(gout, device="cuda")).sum().backward()
    S = numerics.make_S(k).astype(np.float32).astype(np.float64) if k else None
    r = [torch.as_tensor(a, dtype=torch.float64).requires_grad_() for a in arrs]
    ref = diff64.das_tabbed(*r, t0, dt, kind, "tx_and_rx", 0.0, S=S)
    (ref * torch.as_tensor(gout, dtype=torch.float64)).sum().backward()
    o, ro = out.detach().cpu().numpy(), ref.detach().numpy()
    # float32 sample positions x <= t0 + nt dt carry a relative weight error of
    # 2^-23 (t0/dt + nt) that the float64 reference does not have
    tol = max(1e-4, 2.0 ** -23 * (t0 / dt + nt))
    assert np.max(np.abs(o - ro)) <= tol * np.max(np.abs(ro))
    for name, a, b in zip(NAMES, t, r):
        ga, gb = a.grad.cpu().numpy(), b.grad.numpy()
        assert np.max(np.abs(ga - gb)) <= tol * np.max(np.abs(gb)), name


@pytest.mark.parametrize("usf,nt,sumtype", [(2, 160, "tx_and_rx"), (3, 97, "none"),
                                            (4, 160, "rx_only"), (2, 2100, "tx_and_rx"),
                                            (4, 1100, "tx_only")])
def test_nearest_upsampled_adjoint_matches_fp64(usf, nt, sumtype):
    """nearest with usf > 1: the adjoint on the upsampled grid, then the
    transposed spline upsampling and the transposed spline solve, against
    FP64 autograd through the float64 natural spline.  Short traces take the
    per-channel bins (usf (nt-1) + 1 <= 4097), long ones the per-tile path."""
    import torch
    from oracle import diff64
    from pycbf_amd.gpu.autograd import das_tabbed
    print(f"nearest usf={usf} nt={nt} {sumtype}", flush=True)
    rf, tautx, apodtx, taurx, apodrx, t0, dt = _case(40 + usf, ntx=3, nrx=4, nt=nt, np_=300)
    arrs = [rf, tautx, apodtx, taurx, apodrx]
    r = [torch.as_tensor(a, dtype=torch.float64).requires_grad_() for a in arrs]
    ref = diff64.das_tabbed(*r, t0, dt, "nearest", sumtype, 0.0, usf=usf)
    gout = np.random.default_rng(41).standard_normal(tuple(ref.shape)).astype(np.float32)
    (ref * torch.as_tensor(gout, dtype=torch.float64)).sum().backward()
    t = [torch.as_tensor(a, device="cuda").requires_grad_() for a in arrs]
    out = das_tabbed(*t, t0, dt, "nearest", sumtype, 0.0, usf=usf)
    (out * torch.as_tensor(gout, device="cuda")).sum().backward()
    o, ro = out.detach().cpu().numpy(), ref.detach().numpy()
    # the upsampling weights come from float32 knot times t0 + dt j: a relative
    # error of 2^-23 (t0/dt + nt) that the float64 reference does not have
    tol = max(1e-4, 2.0 ** -23 * (t0 / dt + nt))
    assert np.max(np.abs(o - ro)) <= tol * np.max(np.abs(ro))
    for name, a, b in zip(NAMES, t, r):
        ga = a.grad.cpu().numpy()
        if name.startswith("tau"):
            assert np.all(ga == 0.0)
            continue
        gb = b.grad.numpy()
        assert np.max(np.abs(ga - gb)) <= tol * np.max(np.abs(gb)), name


def test_nearest_upsampled_transpose_identity_full_size():
    """<A rf, G> = <rf, A^T G> for nearest with usf = 2 at 31 x 128 x 4096,
    256 x 256 (8191 upsampled samples per channel: the per-tile path)."""
    import torch
    from pycbf_amd import sim
    from pycbf_amd.gpu.autograd import das_tabbed
    print("transpose identity nearest usf=2", flush=True)
    xele = sim.linear_array(128)
    steers = np.radians(np.linspace(-15, 15, 31))
    _, _, pnts = sim.grid(256, 256, (-19e-3, 19e-3), (5e-3, 45e-3))
    tabs = [torch.as_tensor(a, device="cuda") for a in sim.pw_tabs(steers, xele, pnts)]
    g = torch.Generator(device="cuda")
    g.manual_seed(19)
    rf = torch.randn((31, 128, 4096), generator=g, device="cuda", requires_grad=True)
    G = torch.randn((256 * 256,), generator=g, device="cuda")
    out = das_tabbed(rf, *tabs, 0.0, 1.0 / 40e6, "nearest", "tx_and_rx", 0.0, usf=2)
    (grf,) = torch.autograd.grad((out * G).sum(), [rf])
    lhs = float((out.detach().double() * G.double()).sum())
    rhs = float((rf.detach().double() * grf.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + 1e-30), (lhs, rhs)
    assert float(grf.abs().max()) > 0


@pytest.mark.parametrize("mode,interp", [("pw", {"kind": "cubic"}), ("foc", {"kind": "linear"}),
                                         ("div", {"kind": "korder_cubic", "k": 8}),
                                         ("dof", {"kind": "nearest", "usf": 2})])
def test_das_beamformer_differentiates_synthetic_geometry_forward(mode, interp):
    """das_beamformer on a SyntheticBeamformer: the forward equals the
    beamformer's own (in-register geometry) image, and the channel-data
    gradient satisfies <A rf, G> = <rf, A^T G>."""
    import torch
    from pycbf_amd.gpu import SyntheticBeamformer
    from pycbf_amd.gpu.autograd import das_beamformer
    from tests.helpers import small_synthetic
    print(f"das_beamformer {mode} {interp}", flush=True)
    w = small_synthetic(seed=3, mode=mode)
    bf = SyntheticBeamformer(**w["params"], t0=w["t0"], dt=w["dt"], nt=w["nt"], interp=interp,
                             sumtype="tx_and_rx")
    ref = np.asarray(bf(w["rf"]))
    rf = torch.as_tensor(w["rf"].astype(np.float32), device="cuda").requires_grad_()
    out = das_beamformer(rf, bf)
    got = out.detach().cpu().numpy().reshape(ref.shape)
    peak = float(np.abs(ref).max())
    assert peak > 0
    assert float(np.abs(got - ref).max()) <= 1e-6 * peak
    G = torch.as_tensor(np.random.default_rng(5).standard_normal(tuple(out.shape)),
                        dtype=torch.float32, device="cuda")
    (grf,) = torch.autograd.grad((out * G).sum(), [rf])
    lhs = float((out.detach().double() * G.double()).sum())
    rhs = float((rf.detach().double() * grf.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + 1e-30), (lhs, rhs)
