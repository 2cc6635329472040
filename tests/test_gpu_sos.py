"""GPU speed-of-sound path: straight-ray delay kernels and their adjoint
against the FP64 restatement (oracle/diff64.py), and the end-to-end gradient
of the focus loss with respect to the sound-speed map (ray adjoint + DAS
adjoint) against FP64 central differences of the same pipeline
(h = 1e-3 m/s at 32 seeded map entries).

Bars: delays within 1e-6 relative; map gradients of the ray adjoint within
1e-4 of their peak; end-to-end map gradient within 1e-3 relative of the FP64
central difference (or 1e-2 of the gradient peak where the entry's gradient is
near zero)."""
import numpy as np
import pytest

from tests.helpers import float32_inputs

pytestmark = pytest.mark.gpu


def _problem(seed=3, ntx=8, nrx=32, nx=48, nz=48, nt=2048, nzm=16, nxm=16):
    import torch
    from pycbf_amd import sim
    from pycbf_amd.gpu.sos import SlownessGrid, SoSProblem, ray_delays
    xlim, zlim = (-8e-3, 8e-3), (5e-3, 25e-3)
    xele = sim.linear_array(nrx)
    steers = np.radians(np.linspace(-15, 15, ntx))
    _, _, pnts = sim.grid(nx, nz, xlim, zlim)
    geo = SlownessGrid.covering(xlim, zlim, nxm, nzm)
    dev = torch.device("cuda", 0)
    f = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=dev)  # noqa: E731
    ctrue = sim.sos_true_map(nzm, nxm, xlim, zlim, seed=seed, z_layer=10e-3)
    xs, zs, amp = sim.scatterers(30, 3000, (xlim[0] * 0.9, xlim[1] * 0.9), (zlim[0] + 1e-3, zlim[1]),
                                 seed)
    sc = f(np.stack([xs, zs], 1))
    ttx, trx = ray_delays(1.0 / f(ctrue), geo, f(steers), f(sim.pw_origin(steers, xele)), f(xele),
                          sc, 1540.0, 64)
    rf = sim.simulate_rf_delays_torch(ttx, trx, amp, nt, dev, seed=seed)
    prob = SoSProblem(rf=rf, steer=f(steers), xorig=f(sim.pw_origin(steers, xele)), xele=f(xele),
                      pvec=f(pnts), gridshape=(nx, nz), geo=geo, t0=0.0, dt=1.0 / sim.FS)
    return prob, ctrue


def test_ray_delays_match_fp64():
    import torch
    from oracle import diff64
    from pycbf_amd.gpu.sos import ray_delays
    prob, ctrue = _problem()
    slow = 1.0 / torch.as_tensor(ctrue, device="cuda")
    ttx, trx = ray_delays(slow, prob.geo, prob.steer, prob.xorig, prob.xele, prob.pvec, 1540.0, 64)
    rtx, rrx = diff64.ray_delays(slow.double(), prob.geo, prob.steer.double(), prob.xorig.double(),
                                 prob.xele.double(), prob.pvec.double(), 1540.0, 64)
    for a, b in ((ttx, rtx), (trx, rrx)):
        err = float((a.double() - b).abs().max() / b.abs().max())
        assert err <= 1e-6, err


def test_ray_delay_residuals_place_the_delay_beyond_float32():
    """Delay + residual (pycbf_ray_delays' float32 residual outputs) agree with
    the FP64 restatement far below float32 rounding of the delay."""
    import torch
    from oracle import diff64
    from pycbf_amd.gpu.sos import ray_delays
    prob, ctrue = _problem()
    slow = 1.0 / torch.as_tensor(ctrue, device="cuda")
    ttx, trx, tlo, rlo = ray_delays(slow, prob.geo, prob.steer, prob.xorig, prob.xele, prob.pvec,
                                    1540.0, 64, residuals=True)
    rtx, rrx = diff64.ray_delays(slow.double(), prob.geo, prob.steer.double(), prob.xorig.double(),
                                 prob.xele.double(), prob.pvec.double(), 1540.0, 64)
    for hi, lo, ref in ((ttx, tlo, rtx), (trx, rlo, rrx)):
        e_hi = float((hi.double() - ref).abs().max() / ref.abs().max())
        e_lo = float((hi.double() + lo.double() - ref).abs().max() / ref.abs().max())
        print(f"ray delays: float32 {e_hi:.2e}, with residual {e_lo:.2e} (relative)", flush=True)
        assert e_lo <= 1e-8 and e_lo < 0.25 * e_hi


def test_residual_das_and_delay_gradients_match_fp64():
    """pycbf_das_tabbed_residual (samples at delay + residual in double) and the
    delay gradients of pycbf_das_tabbed_bwd with residuals against the FP64
    restatement evaluated at the same delays in float64 (x32=False)."""
    import torch
    from oracle import diff64
    from pycbf_amd.gpu.autograd import das_tabbed
    from pycbf_amd.gpu.sos import ray_delays
    prob, ctrue = _problem()
    slow = 1.0 / torch.as_tensor(ctrue, device="cuda")
    ttx, trx, tlo, rlo = ray_delays(slow, prob.geo, prob.steer, prob.xorig, prob.xele, prob.pvec,
                                    1540.0, 64, residuals=True)
    ttx, trx = ttx.clone().requires_grad_(), trx.clone().requires_grad_()
    ones_t, ones_r = torch.ones_like(ttx), torch.ones_like(trx)
    im = das_tabbed(prob.rf, ttx, ones_t, trx, ones_r, prob.t0, prob.dt, "cubic", "tx_only", 0.0,
                    tautx_lo=tlo, taurx_lo=rlo)
    g = torch.Generator(device="cuda")
    g.manual_seed(8)
    G = torch.randn(im.shape, generator=g, device="cuda")
    gt, gr = torch.autograd.grad((im * G).sum(), [ttx, trx])
    t32, dt32 = float32_inputs(prob.t0, prob.dt)
    t64 = (ttx.detach().double() + tlo.double()).requires_grad_()
    r64 = (trx.detach().double() + rlo.double()).requires_grad_()
    im64 = diff64.das_tabbed(prob.rf.double()[None], t64, ones_t.double(), r64, ones_r.double(),
                             t32, dt32, "cubic", "tx_only", 0.0, x32=False)[0]
    e_im = float((im.double() - im64).abs().max() / im64.abs().max())
    gt64, gr64 = torch.autograd.grad((im64 * G.double()).sum(), [t64, r64])
    e_t = float((gt.double() - gt64).abs().max() / gt64.abs().max())
    e_r = float((gr.double() - gr64).abs().max() / gr64.abs().max())
    print(f"residual DAS vs FP64: image {e_im:.2e}, tx delay grad {e_t:.2e}, rx {e_r:.2e}",
          flush=True)
    assert e_im <= 2e-6 and e_t <= 1e-4 and e_r <= 1e-4


def test_ray_adjoint_matches_fp64_autograd_and_transpose():
    import torch
    from oracle import diff64
    from pycbf_amd.gpu.sos import ray_delays
    prob, ctrue = _problem()
    slow = (1.0 / torch.as_tensor(ctrue, device="cuda")).requires_grad_()
    ttx, trx = ray_delays(slow, prob.geo, prob.steer, prob.xorig, prob.xele, prob.pvec, 1540.0, 64)
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    gt = torch.randn(ttx.shape, generator=g, device="cuda")
    gr = torch.randn(trx.shape, generator=g, device="cuda")
    (gs,) = torch.autograd.grad((ttx * gt).sum() + (trx * gr).sum(), [slow])
    s64 = slow.detach().double().requires_grad_()
    a, b = diff64.ray_delays(s64, prob.geo, prob.steer.double(), prob.xorig.double(),
                             prob.xele.double(), prob.pvec.double(), 1540.0, 64)
    (g64,) = torch.autograd.grad((a * gt.double()).sum() + (b * gr.double()).sum(), [s64])
    assert float((gs.double() - g64).abs().max()) <= 1e-4 * float(g64.abs().max())
    # the delays are affine in the slowness: <R s - R 0, g> = <s, R^T g>
    z0 = torch.zeros_like(slow)
    t0x, t0r = ray_delays(z0, prob.geo, prob.steer, prob.xorig, prob.xele, prob.pvec, 1540.0, 64)
    lhs = float(((ttx - t0x).double() * gt.double()).sum() + ((trx - t0r).double() * gr.double()).sum())
    rhs = float((slow.detach().double() * gs.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_sos_gradient_matches_fp64_central_differences():
    """d loss / d c at 32 seeded map entries: GPU adjoint chain (float32)
    against central differences (h = 1e-3 m/s) of the float64 pipeline."""
    import torch
    from oracle import diff64
    from pycbf_amd.gpu.sos import focus_loss
    prob, ctrue = _problem()
    c0map = torch.full((prob.geo.nzm, prob.geo.nxm), 1540.0, device="cuda")
    with torch.no_grad():
        e0 = float((prob.images(c0map).sum(0) ** 2).mean())
    c = c0map.clone().requires_grad_()
    # float32 kernels (ray delays, DAS and their adjoints); the scalar loss on
    # the images is evaluated in float64
    loss = focus_loss(prob.images(c).double(), e0, prob.kwin)
    (gc,) = torch.autograd.grad(loss, [c])
    gc = gc.double().cpu().numpy()

    rf64 = prob.rf.double()[None]
    st, xo, xe, pv = (t.double() for t in (prob.steer, prob.xorig, prob.xele, prob.pvec))
    nx, nz = prob.gridshape
    t32, dt32, geo32 = float32_inputs(prob.t0, prob.dt, prob.geo)

    def L64(cm):
        with torch.no_grad():
            ttx, trx = diff64.ray_delays(s32 + (1.0 / cm - 1.0 / base), geo32, st, xo, xe, pv,
                                         1540.0, 64)
            im = diff64.das_tabbed(rf64, ttx, torch.ones_like(ttx), trx, torch.ones_like(trx),
                                   t32, dt32, "cubic", "tx_only", 0.0, x32=False)
            return float(focus_loss(im[0].reshape(-1, nx, nz), e0, prob.kwin))

    base = c0map.double()
    # the kernels take the slowness as float32(1 / c); central differences of
    # h = 1e-3 m/s lie far below float32 resolution, so the FP64 pipeline
    # expands around the kernels' base slowness
    s32 = (1.0 / c0map).double()
    rng = np.random.default_rng(5)
    h = 1e-3
    peak = float(np.abs(gc).max())
    errs = []
    for _ in range(32):
        i, k = int(rng.integers(0, prob.geo.nzm)), int(rng.integers(0, prob.geo.nxm))
        cp, cm = base.clone(), base.clone()
        cp[i, k] += h
        cm[i, k] -= h
        fd = (L64(cp) - L64(cm)) / (2 * h)
        errs.append((abs(gc[i, k] - fd) / max(abs(fd), 1e-2 * peak), i, k, gc[i, k], fd))
    errs.sort(reverse=True)
    print("sos gradient vs FP64 central differences: worst rel errs",
          [f"{e[0]:.2e}" for e in errs[:6]], "median", f"{errs[16][0]:.2e}", flush=True)
    assert errs[0][0] <= 1e-3, errs[0]


def test_adam_steps_reduce_the_focus_loss():
    import torch
    prob, ctrue = _problem()
    c = torch.full((prob.geo.nzm, prob.geo.nxm), 1540.0, device="cuda", requires_grad=True)
    with torch.no_grad():
        e0 = float((prob.images(c).sum(0) ** 2).mean())
    opt = torch.optim.Adam([c], lr=2.0)
    losses = []
    for _ in range(15):
        opt.zero_grad()
        loss = prob.loss(c, e0)
        loss.backward()
        opt.step()
        losses.append(float(loss))
    print("losses", [round(v, 5) for v in losses], flush=True)
    assert losses[-1] < losses[0]


def test_angle_loss_gradient_matches_fp64_central_differences():
    """d angle_loss / d c at 24 seeded map entries: GPU chain (per-transmit
    images, sumtype rx_only, float32 kernels) against central differences
    (h = 1e-3 m/s) of the float64 pipeline expanded around the kernels'
    float32 base slowness."""
    import torch
    from oracle import diff64
    prob, ctrue = _problem()
    c0map = torch.full((prob.geo.nzm, prob.geo.nxm), 1510.0, device="cuda")
    c = c0map.clone().requires_grad_()
    loss = prob.angle_loss(c, 3)
    (gc,) = torch.autograd.grad(loss, [c])
    gc = gc.double().cpu().numpy()

    rf64 = prob.rf.double()[None]
    st, xo, xe, pv = (t.double() for t in (prob.steer, prob.xorig, prob.xele, prob.pvec))
    nx, nz = prob.gridshape
    t32, dt32, geo32 = float32_inputs(prob.t0, prob.dt, prob.geo)
    base = c0map.double()
    s32 = (1.0 / c0map).double()

    def L64(cm):
        with torch.no_grad():
            ttx, trx = diff64.ray_delays(s32 + (1.0 / cm - 1.0 / base), geo32, st, xo, xe, pv,
                                         1540.0, 64)
            im = diff64.das_tabbed(rf64, ttx, torch.ones_like(ttx), trx, torch.ones_like(trx),
                                   t32, dt32, "cubic", "rx_only", 0.0, x32=False)[0]
            d = im[3:] - im[:-3]
            return float((d * d).sum() / (im * im).sum())

    rng = np.random.default_rng(9)
    h = 1e-3
    peak = float(np.abs(gc).max())
    errs = []
    for _ in range(24):
        i, k = int(rng.integers(0, prob.geo.nzm)), int(rng.integers(0, prob.geo.nxm))
        cp, cm = base.clone(), base.clone()
        cp[i, k] += h
        cm[i, k] -= h
        fd = (L64(cp) - L64(cm)) / (2 * h)
        errs.append((abs(gc[i, k] - fd) / max(abs(fd), 1e-2 * peak), i, k, gc[i, k], fd))
    errs.sort(reverse=True)
    print("angle-loss gradient vs FP64 central differences: worst rel errs",
          [f"{e[0]:.2e}" for e in errs[:6]], flush=True)
    assert errs[0][0] <= 1e-3, errs[0]


def test_angle_fit_moves_the_map_toward_the_medium():
    """Adam on a depth profile, then on the free map, of the cross-angle loss
    (reduced problem, 1510 m/s start): the loss falls and the map's RMSE
    against the medium falls."""
    import torch
    prob, ctrue = _problem()
    ct = torch.as_tensor(ctrue, dtype=torch.float32, device="cuda")
    rmse = lambda c: float(torch.sqrt(((c - ct) ** 2).mean()))  # noqa: E731
    nzm, nxm = prob.geo.nzm, prob.geo.nxm
    prof = torch.full((nzm, 1), 1510.0, device="cuda", requires_grad=True)
    opt = torch.optim.Adam([prof], lr=5.0)
    with torch.no_grad():
        l0 = float(prob.angle_loss(prof.expand(nzm, nxm)))
    r0 = rmse(prof.detach().expand(nzm, nxm))
    for _ in range(40):
        opt.zero_grad()
        prob.angle_loss(prof.expand(nzm, nxm)).backward()
        opt.step()
    c = prof.detach().expand(nzm, nxm).clone().requires_grad_()
    opt = torch.optim.Adam([c], lr=2.0)
    prob.smooth_weight = 1e-3
    for _ in range(60):
        opt.zero_grad()
        prob.angle_loss(c).backward()
        opt.step()
    prob.smooth_weight = 0.0
    with torch.no_grad():
        l1 = float(prob.angle_loss(c))
    print(f"angle fit: loss {l0:.4f} -> {l1:.4f}, map rmse {r0:.2f} -> {rmse(c.detach()):.2f}",
          flush=True)
    assert l1 < l0 and rmse(c.detach()) < r0
