"""Pins the FP64 differentiable restatement (oracle/diff64.py) that checks the
GPU adjoint kernels: its forward agrees with the float32 oracle, it is linear
in the channel data, its natural spline matches scipy, and its autograd
gradients agree with central finite differences."""
import numpy as np
import pytest
import torch

import oracle
from oracle import diff64


def _case(seed, ntx=2, nrx=3, nt=96, np_=40, nfr=1):
    rng = np.random.default_rng(seed)
    t0, dt = 1e-6, 2.5e-8
    span = dt * (nt - 1)
    rf = rng.standard_normal((nfr, ntx, nrx, nt))
    tautx = rng.uniform(t0 - 0.05 * span, t0 + 0.5 * span, (ntx, np_))
    taurx = rng.uniform(0.0, 0.52 * span, (nrx, np_))
    apodtx = rng.uniform(-1, 1, (ntx, np_))
    apodrx = rng.uniform(-1, 1, (nrx, np_))
    return rf, tautx, apodtx, taurx, apodrx, t0, dt


def test_natural_y2_matches_scipy():
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(1)
    y = rng.standard_normal(50)
    dt = 0.3
    got = diff64.natural_y2(torch.as_tensor(y)[None], dt)[0].numpy()
    ref = CubicSpline(np.arange(50) * dt, y, bc_type="natural")(np.arange(50) * dt, 2)
    np.testing.assert_allclose(got, ref, atol=1e-9)


def test_natural_y2_sweep_matches_scipy():
    """The tridiagonal sweep used for large inputs without autograd."""
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(2)
    y = rng.standard_normal((3, 700))
    dt = 0.7
    rhs = (y[:, 2:] - 2.0 * y[:, 1:-1] + y[:, :-2]) * (6.0 / dt ** 2)
    got = diff64._solve_141(torch.as_tensor(rhs))
    for c in range(3):
        ref = CubicSpline(np.arange(700) * dt, y[c], bc_type="natural")(np.arange(700) * dt, 2)
        np.testing.assert_allclose(got[c].numpy(), ref[1:-1], atol=1e-9)
    big = torch.as_tensor(rng.standard_normal((1100, 4000)))
    sweep = diff64.natural_y2(big, dt)  # above the size where the sweep is taken
    dense = diff64.natural_y2(big[:2].clone().requires_grad_(), dt).detach()
    np.testing.assert_allclose(sweep[:2].numpy(), dense.numpy(), atol=1e-9)


@pytest.mark.parametrize("kind", ["nearest", "linear", "cubic", "korder_cubic"])
def test_forward_agrees_with_float32_oracle(kind):
    from oracle import numerics
    rf, tautx, apodtx, taurx, apodrx, t0, dt = _case(2)
    f32 = [np.float32(a) for a in (rf, tautx, apodtx, taurx, apodrx)]
    S = numerics.make_S(8).astype(np.float32) if kind == "korder_cubic" else None
    got = diff64.das_tabbed(*(torch.as_tensor(a) for a in f32), t0, dt, kind,
                            S=None if S is None else S.astype(np.float64)).numpy()[0]
    ref = oracle.das_tabbed(f32[0][0], t0, dt, *f32[1:], kind=kind, thresh=0.0, S=S)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))


def test_korder_coefs_reproduce_samples_and_cubics():
    """The Hermite pieces pass through the samples (c0 = y_j, c0 + c1 + c2 + c3
    = y_j+1), and a cubic polynomial sampled on the grid is reproduced exactly
    away from the ends (the not-a-knot derivatives of a cubic are exact)."""
    from oracle import numerics
    k, n = 8, 40
    y = torch.as_tensor(np.random.default_rng(9).standard_normal(n))
    C = diff64.korder_coefs(y, numerics.make_S(k))
    torch.testing.assert_close(C[:, 0], y, rtol=0, atol=1e-12)
    torch.testing.assert_close(C[:-1].sum(-1), y[1:], rtol=0, atol=1e-12)
    xs = np.arange(n, dtype=np.float64)
    poly = lambda x: 0.3 - 1.2 * x + 0.05 * x ** 2 - 0.002 * x ** 3  # noqa: E731
    C = diff64.korder_coefs(torch.as_tensor(poly(xs)), numerics.make_S(k))
    for j in (0, 5, 17, n - 3):
        for t in (0.25, 0.5, 0.9):
            v = C[j, 0] + t * (C[j, 1] + t * (C[j, 2] + t * C[j, 3]))
            assert abs(float(v) - poly(j + t)) <= 1e-9 * (1 + abs(poly(j + t)))


@pytest.mark.parametrize("usf", [2, 3, 4])
def test_nearest_upsampled_agrees_with_float32_oracle(usf):
    rf, tautx, apodtx, taurx, apodrx, t0, dt = _case(8)
    f32 = [np.float32(a) for a in (rf, tautx, apodtx, taurx, apodrx)]
    got = diff64.das_tabbed(*(torch.as_tensor(a) for a in f32), t0, dt, "nearest",
                            usf=usf).numpy()[0]
    ref = oracle.das_tabbed(f32[0][0], t0, dt, *f32[1:], kind="nearest", usf=usf, thresh=0.0)
    # the oracle upsamples with its float32 spline recurrence (about 2e-5 of
    # the peak here); a wrong weight or index would be of the order of the peak
    assert np.max(np.abs(got - ref)) <= 1e-4 * np.max(np.abs(ref))


def test_upsample_matches_scipy_natural_spline():
    from scipy.interpolate import CubicSpline
    y = np.random.default_rng(10).standard_normal((1, 40))
    u = diff64.upsample(torch.as_tensor(y), 0.0, 0.25, 4).numpy()[0]
    xs = np.float32(0.0) + np.arange(4 * 39 + 1, dtype=np.float32) * np.float32(np.float32(0.25) / 4)
    ref = CubicSpline(np.arange(40) * 0.25, y[0], bc_type="natural")(xs.astype(np.float64))
    np.testing.assert_allclose(u, ref, atol=1e-12)


def test_linear_in_channel_data():
    rf, tautx, apodtx, taurx, apodrx, t0, dt = _case(3)
    T = [torch.as_tensor(a) for a in (tautx, apodtx, taurx, apodrx)]
    r1 = torch.as_tensor(rf)
    r2 = torch.as_tensor(np.random.default_rng(4).standard_normal(rf.shape))
    f = lambda r: diff64.das_tabbed(r, *T, t0, dt, "cubic")  # noqa: E731
    torch.testing.assert_close(f(2.0 * r1 - 3.0 * r2), 2.0 * f(r1) - 3.0 * f(r2), rtol=1e-12,
                               atol=1e-12)


@pytest.mark.parametrize("method", ["akima", "makima"])
def test_akima_coefs_match_scipy(method):
    """The Hermite pieces of diff64.akima_coefs evaluate to scipy's
    Akima1DInterpolator (same two-point end extension) everywhere, ends
    included, and pass through the samples."""
    from scipy.interpolate import Akima1DInterpolator
    n = 50
    y = np.random.default_rng(21).standard_normal(n)
    C = diff64.akima_coefs(torch.as_tensor(y), method == "makima").numpy()
    xs = np.arange(n - 1)[:, None] + np.array([0.0, 0.2, 0.5, 0.77, 0.999])[None, :]
    t = xs - np.arange(n - 1)[:, None]
    c = C[:-1]
    got = c[:, :1] + t * (c[:, 1:2] + t * (c[:, 2:3] + t * c[:, 3:4]))
    ref = Akima1DInterpolator(np.arange(float(n)), y, method=method)(xs)
    np.testing.assert_allclose(got, ref, atol=1e-12)
    np.testing.assert_allclose(C[:, 0], y, atol=0)


@pytest.mark.parametrize("kind", ["linear", "cubic", "korder_cubic", "akima", "makima",
                                  "nearest_usf3"])
def test_autograd_matches_central_differences(kind):
    from oracle import numerics
    rf, tautx, apodtx, taurx, apodrx, t0, dt = _case(5)
    S = numerics.make_S(8) if kind == "korder_cubic" else None
    usf = 3 if kind == "nearest_usf3" else 1
    kind = "nearest" if kind == "nearest_usf3" else kind
    args = [torch.as_tensor(a, dtype=torch.float64).requires_grad_()
            for a in (rf, tautx, apodtx, taurx, apodrx)]
    gout = torch.as_tensor(np.random.default_rng(6).standard_normal((1, 40)))
    loss = (diff64.das_tabbed(*args, t0, dt, kind, x32=False, S=S, usf=usf) * gout).sum()
    grads = [torch.zeros_like(a) if g is None else g
             for a, g in zip(args, torch.autograd.grad(loss, args, allow_unused=True))]

    def L(vals):
        with torch.no_grad():
            return float((diff64.das_tabbed(*vals, t0, dt, kind, x32=False, S=S, usf=usf)
                          * gout).sum())

    rng = np.random.default_rng(7)
    for ai, h in ((0, 1e-4), (1, 1e-13), (2, 1e-6), (3, 1e-13), (4, 1e-6)):
        for _ in range(6):
            idx = tuple(int(rng.integers(0, s)) for s in args[ai].shape)
            plus = [a.detach().clone() for a in args]
            minus = [a.detach().clone() for a in args]
            plus[ai][idx] += h
            minus[ai][idx] -= h
            fd = (L(plus) - L(minus)) / (2 * h)
            g = float(grads[ai][idx])
            scale = max(abs(fd), 1e-6 * float(grads[ai].abs().max()))
            assert abs(g - fd) <= 1e-4 * scale + 1e-9, (kind, ai, idx, g, fd)


class _Geo:
    def __init__(self, nzm, nxm, xm0, zm0, dxm, dzm):
        self.nzm, self.nxm, self.xm0, self.zm0, self.dxm, self.dzm = nzm, nxm, xm0, zm0, dxm, dzm


def _ray_case(seed=21):
    rng = np.random.default_rng(seed)
    geo = _Geo(8, 10, -0.01, 0.0, 0.002, 0.0025)
    steer = np.radians([-12.0, 0.0, 7.0])
    xele = np.linspace(-0.008, 0.008, 5)
    xorig = np.where(steer <= 0, xele[-1], xele[0])
    pvec = np.stack([rng.uniform(-0.008, 0.008, 30), rng.uniform(0.002, 0.019, 30)], 1)
    return geo, steer, xorig, xele, pvec


def test_ray_delays_homogeneous_closed_form():
    """A uniform map 1/c gives the plane-wave delay <p - o, n>/c and the
    element-to-pixel distance over c."""
    geo, steer, xorig, xele, pvec = _ray_case()
    c = 1540.0
    slow = torch.full((geo.nzm, geo.nxm), 1.0 / c, dtype=torch.float64)
    ttx, trx = diff64.ray_delays(slow, geo, steer, xorig, xele, pvec, c, 16)
    exp_tx = ((pvec[None, :, 0] - xorig[:, None]) * np.sin(steer)[:, None]
              + pvec[None, :, 1] * np.cos(steer)[:, None]) / c
    exp_rx = np.hypot(pvec[None, :, 0] - xele[:, None], pvec[None, :, 1]) / c
    np.testing.assert_allclose(ttx.numpy(), exp_tx, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(trx.numpy(), exp_rx, rtol=1e-12)


def test_ray_delays_linear_slowness_is_integrated_exactly():
    """Slowness linear in depth inside the node range: the midpoint rule of
    the bilinear interpolant is exact, tau_rx = L * s(z_p / 2) for rays that
    stay inside the nodes' depth range."""
    geo = _Geo(12, 4, -0.01, -0.001, 0.005, 0.002)   # node depths 0 .. 0.022
    zc = geo.zm0 + (np.arange(geo.nzm) + 0.5) * geo.dzm
    s = 1 / 1540 + 2e-3 * zc                          # linear in z
    slow = torch.as_tensor(np.repeat(s[:, None], geo.nxm, 1))
    xele = np.array([-0.003, 0.0, 0.004])
    pvec = np.array([[0.001, 0.012], [-0.002, 0.018], [0.0, 0.006]])
    _, trx = diff64.ray_delays(slow, geo, np.zeros(0), np.zeros(0), xele, pvec, 1540.0, 8)
    L = np.hypot(pvec[None, :, 0] - xele[:, None], pvec[None, :, 1])
    # rays start at z = 0 (the first node is at z = 0): mean of s over [0, z_p]
    exp = L * (1 / 1540 + 2e-3 * pvec[None, :, 1] / 2)
    np.testing.assert_allclose(trx.numpy(), exp, rtol=1e-12)


def test_ray_delays_autograd_matches_central_differences():
    geo, steer, xorig, xele, pvec = _ray_case(22)
    rng = np.random.default_rng(23)
    s0 = torch.as_tensor(1 / 1540 + 2e-5 * rng.standard_normal((geo.nzm, geo.nxm)))
    gt = torch.as_tensor(rng.standard_normal((3, 30)))
    gr = torch.as_tensor(rng.standard_normal((5, 30)))

    def L(s):
        a, b = diff64.ray_delays(s, geo, steer, xorig, xele, pvec, 1540.0, 12)
        return (a * gt).sum() + (b * gr).sum()

    s = s0.clone().requires_grad_()
    (g,) = torch.autograd.grad(L(s), [s])
    for _ in range(8):
        i, k = int(rng.integers(0, geo.nzm)), int(rng.integers(0, geo.nxm))
        h = 1e-9
        sp, sm = s0.clone(), s0.clone()
        sp[i, k] += h
        sm[i, k] -= h
        fd = float((L(sp) - L(sm)) / (2 * h))
        assert abs(float(g[i, k]) - fd) <= 1e-6 * max(abs(fd), 1e-3), (i, k)
