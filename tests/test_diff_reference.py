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


@pytest.mark.parametrize("kind", ["nearest", "linear", "cubic"])
def test_forward_agrees_with_float32_oracle(kind):
    rf, tautx, apodtx, taurx, apodrx, t0, dt = _case(2)
    f32 = [np.float32(a) for a in (rf, tautx, apodtx, taurx, apodrx)]
    got = diff64.das_tabbed(*(torch.as_tensor(a) for a in f32), t0, dt, kind).numpy()[0]
    ref = oracle.das_tabbed(f32[0][0], t0, dt, *f32[1:], kind=kind, thresh=0.0)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))


def test_linear_in_channel_data():
    rf, tautx, apodtx, taurx, apodrx, t0, dt = _case(3)
    T = [torch.as_tensor(a) for a in (tautx, apodtx, taurx, apodrx)]
    r1 = torch.as_tensor(rf)
    r2 = torch.as_tensor(np.random.default_rng(4).standard_normal(rf.shape))
    f = lambda r: diff64.das_tabbed(r, *T, t0, dt, "cubic")  # noqa: E731
    torch.testing.assert_close(f(2.0 * r1 - 3.0 * r2), 2.0 * f(r1) - 3.0 * f(r2), rtol=1e-12,
                               atol=1e-12)


@pytest.mark.parametrize("kind", ["linear", "cubic"])
def test_autograd_matches_central_differences(kind):
    rf, tautx, apodtx, taurx, apodrx, t0, dt = _case(5)
    args = [torch.as_tensor(a, dtype=torch.float64).requires_grad_()
            for a in (rf, tautx, apodtx, taurx, apodrx)]
    gout = torch.as_tensor(np.random.default_rng(6).standard_normal((1, 40)))
    loss = (diff64.das_tabbed(*args, t0, dt, kind, x32=False) * gout).sum()
    grads = torch.autograd.grad(loss, args)

    def L(vals):
        with torch.no_grad():
            return float((diff64.das_tabbed(*vals, t0, dt, kind, x32=False) * gout).sum())

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
