"""Pins each oracle function with checks the mathematics fixes (closed forms,
library routines, invariants, brute force on tiny inputs).  Each check is
chosen so a plausible mistake (dropped term, wrong sign or index, transposed
operand) fails it."""
import ctypes as ct
import math

import numpy as np
import pytest
from scipy.interpolate import Akima1DInterpolator, CubicSpline

import oracle
from oracle import numerics


# ----------------------------------------------------------------- spline ---
def test_cubic_matches_scipy_natural_spline():
    """With an exactly representable grid (dt = 0.5, t0 = 1) the reference's
    weights satisfy a + b = 1 and the value is the natural spline."""
    rng = np.random.default_rng(0)
    n, t0, dt = 200, 1.0, 0.5
    y = rng.standard_normal(n).astype(np.float32)
    t = t0 + dt * np.arange(n)
    x = rng.uniform(t[0], t[-2], 5000).astype(np.float32)
    got = oracle.cubic_points(y, t0, dt, x)
    ref = CubicSpline(t, y.astype(np.float64), bc_type="natural")(x.astype(np.float64))
    assert np.max(np.abs(got - ref)) < 2e-5 * np.max(np.abs(y))


def test_spline_y2_matches_scipy_second_derivative():
    rng = np.random.default_rng(3)
    y = rng.standard_normal(100).astype(np.float32)
    got = oracle.spline_y2(y, 0.25)
    ref = CubicSpline(0.25 * np.arange(100), y.astype(np.float64), bc_type="natural")(
        0.25 * np.arange(100), 2)
    assert np.max(np.abs(got - ref)) < 1e-4 * np.max(np.abs(ref))
    assert got[0] == 0.0 and got[-1] == 0.0


def test_cubic_exact_at_knots_and_on_lines():
    n, t0, dt = 64, 0.0, 1.0
    y = (3.0 - 0.25 * np.arange(n)).astype(np.float32)
    assert np.all(oracle.spline_y2(y, dt) == 0.0)
    x = np.linspace(0, n - 1.5, 777).astype(np.float32)
    np.testing.assert_allclose(oracle.cubic_points(y, t0, dt, x), 3.0 - 0.25 * x, atol=2e-6)
    rng = np.random.default_rng(5)
    y = rng.standard_normal(n).astype(np.float32)
    knots = np.arange(n - 1, dtype=np.float32)
    np.testing.assert_array_equal(oracle.cubic_points(y, t0, dt, knots), y[:-1])


def test_cubic_fill_outside_trace():
    y = np.ones(10, np.float32)
    x = np.array([-1e-3, -1e-9, 9.0 + 1e-5, 50.0, np.nan], np.float32)
    assert np.all(oracle.cubic_points(y, 0.0, 1.0, x) == 0.0)
    assert np.all(oracle.index_floor(x, 0.0, 1.0, 10) == -1)


def test_spline_forward_coefficients_reach_fixed_point():
    """The GPU prefilter keeps 32 forward coefficients and reuses the last for
    every later sample: valid because the float32 recurrence is exactly fixed
    from index 7 on."""
    y2f = oracle.spline_forward_coef(200_000)
    assert np.all(y2f[7:] == y2f[31])
    assert y2f[6] != y2f[7]
    assert abs(float(y2f[31]) - (math.sqrt(3) - 2)) < 1e-7


def test_upsample_matches_scipy_on_fine_grid():
    rng = np.random.default_rng(2)
    y = rng.standard_normal(50).astype(np.float32)
    up = oracle.upsample(y, 0.0, 1.0, 4)
    assert up.size == 4 * 49 + 1
    ref = CubicSpline(np.arange(50.0), y.astype(np.float64), bc_type="natural")(
        np.arange(up.size) / 4.0)
    np.testing.assert_allclose(up[:-1], ref[:-1], atol=2e-5 * np.abs(y).max())
    np.testing.assert_array_equal(up[::4], y[: up[::4].size])


# ----------------------------------------------------------------- linear ---
def test_linear_matches_np_interp():
    rng = np.random.default_rng(4)
    y = rng.standard_normal(300).astype(np.float32)
    x = rng.uniform(-5, 305, 4000).astype(np.float32)
    got = oracle.linear_points(y, 0.0, 1.0, x)
    ref = np.interp(x, np.arange(300), y, left=0.0, right=0.0)
    inside = (x >= 0) & (x <= 299)
    np.testing.assert_allclose(got[inside], ref[inside], atol=1e-6)
    assert np.all(got[~inside] == 0)
    assert oracle.linear_points(y, 0.0, 1.0, np.array([299.0], np.float32))[0] == y[-1]


# ---------------------------------------------------------------- nearest ---
def test_nearest_index_brute_force():
    """it = trunc(float(x - t0)/dt + 0.5 in double); valid 0 <= it < n.
    Truncation toward zero sends q in (-1.5, 0.5) to sample 0."""
    rng = np.random.default_rng(9)
    x = rng.uniform(-3.0, 12.0, 20000).astype(np.float32)
    got = oracle.index_nearest(x, 0.0, 1.0, 10)
    q = x.astype(np.float64) + 0.5
    exp = np.trunc(q).astype(np.int64)
    exp[(q <= -1.0) | (q >= 10.0)] = -1
    np.testing.assert_array_equal(got, exp)
    assert oracle.index_nearest(np.array([-1.4], np.float32), 0.0, 1.0, 10)[0] == 0
    assert oracle.index_nearest(np.array([-1.6], np.float32), 0.0, 1.0, 10)[0] == -1
    # 0.49999997 + 0.5 rounds to 1 in float32, but the reference adds in double
    assert oracle.index_nearest(np.array([0.49999997], np.float32), 0.0, 1.0, 10)[0] == 0


# ---------------------------------------------------------------- k-order ---
@pytest.mark.parametrize("k", [4, 8, 16])
def test_make_S_gives_not_a_knot_derivatives(k):
    rng = np.random.default_rng(k)
    y = rng.standard_normal(k)
    d = numerics.make_S(k) @ y
    ref = CubicSpline(np.arange(k), y, bc_type="not-a-knot")(np.arange(k), 1)
    np.testing.assert_allclose(d, ref, atol=1e-10)


def test_korder_exact_for_cubics():
    n, k = 80, 8
    t = np.arange(n, dtype=np.float64)
    y = (0.001 * t ** 3 - 0.05 * t ** 2 + t - 3).astype(np.float32)
    S = numerics.make_S(k).astype(np.float32)
    x = np.linspace(0, n - 1.01, 500).astype(np.float32)
    got = oracle.korder_points(y, 0.0, 1.0, S, x)
    xd = x.astype(np.float64)
    np.testing.assert_allclose(got, 0.001 * xd ** 3 - 0.05 * xd ** 2 + xd - 3,
                               atol=5e-4, rtol=1e-5)


def test_gpu_S_operator_equals_oracle_S():
    from pycbf_amd.gpu import make_S_by_k
    for k in (4, 8, 16, 126):
        np.testing.assert_allclose(make_S_by_k(k), numerics.make_S(k), atol=1e-9)


# --------------------------------------------------------- akima (GPU spec) ---
@pytest.mark.parametrize("method", ["akima", "makima"])
def test_akima_reference_spec_matches_scipy(method):
    """Pins the per-sample Akima derivative formula the GPU prefilter uses
    (pycbf_amd/csrc/spline.hip akima_deriv) by a float64 restatement against
    scipy.interpolate.Akima1DInterpolator."""
    rng = np.random.default_rng(11)
    y = rng.standard_normal(60)
    m = np.diff(y)
    mp = np.concatenate([[3 * m[0] - 2 * m[1], 2 * m[0] - m[1]], m,
                         [2 * m[-1] - m[-2], 3 * m[-1] - 2 * m[-2]]])
    d = np.empty(60)
    for i in range(60):
        mm2, mm1, mp0, mp1 = mp[i], mp[i + 1], mp[i + 2], mp[i + 3]
        w1, w2 = abs(mp1 - mp0), abs(mm1 - mm2)
        if method == "makima":
            w1 += 0.5 * abs(mp1 + mp0)
            w2 += 0.5 * abs(mm1 + mm2)
        d[i] = 0.5 * (mm1 + mp0) if w1 + w2 == 0 else (w1 * mm1 + w2 * mp0) / (w1 + w2)
    ref = Akima1DInterpolator(np.arange(60.0), y, method=method)(np.arange(60.0), 1)
    np.testing.assert_allclose(d, ref, atol=1e-12)


# --------------------------------------------------------------- geometry ---
def _pack(alatx=None, alarx=None):
    from pycbf_amd.gpu import _engine as E
    alatx = np.asarray(alatx if alatx is not None else [0.3], np.float32)
    alarx = np.asarray(alarx if alarx is not None else [0.3], np.float32)
    ntx, nrx = alatx.size, alarx.size
    z2 = lambda n: np.zeros((n, 2), np.float32)  # noqa: E731
    txp = np.zeros((ntx, E.TXPACK_FLOATS), np.float32)
    rxp = np.zeros((nrx, E.RXPACK_FLOATS), np.float32)
    rc = E.lib.pycbf_synthetic_pack(2, ntx, E.ptr(z2(ntx)), E.ptr(z2(ntx)),
                                    E.ptr(np.zeros(ntx, np.float32)), E.ptr(alatx),
                                    E.ptr(np.zeros(ntx, np.float32)), nrx, E.ptr(z2(nrx)),
                                    E.ptr(z2(nrx)), E.ptr(alarx), E.ptr(txp), E.ptr(rxp))
    assert rc == 0
    return txp, rxp


def test_acos_threshold_predicate_is_exact():
    """The device tests 'acos(v) > ala' as 'v < T'; against the oracle's
    acos in double for random, boundary-adjacent and special v."""
    rng = np.random.default_rng(12)
    alas = np.concatenate([rng.uniform(0, math.pi, 200), [0.0, 1e-6, math.pi / 2,
                                                          np.float32(math.pi), 3.2, -0.1]])
    alas = alas.astype(np.float32)
    _, rxp = _pack(alarx=alas)
    for ala, T in zip(alas, rxp[:, 6]):
        c = np.float32(math.cos(float(ala))) if 0 <= ala <= math.pi else np.float32(0)
        near = [c]
        for _ in range(6):
            near += [np.nextafter(near[-1], np.float32(2)), np.nextafter(near[0], np.float32(-2))]
        v = np.concatenate([rng.uniform(-1, 1, 2000).astype(np.float32), near,
                            np.float32([-1, 1, 0, 1.0000001, -1.0000001])]).astype(np.float32)
        dev = (v >= -1) & (v <= 1) & (v < T)
        with np.errstate(invalid="ignore"):
            ref = np.arccos(v.astype(np.float64)) > float(ala)
        assert np.array_equal(dev, ref), (ala, T)


def test_receive_midpoint_rule_matches_float_division():
    """The receive kernels decide RN(proj / mag) in [-1, 1] and < T without the
    division: proj > -(1 + 2^-24) mag and proj < m mag in double, with m the
    midpoint stored at rx pack floats 8-9.  Against numpy's IEEE float32
    quotient for proj one or a few ulps either side of T * mag and -mag, for
    mag over many binades, the special thresholds and non-finite operands."""
    rng = np.random.default_rng(16)
    alas = np.concatenate([rng.uniform(0, math.pi, 60), [0.0, 1e-6, math.pi / 2,
                                                         np.float32(math.pi), 3.2, -0.1,
                                                         np.nan]]).astype(np.float32)
    _, rxp = _pack(alarx=alas)
    m_all = np.ascontiguousarray(rxp[:, 8:10]).view(np.float64)[:, 0]
    lo = -(1.0 + 2.0 ** -24)
    mag = (rng.uniform(1, 2, 400) * 2.0 ** rng.integers(-38, 60, 400)).astype(np.float32)
    mag = np.concatenate([mag, np.float32([1e-12, 1.0000001e-12, 1, 2, np.inf, np.nan])])
    for T, m in zip(rxp[:, 6], m_all):
        anchors = [np.float32(T) if np.isfinite(T) else np.float32(1), np.float32(-1)]
        props = []
        for a in anchors:
            base = (np.float64(a) * mag.astype(np.float64)).astype(np.float32)
            for k in range(-3, 4):
                step = base.copy()
                for _ in range(abs(k)):
                    step = np.nextafter(step, np.float32(np.inf if k > 0 else -np.inf))
                props.append(np.stack([step, mag], 1))
        pm = np.concatenate(props + [np.float32([[np.inf, np.inf], [-np.inf, np.inf],
                                                  [np.nan, 1], [3, np.inf], [-3, np.inf]])])
        proj, mg = pm[:, 0], pm[:, 1]
        with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
            q = proj / mg
            ref = (mg.astype(np.float64) > 1e-12) & (q >= -1) & (q <= 1) & (q < T)
            pd, md = proj.astype(np.float64), mg.astype(np.float64)
            dev = (md > 1e-12) & (pd > md * lo) & (pd < md * m)
        assert np.array_equal(dev, ref), (T, np.flatnonzero(dev != ref)[:5])


def test_geometry_closed_forms():
    """Plane wave tau = <p-o, n>/c, receive tau = |p-e|/c, receive apod from
    the acceptance cone, focused tau = sign(proj)|p-o|/c + t0."""
    rng = np.random.default_rng(13)
    p = np.stack([rng.uniform(-0.02, 0.02, 3000), rng.uniform(0.001, 0.05, 3000)], 1)
    c0 = 1540.0
    o = np.float32([[0.001, 0.0]])
    n = np.float32([[math.sin(0.1), math.cos(0.1)]])
    tt, at, tr, ar = oracle.synthetic_tables(o, n, np.float32([0.0]), np.float32([0.0]),
                                             np.float32([0.0]), o, np.float32([[0, 1]]),
                                             np.float32([0.4]), c0, p)
    d = p - o.astype(np.float64)
    np.testing.assert_allclose(tt[0], d @ n[0].astype(np.float64) / c0, rtol=1e-5, atol=1e-12)
    assert np.all(at == 1)
    mag = np.linalg.norm(d, axis=1)
    np.testing.assert_allclose(tr[0], mag / c0, rtol=1e-6)
    ang = np.arccos(d[:, 1] / mag)
    clear = np.abs(ang - 0.4) > 1e-5
    np.testing.assert_array_equal(ar[0][clear], (ang <= 0.4)[clear].astype(np.float32))
    # diverging point source 5 mm behind the array
    of = np.float32([[0.0, -0.005]])
    tt, at, _, _ = oracle.synthetic_tables(of, np.float32([[0, 1]]), np.float32([-3e-6]),
                                           np.float32([0.5]), np.float32([0.0]), o,
                                           np.float32([[0, 1]]), np.float32([0.4]), c0, p)
    df = p - of.astype(np.float64)
    np.testing.assert_allclose(tt[0], np.linalg.norm(df, axis=1) / c0 - 3e-6, rtol=1e-5,
                               atol=1e-11)
    angf = np.arccos(df[:, 1] / np.linalg.norm(df, axis=1))
    clear = np.abs(angf - 0.5) > 1e-5
    np.testing.assert_array_equal(at[0][clear], (angf <= 0.5)[clear].astype(np.float32))


# ------------------------------------------------------------------- DMAS ---
def test_dmas_brute_force():
    rng = np.random.default_rng(14)
    im = rng.standard_normal((7, 5))
    lags = [1, 3, 9]
    for kind in ("ssr", "power", "pssr"):
        per = np.zeros((3, 5))
        for li, m in enumerate(lags):
            for p in range(5):
                for i in range(7 - m if m < 7 else 0):
                    q = im[i, p] * im[i + m, p]
                    per[li, p] += (math.copysign(math.sqrt(abs(q)), q) if kind == "ssr" else q)
        if kind == "pssr":
            s = per.sum(0)
            exp_sum = np.sign(s) * np.sqrt(np.abs(s))
            exp_per = np.sign(per) * np.sqrt(np.abs(per))
        else:
            exp_sum, exp_per = per.sum(0), per
        np.testing.assert_allclose(numerics.dmas(im, lags, kind, True), exp_sum, atol=1e-12)
        np.testing.assert_allclose(numerics.dmas(im, lags, kind, False), exp_per, atol=1e-12)
    ones = np.ones((6, 2))
    np.testing.assert_array_equal(numerics.dmas(ones, [1, 2], "power", True), [9.0, 9.0])


# ------------------------------------------------------------------- SLSC ---
def test_slsc_special_cases():
    rng = np.random.default_rng(15)
    K, N = 8, 20
    base = rng.standard_normal(K)
    coherent = np.repeat(base[:, None], N, 1)[:, :, None] * rng.uniform(0.5, 2, N)[None, :, None]
    mask = np.ones((N, 1), bool)
    assert abs(numerics.slsc(coherent, mask, [1, 2, 3])[0] - 1.0) < 1e-12
    anti = coherent * ((-1.0) ** np.arange(N))[None, :, None]
    assert abs(numerics.slsc(anti, mask, [1])[0] + 1.0) < 1e-12
    assert abs(numerics.slsc(anti, mask, [2])[0] - 1.0) < 1e-12
    noise = rng.standard_normal((4000, N, 1))
    assert abs(numerics.slsc(noise, mask, [1, 2])[0]) < 0.05
    few = np.zeros((N, 1), bool)
    few[:3] = True
    assert numerics.slsc(coherent, few, [1, 2, 3])[0] == 0.0
    flat = coherent.copy()
    flat[:, ::2] = 1.0  # zero variance on every other element -> lag 1 all NaN
    assert np.isnan(numerics.slsc(flat, mask, [1])[0])
    # lag 2 pairs only flat-flat or live-live: live-live still coherent
    assert abs(numerics.slsc(flat, mask, [2])[0] - 1.0) < 1e-12


def test_slsc_hand_computed():
    X = np.array([[1.0, 2.0, 0.0], [3.0, 1.0, 1.0], [2.0, 0.0, 2.0]])[:, :, None]
    # pairs (0,1), (1,2) at lag 1; population statistics over K = 3 samples
    def corr(a, b):
        a = a - a.mean()
        b = b - b.mean()
        return np.mean(a * b) / (a.std() * b.std())
    r1 = np.mean([corr(X[:, 0, 0], X[:, 1, 0]), corr(X[:, 1, 0], X[:, 2, 0])])
    r2 = corr(X[:, 0, 0], X[:, 2, 0])
    got = numerics.slsc(X, np.ones((3, 1), bool), [1, 2])[0]
    assert abs(got - 0.5 * (r1 + r2)) < 1e-12


def test_slsc_image_closed_forms():
    """The windowed pixel loop: coherent channels give 1, alternating signs -1
    at lag 1, output 0 without enough elements, and the output grid follows
    the downsampling factors."""
    rng = np.random.default_rng(16)
    nrx, nx, nz = 12, 3, 20
    base = rng.standard_normal((nx, nz))
    gain = rng.uniform(0.5, 2.0, nrx)
    im = gain[:, None, None] * base[None]
    out = numerics.slsc_image(im, None, [1, 2], kwin=5, kofs=-2)
    assert out.shape == (nx, nz)
    np.testing.assert_allclose(out, 1.0, atol=1e-12)
    alt = im * ((-1.0) ** np.arange(nrx))[:, None, None]
    np.testing.assert_allclose(numerics.slsc_image(alt, None, [1], 5, -2), -1.0, atol=1e-12)
    np.testing.assert_allclose(numerics.slsc_image(alt, None, [1, 2], 5, -2), 0.0, atol=1e-12)
    assert numerics.slsc_image(im, None, [11, 12], 5, -2).max() == 0.0   # need 13 elements
    ds = numerics.slsc_image(im, None, [1], 4, -1, dsfx=2, dsfz=3)
    assert ds.shape == (2, 7)


def test_slsc_image_pairs_follow_the_selected_list():
    """Lag m pairs elements m apart in the list of selected elements: with
    every other element deselected, alternating-sign data becomes coherent."""
    rng = np.random.default_rng(17)
    nrx, nx, nz = 16, 2, 12
    base = rng.standard_normal((nx, nz))
    alt = ((-1.0) ** np.arange(nrx))[:, None, None] * base[None]
    apod = np.ones((nrx, nx, nz))
    apod[1::2] = 0.0                        # keep the even (same-sign) elements
    np.testing.assert_allclose(numerics.slsc_image(alt, apod, [1], 3, -1), 1.0, atol=1e-12)
    # an element with a single zero apodisation sample inside the window drops out
    apod2 = np.ones((nrx, nx, nz))
    apod2[1::2, :, 6] = 0.0
    got = numerics.slsc_image(alt, apod2, [1], 3, -1)
    assert np.allclose(got[:, 5:8], 1.0) and np.allclose(got[:, :4], -1.0)


def test_slsc_image_window_is_clipped_at_the_edges():
    """At the first and last pixels only the samples inside the image count:
    channels coherent on z 0..2 and z nz-3..nz-1 but not elsewhere give 1 there
    only if the window is clipped."""
    rng = np.random.default_rng(18)
    nrx, nz = 10, 14
    im = rng.standard_normal((nrx, 1, nz))
    shared = rng.standard_normal(nz)
    for z in (0, 1, 2, nz - 3, nz - 2, nz - 1):
        im[:, 0, z] = shared[z] * np.linspace(1, 2, nrx)
    out = numerics.slsc_image(im, None, [1, 3], kwin=5, kofs=-2)
    assert abs(out[0, 0] - 1.0) < 1e-12 and abs(out[0, nz - 1] - 1.0) < 1e-12
    assert abs(out[0, nz // 2]) < 0.9


def test_slsc_image_zero_variance_channels_drop_their_pairs():
    rng = np.random.default_rng(19)
    nrx, nz = 10, 9
    base = rng.standard_normal(nz)
    im = np.tile(base, (nrx, 1))[:, None, :] * np.linspace(1, 3, nrx)[:, None, None]
    im[4] = 2.5                              # constant channel: std 0
    out = numerics.slsc_image(im, None, [1], 3, -1)
    np.testing.assert_allclose(out, 1.0, atol=1e-12)   # pairs (3,4), (4,5) dropped
    im[::2] = 1.0                            # every lag-1 pair now has a flat member
    assert np.all(np.isnan(numerics.slsc_image(im, None, [1], 3, -1)))
