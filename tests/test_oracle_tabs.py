"""The oracle's restatement of pycbf's delay-tab generator (oracle/tabs.py)
against closed forms, the reference's library calls and a per-pixel loop."""
import numpy as np
import pytest

from oracle import tabs

C = 1540.0


def _array(n=9, pitch=0.3e-3):
    return pitch * (np.arange(n) - (n - 1) / 2)


def _grid():
    return np.linspace(-4e-3, 4e-3, 7), np.linspace(2e-3, 30e-3, 11)


def test_rx_tabs_closed_form_and_norm_form():
    xele = _array()
    xout, zout = _grid()
    tau, apod = tabs.rx_tabs(C, xele, xout, zout, 1.5)
    X, Z = np.meshgrid(xout, zout, indexing="ij")
    dx = X.flatten()[None, :] - xele[:, None]
    dz = np.broadcast_to(Z.flatten()[None, :], dx.shape)
    np.testing.assert_allclose(tau, np.hypot(dx, dz) / C, rtol=1e-15, atol=0)
    # the expanded sqrt(dx*dx + dz*dz) is bit for bit numpy's norm along the last axis
    d = np.stack([dx, dz], -1)
    assert np.array_equal(tau, np.linalg.norm(d, axis=-1) / C)
    # f-number acceptance: |x - xe| < z / (2 F)
    assert np.array_equal(apod, np.abs(dx) < dz / 3.0)


def test_plane_wave_broadside_is_depth_over_c():
    xele = _array()
    xout, zout = _grid()
    tau, apod = tabs.tx_tabs_pw(np.array([0.0]), C, C, xele, xout, zout)
    X, Z = np.meshgrid(xout, zout, indexing="ij")
    assert np.array_equal(tau[0], Z.flatten() / C)
    assert np.array_equal(apod[0], (X.flatten() >= xele[0]) & (X.flatten() <= xele[-1]))


def test_plane_wave_steered_origin_and_sign():
    """Steering right (> 0) starts at the first element, left at the last
    (:157); the delay grows along the steered normal."""
    xele = _array()
    xout, zout = _grid()
    th = np.radians(10.0)
    tau, _ = tabs.tx_tabs_pw(np.array([-th, th]), C, C, xele, xout, zout)
    X, Z = np.meshgrid(xout, zout, indexing="ij")
    ref_l = ((X - xele[-1]) * np.sin(-th) + Z * np.cos(-th)) / C
    ref_r = ((X - xele[0]) * np.sin(th) + Z * np.cos(th)) / C
    np.testing.assert_allclose(tau[0], ref_l.flatten(), rtol=0, atol=1e-18)
    np.testing.assert_allclose(tau[1], ref_r.flatten(), rtol=0, atol=1e-18)


def test_plane_wave_sound_speed_correction_refracts_the_angle():
    xele = _array()
    xout, zout = _grid()
    th = np.radians(8.0)
    tau, _ = tabs.tx_tabs_pw(np.array([th]), 1480.0, C, xele, xout, zout)
    st = np.arcsin(C * np.sin(th) / 1480.0)
    X, Z = np.meshgrid(xout, zout, indexing="ij")
    ref = ((X - xele[0]) * np.sin(st) + Z * np.cos(st)) / C
    np.testing.assert_allclose(tau[0], ref.flatten(), rtol=0, atol=1e-18)


def test_divergent_on_axis_reduces_to_depth():
    xele = _array()  # odd count: an element at x = 0
    xout, zout = np.array([0.0]), np.linspace(2e-3, 30e-3, 11)
    tau, apod = tabs.tx_tabs_divergent(np.array([0.0]), -10e-3, C, C, xele, xout, zout)
    np.testing.assert_allclose(tau[0], zout / C, rtol=0, atol=1e-18)
    assert apod.all()


def test_focused_on_axis_and_outside_the_cone():
    xele = _array()
    f = 20e-3
    ovec, nvec, t0, ala = tabs.synthetic_focused(np.array([0.0]), C, C, xele, f)
    assert np.allclose(nvec[0], [0.0, 1.0]) and np.isclose(t0[0], np.hypot(xele[0], f) / C)
    zout = np.linspace(2e-3, 30e-3, 11)
    tau, apod = tabs.tx_tabs_focused(np.array([0.0]), f, C, C, xele, np.array([0.0]), zout)
    np.testing.assert_allclose(tau[0], (zout - f) / C + t0[0], rtol=0, atol=1e-18)
    assert apod.all()  # on axis: arccos(1) = 0 < half-aperture angle
    # at the focal depth, far off axis: arccos(0) = pi/2 is outside the cone
    _, apod2 = tabs.tx_tabs_focused(np.array([0.0]), f, C, C, xele, np.array([15e-3]),
                                    np.array([f]))
    assert not apod2.any()


@pytest.mark.parametrize("r0", [0.0, -12e-3, 18e-3])
def test_vectorised_equals_pixel_loop(r0):
    rng = np.random.default_rng(int(abs(r0) * 1e6) + 3)
    xele = _array(8)
    steers = np.radians(rng.uniform(-12, 12, 4))
    xout, zout = np.linspace(-3e-3, 3e-3, 5), np.linspace(1e-3, 25e-3, 6)
    a = tabs.make_tabs_and_apods_2D(steers, r0, C, C, xele, xout, zout, 1.2)
    b = tabs.pixel_loop(steers, r0, C, C, xele, xout, zout, 1.2)
    for k in ("tautx", "taurx", "apodtx", "apodrx"):
        assert np.array_equal(a[k], b[k]), k


def test_sound_speed_mismatch_raises_for_point_sources():
    xele = _array()
    xout, zout = _grid()
    for r0 in (-10e-3, 10e-3):
        with pytest.raises(ValueError):
            tabs.make_tabs_and_apods_2D(np.array([0.0]), r0, 1500.0, C, xele, xout, zout, 1.5)
