"""Pins the oracle's M4 kernel family (dev_sph_kernels.cu:7-99) with
mathematics that does not reuse its formulas: normalisation, the cumulative
integral, finite-difference derivatives, the textbook M4 softened potential,
and the smoothed-count window's derivative chain."""
import numpy as np
import pytest

W, DW, DWDH, DPHIDH, IW, DIW, IG, DIGDH, DIG = range(9)


@pytest.fixture(scope="module")
def k64(orc):
    return lambda u: orc.kernel_fns(64, u)


def test_normalisation(k64):
    u = np.linspace(0, 2, 200001)
    w = k64(u)[:, W]
    # W = (1/pi) f(u) integrates to one over the sphere of radius 2h
    assert np.trapezoid(4 * np.pi * u**2 * w, u) == pytest.approx(1.0, rel=1e-9)


def test_cumulative_iw_is_integral_of_diw(k64):
    u = np.linspace(0, 2, 40001)
    f = k64(u)
    assert f[-1, IW] == pytest.approx(1.0, abs=1e-12)
    assert f[0, IW] == 0.0
    # diw(u) = 4 pi u^2 W(u) and iw(u) = int_0^u diw
    np.testing.assert_allclose(f[:, DIW], 4 * np.pi * u**2 * f[:, W], rtol=1e-12, atol=1e-15)
    cum = np.concatenate([[0], np.cumsum(0.5 * (f[1:, DIW] + f[:-1, DIW]) * np.diff(u))])
    np.testing.assert_allclose(f[:, IW], cum, atol=2e-7)


def fd(fun, u, h=1e-6):
    return (fun(u + h) - fun(u - h)) / (2 * h)


def test_dw_is_wprime_over_u(k64):
    u = np.concatenate([np.linspace(0.05, 0.99, 50), np.linspace(1.01, 1.99, 50)])
    wp = fd(lambda x: k64(x)[:, W], u)
    np.testing.assert_allclose(k64(u)[:, DW], wp / u, rtol=1e-6, atol=1e-9)


def test_dwdh_is_minus_3w_plus_u_wprime(k64):
    u = np.concatenate([np.linspace(0.0, 0.99, 50), np.linspace(1.01, 1.99, 50)])
    f = k64(u)
    wp = fd(lambda x: k64(x)[:, W], u)
    np.testing.assert_allclose(f[:, DWDH], -(3 * f[:, W] + u * wp), rtol=1e-6, atol=1e-9)


def m4_potential(u):
    """Textbook potential of the M4 cubic spline softening (Monaghan 1992;
    Price & Monaghan 2007, MNRAS 374, 1347, eq. A1), in units of G m / h."""
    u = np.asarray(u, float)
    out = np.empty_like(u)
    a = u < 1
    b = (u >= 1) & (u < 2)
    c = u >= 2
    out[a] = 2 / 3 * u[a]**2 - 0.3 * u[a]**4 + 0.1 * u[a]**5 - 1.4
    out[b] = (4 / 3 * u[b]**2 - u[b]**3 + 0.3 * u[b]**4 - u[b]**5 / 30 - 1.6 + 1 / (15 * u[b]))
    out[c] = -1 / u[c]
    return out


def test_dphidh_is_minus_phi_plus_u_phiprime(k64):
    u = np.concatenate([np.linspace(0.0, 0.99, 50), np.linspace(1.01, 1.99, 50)])
    phi = m4_potential(u)
    dphi = fd(m4_potential, u)
    np.testing.assert_allclose(k64(u)[:, DPHIDH], -(phi + u * dphi), rtol=1e-7, atol=1e-9)


def test_smoothed_count_window(k64):
    u = np.concatenate([np.linspace(0.02, 0.98, 40), np.linspace(1.02, 1.98, 40)])
    f = k64(u)
    iw4u = k64(4 * u)[:, IW]
    iw42u = k64(4 * (2 - u))[:, IW]
    np.testing.assert_allclose(f[:, IG], np.where(u < 1, iw4u, iw42u), rtol=1e-12)
    dig_du = fd(lambda x: k64(x)[:, IG], u)
    np.testing.assert_allclose(f[:, DIGDH], u * dig_du, rtol=1e-6, atol=2e-8)
    np.testing.assert_allclose(f[:, DIG], dig_du / u, rtol=1e-6, atol=2e-8)
    # outside the support everything vanishes and the window is zero at u = 0
    f2 = k64(np.array([0.0, 2.0, 2.5]))
    assert f2[0, IG] == 0 and f2[1, IG] == 0 and f2[1, W] == 0 and f2[2, DPHIDH] == 0


def test_float_instantiation_tracks_double(orc):
    u = np.linspace(0, 2.2, 4001)
    f32 = orc.kernel_fns(32, u.astype(np.float32).astype(np.float64))
    f64 = orc.kernel_fns(64, u.astype(np.float32).astype(np.float64))
    scale = np.abs(f64).max(axis=0) + 1e-30
    assert (np.abs(f32 - f64) / scale).max() < 2e-6
