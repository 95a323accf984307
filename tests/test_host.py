"""CPU-side checks of the product library without a GPU: every symbol declared
in include/cudasph.h is exported, parameter presets, the seeded initial
conditions, and the reference state-file layout."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "cudasph.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"^[\w\s\*]*?\b([A-Za-z_]\w*)\s*\(([^;{]*)\)\s*;", text, flags=re.M):
        if m.group(1) not in ("if", "while", "return", "sizeof"):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol(cs):
    lib = C.CDLL(str(cs.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 39
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    for ref in ("host_solve_range", "host_sph_accelerations", "compute_CUDA_gravity", "allocateCUDAarray",
                "zcopyArrayToDevice", "RadixSort", "get_device_count", "cuda_set_device0"):
        assert ref in names


def test_reference_defaults(cs):
    p = cs.default_params(1000)
    assert p.n_dev == 1024 and p.rel_err == pytest.approx(1e-3) and p.desired_ngb == 24
    assert p.max_iter == 20 and p.av_alpha == 1 and p.av_beta == 2 and p.grav_mode == 0
    assert p.eos_density_unit == 5.89994031541978 and p.eos_temp_base == 1e6
    assert p.ph_period == 100 and p.ph_bits == 18 and p.eos_newton == 0
    assert cs.default_params(256).n_dev == 256 and cs.default_params(257).n_dev == 512


@pytest.mark.parametrize("cfg,n,ngb,eps", [(1, 16384, 32, 0.01), (2, 262144, 50, 0.005),
                                          (3, 1048576, 64, 0.002), (4, 65536, 50, 0.005),
                                          (5, 4194304, 64, 0.001)])
def test_presets(cs, cfg, n, ngb, eps):
    p = cs.preset_params(cfg)
    assert p.n_bodies == n and p.n_dev == n
    assert p.grav_mode == 1 and p.grav_eps == pytest.approx(eps)
    assert p.desired_ngb == pytest.approx(0.7061074 * ngb, rel=1e-6)
    assert p.eos_rad_const == 1e-3 and p.eos_boltzmann == 1 and p.eos_newton == 1


def test_ic_is_seeded_and_deterministic(cs):
    a = cs.generate_ic(2, 4096, 77)
    b = cs.generate_ic(2, 4096, 77)
    c = cs.generate_ic(2, 4096, 78)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])


def test_ic_shapes(cs):
    n = 8192
    pos, vel, hyd = cs.generate_ic(1, n, 1)
    assert pos[:, :3].min() >= 0 and pos[:, :3].max() < 1 and np.all(vel[:, :3] == 0)
    assert np.all(pos[:, :3] * 2**24 == np.round(pos[:, :3] * 2**24)), "24-bit grid"
    assert vel[:, 3].sum() == pytest.approx(1.0) and np.all(hyd[:, 2] == 1.0) and np.all(hyd[:, 3] == 0.6)
    pos, vel, hyd = cs.generate_ic(2, n, 1)
    r = np.linalg.norm(pos[:, :3], axis=1)
    assert r.max() <= 10.0
    # Plummer cumulative mass M(<r) = r^3 / (1 + r^2)^1.5
    for rr in (0.5, 1.0, 2.0):
        assert (r < rr).mean() == pytest.approx(rr**3 / (1 + rr**2) ** 1.5 / (1000 / 101**1.5), abs=0.02)
    vesc = np.sqrt(2.0) * (1 + r * r) ** -0.25
    assert np.all(np.linalg.norm(vel[:, :3], axis=1) <= vesc + 1e-12)
    pos, vel, hyd = cs.generate_ic(3, n, 1)
    c1, c2 = pos[: n // 2, :3].mean(0), pos[n // 2:, :3].mean(0)
    np.testing.assert_allclose(c2 - c1, [-4, -3, 0], atol=0.05)
    v1, v2 = vel[: n // 2, :3].mean(0), vel[n // 2:, :3].mean(0)
    np.testing.assert_allclose(v2 - v1, [0.6, 0.2, 0], atol=1e-12)
    assert np.linalg.norm(pos[: n // 2, :3] - c1, axis=1).max() < 1.0 + 0.05
    pos, vel, hyd = cs.generate_ic(5, 32768, 1)
    assert np.linalg.norm(pos[:, :3], axis=1).max() <= 1.0
    vrms = np.sqrt((vel[:, :3] ** 2).sum(1).mean())
    assert vrms == pytest.approx(2 * np.sqrt(10 / 9), rel=1e-9)


def test_kd_tree_builder_thread_count_does_not_change_the_tree(cs):
    pos, _, _ = cs.generate_ic(3, 100000, 5)
    t1, m1 = cs.build_kd_tree(pos.astype(np.float32), None, 1)
    t8, m8 = cs.build_kd_tree(pos.astype(np.float32), None, 8)
    assert np.array_equal(t1.view(np.uint32), t8.view(np.uint32)) and np.array_equal(m1, m8)
    d1, _ = cs.build_kd_tree(pos, None, 1)
    d8, _ = cs.build_kd_tree(pos, None, 8)
    assert np.array_equal(d1, d8)


def _lane_emden(n_poly=1.5):
    """Independent Lane-Emden solution (scipy, tight tolerances) from the
    series start; returns (xi_1, omega = -xi_1^2 theta'(xi_1), dense sol)."""
    from scipy.integrate import solve_ivp

    x0 = 1e-6

    def rhs(x, y):
        return [y[1], -max(y[0], 0.0) ** n_poly - 2 * y[1] / x]

    def zero(x, y):
        return y[0]
    zero.terminal = True
    sol = solve_ivp(rhs, [x0, 10], [1 - x0**2 / 6, -x0 / 3], events=zero, rtol=1e-12, atol=1e-14,
                    dense_output=True)
    xi1 = sol.t_events[0][0]
    return xi1, -xi1**2 * sol.y_events[0][0][1], sol


def test_polytrope_ic_follows_lane_emden(cs):
    """Config 3's stars: n = 1.5 Lane-Emden profile, R = 1, M = 0.5 each.
    The independent solution reproduces the tabulated xi_1 = 3.65375,
    omega = 2.71406 and rho_c / rho_mean = 5.99071; the sampled radii follow
    its enclosed-mass fraction m(xi)/m(xi_1) (Kolmogorov-Smirnov distance
    below the 99.9% critical value), and u = n K rho_c^(1/n) theta(r) with K
    from the polytrope's mass-radius relation, P = K rho^(5/3)."""
    xi1, omega, sol = _lane_emden()
    assert xi1 == pytest.approx(3.65375, abs=2e-5)
    assert omega == pytest.approx(2.71406, abs=2e-5)
    assert xi1**3 / (3 * omega) == pytest.approx(5.99071, abs=2e-5)
    n = 65536
    pos, vel, hyd = cs.generate_ic(3, n, 21)
    R, M = 1.0, 0.5
    for star, centre in ((0, (2.0, 1.5, 0.0)), (1, (-2.0, -1.5, 0.0))):
        sl = slice(0, n // 2) if star == 0 else slice(n // 2, n)
        r = np.linalg.norm(pos[sl, :3] - np.array(centre), axis=1)
        assert r.max() <= R * (1 + 1e-9)
        xi = np.sort(r) * xi1 / R
        frac = -(xi**2) * sol.sol(np.maximum(xi, 1e-6))[1] / omega     # m(xi) / m(xi_1)
        emp = np.arange(1, xi.size + 1) / xi.size
        assert np.abs(frac - emp).max() < 1.95 / np.sqrt(xi.size)
        # u from the profile: K = 4 pi G alpha^2 rho_c^(1/3) / (n + 1), alpha = R / xi_1
        rho_c = M * xi1**3 / (4 * np.pi * R**3 * omega)
        K = 4 * np.pi * (R / xi1) ** 2 * rho_c ** (1 / 3) / 2.5
        u_c = 1.5 * K * rho_c ** (2 / 3)
        theta = sol.sol(np.maximum(r * xi1 / R, 1e-6))[0]
        want = np.maximum(u_c * np.maximum(theta, 0), 1e-4 * u_c)
        np.testing.assert_allclose(hyd[sl, 2], want, rtol=1e-5, atol=1e-6 * u_c)
    assert np.all(vel[:, 3] == 1.0 / n)


def test_turbulent_ic_spectrum_and_interpolation(cs):
    """Config 5's velocities: each component's 128^3 grid has a shell-averaged
    power spectrum with log-log slope -4 (+-0.1) over k = 2..40, and the
    particle velocities are its trilinear interpolation, scaled to rms Mach 2
    at c_s = sqrt(gamma (gamma - 1) u) with u = 1."""
    import ctypes as C
    n, seed, L = 32768, 4, 128
    lib = cs.lib()
    lib.sph_ic_turbulent_field.argtypes = [C.c_int, C.c_uint64, C.c_void_p]
    field = np.zeros((3, L, L, L))
    assert lib.sph_ic_turbulent_field(n, seed, field.ctypes.data) == 0
    k = np.fft.fftfreq(L, 1.0 / L)
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2)
    shell = np.rint(kk).astype(int).ravel()
    ks = np.arange(2, 41)
    for c in range(3):
        power = np.abs(np.fft.fftn(field[c])) ** 2
        pk = np.bincount(shell, power.ravel()) / np.bincount(shell)
        slope = np.polyfit(np.log(ks), np.log(pk[ks]), 1)[0]
        assert slope == pytest.approx(-4.0, abs=0.1), (c, slope)
    pos, vel, hyd = cs.generate_ic(5, n, seed)
    g = (pos[:, :3] + 1) * 0.5 * L
    i0 = np.minimum(g.astype(int), L - 1)
    f = g - i0
    v = np.zeros((n, 3))
    for a in (0, 1):
        for b in (0, 1):
            for d in (0, 1):
                w = (f[:, 0] if a else 1 - f[:, 0]) * (f[:, 1] if b else 1 - f[:, 1]) * (f[:, 2] if d else 1 - f[:, 2])
                ix, iy, iz = (i0[:, 0] + a) % L, (i0[:, 1] + b) % L, (i0[:, 2] + d) % L
                v += w[:, None] * field[:, ix, iy, iz].T
    cs_ = np.sqrt(5 / 3 * 2 / 3 * 1.0)
    v *= 2 * cs_ / np.sqrt((v**2).sum(1).mean())
    np.testing.assert_allclose(vel[:, :3], v, rtol=1e-9, atol=1e-12 * np.abs(v).max())
    assert np.all(hyd[:, 2] == 1.0)


def test_plummer_thermal_energy_supports_a_tenth_of_the_virial(cs):
    """Config 2's gas: 2U = 0.1 |W| (thermal pressure carries 10% of the
    virial support), U = sum m u, W the direct FP64 sum of -m_i m_j / r_ij;
    the velocities from the Plummer distribution function give 2T ~ |W|.
    Both within sampling noise at N = 8192."""
    pos, vel, hyd = cs.generate_ic(2, 8192, 1)
    x, m = pos[:, :3], vel[:, 3]
    d = np.sqrt(((x[:, None] - x[None]) ** 2).sum(-1))
    np.fill_diagonal(d, np.inf)
    W = -0.5 * (m[:, None] * m[None] / d).sum()
    U = (m * hyd[:, 2]).sum()
    T = 0.5 * (m * (vel[:, :3] ** 2).sum(1)).sum()
    assert U == pytest.approx(0.05 * 3 * np.pi / 32, rel=1e-12)
    assert 0.09 < 2 * U / abs(W) < 0.11
    assert 0.95 < 2 * T / abs(W) < 1.05
