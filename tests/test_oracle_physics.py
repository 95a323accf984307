"""Pins the oracle's density, EOS, viscosity, hydro + reaction, gravity,
integrator, timestep, statistics and Peano-Hilbert keys with checks that do
not reuse its code: FP64 brute force from textbook formulas, conservation
laws, limits, closed forms and error-free-transform identities."""
import numpy as np
import pytest


def m4(q):
    """Textbook M4 cubic spline (Monaghan & Lattanzio 1985), 3-D, h = q-scale:
    W, dW/dq."""
    q = np.asarray(q, float)
    w = np.where(q < 1, 1 - 1.5 * q**2 + 0.75 * q**3, np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0)) / np.pi
    dw = np.where(q < 1, -3 * q + 2.25 * q**2, np.where(q < 2, -0.75 * (2 - q) ** 2, 0.0)) / np.pi
    return w, dw


def m4_cumulative(x):
    """int_0^x 4 pi s^2 W(s) ds, by exact polynomial integration."""
    x = np.asarray(x, float)
    a = 4 * (x**3 / 3 - 0.3 * x**5 + 0.125 * x**6)
    b = 4 * (2 * x**3 / 3 - 0.75 * x**4 + 0.3 * x**5 - x**6 / 24) - 1 / 15
    return np.where(x < 0, 0.0, np.where(x < 1, a, np.where(x < 2, b, 1.0)))


def phi_m4(q):
    """Softened potential kernel of the M4 spline (Price & Monaghan 2007,
    eq. A2-A3): the potential of unit mass at r is phi(r/h)/h, -1/q outside."""
    q = np.asarray(q, float)
    a = 2 / 3 * q**2 - 0.3 * q**4 + 0.1 * q**5 - 1.4
    qs = np.where(q > 0, q, 1.0)
    b = 4 / 3 * q**2 - q**3 + 0.3 * q**4 - q**5 / 30 - 1.6 + 1 / (15 * qs)
    return np.where(q < 1, a, np.where(q < 2, b, -1 / qs))


def dphi_m4(q):
    """d phi / dq of phi_m4 (the force kernel: q^2 times it is the mass
    fraction of the spline inside q)."""
    q = np.asarray(q, float)
    a = 4 / 3 * q - 1.2 * q**3 + 0.5 * q**4
    qs = np.where(q > 0, q, 1.0)
    b = 8 / 3 * q - 3 * q**2 + 1.2 * q**3 - q**4 / 6 - 1 / (15 * qs**2)
    return np.where(q < 1, a, np.where(q < 2, b, 1 / qs**2))


def window(u):
    """smoothed-count window ig(u) and its u-derivative"""
    u = np.asarray(u, float)
    arg = 4 * (1 - np.abs(u - 1))
    x = np.clip(arg, 0, 2)
    w, _ = m4(x)
    d = 4 * np.pi * x**2 * w * 4 * np.where(u < 1, 1.0, -1.0) * (arg < 2) * (u < 2)
    return m4_cumulative(arg) * (u < 2), d


def run_density(cs, orc, cfg, n, bits=64):
    p = cs.preset_params(cfg, n)
    p.fix_phantom = 1
    pos, vel, hyd = cs.generate_ic(cfg, n, 9)
    o = orc.Oracle(bits, p)
    o.load_ic(pos, vel, hyd)
    for pid in (orc.TREE, orc.SOLVE, orc.OFFSETS, orc.FIND, orc.DENSITY):
        o.run_pass(pid)
    return o, p


@pytest.mark.parametrize("cfg", [1, 2])
def test_density_matches_fp64_brute_force(cs, orc, cfg):
    n = 1024
    o, p = run_density(cs, orc, cfg, n)
    pos, vel = o.get("pos")[:n], o.get("vel")[:n]
    h = 0.5 * pos[:, 3]
    dr = pos[:, None, :3] - pos[None, :, :3]
    dv = vel[:, None, :3] - vel[None, :, :3]
    r = np.sqrt((dr**2).sum(-1))
    u = r / h[:, None]
    inside = u < 2
    w, dw = m4(np.where(inside, u, 2.5))
    m = vel[:, 3]
    rho = (m[None, :] * w * inside).sum(1) / h**3
    dwdh = -(3 * w + u * dw) * inside                 # h^4 d/dh [W(r/h)/h^3]
    ig, dig_du = window(np.where(inside, u, 2.5))
    den = -(u * dig_du * inside).sum(1)               # -sum digdh, digdh = u dig/du
    omega = (m[None, :] * dwdh).sum(1) / h**3 / den
    dw_over_u = np.where(u > 0, dw / np.where(u > 0, u, 1), -3 / np.pi) * inside
    mdw = m[:, None] * dw_over_u                      # the reference weights by m_i
    divv = (mdw * (dr * dv).sum(-1)).sum(1) / rho / h**5
    curl = np.cross(dr, dv)
    curlv = np.linalg.norm((mdw[..., None] * curl).sum(1), axis=1) / rho / h**5
    # Psi: h-derivative of the M4 softened potential, phi(q)/h with the
    # textbook phi (Price & Monaghan 2007, eq. A2-A3), normalised like Omega:
    # Psi = sum_j m_j [-(phi + q phi')](u) / h / (-sum digdh)
    dphidh = -(phi_m4(np.where(inside, u, 2.5)) + u * dphi_m4(np.where(inside, u, 2.5))) * inside
    psi = (m[None, :] * dphidh).sum(1) / h / den
    hyd, dots = o.get("hydro")[:n], o.get("dots")[:n]
    np.testing.assert_allclose(hyd[:, 0], rho, rtol=1e-10)
    np.testing.assert_allclose(dots[:, 2], omega, rtol=1e-9)
    np.testing.assert_allclose(dots[:, 3], psi, rtol=1e-9)
    np.testing.assert_allclose(dots[:, 0], divv, rtol=1e-8, atol=1e-10 * np.abs(divv).max())
    np.testing.assert_allclose(dots[:, 1], curlv, rtol=1e-8, atol=1e-10 * np.abs(curlv).max() + 1e-300)


def test_softened_potential_kernel(cs, orc):
    """phi_m4 above is the potential of the M4 density: by Gauss, q^2 phi'(q)
    is the spline's mass inside q; phi is continuous and Keplerian past 2."""
    q = np.linspace(1e-3, 2.5, 2001)
    np.testing.assert_allclose(q**2 * dphi_m4(q), m4_cumulative(q), rtol=1e-12, atol=1e-15)
    for x in (1.0, 2.0):
        assert phi_m4(x - 1e-12) == pytest.approx(phi_m4(x + 1e-12), abs=1e-10)
    h = 1e-6
    np.testing.assert_allclose((phi_m4(q + h) - phi_m4(q - h)) / (2 * h), dphi_m4(q), rtol=1e-6, atol=1e-8)


def test_density_fp32_tracks_fp64(cs, orc):
    n = 2048
    o32, _ = run_density(cs, orc, 3, n, bits=32)
    o64, _ = run_density(cs, orc, 3, n, bits=64)
    same = o32.get("ngb_count")[:n] == o64.get("ngb_count")[:n]
    a, b = o32.get("hydro")[:n, 0], o64.get("hydro")[:n, 0]
    assert (np.abs(a - b) / b)[same].max() < 2e-5


def test_uniform_lattice_density(cs, orc):
    g = (np.arange(24) + 0.5) / 24
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    n = x.size
    p = cs.default_params(n)
    p.fix_phantom = 1
    pos = np.zeros((p.n_dev, 4))
    pos[:n, 0], pos[:n, 1], pos[:n, 2], pos[:n, 3] = x.ravel(), y.ravel(), z.ravel(), 2.6 / 24
    vel = np.zeros((p.n_dev, 4))
    vel[:n, 3] = 1.0 / n
    o = orc.Oracle(64, p)
    o.set("pos", pos)
    o.set("vel", vel)
    for pid in (orc.TREE, orc.OFFSETS, orc.FIND, orc.DENSITY):
        o.run_pass(pid)
    rho = o.get("hydro")[:n, 0]
    interior = np.all((pos[:n, :3] > 0.2) & (pos[:n, :3] < 0.8), axis=1)
    assert np.abs(rho[interior] - 1.0).max() < 0.02


# ---- EOS (dev_compute_pressure.cu) -----------------------------------------

def test_quartic_roots(orc):
    rng = np.random.default_rng(1)
    for _ in range(200):
        q = 10 ** rng.uniform(-3, 3)
        r = -(10 ** rng.uniform(-3, 3))
        x = orc.find_root(64, 1, q, r)
        scale = max(x**4, q * x, -r)
        assert abs(x**4 + q * x + r) <= 1e-13 * scale
        # Ferrari in double is accurate where gas and radiation are comparable
        # (both limits cancel catastrophically)
        if abs(np.log10(q**4 / (-r) ** 3)) < 2:
            xf = orc.find_root(64, 0, q, r)
            assert abs(xf**4 + q * xf + r) <= 1e-8 * scale


def test_eos_limits(cs, orc):
    p = cs.preset_params(2, 1000)
    p.eos_newton = 1
    # gas dominated: P = rho T / mu with u = 1.5 T / mu  ->  P = 2/3 rho u
    assert orc.pressure(64, p, 1.0, 1e-3, 0.6) == pytest.approx(2 / 3 * 1e-3, rel=1e-6)
    # radiation dominated: u = a T^4 / rho, P = a T^4 / 3  ->  P = rho u / 3
    assert orc.pressure(64, p, 1e-12, 1e4, 0.6) == pytest.approx(1e-12 * 1e4 / 3, rel=1e-4)
    # the reference's FP32 Ferrari root loses the gas-dominated regime
    p.eos_newton = 0
    assert orc.pressure(32, p, 0.3, 0.0147, 0.6) < 0.5 * (2 / 3 * 0.3 * 0.0147)


# ---- artificial viscosity (dev_art_visc.cu) --------------------------------

def av(orc, drdv, ds=0.1, divv=(-1.0, -1.0), curl=(0.0, 0.0), alpha=1.0, beta=2.0):
    hi = [1.0, 0.6, 1.0, 0.6]
    hj = [1.2, 0.9, 1.0, 0.6]
    di = [divv[0], curl[0], 0, 0]
    dj = [divv[1], curl[1], 0, 0]
    return orc.art_visc(64, alpha, beta, 0.2, 0.25, ds, drdv, hi, di, hj, dj)


def test_viscosity_limits(orc):
    ci, cj = np.sqrt(5 / 3 * 0.6 / 1.0), np.sqrt(5 / 3 * 0.9 / 1.2)
    pi, vs = av(orc, drdv=+0.05)
    assert pi == 0 and vs == pytest.approx(ci + cj)                  # receding: no viscosity
    pi, vs = av(orc, drdv=-0.05)
    assert pi > 0 and vs == pytest.approx(ci + cj + 2 * 0.05 / 0.1)  # v_sig grows by beta |v_r|
    pi0, _ = av(orc, drdv=-0.05, divv=(0.0, 0.0), curl=(3.0, 3.0))  # pure shear: Balsara -> 0
    assert pi0 == 0
    # linear (alpha) and quadratic (beta) parts scale with the approach speed
    a1, _ = av(orc, drdv=-1e-3, beta=0.0)
    a2, _ = av(orc, drdv=-2e-3, beta=0.0)
    b1, _ = av(orc, drdv=-1e-3, alpha=0.0)
    b2, _ = av(orc, drdv=-2e-3, alpha=0.0)
    assert a2 / a1 == pytest.approx(2, rel=1e-9) and b2 / b1 == pytest.approx(4, rel=1e-9)


# ---- hydro + reaction ------------------------------------------------------

@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("cfg", [1, 3])
def test_hydro_with_reaction_conserves_momentum(cs, orc, cfg, bits):
    n = 4096
    p = cs.preset_params(cfg, n)
    pos, vel, hyd = cs.generate_ic(cfg, n, 13)
    if cfg == 1:   # give the cube a velocity field so viscosity acts
        vel[:, :3] = np.sin(2 * np.pi * pos[:, [1, 2, 0]])
    o = orc.Oracle(bits, p)
    o.load_ic(pos, vel, hyd)
    o.run_pass(orc.SPH_ACC)
    a = o.get("bdots")[:n, :3].astype(np.float64)
    m = o.get("vel")[:n, 3].astype(np.float64)
    x = o.get("pos")[:n, :3].astype(np.float64)
    gross = (m[:, None] * np.abs(a)).sum(0)
    tol = 1e-5 if bits == 32 else 1e-12
    assert np.all(np.abs((m[:, None] * a).sum(0)) <= tol * gross)
    torque = (m[:, None] * np.cross(x, a)).sum(0)
    assert np.all(np.abs(torque) <= tol * (m * np.linalg.norm(x, axis=1) * np.linalg.norm(a, axis=1)).sum())


def hydro_state(cs, orc, cfg, n, alpha=None, beta=None, seed=17):
    """FP64 oracle after the passes of one SPH force evaluation (tree,
    h-solve, lists, density, EOS, hydro, reaction), phantom node removed."""
    p = cs.preset_params(cfg, n)
    p.fix_phantom = 1
    if alpha is not None:
        p.av_alpha, p.av_beta = alpha, beta
    pos, vel, hyd = cs.generate_ic(cfg, n, seed)
    o = orc.Oracle(64, p)
    o.load_ic(pos, vel, hyd)
    for pid in (orc.TREE, orc.SOLVE, orc.OFFSETS, orc.FIND, orc.DENSITY, orc.EOS, orc.HYDRO):
        o.run_pass(pid)
    udot_hydro = o.get("bdots")[:n, 3].copy()
    o.run_pass(orc.REACTION)
    return o, udot_hydro


def rates_along_velocity(pos, vel, n, delta):
    """Central differences in tau of rho_i(tau) = sum_j m_j W(|r_ij + tau v_ij| / h_i) / h_i^3
    and N_i(tau) = sum_j ig(|r_ij + tau v_ij| / h_i), h_i held: the rates of
    change of density and smoothed neighbour count along the velocity field.
    Textbook M4 and window (above), FP64 brute force over all pairs."""
    x, v, m = pos[:n, :3], vel[:n, :3], vel[:n, 3]
    h = 0.5 * pos[:n, 3]

    def sums(tau):
        r = np.sqrt(((x[:, None] - x[None, :] + tau * (v[:, None] - v[None, :])) ** 2).sum(-1))
        u = r / h[:, None]
        inside = u < 2
        w, _ = m4(np.where(inside, u, 2.5))
        ig, _ = window(np.where(inside, u, 2.5))
        return (m[None, :] * w * inside).sum(1) / h**3, (ig * inside).sum(1)

    rp, npl = sums(delta)
    rm, nm = sums(-delta)
    return (rp - rm) / (2 * delta), (npl - nm) / (2 * delta)


@pytest.mark.parametrize("cfg", [2, 5])
def test_hydro_udot_is_the_adiabatic_rate(cs, orc, cfg):
    """With the viscosity off (alpha = beta = 0) the hydro pass's du_i/dt is
    the grad-h form of the first law at fixed h,
        du_i/dt = (P_i / rho_i^2) [d rho_i/dt - Omega_i dN_i/dt],
    rho_i and N_i = sum_j ig(u_ij) rates taken along the velocities by central
    differences of brute-force sums (dev_hydro_dots.cu:76-80).  A wrong power
    of h or a missing Omega in u-dot fails this."""
    n = 1024
    o, udot = hydro_state(cs, orc, cfg, n, alpha=0.0, beta=0.0)
    pos, vel = o.get("pos"), o.get("vel")
    hyd, dots = o.get("hydro")[:n], o.get("dots")[:n]
    vmax = np.abs(vel[:n, :3]).max()
    drho, dN = rates_along_velocity(pos, vel, n, 2e-6 * pos[:n, 3].min() / vmax)
    expect = hyd[:, 1] / hyd[:, 0] ** 2 * (drho - dots[:, 2] * dN)
    scale = np.abs(expect).max()
    err = np.abs(udot - expect) / np.maximum(np.abs(expect), 1e-3 * scale)
    assert err.max() < 1e-6, err.max()
    assert np.abs(udot - expect).max() < 1e-8 * scale


@pytest.mark.parametrize("cfg", [2, 5])
def test_hydro_energy_identity_with_viscosity(cs, orc, cfg):
    """With viscosity on, the pair terms of the acceleration (own loop and
    reaction) and of du/dt cancel in the total energy rate except for the
    Psi (softening) term:
        sum_i m_i (v_i . a_i + du_i/dt) = 1/2 sum_i m_i Psi_i dN_i/dt.
    Holds only if Pi enters a and du/dt with the same factor 1/2, if P/rho^2
    Omega and Psi/2 multiply dig/h^2 as in dev_hydro_dots.cu:64-80, and if
    the reaction adds m_i acc_ij r_ij to j (sph_accelerations.cpp:104-115)."""
    n = 1024
    o, _ = hydro_state(cs, orc, cfg, n)
    pos, vel = o.get("pos"), o.get("vel")
    bd, dots = o.get("bdots")[:n], o.get("dots")[:n]
    m, v = vel[:n, 3], vel[:n, :3]
    vmax = np.abs(v).max()
    _, dN = rates_along_velocity(pos, vel, n, 2e-6 * pos[:n, 3].min() / vmax)
    work = m * (v * bd[:, :3]).sum(1)
    lhs = (work + m * bd[:, 3]).sum()
    rhs = 0.5 * (m * dots[:, 3] * dN).sum()
    gross = np.abs(work).sum() + np.abs(m * bd[:, 3]).sum()
    assert abs(lhs - rhs) <= 1e-11 * gross, (lhs, rhs, gross)
    # the identity is not empty: viscosity and the Psi term both act
    o0, _ = hydro_state(cs, orc, cfg, n, alpha=0.0, beta=0.0)
    assert np.abs(o0.get("bdots")[:n] - bd).max() > 1e-3 * np.abs(bd).max()
    assert abs(rhs) > 1e-6 * gross


def test_viscosity_pure_compression_switch_is_one(orc):
    """Balsara switch |div v| / (|div v| + |curl v| + 1e-4 c / h) -> 1 under
    pure compression (curl 0, |div v| >> c / h): Pi then takes its unswitched
    value (P_i/rho_i^2 + P_j/rho_j^2) (-alpha mu + beta mu^2), with
    mu = v_r / c_bar, v_r = (r . v)/|r| and c_bar = (c_i + c_j)/2
    (dev_art_visc.cu:15-29)."""
    drdv, ds, alpha, beta = -0.05, 0.1, 1.0, 2.0
    hi, hj = [1.0, 0.6, 1.0, 0.6], [1.2, 0.9, 1.0, 0.6]
    ci, cj = np.sqrt(5 / 3 * 0.6 / 1.0), np.sqrt(5 / 3 * 0.9 / 1.2)
    mu = (drdv / ds) / (0.5 * (ci + cj))
    unswitched = (0.6 / 1.0**2 + 0.9 / 1.2**2) * (-alpha * mu + beta * mu**2)
    pi, _ = orc.art_visc(64, alpha, beta, 0.2, 0.25, ds, drdv, hi, [-1e6, 0, 0, 0], hj, [-1e6, 0, 0, 0])
    assert pi == pytest.approx(unswitched, rel=1e-8)
    # a switch of 1/2 on both sides (|curl v| = |div v|, c/h negligible) halves mu
    pi, _ = orc.art_visc(64, alpha, 0.0, 0.2, 0.25, ds, drdv, hi, [-1e6, 1e6, 0, 0], hj, [-1e6, 1e6, 0, 0])
    assert pi == pytest.approx((0.6 + 0.9 / 1.44) * (-alpha * 0.5 * mu), rel=1e-8)


# ---- gravity (sph_grav_kernel.cu) ------------------------------------------

def grav_pair(orc, cs, d, rng_i, rng_j, mode=0, eps=0.0, bits=64):
    p = cs.default_params(2)
    p.grav_mode, p.grav_eps = mode, eps
    o = orc.Oracle(bits, p)
    pos = np.zeros((p.n_dev, 4))
    pos[2:] = [1e10, 1e10, 1e10, 0]
    # off the origin: padding records copy real positions scaled by 1.1
    # (CUDA_gravity.cpp:48-56), which coincide with a particle at the origin
    pos[0] = [0.5, 0.25, 0.125, rng_i]
    pos[1] = [0.5 + d, 0.25, 0.125, rng_j]
    vel = np.zeros((p.n_dev, 4))
    vel[0, 3] = 1.0
    vel[1, 3] = 2.0
    o.set("pos", pos)
    o.set("vel", vel)
    o.run_pass(orc.GRAVITY)
    return o.get("grav")[:2]


def test_gravity_newtonian_far_field(orc, cs):
    g = grav_pair(orc, cs, 3.0, 1.0, 1.5)
    assert g[0, 0] == pytest.approx(2.0 / 9.0, rel=1e-14)          # a_x on 0 toward 1
    assert g[1, 0] == pytest.approx(-1.0 / 9.0, rel=1e-14)
    assert g[0, 3] == pytest.approx(-2.0 / 3.0 - 2.8 * 1.0 / 1.0, rel=1e-12)   # phi + spline self term


def test_spline_softening_is_continuous_and_conservative(orc, cs):
    R = 1.0
    ds = np.concatenate([np.linspace(0.05, 0.995, 60), [0.5 - 1e-7, 0.5 + 1e-7, 1 - 1e-7]])
    acc = np.array([grav_pair(orc, cs, d, R, R)[0, 0] for d in ds])
    phi = np.array([grav_pair(orc, cs, d, R, R)[0, 3] for d in ds])
    # continuity across u = 0.5 and the outer edge (a = m/d^2 at d = range)
    assert acc[-3] == pytest.approx(acc[-2], rel=1e-6)
    assert acc[-1] == pytest.approx(2.0 / (1 - 1e-7) ** 2, rel=1e-5)
    # inside both ranges, a = -d phi / dd with phi the pair potential of j on i
    # (phi also holds the constant self term, whose derivative is zero)
    h = 1e-6
    for d in (0.2, 0.4, 0.6, 0.8, 0.95):
        dphi = (grav_pair(orc, cs, d + h, R, R)[0, 3] - grav_pair(orc, cs, d - h, R, R)[0, 3]) / (2 * h)
        a = grav_pair(orc, cs, d, R, R)[0, 0]
        assert a == pytest.approx(dphi, rel=1e-6)


def test_plummer_closed_form(orc, cs):
    eps = float(np.float32(0.3))     # the parameter block holds eps in single precision
    for d in (0.1, 0.5, 2.0):
        g = grav_pair(orc, cs, d, 1.0, 1.0, mode=1, eps=eps)
        assert g[0, 0] == pytest.approx(2.0 * d / (d * d + eps * eps) ** 1.5, rel=1e-14)
        assert g[0, 3] == pytest.approx(-2.0 / np.sqrt(d * d + eps * eps) - 1.0 / eps, rel=1e-14)


@pytest.mark.parametrize("mode", [0, 1])
def test_gravity_fp64_brute_force_and_fp32(orc, cs, mode):
    n = 300
    p = cs.preset_params(2, n)
    p.grav_mode = mode
    pos, vel, hyd = cs.generate_ic(2, n, 4)
    out = {}
    for bits in (32, 64):
        o = orc.Oracle(bits, p)
        o.load_ic(pos, vel, hyd)
        o.run_pass(orc.GRAVITY)
        out[bits] = o.get("grav")[:n].astype(np.float64)
    x = pos[:, :3]
    m = vel[:, 3]
    if mode == 1:
        dr = x[:, None] - x[None, :]
        d2 = (dr**2).sum(-1) + p.grav_eps**2
        a = -(m[None, :, None] * dr / d2[..., None] ** 1.5).sum(1)
        phi = -(m[None, :] / np.sqrt(d2)).sum(1)
        np.testing.assert_allclose(out[64][:, :3], a, rtol=1e-11, atol=1e-13 * np.abs(a).max())
        np.testing.assert_allclose(out[64][:, 3], phi, rtol=1e-12)
    rel = np.linalg.norm(out[32][:, :3] - out[64][:, :3], axis=1) / np.linalg.norm(out[64][:, :3], axis=1)
    assert np.median(rel) < 1e-6 and rel.max() < 1e-4


def test_dreal_two_sum_is_error_free(orc):
    rng = np.random.default_rng(2)
    for _ in range(2000):
        hi = np.float32(rng.normal() * 10.0 ** rng.integers(-3, 3))
        lo = np.float32(hi * np.float32(rng.uniform(-1, 1)) * np.float32(2**-25))
        y = np.float32(rng.normal() * 10.0 ** rng.integers(-3, 3))
        out = orc.dreal_sub([hi, lo], y)
        exact = float(hi) + float(lo) - float(y)
        assert abs(float(out[0]) + float(out[1]) - exact) <= 2.0**-44 * max(abs(exact), abs(float(hi)), 1e-30)
        assert abs(float(out[1])) <= abs(float(out[0])) * 2.0**-23 + 1e-38


# ---- integrator, timestep, statistics (step.cpp, compute_timestep.cpp) ------

def state_oracle(cs, orc, n=64, bits=64):
    p = cs.default_params(n)
    o = orc.Oracle(bits, p)
    rng = np.random.default_rng(6)
    pos = np.zeros((p.n_dev, 4))
    vel = np.zeros((p.n_dev, 4))
    hyd = np.zeros((p.n_dev, 4))
    bd = np.zeros((p.n_dev, 4))
    pos[:n] = np.c_[rng.normal(size=(n, 3)), rng.uniform(0.1, 0.2, n)]
    vel[:n] = np.c_[rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n) / n]
    hyd[:n] = np.c_[rng.uniform(0.5, 1, n), rng.uniform(0.1, 1, n), rng.uniform(0.5, 2, n), np.full(n, 0.6)]
    bd[:n] = rng.normal(size=(n, 4))
    for nm, a in (("pos", pos), ("vel", vel), ("hydro", hyd), ("bdots", bd)):
        o.set(nm, a)
    return o, pos[:n], vel[:n], hyd[:n], bd[:n]


def test_predictor_corrector_closed_form(cs, orc):
    o, x, v, h, a = state_oracle(cs, orc)
    dt = 0.01
    o.set_scalar("dt", dt)
    o.run_pass(orc.PREDICT)
    n = len(x)
    np.testing.assert_allclose(o.get("pos")[:n, :3], x[:, :3] + v[:, :3] * dt, rtol=1e-15)
    np.testing.assert_allclose(o.get("vel")[:n, :3], v[:, :3] + a[:, :3] * dt / 2, rtol=1e-15)
    np.testing.assert_allclose(o.get("hydro")[:n, 2], h[:, 2] + a[:, 3] * dt / 2, rtol=1e-15)
    bd = o.get("bdots")
    bd[:n] = a          # constant acceleration through the step
    o.set("bdots", bd)
    o.run_pass(orc.CORRECT)  # dt_new = dt_old = dt: kick-drift-kick with constant a
    np.testing.assert_allclose(o.get("vel")[:n, :3], v[:, :3] + a[:, :3] * dt, rtol=1e-14)
    np.testing.assert_allclose(o.get("hydro")[:n, 2], h[:, 2] + a[:, 3] * dt, rtol=1e-14)


def test_timestep_formula(cs, orc):
    o, x, v, h, a = state_oracle(cs, orc)
    h2 = o.get("hydro2")
    n = len(x)
    h2[:n, 0] = np.random.default_rng(8).uniform(0.5, 3, n)
    o.set("hydro2", h2)
    o.run_pass(orc.TIMESTEP)
    dt_v = 0.3 * np.abs(x[:, 3] / h2[:n, 0])
    dt_e = 0.2 * np.abs(h[:, 2] / a[:, 3])
    assert o.scalar("dt") == pytest.approx(np.min(1 / (1 / dt_v + 1 / dt_e)), rel=1e-7)


def test_statistics_formula(cs, orc):
    o, x, v, h, a = state_oracle(cs, orc)
    g = o.get("grav")
    n = len(x)
    g[:n, 3] = -np.linspace(1, 2, n)
    o.set("grav", g)
    o.run_pass(orc.STATS)
    st = o.stats()
    m = v[:, 3]
    T = 0.5 * (m * (v[:, :3] ** 2).sum(1)).sum()
    U = (m * h[:, 2]).sum()
    W = 0.5 * (m * g[:n, 3]).sum()
    J = np.linalg.norm((m[:, None] * np.cross(x[:, :3], v[:, :3])).sum(0))
    assert (st["T"], st["U"], st["W"], st["J"]) == pytest.approx((T, U, W, J), rel=1e-12)
    assert st["E"] == pytest.approx(T + U + W, rel=1e-12)


# ---- Peano-Hilbert keys (sort_bodies.cpp) ------------------------------------

@pytest.mark.parametrize("bits", [1, 2, 3, 4])
def test_ph_key_is_a_hilbert_curve(orc, bits):
    n = 1 << bits
    keys = {}
    for x in range(n):
        for y in range(n):
            for z in range(n):
                keys[orc.ph_key(x, y, z, bits)] = (x, y, z)
    assert sorted(keys) == list(range(n**3)), "bijection onto 0 .. 8^bits - 1"
    for k in range(n**3 - 1):
        a, b = np.array(keys[k]), np.array(keys[k + 1])
        assert np.abs(a - b).sum() == 1, f"keys {k},{k+1} are not face neighbours"
