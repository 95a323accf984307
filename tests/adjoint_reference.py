"""FP64 CPU restatement, in PyTorch, of the map the engine's adjoint
differentiates: init() and K steps of the FP64 form with the smoothing ranges
and timesteps of every force evaluation given (held fixed), then the loss
    L = sum_i m_i |v_i|^2 / 2 + sum_i rho_i^2.
torch.autograd through it gives the reference gradient for the GPU adjoint.

Test infrastructure only.  The formulas restate the reference's passes:
kernel family dev_sph_kernels.cu:7-99, density dev_density_loop.cu:29-111,
EOS dev_compute_pressure.cu:83-111 (with the FP64 Newton root of the quartic),
viscosity dev_art_visc.cu:8-32, hydro dev_hydro_dots.cu:25-113 and the
reaction sph_accelerations.cpp:104-115, predictor/corrector step.cpp:13-147;
gravity is the Plummer-softened direct sum.  Neighbours are found by brute
force (|x_i - x_j| < range_i), which is the walk's predicate.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

INV_PI = 0.318309886183791     # dev_sph_kernels.cu:5
FOUR_PI = 12.5663706143592     # dev_sph_kernels.cu:19


def _pw(u, lo, hi):
    z = torch.zeros_like(u)
    return torch.where(u < 1, lo, torch.where(u < 2, hi, z))


def kern_w(u):
    t = (2 - u).clamp(min=0)
    return _pw(u, INV_PI * (1 - 1.5 * u * u + 0.75 * u ** 3), INV_PI * 0.25 * t ** 3)


def kern_dw(u):      # W'(u)/u; u > 0 on the second branch
    t = (2 - u).clamp(min=0)
    return _pw(u, INV_PI * (-3 + 2.25 * u), -INV_PI * 0.75 * t * t / u)


def kern_dwdh(u):
    return _pw(u, INV_PI * (-3 + 7.5 * u * u - 4.5 * u ** 3), INV_PI * (-6 + 12 * u - 7.5 * u * u + 1.5 * u ** 3))


def kern_dphidh(u):
    return _pw(u, 1.4 - 2 * u * u + 1.5 * u ** 4 - 0.6 * u ** 5,
               1.6 - 4 * u * u + 4 * u ** 3 - 1.5 * u ** 4 + 0.2 * u ** 5)


def kern_diw(y):
    return FOUR_PI * y * y * kern_w(y)


def kern_digdh(u):
    return _pw(u, 4 * u * kern_diw(4 * u), -4 * u * kern_diw(4 * (2 - u)))


def kern_dig(u):     # u > 0
    return _pw(u, 4 / u * kern_diw(4 * u), -4 / u * kern_diw(4 * (2 - u)))


@dataclass
class MapParams:
    alpha: float
    beta: float
    eps: float
    density_unit: float
    energy_unit: float
    pressure_unit: float
    amu: float
    boltzmann: float
    rad_const: float
    gamma: float
    temp_base: float

    @staticmethod
    def of(p) -> "MapParams":
        f32 = lambda v: float(np.float32(v))
        return MapParams(f32(p.av_alpha), f32(p.av_beta), f32(p.grav_eps), p.eos_density_unit,
                         p.eos_energy_unit, p.eos_pressure_unit, p.eos_amu, p.eos_boltzmann, p.eos_rad_const,
                         p.eos_gamma, p.eos_temp_base)


def quartic_root(q, r):
    """x > 0 with x^4 + q x + r = 0 (q >= 0, r < 0): Newton from above,
    then one differentiable Newton step, whose derivative is the implicit
    one, -(x dq + dr) / (4 x^3 + q)."""
    with torch.no_grad():
        x = torch.minimum((-r) ** 0.25, torch.where(q > 0, -r / q, torch.full_like(q, np.inf)))
        for _ in range(200):
            x = x - (x ** 4 + q * x + r) / (4 * x ** 3 + q)
    return x - (x ** 4 + q * x + r) / (4 * x ** 3 + q)


def pressure(rho, u, mu, c: MapParams):
    density, energy, mean_mu = rho * c.density_unit, u * c.energy_unit, mu * c.amu
    tb = c.temp_base
    r = -density * (energy / (tb * tb)) / (c.rad_const * tb * tb)
    q = c.boltzmann / c.rad_const / (c.gamma - 1.0) * density / (mean_mu * tb * tb * tb)
    temp = tb * quartic_root(q, r)
    return (density * (c.boltzmann * temp / mean_mu) + c.rad_const / 3.0 * temp ** 4) / c.pressure_unit


def pairs(x, rng):
    """(i, j) with |x_i - x_j| < range_i, i != j (self terms are added separately)."""
    with torch.no_grad():
        d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
        m = d2 < rng[:, None] ** 2
        m.fill_diagonal_(False)
        return m.nonzero(as_tuple=True)


def forces(x, v, u, m, mu, rng, c: MapParams):
    """(a, udot, rho) of one force evaluation: hydro + reaction + gravity."""
    n = x.shape[0]
    I, J = pairs(x, rng)
    r = x[I] - x[J]
    dv = v[I] - v[J]
    d = torch.sqrt((r * r).sum(1))
    h = 0.5 * rng
    hI = h[I]
    q = d / hI
    mI, mJ = m[I], m[J]
    zero = torch.zeros(n, dtype=x.dtype)
    u0 = torch.zeros(1, dtype=x.dtype)

    def isum(vals):
        return zero.index_add(0, I, vals)

    # density, grad-h terms, div v, curl v (self terms: W(0), dwdh(0), dphidh(0), digdh(0) = 0)
    rho_s = isum(mJ * kern_w(q)) + m * kern_w(u0)
    om = isum(mJ * kern_dwdh(q)) + m * kern_dwdh(u0)
    psi = isum(mJ * kern_dphidh(q)) + m * kern_dphidh(u0)
    den = -isum(kern_digdh(q))
    mdw = mI * kern_dw(q)
    rdv = (r * dv).sum(1)
    divv = isum(mdw * rdv)
    cr = torch.cross(r, dv, dim=1) * mdw[:, None]
    curl = torch.zeros(n, 3, dtype=x.dtype).index_add(0, I, cr)
    h3, h5 = h ** 3, h ** 5
    Om = torch.where(den != 0, om / h3 / den, zero)
    Ps = torch.where(den != 0, psi / h / den, zero)
    rho = rho_s / h3
    D = divv / rho / h5
    C = torch.sqrt((curl * curl).sum(1)) / rho / h5
    P = pressure(rho, u, mu, c)
    # hydro with viscosity
    prho2 = P / (rho * rho)
    cs = torch.sqrt(5.0 / 3.0 * P / rho)
    bal = D.abs() / ((D.abs() + C) + 0.0001 * cs / h)
    csum = cs[I] + cs[J]
    mu_ = rdv / (csum * d) * (bal[I] + bal[J])
    Pi = torch.where(rdv < 0, (prho2[I] + prho2[J]) * (-c.alpha * mu_ + c.beta * mu_ * mu_),
                     torch.zeros_like(rdv))
    w1 = kern_dw(q)
    g = kern_dig(q)
    pv = prho2[I] + 0.5 * Pi
    h2I = hI * hI
    h5I = h5[I]
    A = pv * w1 / h5I - (prho2[I] * Om[I] + 0.5 * Ps[I]) / mJ * g / h2I
    acc = -torch.zeros(n, 3, dtype=x.dtype).index_add(0, I, (mJ * A)[:, None] * r)
    acc = acc + torch.zeros(n, 3, dtype=x.dtype).index_add(0, J, (mI * A)[:, None] * r)
    udot = isum((mJ * pv * w1 / h5I - prho2[I] * Om[I] * g / h2I) * rdv)
    # Plummer gravity, all pairs (the self pair has r = 0)
    R = x[:, None, :] - x[None, :, :]
    s2 = (R * R).sum(-1) + c.eps * c.eps
    acc = acc - ((m[None, :] / s2 ** 1.5)[:, :, None] * R).sum(1)
    return acc, udot, rho


def run_map(x0, v0, u0, m, mu, ranges, dts, c: MapParams, return_state=False):
    """init() + K = len(dts) - 1 steps; ranges[f], dts[f] of force evaluation f."""
    K = len(dts) - 1
    a, udot, rho = forces(x0, v0, u0, m, mu, ranges[0], c)
    x, v, u = x0, v0, u0
    for s in range(1, K + 1):
        dtc = float(dts[s - 1])
        dth = 0.5 * dtc
        xp = x + v * dtc
        vs = v + a * dth
        us = u + udot * dth
        a, udot, rho = forces(xp, vs, us, m, mu, ranges[s], c)
        f = 0.5 * float(dts[s]) + 0.5 * dtc
        v = v + a * f
        u = u + udot * f
        x = xp
    loss = (0.5 * m * (v * v).sum(1)).sum() + (rho * rho).sum()
    if return_state:
        return loss, (x, v, u, rho)
    return loss


def gradient(pos, vel, hyd, ranges, dts, params):
    """Loss and (dL/dx0, dL/dv0, dL/du0) as numpy arrays."""
    c = MapParams.of(params)
    t = lambda a: torch.tensor(np.asarray(a, dtype=np.float64))
    x0 = t(pos[:, :3]).requires_grad_(True)
    v0 = t(vel[:, :3]).requires_grad_(True)
    u0 = t(hyd[:, 2]).requires_grad_(True)
    rng = [t(r) for r in ranges]
    loss = run_map(x0, v0, u0, t(vel[:, 3]), t(hyd[:, 3]), rng, dts, c)
    loss.backward()
    # an input the loss does not depend on (u0 when K = 0) has no .grad
    g = lambda a: np.zeros(tuple(a.shape)) if a.grad is None else a.grad.numpy()
    return loss.item(), g(x0), g(v0), g(u0)
