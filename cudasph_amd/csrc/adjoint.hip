// Reverse-mode (adjoint) form of the FP64 engine.
//
// The differentiated map is one run of the engine from an initial state
// (x0, v0, u0) through init() and K step()s, to the loss
//     L = sum_i 1/2 m_i |v_i|^2 + sum_i rho_i^2
// over the final state.  The smoothing ranges, the per-evaluation timestep
// and the neighbour-count estimates are held at the values the recorded
// forward run produced ("frozen" map), so the map is differentiable: list
// membership still follows the positions, but every pair term of the M4
// family vanishes with its first derivative at u = 2, so a pair entering or
// leaving a list changes nothing to first order.
//
// Forward (adjoint_forward): the ordinary FP64 step with the ranges, counts,
// timesteps and gravity of every force evaluation recorded, and the state
// checkpointed every k steps.
// Backward (adjoint_backward): segments of k steps in reverse; each segment is
// recomputed once from its checkpoint with the recorded ranges, timesteps and
// gravity (no N^2 pass), capturing every step's force-evaluation data (the
// state at the evaluation, the neighbour lists and their sorted transpose),
// and then its steps are reversed from last to first:
//   correct:  a'_bar += v'_bar f, udot'_bar += u'_bar f, f = dt_new/2 + dt_old/2
//   forces:   hydro pairs -> transpose gather -> EOS -> density pairs ->
//             transpose gather -> Plummer gravity (one FP64 N^2 pass)
//   predict:  x_bar = x'_bar,  v_bar = v'_bar + v*_bar + x_bar dt,
//             u_bar = u'_bar + u*_bar,  (a, udot)_bar = (v*, u*)_bar dt/2
// The per-pair j-side adjoints are written to one record per list slot and
// summed per j through the forward's sorted (neighbour, slot) pairs, in
// ascending slot order: the backward is deterministic, like the reaction pass.
// With adj_implicit_h = 1 the ranges are not held: each evaluation adds the
// derivative of its h-solve (see k_adj_range_hydro below), so the gradient is
// that of the map whose ranges are solved from the positions (timesteps and
// counts still recorded).
#include <stdexcept>

#include "engine_impl.h"

namespace cudasph {
namespace {

using D4v = V4<double>;
constexpr int kPair = 10;   // doubles per list slot in the pair record

// ---- derivatives of the M4 family, as f'(u)/u ------------------------------
// (kern_* are the forward functions of sph_math.h; kern_dw(u) = W'(u)/u)

__device__ __forceinline__ double ddw_u(double u) {       // (dw)'(u)/u, u > 0
  if (u < 1.0) return kInvPi * 2.25 / u;
  if (u < 2.0) return kInvPi * 0.75 * (2.0 - u) * (2.0 + u) / (u * u * u);
  return 0.0;
}
__device__ __forceinline__ double ddwdh_u(double u) {     // (dwdh)'(u)/u
  if (u < 1.0) return kInvPi * (15.0 - 13.5 * u);
  if (u < 2.0) return kInvPi * (12.0 - 15.0 * u + 4.5 * u * u) / u;
  return 0.0;
}
__device__ __forceinline__ double ddphidh_u(double u) {   // (dphidh)'(u)/u
  if (u < 1.0) return -4.0 + 6.0 * u * u - 3.0 * u * u * u;
  if (u < 2.0) return -8.0 + 12.0 * u - 6.0 * u * u + u * u * u;
  return 0.0;
}
// diw(y) = 4 pi y^2 W(y)  ->  diw'(y) = 4 pi y (2 W(y) + y^2 dw(y))
__device__ __forceinline__ double ddiw(double y) {
  return kFourPi * y * (2.0 * kern_w<double>(y) + y * y * kern_dw<double>(y));
}
__device__ __forceinline__ double ddigdh(double u) {      // (digdh)'(u)
  if (u <= 0.0) return 0.0;
  if (u < 1.0) return 4.0 * kern_diw<double>(4.0 * u) + 16.0 * u * ddiw(4.0 * u);
  if (u < 2.0) {
    const double y = 4.0 * (2.0 - u);
    return -4.0 * kern_diw<double>(y) + 16.0 * u * ddiw(y);
  }
  return 0.0;
}
// dig(u) = digdh(u)/u^2  ->  dig'(u)/u = digdh'/u^3 - 2 digdh/u^4, u > 0
__device__ __forceinline__ double ddig_u(double u) {
  if (u <= 0.0 || u >= 2.0) return 0.0;
  const double u2 = u * u;
  return ddigdh(u) / (u2 * u) - 2.0 * kern_digdh<double>(u) / (u2 * u2);
}
__device__ __forceinline__ double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

// ---- hydro + artificial viscosity + reaction, reversed ---------------------
// Forward per pair (i, j in list(i), j != i):  r = x_i - x_j, dv = v_i - v_j,
//   A = (P_i/rho_i^2 + Pi/2) dw(u)/h^5 - (P_i/rho_i^2 Omega_i + Psi_i/2) dig(u)/(m_j h^2)
//   a_i -= m_j A r,  a_j += m_i A r,
//   udot_i += (m_j (P_i/rho_i^2 + Pi/2) dw/h^5 - P_i/rho_i^2 Omega_i dig/h^2) (r.dv)
// Given (a_bar, udot_bar) per particle this writes the i-side adjoints of
// x, v, rho, P, divv, curl, Omega, Psi and the j-side ones (x, v, rho, P,
// divv, curl) of every slot.
__global__ __launch_bounds__(kBlock) void k_adj_hydro(const D4v* __restrict__ pos, const D4v* __restrict__ vel,
                                                       const D4v* __restrict__ hydro, const D4v* __restrict__ dots,
                                                       const D4v* __restrict__ abar, const int* __restrict__ offset,
                                                       const int* __restrict__ count, const int* __restrict__ list,
                                                       D4v* __restrict__ xs, D4v* __restrict__ vs,
                                                       D4v* __restrict__ hb1, D4v* __restrict__ hb2,
                                                       double* __restrict__ pb, int n, double alpha, double beta) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const D4v bp = pos[i], bv = vel[i], hd = hydro[i], dd = dots[i], ai = abar[i];
  const double mi = bv.w, h = 0.5 * bp.w, h2 = h * h, h5 = h2 * h2 * h;
  const double rho = hd.x, P = hd.y, prho2 = P / (rho * rho);
  const double Om = dd.z, Ps = dd.w, Di = dd.x, Ci = dd.y;
  const double ci = sqrt(5.0 / 3.0 * P / rho);
  const double deni = (fabs(Di) + Ci) + 0.0001 * ci / h;
  const double bi = fabs(Di) / deni;
  double gx = 0, gy = 0, gz = 0, gvx = 0, gvy = 0, gvz = 0;
  double gP = 0, grho = 0, gOm = 0, gPs = 0, gD = 0, gC = 0, gprho2 = 0;
  const int off = offset[i], cnt = count[i];
  for (int k = off; k < off + cnt; ++k) {
    const int j = list[k];
    double* out = pb + (size_t)k * kPair;
    if (j == i) {   // r = 0 and dv = 0 identically: no dependence
      for (int c = 0; c < kPair; ++c) out[c] = 0.0;
      continue;
    }
    const D4v pj = pos[j], vj = vel[j], hj = hydro[j], dj = dots[j], aj = abar[j];
    const double mj = vj.w;
    const double rx = bp.x - pj.x, ry = bp.y - pj.y, rz = bp.z - pj.z;
    const double dvx = bv.x - vj.x, dvy = bv.y - vj.y, dvz = bv.z - vj.z;
    const double ds = sqrt(rx * rx + ry * ry + rz * rz);
    const double u = ds / h;
    const double drdv = rx * dvx + ry * dvy + rz * dvz;
    // artificial viscosity, forward (sph_math.h art_visc)
    const double rhoj = hj.x, Pj = hj.y, prho2j = Pj / (rhoj * rhoj);
    const double cj = sqrt(5.0 / 3.0 * Pj / rhoj);
    const double hjw = 0.5 * pj.w;
    const bool av = drdv < 0.0;
    double Pi = 0, mu0 = 0, mu = 0, S = 0, bj = 0, denj = 0;
    const double cs = ci + cj;
    if (av) {
      mu0 = drdv / (cs * ds);
      denj = (fabs(dj.x) + dj.y) + 0.0001 * cj / hjw;
      bj = fabs(dj.x) / denj;
      mu = mu0 * (bi + bj);
      S = prho2 + prho2j;
      Pi = S * (-alpha * mu + beta * (mu * mu));
    }
    const double w1 = kern_dw<double>(u), g = kern_dig<double>(u);
    const double pv = prho2 + 0.5 * Pi;
    const double A = pv * w1 / h5 - (prho2 * Om + 0.5 * Ps) / mj * g / h2;
    const double Q = mj * pv * w1 / h5 - prho2 * Om * g / h2;
    // reverse
    const double bx = mi * aj.x - mj * ai.x, by = mi * aj.y - mj * ai.y, bz = mi * aj.z - mj * ai.z;
    const double Ab = bx * rx + by * ry + bz * rz;
    double rbx = A * bx, rby = A * by, rbz = A * bz;
    const double Qb = ai.w * drdv;
    double drdvb = ai.w * Q;
    const double sq = Ab + Qb * mj;
    const double pvb = sq * w1 / h5;
    const double w1b = sq * pv / h5;
    const double gb = -Ab * (prho2 * Om + 0.5 * Ps) / (mj * h2) - Qb * prho2 * Om / h2;
    const double t = (Ab / mj + Qb) * g / h2;
    gprho2 += pvb - t * Om;
    gOm -= t * prho2;
    gPs -= 0.5 * Ab * g / (mj * h2);
    if (ds > 0.0) {
      const double coef = (w1b * ddw_u(u) + gb * ddig_u(u)) / h2;
      rbx += coef * rx;
      rby += coef * ry;
      rbz += coef * rz;
    }
    double jP = 0, jrho = 0, jD = 0, jC = 0;
    if (av) {
      const double Pib = 0.5 * pvb;
      const double Sb = Pib * (-alpha * mu + beta * mu * mu);
      const double mub = Pib * S * (-alpha + 2.0 * beta * mu);
      const double mu0b = mub * (bi + bj);
      const double bb = mub * mu0;
      drdvb += mu0b / (cs * ds);
      const double csb = -mu0b * mu0 / cs;
      const double dsb = -mu0b * mu0 / ds;
      rbx += dsb * rx / ds;
      rby += dsb * ry / ds;
      rbz += dsb * rz / ds;
      double cib = csb, cjb = csb;
      if (deni > 0.0) {
        const double i2 = 1.0 / (deni * deni);
        gD += bb * sgn(Di) * (Ci + 0.0001 * ci / h) * i2;
        gC -= bb * fabs(Di) * i2;
        cib -= bb * fabs(Di) * 0.0001 / h * i2;
      }
      if (denj > 0.0) {
        const double j2 = 1.0 / (denj * denj);
        jD += bb * sgn(dj.x) * (dj.y + 0.0001 * cj / hjw) * j2;
        jC -= bb * fabs(dj.x) * j2;
        cjb -= bb * fabs(dj.x) * 0.0001 / hjw * j2;
      }
      // c = sqrt(5/3 P/rho): dc/dP = 5/(6 c rho), dc/drho = -c/(2 rho)
      if (ci > 0.0) {
        gP += cib * 5.0 / (6.0 * ci * rho);
        grho -= cib * ci / (2.0 * rho);
      }
      if (cj > 0.0) {
        jP += cjb * 5.0 / (6.0 * cj * rhoj);
        jrho -= cjb * cj / (2.0 * rhoj);
      }
      gprho2 += Sb;
      jP += Sb / (rhoj * rhoj);
      jrho -= 2.0 * Sb * Pj / (rhoj * rhoj * rhoj);
    }
    rbx += drdvb * dvx;
    rby += drdvb * dvy;
    rbz += drdvb * dvz;
    const double vbx = drdvb * rx, vby = drdvb * ry, vbz = drdvb * rz;
    gx += rbx;
    gy += rby;
    gz += rbz;
    gvx += vbx;
    gvy += vby;
    gvz += vbz;
    out[0] = -rbx;
    out[1] = -rby;
    out[2] = -rbz;
    out[3] = -vbx;
    out[4] = -vby;
    out[5] = -vbz;
    out[6] = jrho;
    out[7] = jP;
    out[8] = jD;
    out[9] = jC;
  }
  gP += gprho2 / (rho * rho);
  grho -= 2.0 * gprho2 * P / (rho * rho * rho);
  xs[i] = {gx, gy, gz, 0.0};
  vs[i] = {gvx, gvy, gvz, 0.0};
  hb1[i] = {grho, gP, 0.0, 0.0};
  hb2[i] = {gD, gC, gOm, gPs};
}

// j-side sums: for each j, the slots whose neighbour is j, in ascending slot
// order (the forward's sorted pairs); ncomp = 10 after hydro, 6 after density
__global__ __launch_bounds__(kBlock) void k_adj_gather(const KV* __restrict__ pairs, const int* __restrict__ start,
                                                        const int* __restrict__ end, const double* __restrict__ pb,
                                                        D4v* __restrict__ xs, D4v* __restrict__ vs,
                                                        D4v* __restrict__ hb1, D4v* __restrict__ hb2, int n,
                                                        int ncomp) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  double s[kPair] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int p = start[j]; p < end[j]; ++p) {
    const double* r = pb + (size_t)pairs[p].value * kPair;
    for (int c = 0; c < kPair; ++c)
      if (c < ncomp) s[c] += r[c];
  }
  D4v x = xs[j], v = vs[j];
  x.x += s[0];
  x.y += s[1];
  x.z += s[2];
  v.x += s[3];
  v.y += s[4];
  v.z += s[5];
  xs[j] = x;
  vs[j] = v;
  if (ncomp == kPair) {
    D4v a = hb1[j], b = hb2[j];
    a.x += s[6];
    a.y += s[7];
    b.x += s[8];
    b.y += s[9];
    hb1[j] = a;
    hb2[j] = b;
  }
}

// EOS reversed: rho_bar += P_bar dP/drho (+ 2 rho for the loss seed),
// u_bar = P_bar dP/du, through the implicit derivative of the quartic root
// x^4 + q x + r = 0:  dx = -(x dq + dr) / (4 x^3 + q)
__global__ __launch_bounds__(kBlock) void k_adj_eos(const D4v* __restrict__ hydro, D4v* __restrict__ hb1, int n,
                                                     EosConst c, int seed_rho) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const D4v hd = hydro[i];
  D4v b = hb1[i];
  const double density = hd.x * c.density_unit, energy = hd.z * c.energy_unit, mean_mu = hd.w * c.amu;
  const double tb = c.temp_base, a4 = c.rad_const * tb * tb * tb * tb;
  const double r = -density * (energy / (tb * tb)) / (c.rad_const * tb * tb);
  const double kq = c.boltzmann / c.rad_const / (c.gamma - 1.0) / (mean_mu * tb * tb * tb);
  const double q = kq * density;
  const double x = quartic_newton(q, r);
  const double temp = tb * x;
  const double fx = 4.0 * x * x * x + q;
  double dx_dd = 0.0, dx_de = 0.0;
  if (fx > 0.0) {
    dx_dd = -(x * kq - energy / a4) / fx;   // dr/ddensity = -energy / (a tb^4)
    dx_de = (density / a4) / fx;            // dr/denergy = -density / (a tb^4)
  }
  const double dp_dt = (density * c.boltzmann / mean_mu + 4.0 / 3.0 * c.rad_const * temp * temp * temp);
  const double dp_dd = (c.boltzmann * temp / mean_mu + dp_dt * tb * dx_dd) / c.pressure_unit;
  const double dp_de = dp_dt * tb * dx_de / c.pressure_unit;
  b.x += b.y * dp_dd * c.density_unit + (seed_rho ? 2.0 * hd.x : 0.0);
  b.z = b.y * dp_de * c.energy_unit;
  hb1[i] = b;
}

// density, Omega, Psi, div v and |curl v| reversed (forward: k_density)
__global__ __launch_bounds__(kBlock) void k_adj_density(const D4v* __restrict__ pos, const D4v* __restrict__ vel,
                                                         const D4v* __restrict__ hydro, const D4v* __restrict__ dots,
                                                         const D4v* __restrict__ hb1, const D4v* __restrict__ hb2,
                                                         const int* __restrict__ offset, const int* __restrict__ count,
                                                         const int* __restrict__ list, D4v* __restrict__ xs,
                                                         D4v* __restrict__ vs, double* __restrict__ pb, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const D4v bp = pos[i], bv = vel[i];
  const double mi = bv.w, h = 0.5 * bp.w, h2 = h * h, h3 = h2 * h, h5 = h3 * h2;
  const int off = offset[i], cnt = count[i];
  // raw sums of the forward
  double den = 0, cx = 0, cy = 0, cz = 0;
  for (int k = off; k < off + cnt; ++k) {
    const int j = list[k];
    const D4v pj = pos[j], vj = vel[j];
    const double rx = bp.x - pj.x, ry = bp.y - pj.y, rz = bp.z - pj.z;
    const double u = 2.0 * sqrt(rx * rx + ry * ry + rz * rz) / bp.w;
    den -= kern_digdh<double>(u);
    const double dvx = bv.x - vj.x, dvy = bv.y - vj.y, dvz = bv.z - vj.z;
    const double mdw = mi * kern_dw<double>(u);
    cx += mdw * (ry * dvz - rz * dvy);
    cy += mdw * (rz * dvx - rx * dvz);
    cz += mdw * (rx * dvy - ry * dvx);
  }
  const D4v hd = hydro[i], dd = dots[i], b1 = hb1[i], b2 = hb2[i];
  const double rho = hd.x;
  double rhob = b1.x;
  const double sdb = b2.x / (rho * h5);
  rhob -= b2.x * dd.x / rho;
  const double cn = sqrt(cx * cx + cy * cy + cz * cz);
  double cbx = 0, cby = 0, cbz = 0;
  if (cn > 0.0) {
    const double f = b2.y / (cn * rho * h5);
    cbx = f * cx;
    cby = f * cy;
    cbz = f * cz;
  }
  rhob -= b2.y * dd.y / rho;
  double omb = 0, psib = 0, denb = 0;
  if (den != 0.0) {
    omb = b2.z / (h3 * den);
    psib = b2.w / (h * den);
    denb = -(b2.z * dd.z + b2.w * dd.w) / den;
  }
  const double rb3 = rhob / h3;
  double gx = 0, gy = 0, gz = 0, gvx = 0, gvy = 0, gvz = 0;
  for (int k = off; k < off + cnt; ++k) {
    const int j = list[k];
    double* out = pb + (size_t)k * kPair;
    if (j == i) {
      for (int c = 0; c < 6; ++c) out[c] = 0.0;
      continue;
    }
    const D4v pj = pos[j], vj = vel[j];
    const double mj = vj.w;
    const double rx = bp.x - pj.x, ry = bp.y - pj.y, rz = bp.z - pj.z;
    const double dvx = bv.x - vj.x, dvy = bv.y - vj.y, dvz = bv.z - vj.z;
    const double ds = sqrt(rx * rx + ry * ry + rz * rz);
    const double u = 2.0 * ds / bp.w;
    const double w = kern_dw<double>(u);
    double s = mj * (rb3 * w + omb * ddwdh_u(u) + psib * ddphidh_u(u)) - denb * (u > 0.0 ? ddigdh(u) / u : 0.0);
    if (ds > 0.0) {
      const double rdv = rx * dvx + ry * dvy + rz * dvz;
      const double crs = cbx * (ry * dvz - rz * dvy) + cby * (rz * dvx - rx * dvz) + cbz * (rx * dvy - ry * dvx);
      s += mi * ddw_u(u) * (sdb * rdv + crs);
    }
    s /= h2;
    const double mw = mi * w;
    // d/dr [C_bar . (r x dv)] = dv x C_bar;  d/d(dv) = C_bar x r
    const double rbx = s * rx + mw * (sdb * dvx + (dvy * cbz - dvz * cby));
    const double rby = s * ry + mw * (sdb * dvy + (dvz * cbx - dvx * cbz));
    const double rbz = s * rz + mw * (sdb * dvz + (dvx * cby - dvy * cbx));
    const double vbx = mw * (sdb * rx + (cby * rz - cbz * ry));
    const double vby = mw * (sdb * ry + (cbz * rx - cbx * rz));
    const double vbz = mw * (sdb * rz + (cbx * ry - cby * rx));
    gx += rbx;
    gy += rby;
    gz += rbz;
    gvx += vbx;
    gvy += vby;
    gvz += vbz;
    out[0] = -rbx;
    out[1] = -rby;
    out[2] = -rbz;
    out[3] = -vbx;
    out[4] = -vby;
    out[5] = -vbz;
  }
  D4v x = xs[i], v = vs[i];
  x.x += gx;
  x.y += gy;
  x.z += gz;
  v.x += gvx;
  v.y += gvy;
  v.z += gvz;
  xs[i] = x;
  vs[i] = v;
}

// Plummer gravity reversed: a_i = -sum_j m_j r_ij / s^3, s^2 = |r_ij|^2 + eps^2
//   x_bar_i += sum_j T_ij (m_i a_bar_j - m_j a_bar_i),  T = I/s^3 - 3 r r^T / s^5
// One FP64 N^2 pass: j-tiles of (x, m, a_bar) staged in LDS, one i per lane,
// the pair body in fused multiply-adds.  As in the forward FP64 kernel, with
// few i-blocks the j-range is split into gravity64_parts(n) parts whose sums
// (in `part`) are added in part order.
constexpr int kAdjTile = kBlock;
__global__ __launch_bounds__(kBlock) void k_adj_gravity(const D4v* __restrict__ pos, const D4v* __restrict__ vel,
                                                         const D4v* __restrict__ abar, D4v* __restrict__ xs,
                                                         D4v* __restrict__ part, int n, int parts, int chunk,
                                                         double eps2) {
  __shared__ double sx[kAdjTile], sy[kAdjTile], sz[kAdjTile], sm[kAdjTile];
  __shared__ double sax[kAdjTile], say[kAdjTile], saz[kAdjTile];
  const int pt = (int)blockIdx.x % parts, ib = (int)blockIdx.x / parts;
  const int i = ib * kBlock + threadIdx.x;
  D4v pi{}, ai{};
  double mi = 0.0;
  if (i < n) {
    pi = pos[i];
    ai = abar[i];
    mi = vel[i].w;
  }
  const int j_begin = pt * chunk, j_end = min(n, j_begin + chunk);
  double gx = 0, gy = 0, gz = 0;
  for (int j0 = j_begin; j0 < j_end; j0 += kAdjTile) {
    __syncthreads();
    const int j = j0 + threadIdx.x;
    if (j < j_end) {
      const D4v p = pos[j], a = abar[j];
      sx[threadIdx.x] = p.x;
      sy[threadIdx.x] = p.y;
      sz[threadIdx.x] = p.z;
      sm[threadIdx.x] = vel[j].w;
      sax[threadIdx.x] = a.x;
      say[threadIdx.x] = a.y;
      saz[threadIdx.x] = a.z;
    }
    __syncthreads();
    const int m = min(kAdjTile, j_end - j0);
    for (int jj = 0; jj < m; ++jj) {
      const double rx = pi.x - sx[jj], ry = pi.y - sy[jj], rz = pi.z - sz[jj];
      const double d2 = __builtin_fma(rx, rx, __builtin_fma(ry, ry, __builtin_fma(rz, rz, eps2)));
      double inv = __builtin_amdgcn_rsq(d2);
      inv = inv * __builtin_fma(-0.5 * d2, inv * inv, 1.5);   // one Newton step: ~1e-14 relative
      const double inv2 = inv * inv, inv3 = inv2 * inv, inv5 = inv3 * inv2;
      const double mj = sm[jj];
      const double bx = __builtin_fma(mi, sax[jj], -mj * ai.x);
      const double by = __builtin_fma(mi, say[jj], -mj * ai.y);
      const double bz = __builtin_fma(mi, saz[jj], -mj * ai.z);
      const double t = 3.0 * __builtin_fma(rx, bx, __builtin_fma(ry, by, rz * bz)) * inv5;
      gx = __builtin_fma(bx, inv3, __builtin_fma(-t, rx, gx));
      gy = __builtin_fma(by, inv3, __builtin_fma(-t, ry, gy));
      gz = __builtin_fma(bz, inv3, __builtin_fma(-t, rz, gz));
    }
  }
  if (i >= n) return;
  if (parts > 1) {
    part[(size_t)pt * n + i] = {gx, gy, gz, 0.0};
    return;
  }
  D4v x = xs[i];
  x.x += gx;
  x.y += gy;
  x.z += gz;
  xs[i] = x;
}

// the parts' sums in part order, added to x_bar
__global__ __launch_bounds__(kBlock) void k_adj_gravity_combine(const D4v* __restrict__ part, D4v* __restrict__ xs,
                                                                 int n, int parts) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double gx = 0, gy = 0, gz = 0;
  for (int p = 0; p < parts; ++p) {
    const D4v g = part[(size_t)p * n + i];
    gx += g.x;
    gy += g.y;
    gz += g.z;
  }
  D4v x = xs[i];
  x.x += gx;
  x.y += gy;
  x.z += gz;
  xs[i] = x;
}

// corrector reversed: a_bar += v_bar f, udot_bar += u_bar f
__global__ __launch_bounds__(kBlock) void k_adj_correct(D4v* __restrict__ aa, const D4v* __restrict__ av,
                                                         const double* __restrict__ au, double f, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  D4v a = aa[i];
  const D4v v = av[i];
  a.x += v.x * f;
  a.y += v.y * f;
  a.z += v.z * f;
  a.w += au[i] * f;
  aa[i] = a;
}

// predictor reversed (forward: x' = x + v dt, v* = v + a dt/2, u* = u + udot dt/2)
__global__ __launch_bounds__(kBlock) void k_adj_predict(D4v* __restrict__ ax, D4v* __restrict__ av,
                                                         double* __restrict__ au, D4v* __restrict__ aa,
                                                         const D4v* __restrict__ xs, const D4v* __restrict__ vs,
                                                         const D4v* __restrict__ hb1, double dt, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double dth = 0.5 * dt;
  const D4v x0 = ax[i], v0 = av[i], xf = xs[i], vf = vs[i];
  const double uf = hb1[i].z;
  const D4v x{x0.x + xf.x, x0.y + xf.y, x0.z + xf.z, 0.0};
  ax[i] = x;
  av[i] = {v0.x + vf.x + x.x * dt, v0.y + vf.y + x.y * dt, v0.z + vf.z + x.z * dt, 0.0};
  au[i] += uf;
  aa[i] = {vf.x * dth, vf.y * dth, vf.z * dth, uf * dth};
}

// gradient of the initial state: carried adjoints + the init force evaluation's
__global__ __launch_bounds__(kBlock) void k_adj_finish(const D4v* __restrict__ ax, const D4v* __restrict__ av,
                                                        const double* __restrict__ au, const D4v* __restrict__ xs,
                                                        const D4v* __restrict__ vs, const D4v* __restrict__ hb1,
                                                        double* __restrict__ out, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const D4v x = ax[i], v = av[i], a = xs[i], b = vs[i];
  out[3 * i + 0] = x.x + a.x;
  out[3 * i + 1] = x.y + a.y;
  out[3 * i + 2] = x.z + a.z;
  out[3 * (size_t)n + 3 * i + 0] = v.x + b.x;
  out[3 * (size_t)n + 3 * i + 1] = v.y + b.y;
  out[3 * (size_t)n + 3 * i + 2] = v.z + b.z;
  out[6 * (size_t)n + i] = au[i] + hb1[i].z;
}

// loss seed of the final state: v_bar = m v
__global__ __launch_bounds__(kBlock) void k_adj_seed(const D4v* __restrict__ vel, D4v* __restrict__ av, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const D4v v = vel[i];
  av[i] = {v.w * v.x, v.w * v.y, v.w * v.z, 0.0};
}

// loss = sum 1/2 m |v|^2 + sum rho^2: per-block partial sums in a fixed tree
// order, then one block sums the partials in order
__global__ __launch_bounds__(kBlock) void k_adj_loss(const D4v* __restrict__ vel, const D4v* __restrict__ hydro,
                                                      double* __restrict__ partial, int n) {
  __shared__ double s[kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    const D4v w = vel[i];
    const double r = hydro[i].x;
    v = 0.5 * w.w * (w.x * w.x + w.y * w.y + w.z * w.z) + r * r;
  }
  s[threadIdx.x] = v;
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}
__global__ __launch_bounds__(kBlock) void k_adj_loss_final(double* __restrict__ partial, int nb) {
  __shared__ double s[kBlock];
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) v += partial[b];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[nb] = s[0];
}

// x, v or u of real particles from a packed host-layout array (n x 3 or n)
__global__ __launch_bounds__(kBlock) void k_adj_set_xyz(D4v* __restrict__ dst, const double* __restrict__ src,
                                                         int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  D4v d = dst[i];
  d.x = src[3 * i + 0];
  d.y = src[3 * i + 1];
  d.z = src[3 * i + 2];
  dst[i] = d;
}
__global__ __launch_bounds__(kBlock) void k_adj_set_z(D4v* __restrict__ dst, const double* __restrict__ src, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  dst[i].z = src[i];
}

// ---- the ranges' share of the gradient (engine scalar adj_implicit_h) -----
// The frozen map holds each evaluation's ranges r_i; the engine's map solves
// them from the positions: F_i(r_i, x) = sum_j ig(2 |x_ij| / r_i) - target = 0
// (dev_solve_range.cu:17-89).  Its derivative adds, per evaluation,
//   x_bar += r_bar_i dr_i/dx,  dr_i/dx = -(dF_i/dx) / (dF_i/dr_i),
// with r_bar_i = dL/dr_i through everything the evaluation computes from r_i:
// the density outputs (rho, Omega, Psi, div v, |curl v| of i) and the hydro
// pair terms of i's list (u = |x_ij| / h_i, the h^-5 and h^-2 factors, and the
// Balsara switch's 1e-4 c/h of i and, through the viscosity of pair (i, j),
// of j).  Forward derivatives in r_i (h = r / 2, du/dr = -u / r) contracted
// with the adjoints the reverse pass already holds.

// hydro pairs: r_bar_i of i's own terms into rbar[i]; the j-side part (the
// Balsara factor of j) of every slot into pb[k * kPair]
__global__ __launch_bounds__(kBlock) void k_adj_range_hydro(const D4v* __restrict__ pos, const D4v* __restrict__ vel,
                                                             const D4v* __restrict__ hydro,
                                                             const D4v* __restrict__ dots,
                                                             const D4v* __restrict__ abar,
                                                             const int* __restrict__ offset,
                                                             const int* __restrict__ count,
                                                             const int* __restrict__ list, double* __restrict__ rbar,
                                                             double* __restrict__ pb, int n, double alpha,
                                                             double beta) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const D4v bp = pos[i], bv = vel[i], hd = hydro[i], dd = dots[i], ai = abar[i];
  const double mi = bv.w, h = 0.5 * bp.w, h2 = h * h, h5 = h2 * h2 * h;
  const double rho = hd.x, P = hd.y, prho2 = P / (rho * rho);
  const double Om = dd.z, Ps = dd.w, Di = dd.x, Ci = dd.y;
  const double ci = sqrt(5.0 / 3.0 * P / rho);
  const double deni = (fabs(Di) + Ci) + 0.0001 * ci / h;
  const double bi = fabs(Di) / deni;
  const double dh = 0.5;                                   // dh / dr_i
  const double dbi = deni > 0.0 ? fabs(Di) * (0.0001 * ci / (h * h) * dh) / (deni * deni) : 0.0;
  const double dih5 = -5.0 / (h5 * h) * dh, dih2 = -2.0 / (h2 * h) * dh;   // d(h^-5), d(h^-2)
  const double K = prho2 * Om + 0.5 * Ps;
  double rb = 0.0;
  const int off = offset[i], cnt = count[i];
  for (int k = off; k < off + cnt; ++k) {
    const int j = list[k];
    double* out = pb + (size_t)k * kPair;
    if (j == i) {
      out[0] = 0.0;
      continue;
    }
    const D4v pj = pos[j], vj = vel[j], hj = hydro[j], dj = dots[j], aj = abar[j];
    const double mj = vj.w;
    const double rx = bp.x - pj.x, ry = bp.y - pj.y, rz = bp.z - pj.z;
    const double dvx = bv.x - vj.x, dvy = bv.y - vj.y, dvz = bv.z - vj.z;
    const double ds = sqrt(rx * rx + ry * ry + rz * rz);
    const double u = ds / h;
    const double drdv = rx * dvx + ry * dvy + rz * dvz;
    double Pi = 0.0, dPi_i = 0.0, dPi_j = 0.0;
    if (drdv < 0.0) {
      const double rhoj = hj.x, Pj = hj.y, prho2j = Pj / (rhoj * rhoj);
      const double cj = sqrt(5.0 / 3.0 * Pj / rhoj);
      const double hjw = 0.5 * pj.w;
      const double mu0 = drdv / ((ci + cj) * ds);
      const double denj = (fabs(dj.x) + dj.y) + 0.0001 * cj / hjw;
      const double bj = fabs(dj.x) / denj;
      const double dbj = denj > 0.0 ? fabs(dj.x) * (0.0001 * cj / (hjw * hjw) * 0.5) / (denj * denj) : 0.0;
      const double mu = mu0 * (bi + bj), S = prho2 + prho2j;
      Pi = S * (-alpha * mu + beta * (mu * mu));
      const double dPi_dmu = S * (-alpha + 2.0 * beta * mu);
      dPi_i = dPi_dmu * mu0 * dbi;
      dPi_j = dPi_dmu * mu0 * dbj;
    }
    const double w1 = kern_dw<double>(u), g = kern_dig<double>(u);
    const double du = -u / h * dh;
    const double dw1 = u > 0.0 ? u * ddw_u(u) * du : 0.0;
    const double dg = u > 0.0 ? u * ddig_u(u) * du : 0.0;
    const double pv = prho2 + 0.5 * Pi;
    const double dpvw = pv * (dw1 / h5 + w1 * dih5);       // d(pv w1 / h^5) at fixed pv
    const double dgh = dg / h2 + g * dih2;                  // d(g / h^2)
    const double dA_i = 0.5 * dPi_i * w1 / h5 + dpvw - K / mj * dgh;
    const double dQ_i = mj * (0.5 * dPi_i * w1 / h5 + dpvw) - prho2 * Om * dgh;
    const double dA_j = 0.5 * dPi_j * w1 / h5, dQ_j = mj * dA_j;
    const double Ab = (mi * aj.x - mj * ai.x) * rx + (mi * aj.y - mj * ai.y) * ry + (mi * aj.z - mj * ai.z) * rz;
    const double Qb = ai.w * drdv;
    rb += Ab * dA_i + Qb * dQ_i;
    out[0] = Ab * dA_j + Qb * dQ_j;
  }
  rbar[i] = rb;
}

// rbar[j] += the j-side parts of the slots naming j (ascending slot order)
__global__ __launch_bounds__(kBlock) void k_adj_range_gather(const KV* __restrict__ pairs,
                                                             const int* __restrict__ start,
                                                             const int* __restrict__ end,
                                                             const double* __restrict__ pb,
                                                             double* __restrict__ rbar, int n) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  double s = 0.0;
  for (int p = start[j]; p < end[j]; ++p) s += pb[(size_t)pairs[p].value * kPair];
  rbar[j] += s;
}

// density outputs' part of r_bar_i, then the h-solve's implicit derivative:
// x_bar_i += lam sum_j (4 / r^2) dig(u_ij) x_ij, x_bar_j -= lam (4 / r^2)
// dig(u_ij) x_ij (per slot into pb[k * kPair + 0..2]), lam = r_bar_i r_i /
// sum_j digdh(u_ij)  (dF/dr = -sum_j digdh(u) / r, dF/dx_i = sum_j (4 / r^2)
// dig(u) x_ij)
__global__ __launch_bounds__(kBlock) void k_adj_range_density(const D4v* __restrict__ pos,
                                                               const D4v* __restrict__ vel,
                                                               const D4v* __restrict__ hydro,
                                                               const D4v* __restrict__ dots,
                                                               const D4v* __restrict__ hb1,
                                                               const D4v* __restrict__ hb2,
                                                               const int* __restrict__ offset,
                                                               const int* __restrict__ count,
                                                               const int* __restrict__ list,
                                                               const double* __restrict__ rbar, D4v* __restrict__ xs,
                                                               double* __restrict__ pb, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const D4v bp = pos[i], bv = vel[i];
  const double mi = bv.w, r = bp.w, h = 0.5 * r, h3 = h * h * h, h5 = h3 * h * h;
  const int off = offset[i], cnt = count[i];
  double den = 0, cx = 0, cy = 0, cz = 0;   // the forward's normalisation and curl sums
  double rho_d = 0, om_d = 0, psi_d = 0, den_d = 0, dv_d = 0, cx_d = 0, cy_d = 0, cz_d = 0, sdig = 0;
  for (int k = off; k < off + cnt; ++k) {
    const int j = list[k];
    const D4v pj = pos[j], vj = vel[j];
    const double mj = vj.w;
    const double rx = bp.x - pj.x, ry = bp.y - pj.y, rz = bp.z - pj.z;
    const double dvx = bv.x - vj.x, dvy = bv.y - vj.y, dvz = bv.z - vj.z;
    const double u = 2.0 * sqrt(rx * rx + ry * ry + rz * rz) / r;
    den -= kern_digdh<double>(u);
    sdig += kern_digdh<double>(u);
    const double mdw = mi * kern_dw<double>(u);
    const double ex = ry * dvz - rz * dvy, ey = rz * dvx - rx * dvz, ez = rx * dvy - ry * dvx;
    cx += mdw * ex;
    cy += mdw * ey;
    cz += mdw * ez;
    if (u > 0.0) {   // d/dr of each term: f'(u) du/dr, du/dr = -u / r
      const double du = -u / r;
      rho_d += mj * u * kern_dw<double>(u) * du;
      om_d += mj * u * ddwdh_u(u) * du;
      psi_d += mj * u * ddphidh_u(u) * du;
      den_d -= ddigdh(u) * du;
      const double mddw = mi * u * ddw_u(u) * du;
      dv_d += mddw * (rx * dvx + ry * dvy + rz * dvz);
      cx_d += mddw * ex;
      cy_d += mddw * ey;
      cz_d += mddw * ez;
    }
  }
  const D4v hd = hydro[i], dd = dots[i], b1 = hb1[i], b2 = hb2[i];
  const double rho = hd.x;
  // outputs' derivatives in r (dh/dr = 1/2); div v and |curl v| at fixed rho,
  // rho's own derivative carries their dependence through rho_bar below
  const double drho = rho_d / h3 - 1.5 * rho / h;
  const double rhob = b1.x - b2.x * dd.x / rho - b2.y * dd.y / rho;
  double rb = rbar[i] + rhob * drho;
  rb += b2.x * (dv_d / (rho * h5) - 2.5 * dd.x / h);
  const double cn = sqrt(cx * cx + cy * cy + cz * cz);
  if (cn > 0.0) rb += b2.y * ((cx * cx_d + cy * cy_d + cz * cz_d) / (cn * rho * h5) - 2.5 * dd.y / h);
  if (den != 0.0) {
    rb += b2.z * (om_d / (h3 * den) - dd.z * (1.5 / h + den_d / den));
    rb += b2.w * (psi_d / (h * den) - dd.w * (0.5 / h + den_d / den));
  }
  const double lam = sdig != 0.0 ? rb * r / sdig : 0.0;
  const double c4 = 4.0 / (r * r) * lam;
  double gx = 0, gy = 0, gz = 0;
  for (int k = off; k < off + cnt; ++k) {
    const int j = list[k];
    double* out = pb + (size_t)k * kPair;
    const D4v pj = pos[j];
    const double rx = bp.x - pj.x, ry = bp.y - pj.y, rz = bp.z - pj.z;
    const double u = 2.0 * sqrt(rx * rx + ry * ry + rz * rz) / r;
    const double f = j == i ? 0.0 : c4 * kern_dig<double>(u);
    gx += f * rx;
    gy += f * ry;
    gz += f * rz;
    out[0] = -f * rx;
    out[1] = -f * ry;
    out[2] = -f * rz;
  }
  D4v x = xs[i];
  x.x += gx;
  x.y += gy;
  x.z += gz;
  xs[i] = x;
}

inline dim3 grid(int n) { return dim3(ceil_div(n, kBlock)); }

template <class X>
void d2d(X* dst, const X* src, size_t count, hipStream_t s) {
  CS_HIP(hipMemcpyAsync(dst, src, count * sizeof(X), hipMemcpyDeviceToDevice, s));
}

}  // namespace

// ---------------------------------------------------------------------------
// Engine members.  Only the FP64 form has an adjoint; Engine<float> reports
// that as an error.

template <class T>
void Engine<T>::adj_save(DBuf<V4<T>>* dst[4], long slot) {
  DBuf<V4<T>>* src[4] = {&pos_, &vel_, &hydro_, &bdots_};
  for (int c = 0; c < 4; ++c)
    if (dst[c]) d2d(dst[c]->p + (size_t)slot * nd_, src[c]->p, (size_t)nd_, s_);
}

template <class T>
void Engine<T>::adj_load(DBuf<V4<T>>* src[4], long slot) {
  DBuf<V4<T>>* dst[4] = {&pos_, &vel_, &hydro_, &bdots_};
  for (int c = 0; c < 4; ++c)
    if (src[c]) d2d(dst[c]->p, src[c]->p + (size_t)slot * nd_, (size_t)nd_, s_);
}

template <class T>
void Engine<T>::adj_keep_end() {
  DBuf<V4<T>>* src[7] = {&pos_, &vel_, &hydro_, &dots_, &bdots_, &hydro2_, &grav_};
  for (int c = 0; c < 7; ++c) {
    if (adj_end_[c].n < (size_t)nd_) adj_end_[c].alloc(nd_);
    d2d(adj_end_[c].p, src[c]->p, (size_t)nd_, s_);
  }
  if (adj_end_small_.n < 32) adj_end_small_.alloc(32);
  d2d(adj_end_small_.p, dt_cur_.p, 1, s_);
  d2d(adj_end_small_.p + 1, stats6_.p, 6, s_);
  d2d(adj_end_small_.p + 7, force6_.p, 6, s_);
  d2d(adj_end_small_.p + 13, diag_.p, 19, s_);
  adj_end_iter_ = iteration_;
  adj_end_time_ = global_time_;
}

template <class T>
void Engine<T>::adj_put_back_end() {
  DBuf<V4<T>>* dst[7] = {&pos_, &vel_, &hydro_, &dots_, &bdots_, &hydro2_, &grav_};
  for (int c = 0; c < 7; ++c) d2d(dst[c]->p, adj_end_[c].p, (size_t)nd_, s_);
  d2d(dt_cur_.p, adj_end_small_.p, 1, s_);
  d2d(stats6_.p, adj_end_small_.p + 1, 6, s_);
  d2d(force6_.p, adj_end_small_.p + 7, 6, s_);
  d2d(diag_.p, adj_end_small_.p + 13, 19, s_);
  iteration_ = adj_end_iter_;
  global_time_ = adj_end_time_;
}

template <class T>
void Engine<T>::adj_capture(EvalSlot& e) {
  const size_t nd = (size_t)nd_, tot = (size_t)ngb_tot_;
  auto grow = [](auto& b, size_t n) {
    if (b.n < n) b.alloc(n);
  };
  grow(e.pos, nd);
  grow(e.vel, nd);
  grow(e.hyd, nd);
  grow(e.dots, nd);
  grow(e.off, nd + 1);
  grow(e.cnt, nd);
  grow(e.tr_start, nd);
  grow(e.tr_end, nd);
  grow(e.list, std::max<size_t>(tot, 1));
  grow(e.pairs, std::max<size_t>(tot, 1));
  d2d(e.pos.p, pos_.p, nd, s_);
  d2d(e.vel.p, vel_.p, nd, s_);
  d2d(e.hyd.p, hydro_.p, nd, s_);
  d2d(e.dots.p, dots_.p, nd, s_);
  d2d(e.off.p, ngb_offset_.p, nd + 1, s_);
  d2d(e.cnt.p, ngb_count_.p, nd, s_);
  d2d(e.tr_start.p, tr_start_.p, nd, s_);
  d2d(e.tr_end.p, tr_end_.p, nd, s_);
  d2d(e.list.p, list_.p, tot, s_);
  d2d(e.pairs.p, pairs_.p, tot, s_);
  e.tot = ngb_tot_;
}

template <class T>
void Engine<T>::adj_restore(const EvalSlot& e) {
  const size_t nd = (size_t)nd_, tot = (size_t)e.tot;
  d2d(pos_.p, e.pos.p, nd, s_);
  d2d(vel_.p, e.vel.p, nd, s_);
  d2d(hydro_.p, e.hyd.p, nd, s_);
  d2d(dots_.p, e.dots.p, nd, s_);
  d2d(ngb_offset_.p, e.off.p, nd + 1, s_);
  d2d(ngb_count_.p, e.cnt.p, nd, s_);
  d2d(tr_start_.p, e.tr_start.p, nd, s_);
  d2d(tr_end_.p, e.tr_end.p, nd, s_);
  d2d(list_.p, e.list.p, tot, s_);
  d2d(pairs_.p, e.pairs.p, tot, s_);
  ngb_tot_ = e.tot;
}

template <class T>
double Engine<T>::adj_loss() {
  if constexpr (kF32) {
    throw std::runtime_error("adjoint: FP64 form only");
  } else {
    const int nb = ceil_div(n_, kBlock);
    if (adj_partial_.n < (size_t)nb + 1) adj_partial_.alloc(nb + 1);
    hipLaunchKernelGGL(k_adj_loss, dim3(nb), dim3(kBlock), 0, s_, vel_.p, hydro_.p, adj_partial_.p, n_);
    CS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_adj_loss_final, dim3(1), dim3(kBlock), 0, s_, adj_partial_.p, nb);
    CS_CHECK_LAUNCH();
    double out = 0.0;
    CS_HIP(hipMemcpyAsync(&out, adj_partial_.p + nb, 8, hipMemcpyDeviceToHost, s_));
    sync();
    return out;
  }
}

// reverse of one force evaluation: (a_bar, udot_bar) in aa_ ->
// xs_ (x_bar), vs_ (v_bar), hb1_.z (u_bar) of that evaluation's inputs
template <class T>
void Engine<T>::adj_forces_backward(bool seed_rho) {
  if constexpr (kF32) {
    throw std::runtime_error("adjoint: FP64 form only");
  } else {
    if (pairbuf_.n < (size_t)ngb_tot_ * kPair) pairbuf_.alloc((size_t)ngb_tot_ * kPair + 64);
    // gravity's part sums depend only on x, m and a_bar, so with several
    // parts they run on the second stream beside the SPH kernels, and their
    // combine adds them to x_bar after those kernels, as in stream order
    const double eps = (double)p_.grav_eps;
    const int parts = grav_parts_ ? grav_parts_ : gravity64_parts(n_);
    const int chunk = ceil_div(ceil_div(n_, parts), kAdjTile) * kAdjTile;
    const bool beside = overlap_ && parts > 1;
    if (parts > 1 && gpart64_.n < (size_t)parts * nd_) gpart64_.alloc((size_t)parts * nd_);
    if (beside) {
      CS_HIP(hipEventRecord(ev_fork_, s_));
      CS_HIP(hipStreamWaitEvent(gs_, ev_fork_, 0));
      TimerScope tg(timers_, "adjoint_gravity", gs_);
      hipLaunchKernelGGL(k_adj_gravity, dim3(ceil_div(n_, kBlock) * parts), dim3(kBlock), 0, gs_, pos_.p, vel_.p,
                         aa_.p, xs_.p, (D4v*)gpart64_.p, n_, parts, chunk, eps * eps);
      CS_CHECK_LAUNCH();
      CS_HIP(hipEventRecord(ev_join_, gs_));
    }
    TimerScope ts(timers_, "adjoint_sph", s_);
    hipLaunchKernelGGL(k_adj_hydro, grid(n_), dim3(kBlock), 0, s_, pos_.p, vel_.p, hydro_.p, dots_.p, aa_.p,
                       ngb_offset_.p, ngb_count_.p, list_.p, xs_.p, vs_.p, hb1_.p, hb2_.p, pairbuf_.p, n_,
                       (double)p_.av_alpha, (double)p_.av_beta);
    CS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_adj_gather, grid(n_), dim3(kBlock), 0, s_, pairs_.p, tr_start_.p, tr_end_.p, pairbuf_.p,
                       xs_.p, vs_.p, hb1_.p, hb2_.p, n_, kPair);
    CS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_adj_eos, grid(n_), dim3(kBlock), 0, s_, hydro_.p, hb1_.p, n_, eos_const(),
                       seed_rho ? 1 : 0);
    CS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_adj_density, grid(n_), dim3(kBlock), 0, s_, pos_.p, vel_.p, hydro_.p, dots_.p, hb1_.p,
                       hb2_.p, ngb_offset_.p, ngb_count_.p, list_.p, xs_.p, vs_.p, pairbuf_.p, n_);
    CS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_adj_gather, grid(n_), dim3(kBlock), 0, s_, pairs_.p, tr_start_.p, tr_end_.p, pairbuf_.p,
                       xs_.p, vs_.p, hb1_.p, hb2_.p, n_, 6);
    CS_CHECK_LAUNCH();
    if (adj_implicit_h_) {   // the ranges' dependence on the positions (the h-solve)
      if (adj_rbar_.n < (size_t)nd_) adj_rbar_.alloc(nd_);
      hipLaunchKernelGGL(k_adj_range_hydro, grid(n_), dim3(kBlock), 0, s_, pos_.p, vel_.p, hydro_.p, dots_.p, aa_.p,
                         ngb_offset_.p, ngb_count_.p, list_.p, adj_rbar_.p, pairbuf_.p, n_, (double)p_.av_alpha,
                         (double)p_.av_beta);
      CS_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_adj_range_gather, grid(n_), dim3(kBlock), 0, s_, pairs_.p, tr_start_.p, tr_end_.p,
                         pairbuf_.p, adj_rbar_.p, n_);
      CS_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_adj_range_density, grid(n_), dim3(kBlock), 0, s_, pos_.p, vel_.p, hydro_.p, dots_.p,
                         hb1_.p, hb2_.p, ngb_offset_.p, ngb_count_.p, list_.p, adj_rbar_.p, xs_.p, pairbuf_.p, n_);
      CS_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_adj_gather, grid(n_), dim3(kBlock), 0, s_, pairs_.p, tr_start_.p, tr_end_.p, pairbuf_.p,
                         xs_.p, vs_.p, hb1_.p, hb2_.p, n_, 3);
      CS_CHECK_LAUNCH();
    }
    ts.t.end(s_);
    if (beside) {
      CS_HIP(hipStreamWaitEvent(s_, ev_join_, 0));
    } else {
      ts.t.begin("adjoint_gravity", s_);
      hipLaunchKernelGGL(k_adj_gravity, dim3(ceil_div(n_, kBlock) * parts), dim3(kBlock), 0, s_, pos_.p, vel_.p,
                         aa_.p, xs_.p, (D4v*)gpart64_.p, n_, parts, chunk, eps * eps);
      CS_CHECK_LAUNCH();
    }
    if (parts > 1) {
      hipLaunchKernelGGL(k_adj_gravity_combine, grid(n_), dim3(kBlock), 0, s_, (const D4v*)gpart64_.p, xs_.p, n_,
                         parts);
      CS_CHECK_LAUNCH();
    }
  }
}

namespace {
// restores the engine's normal mode when an adjoint call ends, also on error
struct ModeGuard {
  int& mode;
  int& capture;
  ~ModeGuard() {
    mode = 0;
    capture = -1;
  }
};
}  // namespace

template <class T>
double Engine<T>::adjoint_forward(int steps, int ckpt_every) {
  if constexpr (kF32) {
    throw std::runtime_error("adjoint: FP64 form only (create the engine with SPH_FORM_FP64)");
  } else {
    if (p_.grav_mode != 1) throw std::runtime_error("adjoint: Plummer gravity (grav_mode 1) only");
    if (steps < 0 || ckpt_every < 1) throw std::runtime_error("adjoint: steps >= 0 and ckpt_every >= 1 required");
    const int K = steps, ck = ckpt_every, nseg = (K + ck - 1) / ck;
    const size_t nd = (size_t)nd_;
    adj_dt_.assign(K + 1, 0.0);
    adj_range_.alloc((K + 1) * nd);
    adj_nngb_.alloc((K + 1) * nd);
    adj_grav_.alloc((K + 1) * nd);
    for (auto* b : {&adj_ck_pos_, &adj_ck_vel_, &adj_ck_hyd_, &adj_ck_bd_}) b->alloc(std::max(nseg, 1) * nd);
    adj_eval_.clear();
    for (int c = 0; c < ck; ++c) adj_eval_.push_back(std::make_unique<EvalSlot>());
    for (auto* b : {&adj_init_pos_, &adj_init_vel_, &adj_init_hyd_, &adj_final_vel_}) b->alloc(nd);
    for (auto* b : {&ax_, &av_, &aa_, &xs_, &vs_, &hb1_, &hb2_}) b->alloc(nd);
    au_.alloc(nd);
    adj_out_.alloc(7 * nd);
    d2d(adj_init_pos_.p, pos_.p, nd, s_);
    d2d(adj_init_vel_.p, vel_.p, nd, s_);
    d2d(adj_init_hyd_.p, hydro_.p, nd, s_);
    adj_K_ = K;
    adj_k_ = ck;
    adj_iter0_ = (int)iteration_;
    adj_t0_ = global_time_;
    ModeGuard g{adj_mode_, adj_capture_};
    adj_mode_ = kAdjRecord;
    init();
    DBuf<V4<T>>* ckb[4] = {&adj_ck_pos_, &adj_ck_vel_, &adj_ck_hyd_, &adj_ck_bd_};
    adj_save(ckb, 0);
    for (int s = 1; s <= K; ++s) {
      step();
      if (s % ck == 0 && s < K) adj_save(ckb, s / ck);
    }
    d2d(adj_final_vel_.p, vel_.p, nd, s_);
    adj_keep_end();
    adj_ready_ = true;
    return adj_loss();
  }
}

template <class T>
void Engine<T>::adjoint_backward(double* gx, double* gv, double* gu) {
  if constexpr (kF32) {
    throw std::runtime_error("adjoint: FP64 form only");
  } else {
    if (!adj_ready_) throw std::runtime_error("adjoint_backward: run adjoint_forward first");
    const int K = adj_K_, ck = adj_k_, nseg = (K + ck - 1) / ck;
    const size_t nd = (size_t)nd_;
    ModeGuard g{adj_mode_, adj_capture_};
    adj_mode_ = kAdjReplayStoredGravity;
    for (auto* b : {&ax_, &aa_}) CS_HIP(hipMemsetAsync(b->p, 0, b->bytes(), s_));
    CS_HIP(hipMemsetAsync(au_.p, 0, au_.bytes(), s_));
    hipLaunchKernelGGL(k_adj_seed, grid(nd_), dim3(kBlock), 0, s_, adj_final_vel_.p, av_.p, n_);
    CS_CHECK_LAUNCH();
    if (nd_ > n_) CS_HIP(hipMemsetAsync(av_.p + n_, 0, (nd - n_) * sizeof(D4v), s_));
    DBuf<V4<T>>* ckb[4] = {&adj_ck_pos_, &adj_ck_vel_, &adj_ck_hyd_, &adj_ck_bd_};
    for (int sg = nseg - 1; sg >= 0; --sg) {
      const int first = sg * ck + 1, last = std::min((sg + 1) * ck, K);
      // recompute the segment from its checkpoint, capturing each step's
      // force-evaluation data (the state at the evaluation, lists, transpose)
      adj_load(ckb, sg);
      CS_HIP(hipMemcpyAsync(dt_cur_, &adj_dt_[first - 1], 8, hipMemcpyHostToDevice, s_));
      iteration_ = adj_iter0_ + first - 1;
      global_time_ = adj_t0_;   // (time does not enter the map)
      for (int s = first; s <= last; ++s) {
        adj_capture_ = s - first;
        step();
      }
      adj_capture_ = -1;
      // reverse the segment
      for (int s = last; s >= first; --s) {
        adj_restore(*adj_eval_[s - first]);
        const double f = 0.5 * adj_dt_[s] + 0.5 * adj_dt_[s - 1];
        hipLaunchKernelGGL(k_adj_correct, grid(nd_), dim3(kBlock), 0, s_, aa_.p, av_.p, au_.p, f, n_);
        CS_CHECK_LAUNCH();
        adj_forces_backward(s == K);
        hipLaunchKernelGGL(k_adj_predict, grid(nd_), dim3(kBlock), 0, s_, ax_.p, av_.p, au_.p, aa_.p, xs_.p,
                           vs_.p, hb1_.p, adj_dt_[s - 1], n_);
        CS_CHECK_LAUNCH();
      }
    }
    // the initial force evaluation: a_0 = F(x_0, v_0, u_0)
    DBuf<V4<T>>* ini[4] = {&adj_init_pos_, &adj_init_vel_, &adj_init_hyd_, nullptr};
    adj_load(ini, 0);
    iteration_ = adj_iter0_;
    adj_f_ = 0;
    sph_accelerations();
    adj_forces_backward(K == 0);
    hipLaunchKernelGGL(k_adj_finish, grid(nd_), dim3(kBlock), 0, s_, ax_.p, av_.p, au_.p, xs_.p, vs_.p, hb1_.p,
                       adj_out_.p, n_);
    CS_CHECK_LAUNCH();
    sync();
    const size_t n = (size_t)n_;
    if (gx) CS_HIP(hipMemcpy(gx, adj_out_.p, 3 * n * 8, hipMemcpyDeviceToHost));
    if (gv) CS_HIP(hipMemcpy(gv, adj_out_.p + 3 * n, 3 * n * 8, hipMemcpyDeviceToHost));
    if (gu) CS_HIP(hipMemcpy(gu, adj_out_.p + 6 * n, n * 8, hipMemcpyDeviceToHost));
    adj_put_back_end();
    sync();
  }
}

template <class T>
double Engine<T>::adjoint_replay(const double* x0, const double* v0, const double* u0) {
  if constexpr (kF32) {
    throw std::runtime_error("adjoint: FP64 form only");
  } else {
    if (!adj_ready_) throw std::runtime_error("adjoint_replay: run adjoint_forward first");
    const size_t n = (size_t)n_;
    DBuf<V4<T>>* ini[4] = {&adj_init_pos_, &adj_init_vel_, &adj_init_hyd_, nullptr};
    adj_load(ini, 0);
    // host arrays go through adj_out_ (7 n doubles) as staging
    if (x0) {
      CS_HIP(hipMemcpyAsync(adj_out_.p, x0, 3 * n * 8, hipMemcpyHostToDevice, s_));
      hipLaunchKernelGGL(k_adj_set_xyz, grid(n_), dim3(kBlock), 0, s_, pos_.p, adj_out_.p, n_);
      CS_CHECK_LAUNCH();
    }
    if (v0) {
      CS_HIP(hipMemcpyAsync(adj_out_.p + 3 * n, v0, 3 * n * 8, hipMemcpyHostToDevice, s_));
      hipLaunchKernelGGL(k_adj_set_xyz, grid(n_), dim3(kBlock), 0, s_, vel_.p, adj_out_.p + 3 * n, n_);
      CS_CHECK_LAUNCH();
    }
    if (u0) {
      CS_HIP(hipMemcpyAsync(adj_out_.p + 6 * n, u0, n * 8, hipMemcpyHostToDevice, s_));
      hipLaunchKernelGGL(k_adj_set_z, grid(n_), dim3(kBlock), 0, s_, hydro_.p, adj_out_.p + 6 * n, n_);
      CS_CHECK_LAUNCH();
    }
    sync();   // the host arrays may be freed once this returns
    ModeGuard g{adj_mode_, adj_capture_};
    adj_mode_ = kAdjReplay;
    iteration_ = adj_iter0_;
    global_time_ = adj_t0_;
    init();
    for (int s = 1; s <= adj_K_; ++s) step();
    return adj_loss();
  }
}

#define CS_ADJ_INSTANTIATE(T)                                                    \
  template void Engine<T>::adj_save(DBuf<V4<T>>* [4], long);                    \
  template void Engine<T>::adj_load(DBuf<V4<T>>* [4], long);                    \
  template double Engine<T>::adj_loss();                                        \
  template void Engine<T>::adj_capture(EvalSlot&);                              \
  template void Engine<T>::adj_keep_end();                                      \
  template void Engine<T>::adj_put_back_end();                                  \
  template void Engine<T>::adj_restore(const EvalSlot&);                        \
  template void Engine<T>::adj_forces_backward(bool);                           \
  template double Engine<T>::adjoint_forward(int, int);                         \
  template void Engine<T>::adjoint_backward(double*, double*, double*);         \
  template double Engine<T>::adjoint_replay(const double*, const double*, const double*);
CS_ADJ_INSTANTIATE(float)
CS_ADJ_INSTANTIATE(double)

}  // namespace cudasph
