// Seeded initial conditions for the five benchmark configurations.
//
// Everything is generated in FP64 from a xoshiro256** stream seeded by
// splitmix64, using only IEEE arithmetic and libm, so two x86-64 hosts with
// the same libm produce the same bits.  Each generator also sets an initial
// range (2h) from the analytic number density so that the h-solve starts
// inside its bracketing reach.
//
// Records: pos4 = {x, y, z, range}, vel4 = {vx, vy, vz, m}, hydro4 = {rho, P, u, mu}.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

#include "../../include/cudasph.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (auto& v : s) {   // splitmix64
      z += 0x9e3779b97f4a7c15ULL;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
      x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
      v = x ^ (x >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {   // xoshiro256**
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }     // [0,1)
  double uniform24() { return (double)(next() >> 40) * 0x1.0p-24; }   // 24-bit grid
  double gauss() {                                                     // Box-Muller
    double u1 = uniform();
    while (u1 <= 0) u1 = uniform();
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2 * kPi * u2);
  }
  void direction(double& x, double& y, double& z) {
    const double c = 2 * uniform() - 1, ph = 2 * kPi * uniform();
    const double s = std::sqrt(std::max(0.0, 1 - c * c));
    x = s * std::cos(ph);
    y = s * std::sin(ph);
    z = c;
  }
};

// range holding k neighbours at number density n
double range_for(double k, double n) { return std::cbrt(3.0 * k / (4 * kPi * n)); }

void put(double* a, long i, double x, double y, double z, double w) {
  a[4 * i] = x;
  a[4 * i + 1] = y;
  a[4 * i + 2] = z;
  a[4 * i + 3] = w;
}

// config 1: uniform cube [0,1)^3 on a 24-bit grid, v = 0, u = 1
void ic_cube(int n, Rng& rng, double* pos, double* vel, double* hyd) {
  const double range = 2 * std::cbrt(32.0 / n);   // h0 = (32/N)^(1/3)
  for (int i = 0; i < n; ++i) {
    const double x = rng.uniform24(), y = rng.uniform24(), z = rng.uniform24();
    put(pos, i, x, y, z, range);
    put(vel, i, 0, 0, 0, 1.0 / n);
    put(hyd, i, 0, 0, 1.0, 0.6);
  }
}

// configs 2 and 4: Plummer sphere, M = a = G = 1, radii from the cumulative
// mass truncated at 10, speeds from the distribution function by rejection
// (q^2 (1-q^2)^(7/2), v = q v_esc), isotropic directions.  Thermal energy
// U = 0.05 |W| with W = -3 pi / 32, so the gas pressure term of the virial
// theorem (2U for gamma = 5/3) carries 10% of |W|.
void ic_plummer(int n, double ngb, Rng& rng, double* pos, double* vel, double* hyd) {
  const double u_spec = 0.05 * 3 * kPi / 32;
  for (int i = 0; i < n; ++i) {
    double r;
    do {
      double X = rng.uniform();
      while (X <= 0) X = rng.uniform();
      r = 1.0 / std::sqrt(std::pow(X, -2.0 / 3.0) - 1.0);
    } while (!(r <= 10.0));
    double dx, dy, dz;
    rng.direction(dx, dy, dz);
    double q;
    for (;;) {
      q = rng.uniform();
      const double g = q * q * std::pow(1 - q * q, 3.5);
      if (0.1 * rng.uniform() < g) break;
    }
    const double v = q * std::sqrt(2.0) * std::pow(1 + r * r, -0.25);
    double vx, vy, vz;
    rng.direction(vx, vy, vz);
    const double number_density = n * 3.0 / (4 * kPi) * std::pow(1 + r * r, -2.5);
    put(pos, i, r * dx, r * dy, r * dz, range_for(ngb, number_density));
    put(vel, i, v * vx, v * vy, v * vz, 1.0 / n);
    put(hyd, i, 0, 0, u_spec, 0.6);
  }
}

// Lane-Emden n = 1.5: theta'' = -theta^n - 2 theta'/xi, RK4 from the series
// start, tabulated to the first zero xi_1
struct LaneEmden {
  std::vector<double> xi, theta, mass;   // mass(xi) = -xi^2 theta'
  double xi1 = 0;
  explicit LaneEmden(double n_poly) {
    double h = 1e-4, x = 1e-6, t = 1 - x * x / 6, dt = -x / 3;
    auto f = [&](double xx, double tt, double dd, double& a, double& b) {
      a = dd;
      b = -std::pow(std::max(tt, 0.0), n_poly) - 2 * dd / xx;
    };
    xi.push_back(0);
    theta.push_back(1);
    mass.push_back(0);
    while (t > 0) {
      double k1a, k1b, k2a, k2b, k3a, k3b, k4a, k4b;
      f(x, t, dt, k1a, k1b);
      f(x + h / 2, t + h / 2 * k1a, dt + h / 2 * k1b, k2a, k2b);
      f(x + h / 2, t + h / 2 * k2a, dt + h / 2 * k2b, k3a, k3b);
      f(x + h, t + h * k3a, dt + h * k3b, k4a, k4b);
      const double tn = t + h / 6 * (k1a + 2 * k2a + 2 * k3a + k4a);
      const double dn = dt + h / 6 * (k1b + 2 * k2b + 2 * k3b + k4b);
      if (tn <= 0) {   // linear interpolation to the zero
        const double frac = t / (t - tn);
        xi1 = x + frac * h;
        xi.push_back(xi1);
        theta.push_back(0);
        mass.push_back(-xi1 * xi1 * (dt + frac * (dn - dt)));
        break;
      }
      x += h;
      t = tn;
      dt = dn;
      xi.push_back(x);
      theta.push_back(t);
      mass.push_back(-x * x * dt);
    }
  }
  // xi at which the enclosed mass fraction is X (bisection on the table)
  double xi_of_mass(double X, double& th) const {
    const double target = X * mass.back();
    size_t lo = 0, hi = mass.size() - 1;
    while (hi - lo > 1) {
      const size_t mid = (lo + hi) / 2;
      (mass[mid] < target ? lo : hi) = mid;
    }
    const double f = (target - mass[lo]) / std::max(mass[hi] - mass[lo], 1e-300);
    th = theta[lo] + f * (theta[hi] - theta[lo]);
    return xi[lo] + f * (xi[hi] - xi[lo]);
  }
};

// config 3: two n = 1.5 polytropes (R = 1, M = 0.5 each) on a parabolic
// orbit, periastron 0.5, separation 5, approaching; u = n K rho_c^(1/n) theta
void ic_polytropes(int n, double ngb, Rng& rng, double* pos, double* vel, double* hyd) {
  const LaneEmden le(1.5);
  const double R = 1.0, M = 0.5;
  const double rho_c = M * std::pow(le.xi1, 3) / (4 * kPi * R * R * R * le.mass.back());
  const double alpha = R / le.xi1;
  const double K = 4 * kPi * alpha * alpha * std::cbrt(rho_c) / 2.5;   // G = 1, n = 1.5
  const double u_c = 1.5 * K * std::pow(rho_c, 2.0 / 3.0);
  // relative orbit: cos f = -0.8, sin f = -0.6, p = 2 q = 1, G M_tot = 1
  const double rel[3] = {5 * -0.8, 5 * -0.6, 0};
  const double vrel[3] = {0.6, 1 - 0.8, 0};
  const int n1 = n / 2;
  for (int i = 0; i < n; ++i) {
    const int star = i < n1 ? 0 : 1;
    const double sgn = star == 0 ? -0.5 : 0.5;
    double th;
    const double xi = le.xi_of_mass(rng.uniform(), th);
    const double r = xi * alpha;
    double dx, dy, dz;
    rng.direction(dx, dy, dz);
    const double rho = rho_c * std::pow(std::max(th, 0.0), 1.5);
    const double number_density = std::max(rho, 1e-3 * rho_c) / (1.0 / n);
    put(pos, i, sgn * rel[0] + r * dx, sgn * rel[1] + r * dy, sgn * rel[2] + r * dz,
        range_for(ngb, number_density));
    put(vel, i, sgn * vrel[0], sgn * vrel[1], sgn * vrel[2], 1.0 / n);
    put(hyd, i, 0, 0, std::max(u_c * th, 1e-4 * u_c), 0.6);
  }
}

// in-place radix-2 FFT of length L (power of two) with stride
void fft1d(std::complex<double>* a, int L, long stride, int sign) {
  for (int i = 1, j = 0; i < L; ++i) {
    int bit = L >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i * stride], a[j * stride]);
  }
  for (int len = 2; len <= L; len <<= 1) {
    const double ang = sign * 2 * kPi / len;
    const std::complex<double> wl(std::cos(ang), std::sin(ang));
    for (int i = 0; i < L; i += len) {
      std::complex<double> w(1, 0);
      for (int k = 0; k < len / 2; ++k) {
        const std::complex<double> u = a[(i + k) * stride], v = a[(i + k + len / 2) * stride] * w;
        a[(i + k) * stride] = u + v;
        a[(i + k + len / 2) * stride] = u - v;
        w *= wl;
      }
    }
  }
}
void fft3d(std::vector<std::complex<double>>& g, int L, int sign) {
  const long L2 = (long)L * L;
  for (long a = 0; a < L2; ++a) fft1d(&g[a * L], L, 1, sign);                              // z
  for (int x = 0; x < L; ++x)
    for (int z = 0; z < L; ++z) fft1d(&g[x * L2 + z], L, L, sign);                         // y
  for (long yz = 0; yz < L2; ++yz) fft1d(&g[yz], L, L2, sign);                             // x
}

// Gaussian random field with P(k) ~ k^-4 on a periodic L^3 grid (white
// noise filtered by k^-2 in Fourier space), one per velocity component, from
// the stream after the particle positions
void turbulent_field(Rng& rng, int L, std::vector<double> (&field)[3]) {
  const long L3 = (long)L * L * L;
  for (int c = 0; c < 3; ++c) {
    std::vector<std::complex<double>> g(L3);
    for (long k = 0; k < L3; ++k) g[k] = rng.gauss();
    fft3d(g, L, -1);
    for (int a = 0; a < L; ++a)
      for (int b = 0; b < L; ++b)
        for (int d = 0; d < L; ++d) {
          const int ka = a <= L / 2 ? a : a - L, kb = b <= L / 2 ? b : b - L, kd = d <= L / 2 ? d : d - L;
          const double k2 = (double)ka * ka + (double)kb * kb + (double)kd * kd;
          g[((long)a * L + b) * L + d] *= k2 > 0 ? 1.0 / k2 : 0.0;   // sqrt(P) = k^-2
        }
    fft3d(g, L, +1);
    field[c].resize(L3);
    for (long k = 0; k < L3; ++k) field[c][k] = g[k].real();
  }
}

// uniform positions in the unit sphere by rejection from [-1,1]^3
void sphere_positions(int n, double range, Rng& rng, double* pos) {
  for (int i = 0; i < n; ++i) {
    double x, y, z;
    do {
      x = 2 * rng.uniform() - 1;
      y = 2 * rng.uniform() - 1;
      z = 2 * rng.uniform() - 1;
    } while (x * x + y * y + z * z > 1.0);
    put(pos, i, x, y, z, range);
  }
}

constexpr int kTurbGrid = 128;

// config 5: uniform sphere R = 1, M = 1; the k^-4 field on a periodic 128^3
// grid over [-1,1]^3, trilinear to particles, scaled to rms Mach 2 with
// c_s = sqrt(gamma (gamma-1) u), u = 1
void ic_turbulent(int n, double ngb, Rng& rng, double* pos, double* vel, double* hyd, int L = kTurbGrid) {
  const double number_density = n / (4.0 / 3.0 * kPi);
  sphere_positions(n, range_for(ngb, number_density), rng, pos);
  for (int i = 0; i < n; ++i) put(hyd, i, 0, 0, 1.0, 0.6);
  std::vector<double> field[3];
  turbulent_field(rng, L, field);
  double sum2 = 0;
  for (int i = 0; i < n; ++i) {
    double v[3];
    const double gx = (pos[4 * i] + 1) * 0.5 * L, gy = (pos[4 * i + 1] + 1) * 0.5 * L, gz = (pos[4 * i + 2] + 1) * 0.5 * L;
    const int ix = std::min((int)gx, L - 1), iy = std::min((int)gy, L - 1), iz = std::min((int)gz, L - 1);
    const double fx = gx - ix, fy = gy - iy, fz = gz - iz;
    for (int c = 0; c < 3; ++c) {
      double s = 0;
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
          for (int d = 0; d < 2; ++d) {
            const double w = (a ? fx : 1 - fx) * (b ? fy : 1 - fy) * (d ? fz : 1 - fz);
            s += w * field[c][(((long)(ix + a) % L) * L + (iy + b) % L) * L + (iz + d) % L];
          }
      v[c] = s;
    }
    put(vel, i, v[0], v[1], v[2], 1.0 / n);
    sum2 += v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  }
  const double cs = std::sqrt(5.0 / 3.0 * (2.0 / 3.0) * 1.0);
  const double scale = 2.0 * cs / std::sqrt(sum2 / n);
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) vel[4 * i + c] *= scale;
}

}  // namespace

// the velocity grid of config 5 for n particles and this seed (the same
// stream as sph_ic_generate): 3 L^3 doubles, component-major, then x-major
// (index ((x L + y) L + z)), before interpolation and scaling
extern "C" int sph_ic_turbulent_field(int n, uint64_t seed, double* field3) {
  if (n <= 0 || !field3) return -1;
  Rng rng(seed);
  std::vector<double> pos(4 * (size_t)n);
  sphere_positions(n, 0.0, rng, pos.data());
  std::vector<double> field[3];
  turbulent_field(rng, kTurbGrid, field);
  const size_t L3 = (size_t)kTurbGrid * kTurbGrid * kTurbGrid;
  for (int c = 0; c < 3; ++c) std::copy(field[c].begin(), field[c].end(), field3 + c * L3);
  return 0;
}

extern "C" int sph_ic_generate(int config, int n, uint64_t seed, double* pos4, double* vel4, double* hydro4) {
  if (n <= 0) return -1;
  Rng rng(seed);
  switch (config) {
    case 1: ic_cube(n, rng, pos4, vel4, hydro4); return 0;
    case 2: ic_plummer(n, 50, rng, pos4, vel4, hydro4); return 0;
    case 3: ic_polytropes(n, 64, rng, pos4, vel4, hydro4); return 0;
    case 4: ic_plummer(n, 50, rng, pos4, vel4, hydro4); return 0;
    case 5: ic_turbulent(n, 64, rng, pos4, vel4, hydro4); return 0;
    default: return -1;
  }
}
