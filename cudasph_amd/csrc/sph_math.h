// Device-side SPH arithmetic: the M4 kernel family, the gas+radiation EOS and
// the artificial viscosity.
//
// Instantiated with T = float these reproduce the reference's arithmetic
// types: the reference's "float" device code computes in double wherever a
// float operand meets a double literal, then rounds back to float.  Each such
// promotion is written out below as an explicit double temporary so the
// rounding sequence is visible.  All HIP sources are compiled with
// -ffp-contract=off, and float sqrt and division are correctly rounded, so
// these functions return the same bits as the CPU restatement.  With T = double
// they are the FP64 form.
#pragma once

#include "common.h"

namespace cudasph {

constexpr double kInvPi = 0.318309886183791;   // dev_sph_kernels.cu:5
constexpr double kFourPi = 12.5663706143592;   // dev_sph_kernels.cu:19

__device__ __forceinline__ float dsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float dabs(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double dabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float dmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double dmax(double a, double b) { return __builtin_fmax(a, b); }
// float power rounded from a double power: correctly rounded except when the
// double result lies within 1e-16 of a float rounding boundary
__device__ __forceinline__ float dpow(float x, float y) { return (float)pow((double)x, (double)y); }
__device__ __forceinline__ double dpow(double x, double y) { return pow(x, y); }

// ---- M4 cubic spline family, u = r/h (dev_sph_kernels.cu:7-99) ----------

template <class T>
__device__ __forceinline__ T kern_w(T u) {          // W(u)*pi^-1 ... :7-16
  const double du = u;
  if (u < T(0)) return T(kInvPi);
  if (u < T(1)) return T(kInvPi * (1.0 - 1.5 * du * du + 0.75 * du * du * du));
  if (u < T(2)) {
    const double s = (double)(T(2) - u);
    return T(kInvPi * 0.25 * s * s * s);
  }
  return T(0);
}

template <class T>
__device__ __forceinline__ T kern_diw(T u) {        // 4 pi u^2 W, :18-20
  const double du = u;
  return T(kFourPi * du * du * (double)kern_w<T>(u));
}

template <class T>
__device__ __forceinline__ T kern_iw(T u) {         // cumulative mass fraction, :22-34
  if (u < T(0)) return T(0);
  if (u < T(2)) {
    const T c = u * u * u;
    const double dc = c, du = u;
    if (u < T(1)) return T(4.0 * (dc / 3.0 - 0.3 * dc * du * du + 0.125 * dc * dc));
    const T c2 = c * c;                             // single-precision product in the reference
    return T(4.0 * (-1.0 / 60 + 2.0 / 3 * dc - 0.75 * dc * du + 0.3 * dc * du * du - (double)c2 / 24.0));
  }
  return T(1);
}

template <class T>
__device__ __forceinline__ T kern_ig(T u) {         // smoothed-count window, :36-48
  const T v = u - T(1);
  return kern_iw<T>(T(4) * (T(1) - dsqrt(v * v)));
}

template <class T>
__device__ __forceinline__ T kern_digdh(T u) {      // :50-60
  const T k4 = T(4);
  if (u <= T(0)) return T(0);
  if (u < T(1)) return k4 * u * kern_diw<T>(k4 * u);
  if (u < T(2)) return -k4 * u * kern_diw<T>(k4 * (T(2) - u));
  return T(0);
}

template <class T>
__device__ __forceinline__ T kern_dig(T u) {        // :62-72
  const T k4 = T(4);
  if (u <= T(0)) return T(0);
  if (u < T(1)) return k4 / u * kern_diw<T>(k4 * u);
  if (u < T(2)) return -k4 / u * kern_diw<T>(k4 * (T(2) - u));
  return T(0);
}

template <class T>
__device__ __forceinline__ T kern_dw(T u) {         // W'(u)/u, :74-81
  const double du = u;
  if (u < T(1)) return T(kInvPi * (-3.0 + 2.25 * du));
  if (u < T(2)) {
    const double s = (double)(T(2) - u);
    return T(-kInvPi * 0.75 * s * s / du);
  }
  return T(0);
}

template <class T>
__device__ __forceinline__ T kern_dwdh(T u) {       // -(3W + uW'), :83-90
  const double du = u;
  if (u < T(1)) return T(kInvPi * (-3.0 + 7.5 * du * du - 4.5 * du * du * du));
  if (u < T(2)) {
    const T lin = T(-6) + T(12) * u;                // integer-coefficient part stays single
    return T(kInvPi * ((double)lin - 7.5 * du * du + 1.5 * du * du * du));
  }
  return T(0);
}

template <class T>
__device__ __forceinline__ T kern_dphidh(T u) {     // softening d(phi)/dh, :92-99
  const double du = u;
  if (u < T(1)) {
    const T a = T(2) * u * u;
    return T(1.4 - (double)a + 1.5 * du * du * du * du - 0.6 * du * du * du * du * du);
  }
  if (u < T(2)) {
    const T a = T(4) * u * u;
    const T b = T(4) * u * u * u;
    return T(1.6 - (double)a + (double)b - 1.5 * du * du * du * du + 0.2 * du * du * du * du * du);
  }
  return T(0);
}

// ---- the same family in FP32 with FMA polynomials (mixed form) -----------
// u = r/h; the reference's 4 pi * (1/pi) products are folded to 4.  Used only
// by the mixed form, whose pair arithmetic is FP32 by definition; the plain
// form keeps the reference's double-promoted evaluation above.
namespace f32k {
constexpr float kIPi = 0.318309886183791f;
__device__ __forceinline__ float w(float u) {
  if (u < 1.f) return kIPi * __builtin_fmaf(u * u, __builtin_fmaf(0.75f, u, -1.5f), 1.f);
  if (u < 2.f) {
    const float t = 2.f - u;
    return (0.25f * kIPi) * t * t * t;
  }
  return 0.f;
}
// W'(u)/u; iu = 1/u (used on the outer branch only, where u >= 1)
__device__ __forceinline__ float dw(float u, float iu) {
  if (u < 1.f) return kIPi * __builtin_fmaf(2.25f, u, -3.f);
  if (u < 2.f) {
    const float t = 2.f - u;
    return (-0.75f * kIPi) * t * t * iu;
  }
  return 0.f;
}
__device__ __forceinline__ float dwdh(float u) {
  if (u < 1.f) return kIPi * __builtin_fmaf(u * u, __builtin_fmaf(-4.5f, u, 7.5f), -3.f);
  if (u < 2.f) return kIPi * __builtin_fmaf(u, __builtin_fmaf(u, __builtin_fmaf(1.5f, u, -7.5f), 12.f), -6.f);
  return 0.f;
}
__device__ __forceinline__ float dphidh(float u) {
  const float u2 = u * u;
  if (u < 1.f) return __builtin_fmaf(u2, __builtin_fmaf(u2, __builtin_fmaf(-0.6f, u, 1.5f), -2.f), 1.4f);
  if (u < 2.f)
    return __builtin_fmaf(u2, __builtin_fmaf(u, __builtin_fmaf(u, __builtin_fmaf(0.2f, u, -1.5f), 4.f), -4.f), 1.6f);
  return 0.f;
}
// 4 pi y^2 W(y)
__device__ __forceinline__ float diw(float y) {
  const float y2 = y * y;
  if (y < 1.f) return 4.f * y2 * __builtin_fmaf(y2, __builtin_fmaf(0.75f, y, -1.5f), 1.f);
  if (y < 2.f) {
    const float t = 2.f - y;
    return y2 * t * t * t;
  }
  return 0.f;
}
__device__ __forceinline__ float digdh(float u) {
  if (u <= 0.f) return 0.f;
  if (u < 1.f) return 4.f * u * diw(4.f * u);
  if (u < 2.f) return -4.f * u * diw(4.f * (2.f - u));
  return 0.f;
}
// digdh(u) / u^2; iu = 1/u
__device__ __forceinline__ float dig(float u, float iu) {
  if (u <= 0.f) return 0.f;
  if (u < 1.f) return 4.f * iu * diw(4.f * u);
  if (u < 2.f) return -4.f * iu * diw(4.f * (2.f - u));
  return 0.f;
}
}  // namespace f32k

// ---- gas + radiation EOS (dev_compute_pressure.cu:3-111) -----------------

struct EosConst {
  double density_unit, energy_unit, pressure_unit, amu, boltzmann, rad_const, gamma, temp_base;
  int newton;
};

// Ferrari's closed form for x^4 + q x + r = 0, :3-49
template <class T>
__device__ __forceinline__ T quartic_ferrari(T q, T r) {
  const T kh = T(0.5 * 0.125 * (double)q * (double)q);
  const T root = dsqrt(T((double)(kh * kh) - (double)(r * r * r) / 27.0));
  const T p1 = kh + root;
  const T p2 = kh - root;
  const T third = T(1.0 / 3);
  const T y = dpow(p1, third) - dpow(dabs(p2), third);
  const T b2 = dsqrt(T(2.0 * (double)y));
  const T c2 = T((double)y - 0.5 * (double)q / (double)b2);
  T x = T(0.5 * (double)(dsqrt(b2 * b2 - T(4) * c2) - b2));
  if (p1 == -p2) {                                  // radiation dominated
    const T quarter = T(0.25);
    x = dpow(-r - q * dpow(-r - q * dpow(-r, quarter), quarter), quarter);
  }
  if (p2 > T(0)) x = -(r + r * r * r * r / (q * q * q * q)) / q;  // gas dominated
  return x;
}

// FP64 Newton for the same quartic (q >= 0, r < 0).  x^4 + qx + r is convex
// and increasing on x > 0, so Newton started above the root at
// min((-r)^(1/4), -r/q) descends monotonically onto it.
__device__ __forceinline__ double quartic_newton(double q, double r) {
  if (!(r < 0.0)) return 0.0;
  double x = pow(-r, 0.25);
  if (q > 0.0) x = fmin(x, -r / q);
  for (int it = 0; it < 100; ++it) {
    const double x2 = x * x;
    const double dx = (x2 * x2 + q * x + r) / (4.0 * x2 * x + q);
    x -= dx;
    if (fabs(dx) <= 1e-15 * x) break;
  }
  return x;
}

// pressure from (density, energy, mean_mu) in the EOS's own units, :83-96
template <class T>
__device__ __forceinline__ T eos_pressure(T density, T energy, T mean_mu, const EosConst& c) {
  const T tb = T(c.temp_base);
  const T r = T((double)(-density * (energy / (tb * tb))) / (c.rad_const * (double)tb * (double)tb));
  const T q = T(c.boltzmann / c.rad_const / (c.gamma - 1.0) * (double)density /
                (double)(mean_mu * tb * tb * tb));
  const T x = c.newton ? T(quartic_newton((double)q, (double)r)) : quartic_ferrari<T>(q, r);
  const T temp = tb * x;
  const T p_gas = T((double)density * (c.boltzmann * (double)temp / (double)mean_mu));
  const T p_rad = T(c.rad_const / 3.0 * (double)temp * (double)temp * (double)temp * (double)temp);
  return T((double)(p_gas + p_rad) / c.pressure_unit);
}

// P for one hydro record {rho, P, u, mu/amu}, :99-111
template <class T>
__device__ __forceinline__ T eos_record(const V4<T>& hd, const EosConst& c) {
  const T density = T((double)hd.x * c.density_unit);
  const T energy = T((double)hd.z * c.energy_unit);
  const T mean_mu = T((double)hd.w * c.amu);
  return eos_pressure<T>(density, energy, mean_mu, c);
}

// FP64 pressure (mixed form): every step in double, Newton root
__device__ __forceinline__ double eos_pressure_fp64(double rho, double u, double mu, const EosConst& c) {
  const double density = rho * c.density_unit, energy = u * c.energy_unit, mean_mu = mu * c.amu;
  const double tb = c.temp_base;
  const double r = -density * (energy / (tb * tb)) / (c.rad_const * tb * tb);
  const double q = c.boltzmann / c.rad_const / (c.gamma - 1.0) * density / (mean_mu * tb * tb * tb);
  const double temp = tb * quartic_newton(q, r);
  const double p_gas = density * (c.boltzmann * temp / mean_mu);
  const double p_rad = c.rad_const / 3.0 * temp * temp * temp * temp;
  return (p_gas + p_rad) / c.pressure_unit;
}

// ---- artificial viscosity (dev_art_visc.cu:8-32) -------------------------
// returns {Pi_ij, v_sig_ij}
template <class T>
struct Visc {
  T pi, vsig;
};

template <class T>
__device__ __forceinline__ Visc<T> art_visc(T range_i, T range_j, T ds, T drdv, const V4<T>& hi,
                                            const V4<T>& di, const V4<T>& hj, const V4<T>& dj,
                                            T alpha, T beta) {
  T ci = dsqrt(T(5.0 / 3.0 * (double)hi.y / (double)hi.x));
  T cj = dsqrt(T(5.0 / 3.0 * (double)hj.y / (double)hj.x));
  Visc<T> out{T(0), ci + cj};
  if (drdv < T(0)) {
    out.vsig -= beta * drdv / ds;
    T mu = drdv / ((ci + cj) * ds);
    // Balsara factors |div v| / (|div v| + |curl v| + 1e-4 c/h)
    ci = T((double)dabs(di.x) / ((double)(dabs(di.x) + di.y) + 0.0001 * (double)ci / (0.5 * (double)range_i)));
    cj = T((double)dabs(dj.x) / ((double)(dabs(dj.x) + dj.y) + 0.0001 * (double)cj / (0.5 * (double)range_j)));
    mu *= (ci + cj);
    out.pi = (hi.y / (hi.x * hi.x) + hj.y / (hj.x * hj.x)) * (-alpha * mu + beta * (mu * mu));
  }
  return out;
}

}  // namespace cudasph
