/*
 * CPU oracle for the SPH hot path of 3gx/CUDAsph.
 *
 * TEST INFRASTRUCTURE ONLY -- see sph_oracle.h.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
 *
 * Every function restates one piece of the reference (file:line cited at the
 * function).  The instantiation with T = float reproduces the reference's
 * arithmetic types exactly: wherever a float operand meets a double literal
 * the reference computes in double and rounds back to float, and so does this
 * code (the promotions are written out as explicit double temporaries).  The
 * library is built with -ffp-contract=off so no FMA changes a rounding.  The
 * instantiation with T = double is the FP64 oracle: the same steps in double.
 *
 * Deviations from the reference, each deliberate and mirrored by the GPU path:
 *  - powf is evaluated as (float)pow((double)x,(double)y): a correctly rounded
 *    float power (the reference's device powf is not correctly rounded and its
 *    bits are not available), so CPU and GPU agree bit for bit.
 *  - rsqrtf in the gravity kernel is 1.0f/sqrtf(x) (correctly rounded).
 *  - the density/hydro/reaction loops run over the neighbour count that
 *    find_ngb actually wrote, not the h-solve count (they differ in a rare
 *    bracket edge case, where the reference reads stale list slots).
 *  - the walk's read of node n+1 (phantom node, see walk_sums) returns a zero
 *    node, as the reference's out-of-range texture fetch does; fix_phantom=1
 *    removes it instead.
 *
 * Pinning (what checks each function; tests/test_oracle_*.py):
 *  kernel functions: closed forms, normalisation, derivative identities.
 *  tree + walk + find_ngb: brute-force O(N^2) neighbour sets, heap invariants.
 *  solve_range: convergence invariant |f| <= |f'| rel_err range, brute count.
 *  density: FP64 brute-force pair sums, uniform lattice rho ~ rho0.
 *  EOS: quartic residual, gas and radiation limits, FP64 Newton root.
 *  art_visc: two-particle closed form, Balsara limits.
 *  hydro + reaction: total momentum conservation, pair antisymmetry.
 *  gravity: Newtonian far field, softening continuity, force = -dphi/dr,
 *           Plummer closed form, FP64 brute force, dreal error-free property.
 *  PH key: bijection and unit-step adjacency.
 *  step / timestep / statistics: closed forms (free particle, constant
 *           acceleration), timestep formula, diagnostics formula.
 */
#include "sph_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <class T> struct R4 { T x, y, z, w; };
template <class T> struct R3 { T x, y, z; };

constexpr double kInvPi = 0.318309886183791;   // dev_sph_kernels.cu:5
constexpr double kFourPi = 12.5663706143592;   // dev_sph_kernels.cu:19

/* ------------------------------------------------------------------------ */
/* float power, correctly rounded through double (see header comment)       */
inline float pw(float x, float y) { return (float)std::pow((double)x, (double)y); }
inline double pw(double x, double y) { return std::pow(x, y); }

/* ------------------------------------------------------------------------ */
/* M4 kernel family, dev_sph_kernels.cu:7-99.  u = r/h, support u < 2.      */
template <class T> T k_w(T u) {                      // :7-16
  const double d = u;
  if (u < 0) return T(kInvPi);
  if (u < 1) return T(kInvPi * (1 - 1.5 * d * d + 0.75 * d * d * d));
  if (u < 2) {
    const double t = T(2 - u);
    return T(kInvPi * 0.25 * t * t * t);
  }
  return 0;
}
template <class T> T k_diw(T u) {                    // :18-20
  const double d = u;
  return T(kFourPi * d * d * double(k_w<T>(u)));
}
template <class T> T k_iw(T u) {                     // :22-34
  if (u < 0) return 0;
  if (u < 1) {
    const T u3 = u * u * u;
    const double a = u3, b = u;
    return T(4 * (a / 3.0 - 0.3 * a * b * b + 0.125 * a * a));
  }
  if (u < 2) {
    const T u3 = u * u * u;
    const T u6 = u3 * u3;                            // float product in the reference
    const double a = u3, b = u;
    return T(4 * (-1.0 / 60 + 2.0 / 3 * a - 0.75 * a * b + 0.3 * a * b * b - double(u6) / 24.0));
  }
  return 1;
}
template <class T> T k_ig(T u) {                     // :36-48
  const T t = u - 1;
  return k_iw<T>(T(4) * (1 - std::sqrt(t * t)));
}
template <class T> T k_digdh(T u) {                  // :50-60
  const T a = 4;
  if (u <= 0) return 0;
  if (u < 1) return a * u * k_diw<T>(a * u);
  if (u < 2) return -a * u * k_diw<T>(a * (2 - u));
  return 0;
}
template <class T> T k_dig(T u) {                    // :62-72
  const T a = 4;
  if (u <= 0) return 0;
  if (u < 1) return a / u * k_diw<T>(a * u);
  if (u < 2) return -a / u * k_diw<T>(a * (2 - u));
  return 0;
}
template <class T> T k_dw(T u) {                     // :74-81
  const double d = u;
  if (u < 1) return T(kInvPi * (-3 + 2.25 * d));
  if (u < 2) {
    const double t = T(2 - u);
    return T(-kInvPi * 0.75 * t * t / d);
  }
  return 0;
}
template <class T> T k_dwdh(T u) {                   // :83-90
  const double d = u;
  if (u < 1) return T(kInvPi * (-3 + 7.5 * d * d - 4.5 * d * d * d));
  if (u < 2) {
    const T lin = T(-6) + T(12) * u;                 // int-coefficient terms stay float
    return T(kInvPi * (double(lin) - 7.5 * d * d + 1.5 * d * d * d));
  }
  return 0;
}
template <class T> T k_dphidh(T u) {                 // :92-99
  const double d = u;
  if (u < 1) {
    const T q2 = T(2) * u * u;
    return T(1.4 - double(q2) + 1.5 * d * d * d * d - 0.6 * d * d * d * d * d);
  }
  if (u < 2) {
    const T q2 = T(4) * u * u, q3 = T(4) * u * u * u;
    return T(1.6 - double(q2) + double(q3) - 1.5 * d * d * d * d + 0.2 * d * d * d * d * d);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* kd-tree node word: int bits (body_id | split << 24), sph_accelerations.cpp:30-31 */
inline float meta_word(float, int bits) { float f; std::memcpy(&f, &bits, 4); return f; }
inline double meta_word(double, int bits) { return (double)bits; }
inline int meta_bits(float w) { int b; std::memcpy(&b, &w, 4); return b; }
inline int meta_bits(double w) { return (int)w; }

template <class T> struct Elem { int id; T p[3]; };

/* first-element-pivot quicksort, kd_tree.cpp:4-31.  The reference recurses
 * on both sides; the two index ranges are disjoint, so recursing on the
 * smaller one and looping on the larger one gives the same array and bounds
 * the recursion depth. */
template <class T> void qsort_ref(Elem<T>* d, int dim, int left, int right) {
  while (left < right) {
    const int lo = left, hi = right;
    const Elem<T> pivot = d[left];
    while (left < right) {
      while (d[right].p[dim] >= pivot.p[dim] && left < right) right--;
      if (left != right) { d[left] = d[right]; left++; }
      while (d[left].p[dim] <= pivot.p[dim] && left < right) left++;
      if (left != right) { d[right] = d[left]; right--; }
    }
    d[left] = pivot;
    const int ip = left;
    if (ip - lo < hi - ip) {
      if (lo < ip) qsort_ref(d, dim, lo, ip - 1);
      left = ip + 1; right = hi;
    } else {
      if (hi > ip) qsort_ref(d, dim, ip + 1, hi);
      left = lo; right = ip - 1;
    }
  }
}

/* left-balanced median split, kd_tree.cpp:39-95 */
template <class T>
void build_node(R4<T>* tree, int node, Elem<T>* d, int depth, int n) {
  const int split = depth % 3;
  qsort_ref(d, split, 0, n - 1);
  int m = 1;
  while (m <= n) m <<= 1;
  m >>= 1;                                      // largest power of two <= n
  const int r = n - (m - 1);                    // nodes on the last level
  const int lt = (r <= m / 2) ? (m - 2) / 2 + r : (m - 2) / 2 + m / 2;
  const int n_left = lt, n_right = n - lt - 1;
  const Elem<T>& e = d[lt];
  tree[node].x = e.p[0];
  tree[node].y = e.p[1];
  tree[node].z = e.p[2];
  tree[node].w = meta_word(T(0), e.id | (split << 24));
  if (n_left > 0) {
#pragma omp task if (n_left > 16384)
    build_node(tree, 2 * node, d, depth + 1, n_left);
  }
  if (n_right > 0) {
#pragma omp task if (n_right > 16384)
    build_node(tree, 2 * node + 1, d + lt + 1, depth + 1, n_right);
  }
#pragma omp taskwait
}

/* ------------------------------------------------------------------------ */
/* tree walk, dev_solve_range.cu:17-89 (sums) and dev_find_ngb.cu:13-73
 * (list).  Depth-first: pop a node, push its left then right child when the
 * query sphere reaches across the split plane (so the right child is visited
 * first), then test the popped node's own particle.  Children are pushed for
 * any node with 2p <= n_nodes, which for even n_nodes includes node n+1: the
 * reference fetches it out of range and gets a zero node (origin, id 0,
 * split 0). */
template <class T, class Visit>
int walk(const R4<T>& q, const R4<T>* tree, int n_nodes, bool fix_phantom, Visit&& visit, int* phantom = nullptr) {
  int stack[64];
  int top = 0;
  stack[0] = 1;
  int n_walk = 0;
  while (top > -1) {
    n_walk++;
    int id = stack[top];
    const bool ph = phantom && id == n_nodes + 1;   // bit 0: node n+1 read, bit 1: within range
    const R4<T> node = tree[id];
    const int split = (meta_bits(node.w) & (3 << 24)) >> 24;
    id *= 2;
    if (id <= n_nodes) {
      const bool has_right = !fix_phantom || id + 1 <= n_nodes;
      const T qc = split == 0 ? q.x : split == 1 ? q.y : q.z;
      const T nc = split == 0 ? node.x : split == 1 ? node.y : node.z;
      if (split < 3) {
        if (qc - q.w < nc) stack[top++] = id;
        if (qc + q.w > nc && has_right) stack[top++] = id + 1;
      }
    }
    top--;
    const T dx = q.x - node.x, dy = q.y - node.y, dz = q.z - node.z;
    const T ds2 = dx * dx + dy * dy + dz * dz;
    if (ds2 < q.w * q.w) visit(ds2, meta_bits(node.w) & ~(3 << 24));
    if (ph) *phantom |= ds2 < q.w * q.w ? 3 : 1;
  }
  return n_walk;
}

template <class T> struct Ngb { T f, step, count; };

template <class T>
Ngb<T> compute_ngb(const R4<T>& q, const R4<T>* tree, int n_nodes, const orc_params& p) {
  T f = 0, df = 0, cnt = 0;
  walk<T>(q, tree, n_nodes, p.fix_phantom != 0, [&](T ds2, int) {
    const T u = T(2) * std::sqrt(ds2) / q.w;
    f += k_ig<T>(u);
    df -= k_digdh<T>(u);
    cnt += 1;
  });
  f -= T(p.desired_ngb);
  df = T(double(df) * (1.0 / double(q.w)));
  if (df != 0) df = f / df;
  else df = (f < 0) ? -q.w : q.w;
  return {f, df, cnt};
}

template <class T> inline T tabs(T v) { return std::fabs(v); }

/* bracketing, dev_solve_range.cu:92-135 (f_scale = 0.1 is a double literal) */
template <class T>
R3<T> bracket_root(R4<T> q, const R4<T>* tree, int n_nodes, const orc_params& p) {
  const T rel = T(p.rel_err);
  Ngb<T> f = compute_ngb<T>(q, tree, n_nodes, p);
  T dx = f.step;
  if (tabs(dx) < rel * q.w)
    return {T(q.w - 0.5 * double(tabs(dx))), T(q.w + 0.5 * double(tabs(dx))), f.count};
  if (double(tabs(dx)) > 0.1 * double(q.w)) dx = T((dx < 0 ? -0.1 : 0.1) * double(q.w));
  T x_left = q.w, f_try = f.f;
  int iter = 0;
  while ((f_try * f.f >= 0) && (iter++ < p.max_iter)) {
    x_left = q.w;
    f_try = f.f;
    q.w -= dx;
    f = compute_ngb<T>(q, tree, n_nodes, p);
    dx = f.step;
    if (tabs(dx) < rel * q.w)
      return {T(q.w - 0.5 * double(tabs(dx))), T(q.w + 0.5 * double(tabs(dx))), f.count};
    if (double(tabs(dx)) > 0.1 * double(q.w)) dx = T((dx < 0 ? -0.1 : 0.1) * double(q.w));
  }
  if (x_left < q.w) return {x_left, q.w, f.count};
  return {q.w, x_left, f.count};
}

/* safeguarded Newton, dev_solve_range.cu:137-172; returns (range, count) */
template <class T>
std::pair<T, T> solve_range(R4<T> q, const R4<T>* tree, int n_nodes, const orc_params& p) {
  const T rel = T(p.rel_err);
  int iter = 0;
  bool keep = true;
  R3<T> root = bracket_root<T>(q, tree, n_nodes, p);
  q.w = T(0.5 * double(root.x + root.y));
  T count = root.z, dx = 0;
  if (tabs(root.x - root.y) > rel * q.w) {
    while (keep && (iter++ < p.max_iter)) {
      q.w -= dx;
      if ((q.w - root.y) * (q.w - root.x) >= 0) {
        dx = T(0.5 * double(root.y - root.x));
        q.w = T(0.5 * double(root.x + root.y));
      }
      const Ngb<T> f = compute_ngb<T>(q, tree, n_nodes, p);
      count = f.count;
      if (f.f < 0) root.x = q.w;
      else root.y = q.w;
      dx = f.step;
      keep = tabs(dx) > rel * q.w;
    }
  }
  return {q.w, count};
}

inline int n_norm(int n, int j) {             // CUDAsph.h:16-20
  n = ((n - 1) / j) * j + j;
  if (n == 0) n = j;
  return n;
}

/* ------------------------------------------------------------------------ */
/* gas + radiation EOS, dev_compute_pressure.cu:3-111                        */
template <class T> T find_root_ferrari(T q, T r) {     // :3-49
  const T kh = T(0.5 * 0.125 * double(q) * double(q));
  const double disc = double(kh * kh) - double(r * r * r) / 27.0;
  const T s = std::sqrt(T(disc));
  const T piece_1 = kh + s;
  const T piece_2 = kh - s;
  const T third = T(1.0 / 3);
  const T y = pw(piece_1, third) - pw(tabs(piece_2), third);
  const T b_2 = std::sqrt(T(2.0 * double(y)));
  const T c_2 = T(double(y) - 0.5 * double(q) / double(b_2));
  T x_3 = T(0.5 * double(std::sqrt(b_2 * b_2 - T(4) * c_2) - b_2));
  if (piece_1 == -piece_2) {                             // radiation dominated
    const T qq = T(0.25);
    x_3 = pw(-r - q * pw(-r - q * pw(-r, qq), qq), qq);
  }
  if (piece_2 > 0) x_3 = -(r + r * r * r * r / (q * q * q * q)) / q;
  return x_3;
}

/* FP64 Newton root of x^4 + q x + r = 0 (q >= 0, r < 0): the quartic is convex
 * and increasing for x > 0, so Newton from the upper bound
 * min((-r)^(1/4), -r/q) decreases monotonically onto the unique positive
 * root.  No reference counterpart (the reference uses Ferrari's formula). */
inline double find_root_newton(double q, double r) {
  if (!(r < 0)) return 0;
  double x = std::pow(-r, 0.25);
  if (q > 0) x = std::min(x, -r / q);
  for (int it = 0; it < 100; ++it) {
    const double x2 = x * x;
    const double f = x2 * x2 + q * x + r;
    const double fp = 4 * x2 * x + q;
    const double dx = f / fp;
    x -= dx;
    if (std::fabs(dx) <= 1e-15 * x) break;
  }
  return x;
}

template <class T> T get_pressure(T density, T energy, T mean_mu, const orc_params& p) {  // :83-96
  const T tb = T(p.eos_temp_base);
  const T r = T(double(-density * (energy / (tb * tb))) / (p.eos_rad_const * double(tb) * double(tb)));
  const T q = T(p.eos_boltzmann / p.eos_rad_const / (p.eos_gamma - 1) * double(density) /
                double(mean_mu * tb * tb * tb));
  T x;
  if (p.eos_newton) x = T(find_root_newton(double(q), double(r)));
  else x = find_root_ferrari<T>(q, r);
  const T temperature = tb * x;
  const T p_gas = T(double(density) * (p.eos_boltzmann * double(temperature) / double(mean_mu)));
  const T p_rad = T(p.eos_rad_const / 3.0 * double(temperature) * double(temperature) *
                    double(temperature) * double(temperature));
  return T(double(p_gas + p_rad) / p.eos_pressure_unit);
}

template <class T> void eos_one(R4<T>& hd, const orc_params& p) {      // :99-111
  const T density = T(double(hd.x) * p.eos_density_unit);
  const T energy = T(double(hd.z) * p.eos_energy_unit);
  const T mean_mu = T(double(hd.w) * p.eos_amu);
  hd.y = get_pressure<T>(density, energy, mean_mu, p);
}

/* ------------------------------------------------------------------------ */
/* artificial viscosity, dev_art_visc.cu:8-32; returns (Pi_ij, v_sig_ij)     */
template <class T>
std::pair<T, T> art_visc(T range_i, T range_j, T ds, T drdv, const R4<T>& hi, const R4<T>& di,
                         const R4<T>& hj, const R4<T>& dj, T alpha, T beta) {
  T c_i = std::sqrt(T(5.0 / 3.0 * double(hi.y) / double(hi.x)));
  T c_j = std::sqrt(T(5.0 / 3.0 * double(hj.y) / double(hj.x)));
  T visc = 0;
  T v_sig = c_i + c_j;
  if (drdv < 0) {
    v_sig -= beta * drdv / ds;
    visc = drdv / ((c_i + c_j) * ds);
    c_i = T(double(tabs(di.x)) /
            (double(tabs(di.x) + di.y) + 0.0001 * double(c_i) / (0.5 * double(range_i))));
    c_j = T(double(tabs(dj.x)) /
            (double(tabs(dj.x) + dj.y) + 0.0001 * double(c_j) / (0.5 * double(range_j))));
    visc *= (c_i + c_j);
    visc = (hi.y / (hi.x * hi.x) + hj.y / (hj.x * hj.x)) * (-alpha * visc + beta * (visc * visc));
  }
  return {visc, v_sig};
}

/* ------------------------------------------------------------------------ */
/* double-float accumulator, dreal.h:43-52 (Knuth two-sum + renormalise)    */
struct DReal { float hi, lo; };
inline DReal dsub(DReal x, float y) {
  const float t1 = x.hi - y;
  float e = t1 - x.hi;
  const float t2 = ((-y - e) + (x.hi - (t1 - e))) + x.lo;
  DReal c;
  c.hi = e = t1 + t2;
  c.lo = t2 - (e - t1);
  return c;
}
/* accumulation policy: dreal for the FP32 form, plain double for FP64 */
template <class T> struct Acc;
template <> struct Acc<float> {
  DReal v{0.f, 0.f};
  void sub(float y) { v = dsub(v, y); }
  float get() const { return v.hi; }
};
template <> struct Acc<double> {
  double v = 0;
  void sub(double y) { v -= y; }
  double get() const { return v; }
};

/* softened pair, sph_grav_kernel.cu:14-111.  pj = (x,y,z,m_j), vi.w/vj.w =
 * range^2.  Accumulates into a[0..3] = (ax, ay, az, phi). */
template <class T>
void grav_pair_spline(Acc<T>* a, const R4<T>& pi, const R4<T>& vi, const R4<T>& pj, const R4<T>& vj) {
  const T rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
  const T d2 = rx * rx + ry * ry + rz * rz;
  const T inv = T(1) / std::sqrt(d2);
  const T inv3 = inv * inv * inv;
  T s3 = pj.w * inv3;
  if (d2 >= vi.w && d2 >= vj.w) {
    a[0].sub(rx * s3); a[1].sub(ry * s3); a[2].sub(rz * s3);
    a[3].sub(pj.w * inv);
    return;
  }
  s3 = T(0.5f) * s3;
  if (d2 < vi.w) {
    const T u2 = d2 / vi.w, h = std::sqrt(vi.w), u = std::sqrt(u2), u3 = u2 * u;
    T gacc, gpot;
    if (u < T(0.5)) {
      gacc = T(32.0 / 3.0) + u2 * (T(32) * u - T(38.4));
      gpot = T(-2.8) + u2 * (T(16.0 / 3.0) + u2 * (T(6.4) * u - T(9.6)));
    } else {
      gacc = T(64.0 / 3.0) - T(48) * u + T(38.4) * u2 - T(32.0 / 3.0) * u3 - T(1.0 / 15.0) / u3;
      gpot = T(-3.2) + T(1.0 / 15.0) / u +
             u2 * (T(32.0 / 3.0) + u * (T(-16) + u * (T(9.6) - T(32.0 / 15.0) * u)));
    }
    gacc = T(0.5 * double(pj.w) * double(gacc) / double(vi.w * h));
    gpot = -gpot;
    a[0].sub(rx * gacc); a[1].sub(ry * gacc); a[2].sub(rz * gacc);
    a[3].sub(pj.w * gpot / h);
  } else {
    a[0].sub(rx * s3); a[1].sub(ry * s3); a[2].sub(rz * s3);
    a[3].sub(pj.w * inv);
  }
  if (d2 < vj.w) {
    const T u2 = d2 / vj.w, h = std::sqrt(vj.w), u = std::sqrt(u2), u3 = u2 * u;
    T gacc;
    if (u < T(0.5)) gacc = T(32.0 / 3.0) + u2 * (T(32) * u - T(38.4));
    else gacc = T(64.0 / 3.0) - T(48) * u + T(38.4) * u2 - T(32.0 / 3.0) * u3 - T(1.0 / 15.0) / u3;
    gacc = T(0.5 * double(pj.w) * double(gacc) / double(vj.w * h));
    a[0].sub(rx * gacc); a[1].sub(ry * gacc); a[2].sub(rz * gacc);
  } else {
    a[0].sub(rx * s3); a[1].sub(ry * s3); a[2].sub(rz * s3);
  }
}

/* Plummer-softened pair (the configs' softening; no reference counterpart):
 * a -= m r / (r^2+eps^2)^(3/2),  phi -= m / (r^2+eps^2)^(1/2).  The self pair
 * contributes -m/eps to phi and nothing to a. */
template <class T>
void grav_pair_plummer(Acc<T>* a, const R4<T>& pi, const R4<T>& pj, T eps2) {
  const T rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
  const T d2 = rx * rx + ry * ry + rz * rz + eps2;
  const T inv = T(1) / std::sqrt(d2);
  const T inv3 = inv * inv * inv;
  const T s3 = pj.w * inv3;
  a[0].sub(rx * s3); a[1].sub(ry * s3); a[2].sub(rz * s3);
  a[3].sub(pj.w * inv);
}

/* ------------------------------------------------------------------------ */
/* Peano-Hilbert key, sort_bodies.cpp:11-105 (Gadget-2's state tables; the
 * table values fix the curve and so the particle order) */
const int kPhQuad[24][2][2][2] = {
    {{{0, 7}, {1, 6}}, {{3, 4}, {2, 5}}}, {{{7, 4}, {6, 5}}, {{0, 3}, {1, 2}}},
    {{{4, 3}, {5, 2}}, {{7, 0}, {6, 1}}}, {{{3, 0}, {2, 1}}, {{4, 7}, {5, 6}}},
    {{{1, 0}, {6, 7}}, {{2, 3}, {5, 4}}}, {{{0, 3}, {7, 4}}, {{1, 2}, {6, 5}}},
    {{{3, 2}, {4, 5}}, {{0, 1}, {7, 6}}}, {{{2, 1}, {5, 6}}, {{3, 0}, {4, 7}}},
    {{{6, 1}, {7, 0}}, {{5, 2}, {4, 3}}}, {{{1, 2}, {0, 3}}, {{6, 5}, {7, 4}}},
    {{{2, 5}, {3, 4}}, {{1, 6}, {0, 7}}}, {{{5, 6}, {4, 7}}, {{2, 1}, {3, 0}}},
    {{{7, 6}, {0, 1}}, {{4, 5}, {3, 2}}}, {{{6, 5}, {1, 2}}, {{7, 4}, {0, 3}}},
    {{{5, 4}, {2, 3}}, {{6, 7}, {1, 0}}}, {{{4, 7}, {3, 0}}, {{5, 6}, {2, 1}}},
    {{{6, 7}, {5, 4}}, {{1, 0}, {2, 3}}}, {{{7, 0}, {4, 3}}, {{6, 1}, {5, 2}}},
    {{{0, 1}, {3, 2}}, {{7, 6}, {4, 5}}}, {{{1, 6}, {2, 5}}, {{0, 7}, {3, 4}}},
    {{{2, 3}, {1, 0}}, {{5, 4}, {6, 7}}}, {{{3, 4}, {0, 7}}, {{2, 5}, {1, 6}}},
    {{{4, 5}, {7, 6}}, {{3, 2}, {0, 1}}}, {{{5, 2}, {6, 1}}, {{4, 3}, {7, 0}}}};
const int kPhRotX[24] = {4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 0, 1, 2, 3, 17, 18, 19, 16, 23, 20, 21, 22};
const int kPhRotY[24] = {1, 2, 3, 0, 16, 17, 18, 19, 11, 8, 9, 10, 22, 23, 20, 21, 14, 15, 12, 13, 4, 5, 6, 7};
const int kPhQx[8] = {3, 0, 0, 2, 2, 0, 0, 1};
const int kPhQy[8] = {0, 1, 1, 2, 2, 3, 3, 0};
const int kPhSense[8] = {-1, -1, -1, +1, +1, -1, -1, -1};

long long ph_key(int x, int y, int z, int bits) {
  long long key = 0;
  int rot = 0, sense = 1;
  for (int b = bits - 1; b >= 0; --b) {
    const int q = kPhQuad[rot][(x >> b) & 1][(y >> b) & 1][(z >> b) & 1];
    key = (key << 3) + (sense == 1 ? q : 7 - q);
    for (int k = 0; k < kPhQx[q]; ++k) rot = kPhRotX[rot];
    for (int k = 0; k < kPhQy[q]; ++k) rot = kPhRotY[rot];
    sense *= kPhSense[q];
  }
  return key;
}

/* ------------------------------------------------------------------------ */
/* the simulation state and the passes of step()                            */
struct Stats { double T, U, W, E, J, F[3], Tq[3]; };

struct Base {
  virtual ~Base() {}
  virtual long get(const std::string& name, void* dst, long max_bytes) = 0;
  virtual long set(const std::string& name, const void* src, long n_bytes) = 0;
  virtual int pass(int id) = 0;
  virtual void gravity_range(int i0, int i1) = 0;
  double dt = 0, global_time = 0;
  int iteration = 0, n_overflow_reruns = 0, n_count_mismatch = 0, max_count = 0;
  int phantom_reads = 0, phantom_ngb = 0;
  long ngb_tot = 0;
  Stats st{};
  /* the rest of system_statistics.cpp:19-64: box, mass, centre of mass and
   * its velocity, J components, h-solve count mean / sigma / min / max */
  double r_min[3] = {0, 0, 0}, r_max[3] = {0, 0, 0}, m_tot = 0, r_com[3] = {0, 0, 0}, v_com[3] = {0, 0, 0};
  double Jv[3] = {0, 0, 0}, ngb_mean = 0, ngb_sigma = 0;
  int ngb_min = 0, ngb_max = 0;
};

template <class T> struct State : Base {
  orc_params p;
  int n, nd;
  std::vector<R4<T>> pos, vel, hydro, dots, bdots, hydro2, grav, tree, gpos, gvel;
  std::vector<uint32_t> hash;                   // (key, value) pairs
  std::vector<int> n_ngb, ngb_count, ngb_offset, ngb_list;
  std::vector<T> ngb_acc;
  std::vector<R3<T>> old_vel;
  std::vector<T> old_u;
  long cap;

  explicit State(const orc_params& pp) : p(pp), n(pp.n_bodies), nd(pp.n_dev) {
    pos.assign(nd, {}); vel.assign(nd, {}); hydro.assign(nd, {}); dots.assign(nd, {});
    bdots.assign(nd, {}); hydro2.assign(nd, {}); grav.assign(nd, {});
    tree.assign(nd + 2, {});
    hash.assign(2 * (size_t)nd, 0);
    n_ngb.assign(nd, 0); ngb_count.assign(nd, 0); ngb_offset.assign(nd + 1, 0);
    cap = 100L * nd;                            // ngb_tot_max, read_state.cpp:29
    ngb_list.assign(cap, 0); ngb_acc.assign(cap, 0);
    old_vel.assign(nd, {}); old_u.assign(nd, 0);
    for (int i = 0; i < nd; ++i) { hash[2 * i] = i; hash[2 * i + 1] = i; }
    for (int i = n; i < nd; ++i) {              // padding, read_state.cpp:75-89
      hash[2 * i] = 1u << 24;
      vel[i] = {T(1e10), T(1e10), T(1e10), 0};
      pos[i] = {T(1e10), T(1e10), T(1e10), 0};
      hydro[i].w = 0;
    }
#ifdef _OPENMP
    if (p.n_threads > 0) omp_set_num_threads(p.n_threads);
#endif
  }

  template <class V> long xfer(V& v, void* dst, const void* src, long nb) {
    const long have = (long)(v.size() * sizeof(v[0]));
    if (dst) { const long c = std::min(have, nb); std::memcpy(dst, v.data(), c); return c; }
    const long c = std::min(have, nb);
    std::memcpy(v.data(), src, c);
    return c;
  }
  long io(const std::string& nm, void* dst, const void* src, long nb) {
    if (nm == "pos") return xfer(pos, dst, src, nb);
    if (nm == "vel") return xfer(vel, dst, src, nb);
    if (nm == "hydro") return xfer(hydro, dst, src, nb);
    if (nm == "dots") return xfer(dots, dst, src, nb);
    if (nm == "bdots") return xfer(bdots, dst, src, nb);
    if (nm == "hydro2") return xfer(hydro2, dst, src, nb);
    if (nm == "grav") return xfer(grav, dst, src, nb);
    if (nm == "tree") return xfer(tree, dst, src, nb);
    if (nm == "hash") return xfer(hash, dst, src, nb);
    if (nm == "n_ngb") return xfer(n_ngb, dst, src, nb);
    if (nm == "ngb_count") return xfer(ngb_count, dst, src, nb);
    if (nm == "ngb_offset") return xfer(ngb_offset, dst, src, nb);
    if (nm == "ngb_list") return xfer(ngb_list, dst, src, nb);
    if (nm == "ngb_acc") return xfer(ngb_acc, dst, src, nb);
    return -1;
  }
  long get(const std::string& nm, void* dst, long mb) override { return io(nm, dst, nullptr, mb); }
  long set(const std::string& nm, const void* src, long nb) override { return io(nm, nullptr, src, nb); }

  /* kd_tree.cpp:97-114 + packing sph_accelerations.cpp:21-33 */
  void build_tree() {
    std::vector<Elem<T>> d(n);
    for (int i = 0; i < n; ++i) d[i] = {i, {pos[i].x, pos[i].y, pos[i].z}};
    const T big = T(1e10);
    for (int i = 0; i < nd + 2; ++i) tree[i] = {big, big, big, meta_word(T(0), -1)};
    tree[0] = {0, 0, 0, 0};
    tree[nd + 1] = {0, 0, 0, 0};                // out-of-range fetch reads zeros
#pragma omp parallel
#pragma omp single
    build_node<T>(tree.data(), 1, d.data(), 0, n);
  }

  /* dev_solve_range.cu:174-189 */
  void solve() {
#pragma omp parallel for schedule(dynamic, 64)
    for (int t = 0; t < nd; ++t) {
      const int i = (int)hash[2 * t + 1];
      R4<T> q = pos[i];
      std::pair<T, T> r{0, 0};
      if (q.w > 0) r = solve_range<T>(q, tree.data(), n, p);
      pos[i].w = r.first;
      n_ngb[i] = (int)r.second;
    }
  }

  /* sph_accelerations.cpp:58-72 */
  void offsets(const std::vector<int>& cnt) {
    ngb_offset[0] = 0;
    for (int i = 0; i < nd; ++i) ngb_offset[i + 1] = ngb_offset[i] + n_norm(cnt[i], 4);
    ngb_tot = ngb_offset[nd];
    if (ngb_tot > cap) {
      std::fprintf(stderr, "oracle: neighbour list overflow %ld > %ld\n", ngb_tot, cap);
      std::abort();
    }
  }

  /* dev_find_ngb.cu:75-87; writes at most the slot capacity and records the
   * true count; on overflow the offsets are rebuilt from the true counts */
  void find() {
    for (int pass = 0; pass < 2; ++pass) {
      int overflow = 0, reads = 0, took = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : overflow) reduction(+ : reads, took)
      for (int t = 0; t < nd; ++t) {
        const int i = (int)hash[2 * t + 1];
        const int off = ngb_offset[i];
        const int cap_i = ngb_offset[i + 1] - off;
        int c = 0, ph = 0;
        const int nw = walk<T>(pos[i], tree.data(), n, p.fix_phantom != 0, [&](T, int id) {
          if (c < cap_i) ngb_list[off + c] = id;
          c++;
        }, &ph);
        reads += ph & 1;
        took += (ph >> 1) & 1;
        hash[2 * t] = (uint32_t)nw;
        ngb_count[i] = c;
        if (c > cap_i) overflow = 1;
      }
      phantom_reads = reads;
      phantom_ngb = took;
      if (!overflow) break;
      n_overflow_reruns++;
      offsets(ngb_count);
    }
    int mm = 0, mx = 0;
    for (int i = 0; i < n; ++i) { mm += ngb_count[i] != n_ngb[i]; mx = std::max(mx, ngb_count[i]); }
    n_count_mismatch = mm;
    max_count = mx;
  }

  /* dev_density_loop.cu:29-111 */
  void density() {
#pragma omp parallel for schedule(dynamic, 64)
    for (int t = 0; t < nd; ++t) {
      const int i = (int)hash[2 * t + 1];
      const R4<T> bp = pos[i], bv = vel[i];
      T rho = 0, om = 0, psi = 0, den = 0, divv = 0, cx = 0, cy = 0, cz = 0;
      const int off = ngb_offset[i], c = ngb_count[i];
      for (int k = off; k < off + c; ++k) {
        const int j = ngb_list[k];
        const R4<T> pj = pos[j], vj = vel[j];
        const T dx = bp.x - pj.x, dy = bp.y - pj.y, dz = bp.z - pj.z;
        const T u = T(2) * std::sqrt(dx * dx + dy * dy + dz * dz) / bp.w;
        rho += vj.w * k_w<T>(u);
        om += vj.w * k_dwdh<T>(u);
        psi += vj.w * k_dphidh<T>(u);
        den -= k_digdh<T>(u);
        const T dvx = bv.x - vj.x, dvy = bv.y - vj.y, dvz = bv.z - vj.z;
        const T mdw = bv.w * k_dw<T>(u);
        divv += mdw * (dx * dvx + dy * dvy + dz * dvz);
        cx += mdw * (dy * dvz - dz * dvy);
        cy += mdw * (dz * dvx - dx * dvz);
        cz += mdw * (dx * dvy - dy * dvx);
      }
      const T h = T(0.5 * double(bp.w));
      const T h3 = h * h * h, h5 = h * h * h * h * h;
      if (den != 0) {
        om = T(double(om) * (1.0 / double(h3) / double(den)));
        psi = T(double(psi) * (1.0 / double(h) / double(den)));
      } else {
        om = psi = 0;
      }
      rho = T(double(rho) * (1.0 / double(h3)));
      divv = T(double(divv) * (1.0 / double(rho) / double(h5)));
      const T curl = std::sqrt(cx * cx + cy * cy + cz * cz) / rho / h5;
      hydro[i].x = rho;
      dots[i] = {divv, curl, om, psi};
    }
  }

  void eos() {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nd; ++i) eos_one<T>(hydro[i], p);
  }

  /* dev_hydro_dots.cu:25-113 */
  void hydro_pass() {
    const T alpha = T(p.av_alpha), beta = T(p.av_beta);
#pragma omp parallel for schedule(dynamic, 64)
    for (int t = 0; t < nd; ++t) {
      const int i = (int)hash[2 * t + 1];
      const R4<T> bp = pos[i], bv = vel[i], hd = hydro[i], dd = dots[i];
      const T prho2 = hd.y / (hd.x * hd.x);
      const double hw = 0.5 * double(bp.w);
      const double pow5 = hw * 0.5 * double(bp.w) * 0.5 * double(bp.w) * 0.5 * double(bp.w) * 0.5 * double(bp.w);
      const double pow2 = hw * 0.5 * double(bp.w);
      T ax = 0, ay = 0, az = 0, ud = 0, vsig = 0;
      const int off = ngb_offset[i], c = ngb_count[i];
      for (int k = off; k < off + c; ++k) {
        const int j = ngb_list[k];
        const R4<T> pj = pos[j], vj = vel[j], hj = hydro[j], dj = dots[j];
        const T dx = bp.x - pj.x, dy = bp.y - pj.y, dz = bp.z - pj.z;
        const T dvx = bv.x - vj.x, dvy = bv.y - vj.y, dvz = bv.z - vj.z;
        const T ds = std::sqrt(dx * dx + dy * dy + dz * dz);
        const T u = T(double(ds) / hw);
        const T drdv = dx * dvx + dy * dvy + dz * dvz;
        const std::pair<T, T> av = art_visc<T>(bp.w, pj.w, ds, drdv, hd, dd, hj, dj, alpha, beta);
        vsig = std::fmax(vsig, av.second);
        const double wdw = double(k_dw<T>(u)), wdig = double(k_dig<T>(u));
        const double pv = double(prho2) + 0.5 * double(av.first);
        T acc = T(pv * wdw / pow5);
        acc = T(double(acc) - (double(prho2 * dd.z) + 0.5 * double(dd.w)) / double(vj.w) * wdig / pow2);
        ax -= vj.w * acc * dx;
        ay -= vj.w * acc * dy;
        az -= vj.w * acc * dz;
        ngb_acc[k] = bv.w * acc;
        T udot = T(double(vj.w) * pv * wdw / pow5);
        udot = T(double(udot) - double(prho2 * dd.z * T(wdig)) / pow2);
        ud += udot * drdv;
      }
      bdots[i] = {ax, ay, az, ud};
      hydro2[i].x = vsig;
    }
  }

  /* host reaction scatter, sph_accelerations.cpp:104-115 (serial order) */
  void reaction() {
    for (int i = 0; i < n; ++i) {
      const int off = ngb_offset[i], c = ngb_count[i];
      for (int k = off; k < off + c; ++k) {
        const int j = ngb_list[k];
        const T acc = ngb_acc[k];
        bdots[j].x += acc * (pos[i].x - pos[j].x);
        bdots[j].y += acc * (pos[i].y - pos[j].y);
        bdots[j].z += acc * (pos[i].z - pos[j].z);
      }
    }
  }

  /* CUDA_gravity.cpp:36-56 packing + sph_grav_kernel.cu:179-195 */
  void gravity(int i_begin = 0, int i_end = -1) {
    if (i_end < 0) i_end = nd;
    gpos.assign(nd, {}); gvel.assign(nd, {});
    for (int i = 0; i < n; ++i) {
      gpos[i] = {pos[i].x, pos[i].y, pos[i].z, vel[i].w};
      gvel[i] = {vel[i].x, vel[i].y, vel[i].z, pos[i].w * pos[i].w};
    }
    for (int i = n, pc = 0; i < nd; ++i) {
      gpos[i] = {T(double(pos[pc].x) * 1.1), T(double(pos[pc].y) * 1.1), T(double(pos[pc].z) * 1.1), 0};
      if (++pc >= n) pc = 0;
    }
    const T eps2 = T(p.grav_eps) * T(p.grav_eps);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = i_begin; i < i_end; ++i) {
      Acc<T> a[4];
      const R4<T> pi = gpos[i], vi = gvel[i];
      if (p.grav_mode == 0)
        for (int j = 0; j < nd; ++j) grav_pair_spline<T>(a, pi, vi, gpos[j], gvel[j]);
      else
        for (int j = 0; j < nd; ++j) grav_pair_plummer<T>(a, pi, gpos[j], eps2);
      grav[i] = {a[0].get(), a[1].get(), a[2].get(), a[3].get()};
    }
  }

  /* step.cpp:97-118 */
  void add_gravity() {
    double F[3] = {0, 0, 0}, Tq[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
      R4<T>& d = bdots[i];
      d.x += grav[i].x; d.y += grav[i].y; d.z += grav[i].z;
      const T m = vel[i].w;
      F[0] += m * d.x; F[1] += m * d.y; F[2] += m * d.z;
      Tq[0] += m * (pos[i].y * d.z - pos[i].z * d.y);
      Tq[1] += m * (pos[i].z * d.x - pos[i].x * d.z);
      Tq[2] += m * (pos[i].x * d.y - pos[i].y * d.x);
    }
    for (int k = 0; k < 3; ++k) { st.F[k] = F[k]; st.Tq[k] = Tq[k]; }
  }

  /* compute_timestep.cpp:3-25; the FP32 form returns the float-rounded dt */
  double timestep() {
    const T cn1 = T(p.cn_vel), cn2 = T(p.cn_eth);
    double dt_min = 1.0e30;
    for (int i = 0; i < n; ++i) {
      const T v_sig = hydro2[i].x;
      const double dt_vel = cn1 * tabs(pos[i].w / v_sig);
      const double dt_eth = cn2 * tabs(hydro[i].z / bdots[i].w);
      dt_min = std::min(dt_min, 1.0 / (1.0 / dt_vel + 1.0 / dt_eth));
    }
    return double(T(dt_min));
  }

  /* step.cpp:37-61 */
  void predict() {
    const T dth = T(0.5 * dt);
    for (int i = 0; i < n; ++i) {
      pos[i].x = T(double(pos[i].x) + double(vel[i].x) * dt);
      pos[i].y = T(double(pos[i].y) + double(vel[i].y) * dt);
      pos[i].z = T(double(pos[i].z) + double(vel[i].z) * dt);
      old_vel[i] = {vel[i].x, vel[i].y, vel[i].z};
      vel[i].x += bdots[i].x * dth;
      vel[i].y += bdots[i].y * dth;
      vel[i].z += bdots[i].z * dth;
      old_u[i] = hydro[i].z;
      hydro[i].z += bdots[i].w * dth;
      bdots[i].x = bdots[i].y = bdots[i].z = 0;
      grav[i].x = grav[i].y = grav[i].z = 0;
    }
  }

  /* step.cpp:128-134; dth is the half of the step just taken */
  void correct(double dt_new, T dth) {
    const double f = 0.5 * dt_new + double(dth);
    for (int i = 0; i < n; ++i) {
      vel[i].x = T(double(old_vel[i].x) + double(bdots[i].x) * f);
      vel[i].y = T(double(old_vel[i].y) + double(bdots[i].y) * f);
      vel[i].z = T(double(old_vel[i].z) + double(bdots[i].z) * f);
      hydro[i].z = T(double(old_u[i]) + double(bdots[i].w) * f);
    }
  }

  /* system_statistics.cpp:4-84, accumulated in double */
  void stats() {
    double Tk = 0, U = 0, W = 0, J[3] = {0, 0, 0};
    double lo[3] = {double(pos[0].x), double(pos[0].y), double(pos[0].z)};
    double hi[3] = {lo[0], lo[1], lo[2]};
    double M = 0, mr[3] = {0, 0, 0}, mv[3] = {0, 0, 0}, sum = 0, sq = 0;
    int kmin = n, kmax = 0;
    for (int i = 0; i < n; ++i) {
      const double m = vel[i].w;
      const double x = pos[i].x, y = pos[i].y, z = pos[i].z;
      const double vx = vel[i].x, vy = vel[i].y, vz = vel[i].z;
      U += m * double(hydro[i].z);
      W += 0.5 * m * double(grav[i].w);
      Tk += 0.5 * m * (vx * vx + vy * vy + vz * vz);
      J[0] += m * (y * vz - z * vy);
      J[1] += m * (z * vx - x * vz);
      J[2] += m * (x * vy - y * vx);
      const double r[3] = {x, y, z}, v[3] = {vx, vy, vz};
      for (int c = 0; c < 3; ++c) {
        lo[c] = std::min(lo[c], r[c]);
        hi[c] = std::max(hi[c], r[c]);
        mr[c] += m * r[c];
        mv[c] += m * v[c];
      }
      M += m;
      kmin = std::min(kmin, n_ngb[i]);
      kmax = std::max(kmax, n_ngb[i]);
      sum += n_ngb[i];
      sq += double(n_ngb[i]) * n_ngb[i];
    }
    st.T = Tk; st.U = U; st.W = W; st.E = Tk + U + W;
    st.J = std::sqrt(J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
    for (int c = 0; c < 3; ++c) {
      r_min[c] = lo[c]; r_max[c] = hi[c]; Jv[c] = J[c];
      r_com[c] = mr[c] / M; v_com[c] = mv[c] / M;
    }
    m_tot = M;
    ngb_mean = sum / n;
    ngb_sigma = std::sqrt(sq / n - ngb_mean * ngb_mean);
    ngb_min = kmin; ngb_max = kmax;
  }

  /* sort_bodies.cpp:119-186: PH keys on the bounding cube, stable sort,
   * permute the seven state arrays */
  void ph_sort() {
    R4<T> lo = pos[0], hi = pos[0];
    for (int i = 0; i < n; ++i) {
      lo.x = std::min(lo.x, pos[i].x); lo.y = std::min(lo.y, pos[i].y); lo.z = std::min(lo.z, pos[i].z);
      hi.x = std::max(hi.x, pos[i].x); hi.y = std::max(hi.y, pos[i].y); hi.z = std::max(hi.z, pos[i].z);
    }
    const T size = std::max(hi.z - lo.z, std::max(hi.y - lo.y, hi.x - lo.x));
    const T fac = T(1.0 / double(size) * double(1LL << p.ph_bits));
    std::vector<std::pair<long long, int>> keys(n);
    for (int i = 0; i < n; ++i) {
      const int x = (int)((pos[i].x - lo.x) * fac), y = (int)((pos[i].y - lo.y) * fac),
                z = (int)((pos[i].z - lo.z) * fac);
      keys[i] = {ph_key(x, y, z, p.ph_bits), i};
    }
    std::stable_sort(keys.begin(), keys.end(),
                     [](const std::pair<long long, int>& a, const std::pair<long long, int>& b) {
                       return a.first < b.first;
                     });
    auto perm = [&](std::vector<R4<T>>& v) {
      std::vector<R4<T>> c(v.begin(), v.begin() + n);
      for (int i = 0; i < n; ++i) v[i] = c[keys[i].second];
    };
    perm(pos); perm(vel); perm(bdots); perm(hydro); perm(dots); perm(hydro2); perm(grav);
  }

  void sph_acc() {
    build_tree();
    solve();
    offsets(n_ngb);
    find();
    density();
    eos();
    hydro_pass();
    reaction();
  }

  /* step.cpp:13-147 */
  void step() {
    iteration++;
    if (p.ph_period > 0 && iteration % p.ph_period == 0) ph_sort();
    predict();
    const T dth = T(0.5 * dt);
    sph_acc();
    gravity();
    add_gravity();
    global_time += dt;
    stats();
    dt = timestep();
    correct(dt, dth);
  }

  void gravity_range(int i0, int i1) override { gravity(i0, i1); }

  int pass(int id) override {
    switch (id) {
      case ORC_TREE: build_tree(); break;
      case ORC_SOLVE: solve(); break;
      case ORC_OFFSETS: offsets(n_ngb); break;
      case ORC_FIND: find(); break;
      case ORC_DENSITY: density(); break;
      case ORC_EOS: eos(); break;
      case ORC_HYDRO: hydro_pass(); break;
      case ORC_REACTION: reaction(); break;
      case ORC_GRAVITY: gravity(); break;
      case ORC_ADD_GRAVITY: add_gravity(); break;
      case ORC_TIMESTEP: dt = timestep(); break;
      case ORC_PREDICT: predict(); break;
      case ORC_CORRECT: correct(dt, T(0.5 * dt)); break;
      case ORC_STATS: stats(); break;
      case ORC_PH_SORT: ph_sort(); break;
      case ORC_SPH_ACC: sph_acc(); break;
      case ORC_INIT: sph_acc(); gravity(); add_gravity(); stats(); dt = timestep(); break;
      case ORC_STEP: step(); break;
      default: return -1;
    }
    return 0;
  }
};

}  // namespace

/* ======================================================================== */
extern "C" {

void* orc_create(int real_bits, const orc_params* p) {
  if (real_bits == 32) return static_cast<Base*>(new State<float>(*p));
  if (real_bits == 64) return static_cast<Base*>(new State<double>(*p));
  return nullptr;
}
void orc_destroy(void* h) { delete static_cast<Base*>(h); }
long orc_get(void* h, const char* name, void* dst, long mb) { return static_cast<Base*>(h)->get(name, dst, mb); }
long orc_set(void* h, const char* name, const void* src, long nb) { return static_cast<Base*>(h)->set(name, src, nb); }
int orc_pass(void* h, int pass) { return static_cast<Base*>(h)->pass(pass); }

double orc_get_scalar(void* h, const char* name) {
  Base* b = static_cast<Base*>(h);
  const std::string s(name);
  if (s == "dt") return b->dt;
  if (s == "global_time") return b->global_time;
  if (s == "iteration") return b->iteration;
  if (s == "ngb_tot") return (double)b->ngb_tot;
  if (s == "overflow_reruns") return b->n_overflow_reruns;
  if (s == "count_mismatch") return b->n_count_mismatch;
  if (s == "max_count") return b->max_count;
  if (s == "phantom_reads") return b->phantom_reads;
  if (s == "phantom_ngb") return b->phantom_ngb;
  return NAN;
}
void orc_set_scalar(void* h, const char* name, double v) {
  Base* b = static_cast<Base*>(h);
  const std::string s(name);
  if (s == "dt") b->dt = v;
  else if (s == "global_time") b->global_time = v;
  else if (s == "iteration") b->iteration = (int)v;
}
void orc_gravity_range(void* h, int i0, int i1) { static_cast<Base*>(h)->gravity_range(i0, i1); }

void orc_stats(void* h, double* o) {
  const Stats& s = static_cast<Base*>(h)->st;
  o[0] = s.T; o[1] = s.U; o[2] = s.W; o[3] = s.E; o[4] = s.J;
  for (int k = 0; k < 3; ++k) { o[5 + k] = s.F[k]; o[8 + k] = s.Tq[k]; }
}

void orc_statistics(void* h, double* o) {
  const Base* b = static_cast<Base*>(h);
  orc_stats(h, o);
  double r = 0, v = 0;
  for (int c = 0; c < 3; ++c) {
    o[11 + c] = b->Jv[c];
    o[14 + c] = b->r_min[c];
    o[17 + c] = b->r_max[c];
    o[21 + c] = b->r_com[c];
    o[24 + c] = b->v_com[c];
    r += b->r_com[c] * b->r_com[c];
    v += b->v_com[c] * b->v_com[c];
  }
  o[20] = b->m_tot;
  o[27] = std::sqrt(r);
  o[28] = std::sqrt(v);
  o[29] = b->ngb_mean; o[30] = b->ngb_sigma; o[31] = b->ngb_min; o[32] = b->ngb_max;
  o[33] = b->n_count_mismatch; o[34] = b->max_count;
  o[35] = b->phantom_reads; o[36] = b->phantom_ngb;
  o[37] = b->n_overflow_reruns; o[38] = b->iteration; o[39] = b->global_time;
}

void orc_kernel_fns(int real_bits, int n, const double* u, double* out) {
  for (int i = 0; i < n; ++i) {
    double* o = out + 9 * (size_t)i;
    if (real_bits == 32) {
      const float x = (float)u[i];
      o[0] = k_w(x); o[1] = k_dw(x); o[2] = k_dwdh(x); o[3] = k_dphidh(x); o[4] = k_iw(x);
      o[5] = k_diw(x); o[6] = k_ig(x); o[7] = k_digdh(x); o[8] = k_dig(x);
    } else {
      const double x = u[i];
      o[0] = k_w(x); o[1] = k_dw(x); o[2] = k_dwdh(x); o[3] = k_dphidh(x); o[4] = k_iw(x);
      o[5] = k_diw(x); o[6] = k_ig(x); o[7] = k_digdh(x); o[8] = k_dig(x);
    }
  }
}

double orc_pressure(int real_bits, const orc_params* p, double rho, double u, double mu) {
  if (real_bits == 32) {
    R4<float> hd{(float)rho, 0, (float)u, (float)mu};
    eos_one<float>(hd, *p);
    return hd.y;
  }
  R4<double> hd{rho, 0, u, mu};
  eos_one<double>(hd, *p);
  return hd.y;
}

double orc_find_root(int real_bits, int newton, double q, double r) {
  if (newton) return find_root_newton(q, r);
  if (real_bits == 32) return find_root_ferrari<float>((float)q, (float)r);
  return find_root_ferrari<double>(q, r);
}

void orc_art_visc(int real_bits, float alpha, float beta, double range_i, double range_j, double ds,
                  double drdv, const double* hi, const double* di, const double* hj, const double* dj,
                  double* out) {
  if (real_bits == 32) {
    auto c = [](const double* v) { return R4<float>{(float)v[0], (float)v[1], (float)v[2], (float)v[3]}; };
    auto r = art_visc<float>((float)range_i, (float)range_j, (float)ds, (float)drdv, c(hi), c(di), c(hj),
                             c(dj), alpha, beta);
    out[0] = r.first; out[1] = r.second;
  } else {
    auto c = [](const double* v) { return R4<double>{v[0], v[1], v[2], v[3]}; };
    auto r = art_visc<double>(range_i, range_j, ds, drdv, c(hi), c(di), c(hj), c(dj), alpha, beta);
    out[0] = r.first; out[1] = r.second;
  }
}

void orc_dreal_sub(const float* x2, float y, float* out2) {
  const DReal r = dsub(DReal{x2[0], x2[1]}, y);
  out2[0] = r.hi; out2[1] = r.lo;
}

long long orc_ph_key(int x, int y, int z, int bits) { return ph_key(x, y, z, bits); }

void orc_brute_ngb(int n, const float* p4, int* count) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    int c = 0;
    const float r = p4[4 * i + 3];
    for (int j = 0; j < n; ++j) {
      const float dx = p4[4 * i] - p4[4 * j], dy = p4[4 * i + 1] - p4[4 * j + 1], dz = p4[4 * i + 2] - p4[4 * j + 2];
      if (dx * dx + dy * dy + dz * dz < r * r) c++;
    }
    count[i] = c;
  }
}

}  // extern "C"
