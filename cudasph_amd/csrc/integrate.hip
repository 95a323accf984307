// Integration and bookkeeping kernels: gravity packing, predictor and
// corrector (step.cpp:37-61, 128-134), adding gravity with the force and
// torque diagnostics (step.cpp:97-118), the timestep minimum
// (compute_timestep.cpp:3-25) and the energy / angular momentum sums
// (system_statistics.cpp:4-84).
//
// The plain form keeps the reference's types: FP32 state, a double dt, and a
// float half step dth.  The mixed form integrates an FP64 copy of x, v and u
// and refreshes the FP32 arrays the SPH passes read (positions as hi + lo
// words).  Reductions are two-level (per-block partials, then one block in
// fixed order) and accumulate in FP64.
#include "kernels.h"
#include "sph_math.h"

namespace cudasph {

constexpr int kRedBlock = 256;

// CUDA_gravity.cpp:36-56: (x,y,z,m) and (vx,vy,vz,range^2); padding records
// copy real positions scaled by 1.1 with zero mass
template <class T>
__global__ __launch_bounds__(kBlock) void k_gravity_pack(const V4<T>* __restrict__ pos, const V4<T>* __restrict__ vel,
                                                          V4<T>* __restrict__ gpos, V4<T>* __restrict__ gvel, int n,
                                                          int n_dev) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n_dev) return;
  if (i < n) {
    const V4<T> p = pos[i], v = vel[i];
    gpos[i] = {p.x, p.y, p.z, v.w};
    gvel[i] = {v.x, v.y, v.z, p.w * p.w};
  } else {
    const V4<T> p = pos[(i - n) % n];
    gpos[i] = {T((double)p.x * 1.1), T((double)p.y * 1.1), T((double)p.z * 1.1), T(0)};
    gvel[i] = {T(0), T(0), T(0), T(0)};
  }
}

template <class T>
void launch_gravity_pack(const V4<T>* pos, const V4<T>* vel, V4<T>* gpos, V4<T>* gvel, int n, int n_dev,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_gravity_pack<T>, dim3(ceil_div(n_dev, kBlock)), dim3(kBlock), 0, s, pos, vel, gpos, gvel, n,
                     n_dev);
  CS_CHECK_LAUNCH();
}
template void launch_gravity_pack<float>(const F4*, const F4*, F4*, F4*, int, int, hipStream_t);
template void launch_gravity_pack<double>(const D4*, const D4*, D4*, D4*, int, int, hipStream_t);

__global__ __launch_bounds__(kBlock) void k_gravity_pack_mixed(const D4* __restrict__ pos64, const F4* __restrict__ pos,
                                                                const F4* __restrict__ vel, F4* __restrict__ hi,
                                                                F4* __restrict__ lo, F4* __restrict__ gvel, int n,
                                                                int n_dev) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n_dev) return;
  if (i < n) {
    const D4 x = pos64[i];
    const F4 p = pos[i];
    const float hx = (float)x.x, hy = (float)x.y, hz = (float)x.z;
    hi[i] = {hx, hy, hz, vel[i].w};
    lo[i] = {(float)(x.x - hx), (float)(x.y - hy), (float)(x.z - hz), 0.f};
    gvel[i] = {0.f, 0.f, 0.f, p.w * p.w};
  } else {
    // padding: massless and far away (r^2 overflows in FP32, so rsq gives 0
    // and the record adds exactly nothing, with or without its zero mass)
    hi[i] = {1e30f, 1e30f, 1e30f, 0.f};
    lo[i] = {0.f, 0.f, 0.f, 0.f};
    gvel[i] = {0.f, 0.f, 0.f, 0.f};
  }
}

void launch_gravity_pack_mixed(const D4* pos64, const F4* pos, const F4* vel, F4* hi, F4* lo, F4* gvel, int n,
                               int n_dev, hipStream_t s) {
  hipLaunchKernelGGL(k_gravity_pack_mixed, dim3(ceil_div(n_dev, kBlock)), dim3(kBlock), 0, s, pos64, pos, vel, hi,
                     lo, gvel, n, n_dev);
  CS_CHECK_LAUNCH();
}

// step.cpp:37-61.  old_vu keeps (vx, vy, vz, u) for the corrector.
template <class T>
__global__ __launch_bounds__(kBlock) void k_predict(V4<T>* __restrict__ pos, V4<T>* __restrict__ vel,
                                                     V4<T>* __restrict__ hydro, V4<T>* __restrict__ bdots,
                                                     V4<T>* __restrict__ grav, V4<T>* __restrict__ old_vu,
                                                     const double* __restrict__ dt_p, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double dt = *dt_p;
  const T dth = T(0.5 * dt);
  V4<T> p = pos[i], v = vel[i], h = hydro[i], d = bdots[i];
  p.x = T((double)p.x + (double)v.x * dt);
  p.y = T((double)p.y + (double)v.y * dt);
  p.z = T((double)p.z + (double)v.z * dt);
  old_vu[i] = {v.x, v.y, v.z, h.z};
  v.x += d.x * dth;
  v.y += d.y * dth;
  v.z += d.z * dth;
  h.z += d.w * dth;
  d.x = d.y = d.z = T(0);
  pos[i] = p;
  vel[i] = v;
  hydro[i] = h;
  bdots[i] = d;
  V4<T> g = grav[i];
  g.x = g.y = g.z = T(0);
  grav[i] = g;
}

template <class T>
void launch_predict(V4<T>* pos, V4<T>* vel, V4<T>* hydro, V4<T>* bdots, V4<T>* grav, V4<T>* old_vu,
                    const double* dt, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_predict<T>, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, s, pos, vel, hydro, bdots, grav,
                     old_vu, dt, n);
  CS_CHECK_LAUNCH();
}
template void launch_predict<float>(F4*, F4*, F4*, F4*, F4*, F4*, const double*, int, hipStream_t);
template void launch_predict<double>(D4*, D4*, D4*, D4*, D4*, D4*, const double*, int, hipStream_t);

// step.cpp:128-134: v = v_old + a (dt_new/2 + dth_old)
template <class T>
__global__ __launch_bounds__(kBlock) void k_correct(V4<T>* __restrict__ vel, V4<T>* __restrict__ hydro,
                                                     const V4<T>* __restrict__ bdots, const V4<T>* __restrict__ old_vu,
                                                     const double* __restrict__ dt_new,
                                                     const double* __restrict__ dt_old, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double f = 0.5 * (*dt_new) + (double)T(0.5 * (*dt_old));
  const V4<T> o = old_vu[i], d = bdots[i];
  V4<T> v = vel[i];
  v.x = T((double)o.x + (double)d.x * f);
  v.y = T((double)o.y + (double)d.y * f);
  v.z = T((double)o.z + (double)d.z * f);
  vel[i] = v;
  hydro[i].z = T((double)o.w + (double)d.w * f);
}

template <class T>
void launch_correct(V4<T>* vel, V4<T>* hydro, const V4<T>* bdots, const V4<T>* old_vu, const double* dt_new,
                    const double* dt_old, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_correct<T>, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, s, vel, hydro, bdots, old_vu, dt_new,
                     dt_old, n);
  CS_CHECK_LAUNCH();
}
template void launch_correct<float>(F4*, F4*, const F4*, const F4*, const double*, const double*, int, hipStream_t);
template void launch_correct<double>(D4*, D4*, const D4*, const D4*, const double*, const double*, int,
                                     hipStream_t);

// FP32 views of the FP64 state: positions as hi + lo words
__device__ __forceinline__ void refresh_one(const D4& x, const D4& v, double u, F4& pos, F4& lo, F4& vel,
                                            F4& hydro) {
  const float hx = (float)x.x, hy = (float)x.y, hz = (float)x.z;
  pos.x = hx;
  pos.y = hy;
  pos.z = hz;
  lo = {(float)(x.x - hx), (float)(x.y - hy), (float)(x.z - hz), 0.f};
  vel.x = (float)v.x;
  vel.y = (float)v.y;
  vel.z = (float)v.z;
  hydro.z = (float)u;
}

__global__ __launch_bounds__(kBlock) void k_predict_mixed(D4* __restrict__ pos64, D4* __restrict__ vel64,
                                                           double* __restrict__ u64, D4* __restrict__ old64,
                                                           F4* __restrict__ pos, F4* __restrict__ pos_lo,
                                                           F4* __restrict__ vel, F4* __restrict__ hydro,
                                                           F4* __restrict__ bdots, F4* __restrict__ grav,
                                                           const double* __restrict__ dt_p, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double dt = *dt_p, dth = 0.5 * dt;
  D4 x = pos64[i], v = vel64[i];
  double u = u64[i];
  F4 d = bdots[i];
  x.x += v.x * dt;
  x.y += v.y * dt;
  x.z += v.z * dt;
  old64[i] = {v.x, v.y, v.z, u};
  v.x += (double)d.x * dth;
  v.y += (double)d.y * dth;
  v.z += (double)d.z * dth;
  u += (double)d.w * dth;
  pos64[i] = x;
  vel64[i] = v;
  u64[i] = u;
  F4 p = pos[i], vv = vel[i], h = hydro[i], lo;
  refresh_one(x, v, u, p, lo, vv, h);
  pos[i] = p;
  pos_lo[i] = lo;
  vel[i] = vv;
  hydro[i] = h;
  d.x = d.y = d.z = 0.f;
  bdots[i] = d;
  F4 g = grav[i];
  g.x = g.y = g.z = 0.f;
  grav[i] = g;
}

void launch_predict_mixed(D4* pos64, D4* vel64, double* u64, D4* old64, F4* pos, F4* pos_lo, F4* vel, F4* hydro,
                          F4* bdots, F4* grav, const double* dt, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_predict_mixed, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, s, pos64, vel64, u64, old64, pos,
                     pos_lo, vel, hydro, bdots, grav, dt, n);
  CS_CHECK_LAUNCH();
}

__global__ __launch_bounds__(kBlock) void k_correct_mixed(D4* __restrict__ vel64, double* __restrict__ u64,
                                                           const D4* __restrict__ old64, F4* __restrict__ vel,
                                                           F4* __restrict__ hydro, const F4* __restrict__ bdots,
                                                           const double* __restrict__ dt_new,
                                                           const double* __restrict__ dt_old, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double f = 0.5 * (*dt_new) + 0.5 * (*dt_old);
  const D4 o = old64[i];
  const F4 d = bdots[i];
  D4 v = vel64[i];
  v.x = o.x + (double)d.x * f;
  v.y = o.y + (double)d.y * f;
  v.z = o.z + (double)d.z * f;
  const double u = o.w + (double)d.w * f;
  vel64[i] = v;
  u64[i] = u;
  F4 vv = vel[i];
  vv.x = (float)v.x;
  vv.y = (float)v.y;
  vv.z = (float)v.z;
  vel[i] = vv;
  hydro[i].z = (float)u;
}

void launch_correct_mixed(D4* vel64, double* u64, const D4* old64, F4* vel, F4* hydro, const F4* bdots,
                          const double* dt_new, const double* dt_old, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_correct_mixed, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, s, vel64, u64, old64, vel, hydro,
                     bdots, dt_new, dt_old, n);
  CS_CHECK_LAUNCH();
}

__global__ __launch_bounds__(kBlock) void k_refresh_mixed(const D4* __restrict__ pos64, const D4* __restrict__ vel64,
                                                           const double* __restrict__ u64, F4* __restrict__ pos,
                                                           F4* __restrict__ pos_lo, F4* __restrict__ vel,
                                                           F4* __restrict__ hydro, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  F4 p = pos[i], vv = vel[i], h = hydro[i], lo;
  refresh_one(pos64[i], vel64[i], u64[i], p, lo, vv, h);
  vv.w = (float)vel64[i].w;
  pos[i] = p;
  pos_lo[i] = lo;
  vel[i] = vv;
  hydro[i] = h;
}

void launch_refresh_mixed(const D4* pos64, const D4* vel64, const double* u64, F4* pos, F4* pos_lo, F4* vel,
                          F4* hydro, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_refresh_mixed, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, s, pos64, vel64, u64, pos, pos_lo,
                     vel, hydro, n);
  CS_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// reductions: each block reduces 6 doubles into partials[6*block + c]; one
// block then sums the partials in index order.
__device__ __forceinline__ void block_sum6(double v[6], double* partials) {
  __shared__ double s[6][kRedBlock];
  for (int c = 0; c < 6; ++c) s[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int off = kRedBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int c = 0; c < 6; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x < 6) partials[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(kRedBlock) void k_final_sum6(const double* __restrict__ partials, int nb,
                                                           double* __restrict__ out6) {
  __shared__ double s[6][kRedBlock];
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += kRedBlock)
    for (int c = 0; c < 6; ++c) v[c] += partials[6 * b + c];
  for (int c = 0; c < 6; ++c) s[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int off = kRedBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int c = 0; c < 6; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x < 6) out6[threadIdx.x] = s[threadIdx.x][0];
}

long reduce_partials(int n) { return 6L * ceil_div(n, kRedBlock) + 6; }

template <class T>
__global__ __launch_bounds__(kRedBlock) void k_add_gravity(V4<T>* __restrict__ bdots, const V4<T>* __restrict__ grav,
                                                            const V4<T>* __restrict__ pos,
                                                            const V4<T>* __restrict__ vel, int n,
                                                            double* __restrict__ partials) {
  const int i = blockIdx.x * kRedBlock + threadIdx.x;
  double v[6] = {0, 0, 0, 0, 0, 0};
  if (i < n) {
    V4<T> d = bdots[i];
    const V4<T> g = grav[i], p = pos[i];
    d.x += g.x;
    d.y += g.y;
    d.z += g.z;
    bdots[i] = d;
    const T m = vel[i].w;
    v[0] = m * d.x;
    v[1] = m * d.y;
    v[2] = m * d.z;
    v[3] = m * (p.y * d.z - p.z * d.y);
    v[4] = m * (p.z * d.x - p.x * d.z);
    v[5] = m * (p.x * d.y - p.y * d.x);
  }
  block_sum6(v, partials);
}

template <class T>
void launch_add_gravity(V4<T>* bdots, const V4<T>* grav, const V4<T>* pos, const V4<T>* vel, int n,
                        double* partials, double* out6, hipStream_t s) {
  const int nb = ceil_div(n, kRedBlock);
  hipLaunchKernelGGL(k_add_gravity<T>, dim3(nb), dim3(kRedBlock), 0, s, bdots, grav, pos, vel, n, partials);
  CS_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_final_sum6, dim3(1), dim3(kRedBlock), 0, s, partials, nb, out6);
  CS_CHECK_LAUNCH();
}
template void launch_add_gravity<float>(F4*, const F4*, const F4*, const F4*, int, double*, double*, hipStream_t);
template void launch_add_gravity<double>(D4*, const D4*, const D4*, const D4*, int, double*, double*, hipStream_t);

template <class T>
__global__ __launch_bounds__(kRedBlock) void k_stats(const V4<T>* __restrict__ pos, const V4<T>* __restrict__ vel,
                                                      const V4<T>* __restrict__ hydro, const V4<T>* __restrict__ grav,
                                                      int n, double* __restrict__ partials) {
  const int i = blockIdx.x * kRedBlock + threadIdx.x;
  double v[6] = {0, 0, 0, 0, 0, 0};
  if (i < n) {
    const V4<T> p = pos[i], u = vel[i];
    const double m = u.w;
    v[0] = 0.5 * m * ((double)u.x * u.x + (double)u.y * u.y + (double)u.z * u.z);
    v[1] = m * (double)hydro[i].z;
    v[2] = 0.5 * m * (double)grav[i].w;
    v[3] = m * ((double)p.y * u.z - (double)p.z * u.y);
    v[4] = m * ((double)p.z * u.x - (double)p.x * u.z);
    v[5] = m * ((double)p.x * u.y - (double)p.y * u.x);
  }
  block_sum6(v, partials);
}

template <class T>
void launch_stats(const V4<T>* pos, const V4<T>* vel, const V4<T>* hydro, const V4<T>* grav, int n,
                  double* partials, double* out6, hipStream_t s) {
  const int nb = ceil_div(n, kRedBlock);
  hipLaunchKernelGGL(k_stats<T>, dim3(nb), dim3(kRedBlock), 0, s, pos, vel, hydro, grav, n, partials);
  CS_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_final_sum6, dim3(1), dim3(kRedBlock), 0, s, partials, nb, out6);
  CS_CHECK_LAUNCH();
}
template void launch_stats<float>(const F4*, const F4*, const F4*, const F4*, int, double*, double*, hipStream_t);
template void launch_stats<double>(const D4*, const D4*, const D4*, const D4*, int, double*, double*, hipStream_t);

__global__ __launch_bounds__(kRedBlock) void k_stats_mixed(const D4* __restrict__ pos64, const D4* __restrict__ vel64,
                                                            const double* __restrict__ u64,
                                                            const F4* __restrict__ grav, int n,
                                                            double* __restrict__ partials) {
  const int i = blockIdx.x * kRedBlock + threadIdx.x;
  double v[6] = {0, 0, 0, 0, 0, 0};
  if (i < n) {
    const D4 p = pos64[i], u = vel64[i];
    const double m = u.w;
    v[0] = 0.5 * m * (u.x * u.x + u.y * u.y + u.z * u.z);
    v[1] = m * u64[i];
    v[2] = 0.5 * m * (double)grav[i].w;
    v[3] = m * (p.y * u.z - p.z * u.y);
    v[4] = m * (p.z * u.x - p.x * u.z);
    v[5] = m * (p.x * u.y - p.y * u.x);
  }
  block_sum6(v, partials);
}

void launch_stats_mixed(const D4* pos64, const D4* vel64, const double* u64, const F4* grav, int n,
                        double* partials, double* out6, hipStream_t s) {
  const int nb = ceil_div(n, kRedBlock);
  hipLaunchKernelGGL(k_stats_mixed, dim3(nb), dim3(kRedBlock), 0, s, pos64, vel64, u64, grav, n, partials);
  CS_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_final_sum6, dim3(1), dim3(kRedBlock), 0, s, partials, nb, out6);
  CS_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// The rest of system_statistics.cpp:19-64: bounding box, mass-weighted sums
// for the centre of mass, and the h-solve counts n_ngb (min, max, sum, sum of
// squares), plus the particles whose list length ngb_count differs from
// n_ngb and the longest list.  Each block writes kDiag values to its partials
// row; one block folds the rows in a fixed order (sums of integers are exact
// in doubles, min and max in any order), so the result is deterministic.
constexpr int kDiag = 19;
enum { kDgMin = 0, kDgMax = 3, kDgM = 6, kDgMx = 7, kDgMv = 10, kDgSum = 13, kDgSq = 14, kDgMis = 15,
       kDgNmin = 16, kDgNmax = 17, kDgCmax = 18 };

__device__ __forceinline__ bool diag_is_min(int c) { return c < kDgMax || c == kDgNmin; }
__device__ __forceinline__ bool diag_is_max(int c) { return (c >= kDgMax && c < kDgM) || c == kDgNmax || c == kDgCmax; }

__device__ __forceinline__ double diag_fold(int c, double a, double b) {
  if (diag_is_min(c)) return fmin(a, b);
  if (diag_is_max(c)) return fmax(a, b);
  return a + b;
}

// fold v[kDiag] over the block: within each wavefront by shuffles, then the
// wavefront results in LDS in wavefront order; thread 0 holds the result
__device__ __forceinline__ void block_diag(double v[kDiag]) {
  __shared__ double s[kRedBlock / 64][kDiag];
  for (int c = 0; c < kDiag; ++c)
    for (int o = 32; o > 0; o >>= 1) v[c] = diag_fold(c, v[c], __shfl_xor(v[c], o, 64));
  const int w = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0)
    for (int c = 0; c < kDiag; ++c) s[w][c] = v[c];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int c = 0; c < kDiag; ++c) {
      double a = s[0][c];
      for (int k = 1; k < kRedBlock / 64; ++k) a = diag_fold(c, a, s[k][c]);
      v[c] = a;
    }
}

__device__ __forceinline__ void diag_identity(double v[kDiag]) {
  for (int c = 0; c < kDiag; ++c) v[c] = diag_is_min(c) ? 1e300 : diag_is_max(c) ? -1e300 : 0.0;
}

template <class P>
__global__ __launch_bounds__(kRedBlock) void k_diagnostics(const V4<P>* __restrict__ pos,
                                                            const V4<P>* __restrict__ vel,
                                                            const int* __restrict__ n_ngb,
                                                            const int* __restrict__ ngb_count, int n,
                                                            double* __restrict__ partials) {
  const int i = blockIdx.x * kRedBlock + threadIdx.x;
  double v[kDiag];
  diag_identity(v);
  if (i < n) {
    const V4<P> p = pos[i], u = vel[i];
    const double m = u.w, x = p.x, y = p.y, z = p.z;
    v[kDgMin] = v[kDgMax] = x;
    v[kDgMin + 1] = v[kDgMax + 1] = y;
    v[kDgMin + 2] = v[kDgMax + 2] = z;
    v[kDgM] = m;
    v[kDgMx] = m * x;
    v[kDgMx + 1] = m * y;
    v[kDgMx + 2] = m * z;
    v[kDgMv] = m * (double)u.x;
    v[kDgMv + 1] = m * (double)u.y;
    v[kDgMv + 2] = m * (double)u.z;
    const int k = n_ngb[i], c = ngb_count[i];
    v[kDgSum] = k;
    v[kDgSq] = (double)k * k;
    v[kDgMis] = c != k;
    v[kDgNmin] = v[kDgNmax] = k;
    v[kDgCmax] = c;
  }
  block_diag(v);
  if (threadIdx.x == 0)
    for (int c = 0; c < kDiag; ++c) partials[(size_t)kDiag * blockIdx.x + c] = v[c];
}

__global__ __launch_bounds__(kRedBlock) void k_diag_final(const double* __restrict__ partials, int nb,
                                                           double* __restrict__ out) {
  double v[kDiag];
  diag_identity(v);
  for (int b = threadIdx.x; b < nb; b += kRedBlock)
    for (int c = 0; c < kDiag; ++c) v[c] = diag_fold(c, v[c], partials[(size_t)kDiag * b + c]);
  block_diag(v);
  if (threadIdx.x == 0)
    for (int c = 0; c < kDiag; ++c) out[c] = v[c];
}

long diag_partials(int n) { return (long)kDiag * (ceil_div(n, kRedBlock) + 1); }

template <class P>
void launch_diagnostics(const V4<P>* pos, const V4<P>* vel, const int* n_ngb, const int* ngb_count, int n,
                        double* partials, double* out, hipStream_t s) {
  const int nb = ceil_div(n, kRedBlock);
  hipLaunchKernelGGL(k_diagnostics<P>, dim3(nb), dim3(kRedBlock), 0, s, pos, vel, n_ngb, ngb_count, n, partials);
  CS_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(kRedBlock), 0, s, partials, nb, out);
  CS_CHECK_LAUNCH();
}
template void launch_diagnostics<float>(const F4*, const F4*, const int*, const int*, int, double*, double*,
                                        hipStream_t);
template void launch_diagnostics<double>(const D4*, const D4*, const int*, const int*, int, double*, double*,
                                         hipStream_t);

// ---------------------------------------------------------------------------
// timestep: per-block minimum, then one block's minimum (exact in any order)
__device__ __forceinline__ void block_min(double v, double* out) {
  __shared__ double s[kRedBlock];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = kRedBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] = fmin(s[threadIdx.x], s[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

template <class T>
__global__ __launch_bounds__(kRedBlock) void k_timestep(const V4<T>* __restrict__ pos, const V4<T>* __restrict__ hydro,
                                                         const V4<T>* __restrict__ hydro2,
                                                         const V4<T>* __restrict__ bdots, int n, T cn_vel, T cn_eth,
                                                         double* __restrict__ partials) {
  const int i = blockIdx.x * kRedBlock + threadIdx.x;
  double v = 1.0e30;
  if (i < n) {
    const T vsig = hydro2[i].x;
    const double dt_vel = cn_vel * dabs(pos[i].w / vsig);
    const double dt_eth = cn_eth * dabs(hydro[i].z / bdots[i].w);
    v = 1.0 / (1.0 / dt_vel + 1.0 / dt_eth);
    if (!(v < 1.0e30)) v = 1.0e30;     // NaN and +inf leave the running minimum unchanged
  }
  block_min(v, partials + blockIdx.x);
}

__global__ __launch_bounds__(kRedBlock) void k_timestep_mixed(const F4* __restrict__ pos,
                                                               const double* __restrict__ u64,
                                                               const F4* __restrict__ hydro2,
                                                               const F4* __restrict__ bdots, int n, float cn_vel,
                                                               float cn_eth, double* __restrict__ partials) {
  const int i = blockIdx.x * kRedBlock + threadIdx.x;
  double v = 1.0e30;
  if (i < n) {
    const double dt_vel = (double)cn_vel * fabs((double)pos[i].w / (double)hydro2[i].x);
    const double dt_eth = (double)cn_eth * fabs(u64[i] / (double)bdots[i].w);
    v = 1.0 / (1.0 / dt_vel + 1.0 / dt_eth);
    if (!(v < 1.0e30)) v = 1.0e30;
  }
  block_min(v, partials + blockIdx.x);
}

__global__ __launch_bounds__(kRedBlock) void k_final_min(const double* __restrict__ partials, int nb,
                                                          int round_float, double* __restrict__ out) {
  double v = 1.0e30;
  for (int b = threadIdx.x; b < nb; b += kRedBlock) v = fmin(v, partials[b]);
  __shared__ double s[kRedBlock];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = kRedBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] = fmin(s[threadIdx.x], s[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = round_float ? (double)(float)s[0] : s[0];
}

template <class T>
void launch_timestep(const V4<T>* pos, const V4<T>* hydro, const V4<T>* hydro2, const V4<T>* bdots, int n,
                     float cn_vel, float cn_eth, int round_float, double* partials, double* dt_out,
                     hipStream_t s) {
  const int nb = ceil_div(n, kRedBlock);
  hipLaunchKernelGGL(k_timestep<T>, dim3(nb), dim3(kRedBlock), 0, s, pos, hydro, hydro2, bdots, n, T(cn_vel),
                     T(cn_eth), partials);
  CS_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_final_min, dim3(1), dim3(kRedBlock), 0, s, partials, nb, round_float, dt_out);
  CS_CHECK_LAUNCH();
}
template void launch_timestep<float>(const F4*, const F4*, const F4*, const F4*, int, float, float, int, double*,
                                     double*, hipStream_t);
template void launch_timestep<double>(const D4*, const D4*, const D4*, const D4*, int, float, float, int, double*,
                                      double*, hipStream_t);

void launch_timestep_mixed(const F4* pos, const double* u64, const F4* hydro2, const F4* bdots, int n,
                           float cn_vel, float cn_eth, double* partials, double* dt_out, hipStream_t s) {
  const int nb = ceil_div(n, kRedBlock);
  hipLaunchKernelGGL(k_timestep_mixed, dim3(nb), dim3(kRedBlock), 0, s, pos, u64, hydro2, bdots, n, cn_vel, cn_eth,
                     partials);
  CS_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_final_min, dim3(1), dim3(kRedBlock), 0, s, partials, nb, 0, dt_out);
  CS_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(kBlock) void k_gather(const R* __restrict__ src, R* __restrict__ dst,
                                                    const KV* __restrict__ perm, int n) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t < n) dst[t] = src[perm[t].value];
}

template <class R>
void launch_gather(const R* src, R* dst, const KV* perm, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_gather<R>, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, s, src, dst, perm, n);
  CS_CHECK_LAUNCH();
}
template void launch_gather<F4>(const F4*, F4*, const KV*, int, hipStream_t);
template void launch_gather<D4>(const D4*, D4*, const KV*, int, hipStream_t);
template void launch_gather<double>(const double*, double*, const KV*, int, hipStream_t);
template void launch_gather<uint32_t>(const uint32_t*, uint32_t*, const KV*, int, hipStream_t);

// plain 16-byte-per-lane copy: the reference point for scaling memory
// counters (a copy of B bytes moves exactly B in and B out)
__global__ __launch_bounds__(kBlock) void k_copy16(const F4* __restrict__ src, F4* __restrict__ dst, long n) {
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  if (t < n) dst[t] = src[t];
}
void launch_copy16(const F4* src, F4* dst, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_copy16, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, s, src, dst, n);
  CS_CHECK_LAUNCH();
}

// FP32 FMA issue-rate probe: 8 independent chains per lane with runtime
// coefficients (nothing folds), single-lane v_fma_f32 (the library is
// compiled without packed FP32), 2 flops per FMA; each lane writes its sum
template <int FORM>
__global__ __launch_bounds__(kBlock) void k_fma_peak(float* __restrict__ out, int iters, float b, float c) {
  constexpr int kChains = FORM ? 16 : 8;
  float x[kChains];
#pragma unroll
  for (int k = 0; k < kChains; ++k) x[k] = (float)(threadIdx.x + k) * 1e-3f;
  // per-lane coefficients, so the accumulate form takes vector operands
  const float y = b + (float)(threadIdx.x & 7) * 1e-6f, z = c + (float)(threadIdx.x & 3) * 1e-9f;
#pragma unroll 4
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < kChains; ++k) x[k] = FORM ? __builtin_fmaf(y, z, x[k]) : __builtin_fmaf(x[k], b, c);
  }
  float sum = 0.f;
#pragma unroll
  for (int k = 0; k < kChains; ++k) sum += x[k];
  out[(long)blockIdx.x * kBlock + threadIdx.x] = sum;
}
// The gravity kernel's expanded equal-mass pair body alone (11 VALU + rsq
// per pair) on 8 records held in registers, 16 pairs per round: no LDS, no
// memory in the loop, 8 wavefronts per SIMD.
// The records are wave-uniform, as in the kernel; d_i moves a little each
// round so nothing is loop-invariant.  VARIANT 1 replaces the rsq by a
// multiply (12 VALU, no transcendental).  Each lane writes its sums.
template <int VARIANT>
__global__ __launch_bounds__(kBlock) void k_pair_body(float* __restrict__ out, int rounds, float4 r0, float4 r1,
                                                      float4 r2, float4 r3) {
  float4 rec[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float4 b = (k & 3) == 0 ? r0 : (k & 3) == 1 ? r1 : (k & 3) == 2 ? r2 : r3;
    rec[k] = make_float4(b.x + 0.01f * k, b.y - 0.02f * k, b.z + 0.03f * k, b.w + 0.001f * k);
  }
  float dx = 1.f + 1e-3f * (threadIdx.x & 63), dy = 0.5f, dz = -0.25f;
  float sa = 0.f, mx = 0.f, my = 0.f, mz = 0.f, pot = 0.f;
  for (int it = 0; it < rounds; ++it) {
    const float ci = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, 4e-6f)));
    const float bx = -2.f * dx, by = -2.f * dy, bz = -2.f * dz;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const float4 pj = rec[k & 7];
      const float d2 = __builtin_fmaf(pj.x, bx, __builtin_fmaf(pj.y, by, __builtin_fmaf(pj.z, bz, ci + pj.w)));
      const float inv = VARIANT == 0 ? __builtin_amdgcn_rsqf(d2) : d2 * 0.5f;
      const float s3 = inv * (inv * inv);
      sa += s3;
      mx = __builtin_fmaf(pj.x, s3, mx);
      my = __builtin_fmaf(pj.y, s3, my);
      mz = __builtin_fmaf(pj.z, s3, mz);
      pot -= inv;
    }
    dx += 1e-7f;
  }
  out[(long)blockIdx.x * kBlock + threadIdx.x] = sa + mx + my + mz + pot;
}
void launch_pair_body(float* out, int blocks, int rounds, int variant, hipStream_t s) {
  const float4 r0 = make_float4(0.1f, 0.2f, 0.3f, 0.14f), r1 = make_float4(-0.2f, 0.1f, 0.05f, 0.0525f),
               r2 = make_float4(0.3f, -0.1f, 0.2f, 0.14f), r3 = make_float4(0.05f, 0.05f, -0.3f, 0.095f);
  if (variant)
    hipLaunchKernelGGL(k_pair_body<1>, dim3(blocks), dim3(kBlock), 0, s, out, rounds, r0, r1, r2, r3);
  else
    hipLaunchKernelGGL(k_pair_body<0>, dim3(blocks), dim3(kBlock), 0, s, out, rounds, r0, r1, r2, r3);
  CS_CHECK_LAUNCH();
}

void launch_fma_peak(float* out, int blocks, int iters, int form, hipStream_t s) {
  if (form)
    hipLaunchKernelGGL(k_fma_peak<1>, dim3(blocks), dim3(kBlock), 0, s, out, iters / 2, 0.999f, 1e-3f);
  else
    hipLaunchKernelGGL(k_fma_peak<0>, dim3(blocks), dim3(kBlock), 0, s, out, iters, 0.999f, 1e-3f);
  CS_CHECK_LAUNCH();
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_set_range(V4<T>* __restrict__ pos, const double* __restrict__ r, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) pos[i].w = T(r[i]);
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_get_range(const V4<T>* __restrict__ pos, double* __restrict__ r, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) r[i] = (double)pos[i].w;
}
template <class T>
void launch_set_range(V4<T>* pos, const double* range, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_set_range<T>, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, s, pos, range, n);
  CS_CHECK_LAUNCH();
}
template <class T>
void launch_get_range(const V4<T>* pos, double* range, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_get_range<T>, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, s, pos, range, n);
  CS_CHECK_LAUNCH();
}
template void launch_set_range<float>(F4*, const double*, int, hipStream_t);
template void launch_set_range<double>(D4*, const double*, int, hipStream_t);
template void launch_get_range<float>(const F4*, double*, int, hipStream_t);
template void launch_get_range<double>(const D4*, double*, int, hipStream_t);

__global__ __launch_bounds__(kBlock) void k_iota_kv(KV* kv, int n) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t < n) kv[t].value = (uint32_t)t;
}
void launch_iota_kv(KV* kv, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_iota_kv, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, s, kv, n);
  CS_CHECK_LAUNCH();
}

}  // namespace cudasph
