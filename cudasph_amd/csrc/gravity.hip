// Softened direct-N gravity (sph_grav_kernel.cu:14-195).
//
// Layout: one lane per i-particle, 256 i per block.  The block stages j
// records in LDS tiles of 512 with coalesced loads; all lanes of a wave then
// read the same j from LDS (a broadcast), so the loop body is VALU work plus
// one LDS read per j per wave.  (Reading j through scalar loads instead was
// measured ~4% slower at N = 2^20: the scalar cache misses once the j array
// outgrows L2.)  j runs in ascending order over all records, the reference's
// summation order.
//
// Variants (template parameters):
//   SOFT  0: M4 spline at the SPH range, half from each side (reference)
//         1: Plummer eps
//   ACC   0: per-pair double-float (dreal) accumulation (reference, dreal.h)
//         1: blocked compensation -- a plain FP32 partial over each tile of
//            kTile j, folded into a two-sum (hi,lo) total (mixed form)
//         2: plain accumulation in T (FP64 form)
//   EXACT 1: 1/sqrt correctly rounded (bit-exact with the CPU restatement);
//         0: hardware reciprocal square root (the reference uses rsqrtf)
//   HILO  positions carry low words (mixed form): r_ij = (hi_i-hi_j)+(lo_i-lo_j)
#include <type_traits>

#include "kernels.h"
#include "sph_math.h"

namespace cudasph {

struct DR {
  float hi, lo;
};
// two-sum subtraction with renormalisation (Knuth; the reference's dsub,
// dreal.h:43-52)
__device__ __forceinline__ void dr_sub(DR& x, float y) {
  const float t1 = x.hi - y;
  const float e = t1 - x.hi;
  const float t2 = ((-y - e) + (x.hi - (t1 - e))) + x.lo;
  const float s = t1 + t2;
  x.lo = t2 - (s - t1);
  x.hi = s;
}
__device__ __forceinline__ void dr_add(DR& x, float y) { dr_sub(x, -y); }

template <class T, int ACC>
struct GAcc;
template <>
struct GAcc<float, 0> {                      // per-pair dreal
  DR v{0.f, 0.f};
  __device__ __forceinline__ void sub(float y) { dr_sub(v, y); }
  __device__ __forceinline__ void tile_end() {}
  __device__ __forceinline__ float get() const { return v.hi; }
};
template <>
struct GAcc<float, 1> {                      // blocked compensation
  DR v{0.f, 0.f};
  float part = 0.f;
  __device__ __forceinline__ void sub(float y) { part -= y; }
  __device__ __forceinline__ void fsub(float a, float b) { part = __builtin_fmaf(-a, b, part); }
  // fold the block's partial into (hi, lo): Knuth's two-sum of hi and the
  // partial, its exact error added to lo (lo stays a small FP32 sum of the
  // rounding errors; no renormalisation is needed until get())
  __device__ __forceinline__ void tile_end() {
    const float s = v.hi + part;
    const float bp = s - v.hi;
    v.lo += (v.hi - (s - bp)) + (part - bp);
    v.hi = s;
    part = 0.f;
  }
  __device__ __forceinline__ float get() const { return v.hi + v.lo; }
};
template <>
struct GAcc<double, 2> {
  double v = 0.0;
  __device__ __forceinline__ void sub(double y) { v -= y; }
  __device__ __forceinline__ void tile_end() {}
  __device__ __forceinline__ double get() const { return v; }
};

template <bool EXACT>
__device__ __forceinline__ float inv_sqrt(float x) {
  if constexpr (EXACT) return 1.0f / __builtin_sqrtf(x);
  else return __builtin_amdgcn_rsqf(x);
}
template <bool EXACT>
// FP64: hardware reciprocal square root refined by two Newton steps (each
// squares the relative error; the result is within a few ulp of 1/sqrt(x)),
// in place of the long correctly rounded sqrt + divide sequences
__device__ __forceinline__ double inv_sqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * (1.5 - 0.5 * x * y * y);
  return y * (1.5 - 0.5 * x * y * y);
}

// spline-softened pair (sph_grav_kernel.cu:14-111); pj = (x,y,z,m),
// vi.w / vj.w = range^2
template <class T, bool EXACT, class A>
__device__ __forceinline__ void pair_spline(A* a, T rx, T ry, T rz, T ri2, T mj, T rj2) {
  const T d2 = rx * rx + ry * ry + rz * rz;
  const T inv = inv_sqrt<EXACT>(d2);
  const T inv3 = inv * inv * inv;
  T s3 = mj * inv3;
  if (d2 >= ri2 && d2 >= rj2) {
    a[0].sub(rx * s3);
    a[1].sub(ry * s3);
    a[2].sub(rz * s3);
    a[3].sub(mj * inv);
    return;
  }
  s3 = T(0.5f) * s3;
  if (d2 < ri2) {
    const T u2 = d2 / ri2, h = dsqrt(ri2), u = dsqrt(u2), u3 = u2 * u;
    T gacc, gpot;
    if (u < T(0.5)) {
      gacc = T(32.0 / 3.0) + u2 * (T(32) * u - T(38.4));
      gpot = T(-2.8) + u2 * (T(16.0 / 3.0) + u2 * (T(6.4) * u - T(9.6)));
    } else {
      gacc = T(64.0 / 3.0) - T(48) * u + T(38.4) * u2 - T(32.0 / 3.0) * u3 - T(1.0 / 15.0) / u3;
      gpot = T(-3.2) + T(1.0 / 15.0) / u + u2 * (T(32.0 / 3.0) + u * (T(-16) + u * (T(9.6) - T(32.0 / 15.0) * u)));
    }
    gacc = T(0.5 * (double)mj * (double)gacc / (double)(ri2 * h));
    gpot = -gpot;
    a[0].sub(rx * gacc);
    a[1].sub(ry * gacc);
    a[2].sub(rz * gacc);
    a[3].sub(mj * gpot / h);
  } else {
    a[0].sub(rx * s3);
    a[1].sub(ry * s3);
    a[2].sub(rz * s3);
    a[3].sub(mj * inv);
  }
  if (d2 < rj2) {
    const T u2 = d2 / rj2, h = dsqrt(rj2), u = dsqrt(u2), u3 = u2 * u;
    T gacc;
    if (u < T(0.5)) gacc = T(32.0 / 3.0) + u2 * (T(32) * u - T(38.4));
    else gacc = T(64.0 / 3.0) - T(48) * u + T(38.4) * u2 - T(32.0 / 3.0) * u3 - T(1.0 / 15.0) / u3;
    gacc = T(0.5 * (double)mj * (double)gacc / (double)(rj2 * h));
    a[0].sub(rx * gacc);
    a[1].sub(ry * gacc);
    a[2].sub(rz * gacc);
  } else {
    a[0].sub(rx * s3);
    a[1].sub(ry * s3);
    a[2].sub(rz * s3);
  }
}

// Plummer pair: a -= m r (r^2+eps^2)^-3/2, phi -= m (r^2+eps^2)^-1/2
template <class T, bool EXACT, class A>
__device__ __forceinline__ void pair_plummer(A* a, T rx, T ry, T rz, T mj, T eps2) {
  const T d2 = rx * rx + ry * ry + rz * rz + eps2;
  const T inv = inv_sqrt<EXACT>(d2);
  const T inv3 = inv * inv * inv;
  const T s3 = mj * inv3;
  a[0].sub(rx * s3);
  a[1].sub(ry * s3);
  a[2].sub(rz * s3);
  a[3].sub(mj * inv);
}

constexpr int kGravBlock = 256;
constexpr int kTile = 64;        // j per blocked-compensation partial (mixed form)
constexpr int kJTile = 512;      // j records staged in LDS per block iteration (reference-layout kernels)
constexpr int kGravParts = 8;    // j-parts of the mixed Plummer kernel, one per XCD (at most)
// j records per LDS tile of the mixed Plummer kernel: one 64-record sub-tile
// per wavefront.  Its 12 KB of LDS lets 8 blocks share a CU (the kernel's 63
// VGPRs allow 8 waves per SIMD), so the 4096 x 8 blocks at N = 2^20 run in
// exactly 16 rounds; 512-record tiles (24 KB) fit 6 blocks and left the last
// of 21 1/3 rounds a third full (measured 2 % slower)
constexpr int kJTileMixed = kGravBlock;

// Block-cooperative staging of j records into LDS.  Every lane of a wave then
// reads the same LDS address (a broadcast read), so one j record serves the
// whole wave, and the block's 256 lanes share one coalesced global read of
// the tile.
template <class R>
__device__ __forceinline__ int stage_tile(R* __restrict__ s, const R* __restrict__ g, int t0, int n_j) {
  const int cnt = min(kJTile, n_j - t0);
  __syncthreads();
  for (int q = threadIdx.x; q < cnt; q += kGravBlock) s[q] = g[t0 + q];
  __syncthreads();
  return cnt;
}

// reference-layout kernel (plain and FP64 forms): gpos = (x,y,z,m),
// gvel.w = range^2
template <class T, int SOFT, int ACC, bool EXACT>
__global__ __launch_bounds__(kGravBlock) void k_gravity(const V4<T>* __restrict__ gpos,
                                                        const V4<T>* __restrict__ gvel, V4<T>* __restrict__ out,
                                                        int n_i, int n_j, T eps2) {
  __shared__ V4<T> s_pos[kJTile];
  __shared__ T s_r2[SOFT == 0 ? kJTile : 1];
  const int i = blockIdx.x * kGravBlock + threadIdx.x;
  const int ii = i < n_i ? i : n_i - 1;
  const V4<T> pi = gpos[ii];
  const T ri2 = SOFT == 0 ? gvel[ii].w : T(0);
  GAcc<T, ACC> a[4];
  for (int t0 = 0; t0 < n_j; t0 += kJTile) {
    const int cnt = stage_tile(s_pos, gpos, t0, n_j);
    if constexpr (SOFT == 0) {
      for (int q = threadIdx.x; q < cnt; q += kGravBlock) s_r2[q] = gvel[t0 + q].w;
      __syncthreads();
    }
    auto body = [&](int j) {
      const V4<T> pj = s_pos[j];
      const T rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
      if constexpr (SOFT == 0) pair_spline<T, EXACT>(a, rx, ry, rz, ri2, pj.w, s_r2[j]);
      else pair_plummer<T, EXACT>(a, rx, ry, rz, pj.w, eps2);
    };
    // j in ascending order (the reference's summation order).  The loop is
    // left rolled: unrolling it measured slower for these longer pair bodies
    // (dreal accumulation, FP64)
    for (int j = 0; j < cnt; ++j) body(j);
  }
  if (i < n_i) out[i] = {a[0].get(), a[1].get(), a[2].get(), a[3].get()};
}

// Plain form, Plummer softening (the presets' softening).  A particle's four
// per-pair dreal sums (ax, ay, az, potential) are serial chains from record
// to record, and the issue is in order, so both kernels below are
// software-pipelined: the terms of the next G records are computed in the
// same block as the chain over the current G, which gives the issue
// independent work between the chain's dependent steps.  The terms enter the
// chains in ascending j and each term is pair_plummer's expression, so the
// sums are pair_plummer's bits.
template <int C, bool EXACT>
__device__ __forceinline__ float plummer_term(const F4& pj, const F4& pi, float eps2) {
  const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
  const float d2 = rx * rx + ry * ry + rz * rz + eps2;
  const float inv = inv_sqrt<EXACT>(d2);
  if constexpr (C == 3) {
    return pj.w * inv;
  } else {
    const float s3 = pj.w * (inv * inv * inv);
    return (C == 0 ? rx : C == 1 ? ry : rz) * s3;
  }
}

template <int G, class Term, class Chain>
__device__ __forceinline__ void pipelined(int cnt, Term term, Chain chain) {
  using Y = decltype(term(0));
  int j = 0;
  if (cnt >= G) {
    Y y[G];
#pragma unroll
    for (int k = 0; k < G; ++k) y[k] = term(k);
    for (j = G; j + G <= cnt; j += G) {
      Y yn[G];
#pragma unroll
      for (int k = 0; k < G; ++k) yn[k] = term(j + k);
#pragma unroll
      for (int k = 0; k < G; ++k) chain(y[k]);
#pragma unroll
      for (int k = 0; k < G; ++k) y[k] = yn[k];
    }
#pragma unroll
    for (int k = 0; k < G; ++k) chain(y[k]);   // records j - G .. j - 1
  }
  for (; j < cnt; ++j) chain(term(j));
}

// One lane per particle, its four chains side by side, one record ahead
// (measured against the same loop without the pipelining: 8.66 -> 7.38 ms at
// N = 65 536, 71.5 -> 66.9 ms at 262 144, 1037 -> 1012 ms at 2^20; two
// records ahead gained less).
template <bool EXACT>
__global__ __launch_bounds__(kGravBlock) void k_gravity_lane(const F4* __restrict__ gpos, F4* __restrict__ out,
                                                             int n_i, int n_j, float eps2) {
  __shared__ F4 s_pos[kJTile];
  const int i = blockIdx.x * kGravBlock + threadIdx.x;
  const F4 pi = gpos[i < n_i ? i : n_i - 1];
  DR a0{0.f, 0.f}, a1{0.f, 0.f}, a2{0.f, 0.f}, a3{0.f, 0.f};
  for (int t0 = 0; t0 < n_j; t0 += kJTile) {
    const int cnt = stage_tile(s_pos, gpos, t0, n_j);
    pipelined<1>(
        cnt,
        [&](int j) {
          const F4 pj = s_pos[j];
          return F4{plummer_term<0, EXACT>(pj, pi, eps2), plummer_term<1, EXACT>(pj, pi, eps2),
                    plummer_term<2, EXACT>(pj, pi, eps2), plummer_term<3, EXACT>(pj, pi, eps2)};
        },
        [&](const F4& y) {
          dr_sub(a0, y.x);
          dr_sub(a1, y.y);
          dr_sub(a2, y.z);
          dr_sub(a3, y.w);
        });
  }
  if (i < n_i) out[i] = {a0.hi, a1.hi, a2.hi, a3.hi};
}

// Few particles: one wavefront per sum.  A block takes 64 particles and its
// wavefront c keeps sum c of each, repeating the pair's distance and rsq and
// running one two-sum per pair instead of four (~23 VALU + rsq per pair and
// lane against ~58 + rsq), on four times as many wavefronts: at N = 16 384
// the one-lane kernel has a wavefront for 256 of the GPU's 1024 SIMDs, this
// one fills all of them, 8 records ahead (16 384: 2.19 -> 0.89 ms).
constexpr int kSumParticles = kGravBlock / 4;   // particles per block of k_gravity_sums

template <int C, bool EXACT>
__device__ __forceinline__ void plummer_sum(DR& acc, const F4* __restrict__ s_pos, int cnt, const F4& pi,
                                            float eps2) {
  pipelined<8>(
      cnt, [&](int j) { return plummer_term<C, EXACT>(s_pos[j], pi, eps2); }, [&](float y) { dr_sub(acc, y); });
}

template <bool EXACT>
__global__ __launch_bounds__(kGravBlock) void k_gravity_sums(const F4* __restrict__ gpos, F4* __restrict__ out,
                                                             int n_i, int n_j, float eps2) {
  __shared__ F4 s_pos[kJTile];
  const int c = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);   // the wavefront's sum
  const int i = blockIdx.x * kSumParticles + (threadIdx.x & 63);
  const F4 pi = gpos[i < n_i ? i : n_i - 1];
  DR acc{0.f, 0.f};
  for (int t0 = 0; t0 < n_j; t0 += kJTile) {
    const int cnt = stage_tile(s_pos, gpos, t0, n_j);
    switch (c) {
      case 0: plummer_sum<0, EXACT>(acc, s_pos, cnt, pi, eps2); break;
      case 1: plummer_sum<1, EXACT>(acc, s_pos, cnt, pi, eps2); break;
      case 2: plummer_sum<2, EXACT>(acc, s_pos, cnt, pi, eps2); break;
      default: plummer_sum<3, EXACT>(acc, s_pos, cnt, pi, eps2); break;
    }
  }
  if (i < n_i) reinterpret_cast<float*>(out + i)[c] = acc.hi;
}

// Between the two: two wavefronts per 64 particles, one keeping ax and ay,
// the other az and the potential, each repeating the pair's distance and
// rsq once for its two two-sums, 4 records ahead.  Measured alone (ms per
// launch, one lane / four wavefronts / two wavefronts): 16 384 particles
// 1.85 / 0.89 / 1.23; 20 480: 2.31 / 1.56 / 1.53; 32 768: 3.71 / 2.54 / 2.46;
// 49 152: 5.56 / 5.16 / 4.42; 65 536: 7.43 / 8.9 / 5.9.  In the plain step,
// beside the SPH passes (whose walks take the higher issue priority), ms per
// step: 16 384: 2.05 / 1.33 / 1.45; 32 768: 4.10 / 3.84 / 3.38; 49 152:
// 5.72 / - / 5.26; 65 536: 7.62 / - / 6.72; 81 920: 10.47 / - / 13.2.
template <int C0, int C1, bool EXACT>
__device__ __forceinline__ void plummer_two(DR& a0, DR& a1, const F4* __restrict__ s_pos, int cnt, const F4& pi,
                                            float eps2) {
  pipelined<4>(
      cnt,
      [&](int j) {
        const F4 pj = s_pos[j];
        return float2{plummer_term<C0, EXACT>(pj, pi, eps2), plummer_term<C1, EXACT>(pj, pi, eps2)};
      },
      [&](const float2& y) {
        dr_sub(a0, y.x);
        dr_sub(a1, y.y);
      });
}

constexpr int kTwoParticles = kGravBlock / 2;   // particles per block of k_gravity_two

template <bool EXACT>
__global__ __launch_bounds__(kGravBlock) void k_gravity_two(const F4* __restrict__ gpos, F4* __restrict__ out,
                                                            int n_i, int n_j, float eps2) {
  __shared__ F4 s_pos[kJTile];
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int half = w & 1;   // 0: ax, ay; 1: az, potential
  const int i = blockIdx.x * kTwoParticles + (w >> 1) * 64 + (threadIdx.x & 63);
  const F4 pi = gpos[i < n_i ? i : n_i - 1];
  DR a0{0.f, 0.f}, a1{0.f, 0.f};
  for (int t0 = 0; t0 < n_j; t0 += kJTile) {
    const int cnt = stage_tile(s_pos, gpos, t0, n_j);
    if (half == 0) plummer_two<0, 1, EXACT>(a0, a1, s_pos, cnt, pi, eps2);
    else plummer_two<2, 3, EXACT>(a0, a1, s_pos, cnt, pi, eps2);
  }
  if (i < n_i) {
    float* o = reinterpret_cast<float*>(out + i) + 2 * half;
    o[0] = a0.hi;
    o[1] = a1.hi;
  }
}

// ---------------------------------------------------------------------------
// Multipole far field of the mixed Plummer kernel (optional, flags bit 4).
//
// Each 64-record sub-tile of the j-set (records [64 s, 64 s + 64), the
// sub-tiles the kernel below stages) gets its moments up to order 4 about its
// centre of mass, for the softened kernel phi(y) = (|y|^2 + eps^2)^(-1/2):
//   sum_j m_j phi(y - e_j) = sum_n (-1)^n / n! M^(n) . D^(n)(y),
//   M^(n) = sum_j m_j e_j^(x n),  D^(n) = grad^n phi,  e_j = x_j - COM.
// phi depends on u = |y|^2 / 2 only, so D^(n) is a sum over pairings of its
// indices: g_{n-m} times y^(n-2m) and m Kronecker deltas, with
// g_k = d^k phi / du^k = (-1)^k (2k-1)!! s^-(2k+1), s^2 = |y|^2 + eps^2.
// Contracted with the (not traceless: phi is not harmonic under softening)
// moments this needs M2, M3, M4, their traces tr M2, t3 = tr M3 and
// Q = tr M4, tr Q.  M^(1) = 0 about the centre of mass.  Truncation after
// order 4 leaves a relative error ~ (R / |y|)^5 of the sub-tile's force
// (R = max |e_j|): the far field is used for a (wavefront, sub-tile) only
// when every lane has |y| >= q R.  The default q = 8 (the engine's
// grav_far_ratio) bounds that error near 1e-6 of one sub-tile's force, below
// the FP32 pair arithmetic's: measured against FP64 at N = 2^20 the median
// per-particle error rises by 5 % over the pairwise kernel's, the maximum
// not at all; at q = 5 the maximum passes 1e-5.
constexpr int kMomWords = 52;     // floats per sub-tile record, see k_subtile_moments

__device__ __forceinline__ float wave_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// record of sub-tile s: [0..2] centre-of-mass offset from the origin record
// (the middle real record; its coordinates in [48..50]), [3] R^2, [4] M0
// (0: no real record, the sub-tile is skipped), [5] tr M2,
// [6] tr Q, [8..13] M2 (xx xy xz yy yz zz), [14..16] t3, [17..22] Q,
// [23..32] M3 (xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz), [33..47] M4
// (xxxx xxxy xxxz xxyy xxyz xxzz xyyy xyyz xyzz xzzz yyyy yyyz yyzz yzzz
// zzzz).  Masses as the kernel uses them (1 when every mass is equal);
// padding records (j >= n_real) are left out.  One wavefront per sub-tile.
__global__ __launch_bounds__(kGravBlock) void k_subtile_moments(const F4* __restrict__ hi, const F4* __restrict__ lo,
                                                                float* __restrict__ mom, int n_j, int n_real,
                                                                int n_sub, int G, int flags,
                                                                const float* __restrict__ mass_range) {
  const bool uniform = (flags & 2) == 0 && mass_range[0] == mass_range[1];
  const int s = blockIdx.x * (kGravBlock / 64) + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (s >= n_sub) return;   // whole wavefronts
  // a group of G <= 64 records; lanes past it add nothing (a group past
  // n_j, possible for G < 64, is empty: M0 = 0, skipped where it is read)
  const int j0 = s * G, nj = min(G, n_j - j0);
  // origin: the middle real record (a padding record sits at 1e30)
  const int nr = max(0, min(nj, n_real - j0));
  const F4 c = hi[min(j0 + nr / 2, n_j - 1)];
  const int j = j0 + lane;
  const bool valid = lane < nj && j < n_real;
  float dx = 0.f, dy = 0.f, dz = 0.f, m = 0.f;
  if (valid) {
    const F4 h = hi[j], l = lo[j];
    dx = (h.x - c.x) + l.x;
    dy = (h.y - c.y) + l.y;
    dz = (h.z - c.z) + l.z;
    m = uniform ? 1.f : h.w;
  }
  const float M0 = wave_sum(m);
  const float sx = wave_sum(m * dx), sy = wave_sum(m * dy), sz = wave_sum(m * dz);
  const float mx = M0 > 0.f ? sx / M0 : 0.f, my = M0 > 0.f ? sy / M0 : 0.f, mz = M0 > 0.f ? sz / M0 : 0.f;
  const float ex = valid ? dx - mx : 0.f, ey = valid ? dy - my : 0.f, ez = valid ? dz - mz : 0.f;
  float r2 = __builtin_fmaf(ex, ex, __builtin_fmaf(ey, ey, ez * ez));
  for (int o = 32; o > 0; o >>= 1) r2 = fmaxf(r2, __shfl_xor(r2, o));
  const float mex = m * ex, mey = m * ey, mez = m * ez;
  const float xx = mex * ex, xy = mex * ey, xz = mex * ez, yy = mey * ey, yz = mey * ez, zz = mez * ez;
  const float xxx = xx * ex, xxy = xx * ey, xxz = xx * ez, xyy = xy * ey, xyz = xy * ez, xzz = xz * ez,
              yyy = yy * ey, yyz = yy * ez, yzz = yz * ez, zzz = zz * ez;
  const float p4[15] = {xxx * ex, xxx * ey, xxx * ez, xxy * ey, xxy * ez, xxz * ez, xyy * ey, xyy * ez,
                        xyz * ez, xzz * ez, yyy * ey, yyy * ez, yyz * ez, yzz * ez, zzz * ez};
  const float p2[6] = {xx, xy, xz, yy, yz, zz};
  const float p3[10] = {xxx, xxy, xxz, xyy, xyz, xzz, yyy, yyz, yzz, zzz};
  float M2[6], M3[10], M4[15];
#pragma unroll
  for (int k = 0; k < 6; ++k) M2[k] = wave_sum(p2[k]);
#pragma unroll
  for (int k = 0; k < 10; ++k) M3[k] = wave_sum(p3[k]);
#pragma unroll
  for (int k = 0; k < 15; ++k) M4[k] = wave_sum(p4[k]);
  if (lane == 0) {
    float* r = mom + (size_t)s * kMomWords;
    const float Q[6] = {M4[0] + M4[3] + M4[5], M4[1] + M4[6] + M4[8], M4[2] + M4[7] + M4[9],
                        M4[3] + M4[10] + M4[12], M4[4] + M4[11] + M4[13], M4[5] + M4[12] + M4[14]};
    r[0] = mx;
    r[1] = my;
    r[2] = mz;
    r[3] = r2;
    r[4] = M0;
    r[5] = M2[0] + M2[3] + M2[5];
    r[6] = Q[0] + Q[3] + Q[5];
    r[7] = 0.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) r[8 + k] = M2[k];
    r[14] = M3[0] + M3[3] + M3[5];
    r[15] = M3[1] + M3[6] + M3[8];
    r[16] = M3[2] + M3[7] + M3[9];
#pragma unroll
    for (int k = 0; k < 6; ++k) r[17 + k] = Q[k];
#pragma unroll
    for (int k = 0; k < 10; ++k) r[23 + k] = M3[k];
#pragma unroll
    for (int k = 0; k < 15; ++k) r[33 + k] = M4[k];
    r[48] = c.x;
    r[49] = c.y;
    r[50] = c.z;
    r[51] = 0.f;
  }
}

// Cells of C = 256, 1024, 4096 and 16384 records (the hierarchy above the
// sub-tiles): one block per cell, the same record layout.  Sums in FP64, the
// whole cell at once: a cell holds up to 16384 records, whose FP32 sums
// would round at ~1e-5.
constexpr int kFarLevels = 4;
constexpr int kCellSize[kFarLevels] = {16384, 4096, 1024, 256};   // coarse to fine
// squared acceptance ratios: [L] for the cell level L, [kFarLevels] for
// the 64-record sub-tiles, [kFarLevels + 1] for their 32-record halves
struct FarRatios {
  float q2[kFarLevels + 2];
};

__device__ __forceinline__ double block_sum(double v, double* s_red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < kGravBlock / 64; ++w) t += s_red[w];   // fixed order, every thread the same
  return t;
}

__global__ __launch_bounds__(kGravBlock) void k_cell_moments(const F4* __restrict__ hi, const F4* __restrict__ lo,
                                                             float* __restrict__ mom, int n_j, int n_real, int C,
                                                             int flags, const float* __restrict__ mass_range) {
  __shared__ double s_red[kGravBlock / 64];
  const bool uniform = (flags & 2) == 0 && mass_range[0] == mass_range[1];
  const int cell = blockIdx.x;
  const int j0 = cell * C, nj = min(C, n_j - j0);
  // origin: the middle real record (a padding record sits at 1e30)
  const int nr = max(0, min(nj, n_real - j0));
  const F4 c = hi[j0 + nr / 2];
  // pass 1: mass and centre of mass
  double m0 = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
  for (int q = threadIdx.x; q < nj; q += kGravBlock) {
    const int j = j0 + q;
    if (j >= n_real) continue;
    const F4 h = hi[j], l = lo[j];
    const double m = uniform ? 1.0 : (double)h.w;
    m0 += m;
    sx += m * (double)((h.x - c.x) + l.x);
    sy += m * (double)((h.y - c.y) + l.y);
    sz += m * (double)((h.z - c.z) + l.z);
  }
  m0 = block_sum(m0, s_red);
  sx = block_sum(sx, s_red);
  sy = block_sum(sy, s_red);
  sz = block_sum(sz, s_red);
  const double mx = m0 > 0.0 ? sx / m0 : 0.0, my = m0 > 0.0 ? sy / m0 : 0.0, mz = m0 > 0.0 ? sz / m0 : 0.0;
  // pass 2: radius and central moments
  double acc[31];
  for (int k = 0; k < 31; ++k) acc[k] = 0.0;
  float r2 = 0.f;
  for (int q = threadIdx.x; q < nj; q += kGravBlock) {
    const int j = j0 + q;
    if (j >= n_real) continue;
    const F4 h = hi[j], l = lo[j];
    const double m = uniform ? 1.0 : (double)h.w;
    const double ex = (double)((h.x - c.x) + l.x) - mx, ey = (double)((h.y - c.y) + l.y) - my,
                 ez = (double)((h.z - c.z) + l.z) - mz;
    r2 = fmaxf(r2, (float)(ex * ex + ey * ey + ez * ez));
    const double xx = m * ex * ex, xy = m * ex * ey, xz = m * ex * ez, yy = m * ey * ey, yz = m * ey * ez,
                 zz = m * ez * ez;
    const double xxx = xx * ex, xxy = xx * ey, xxz = xx * ez, xyy = xy * ey, xyz = xy * ez, xzz = xz * ez,
                 yyy = yy * ey, yyz = yy * ez, yzz = yz * ez, zzz = zz * ez;
    const double t[31] = {xx, xy, xz, yy, yz, zz, xxx, xxy, xxz, xyy, xyz, xzz, yyy, yyz, yzz, zzz,
                          xxx * ex, xxx * ey, xxx * ez, xxy * ey, xxy * ez, xxz * ez, xyy * ey, xyy * ez,
                          xyz * ez, xzz * ez, yyy * ey, yyy * ez, yyz * ez, yzz * ez, zzz * ez};
    for (int k = 0; k < 31; ++k) acc[k] += t[k];
  }
  for (int k = 0; k < 31; ++k) acc[k] = block_sum(acc[k], s_red);
  for (int o = 32; o > 0; o >>= 1) r2 = fmaxf(r2, __shfl_xor(r2, o));
  __syncthreads();
  float* s_r2 = reinterpret_cast<float*>(s_red);
  if ((threadIdx.x & 63) == 0) s_r2[threadIdx.x >> 6] = r2;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kGravBlock / 64; ++w) r2 = fmaxf(r2, s_r2[w]);
    const double* M2 = acc;
    const double* M3 = acc + 6;
    const double* M4 = acc + 16;
    const double Q[6] = {M4[0] + M4[3] + M4[5], M4[1] + M4[6] + M4[8], M4[2] + M4[7] + M4[9],
                         M4[3] + M4[10] + M4[12], M4[4] + M4[11] + M4[13], M4[5] + M4[12] + M4[14]};
    float* r = mom + (size_t)cell * kMomWords;
    r[0] = (float)mx;
    r[1] = (float)my;
    r[2] = (float)mz;
    r[3] = r2;
    r[4] = (float)m0;
    r[5] = (float)(M2[0] + M2[3] + M2[5]);
    r[6] = (float)(Q[0] + Q[3] + Q[5]);
    r[7] = 0.f;
    for (int k = 0; k < 6; ++k) r[8 + k] = (float)M2[k];
    r[14] = (float)(M3[0] + M3[3] + M3[5]);
    r[15] = (float)(M3[1] + M3[6] + M3[8]);
    r[16] = (float)(M3[2] + M3[7] + M3[9]);
    for (int k = 0; k < 6; ++k) r[17 + k] = (float)Q[k];
    for (int k = 0; k < 10; ++k) r[23 + k] = (float)M3[k];
    for (int k = 0; k < 15; ++k) r[33 + k] = (float)M4[k];
    r[48] = c.x;
    r[49] = c.y;
    r[50] = c.z;
    r[51] = 0.f;
  }
}

// acceleration G = sum_j m_j (x_j - x_i) s^-3 and potential -sum_j m_j / s of
// one sub-tile from its record M, at y = x_i - COM (order-4 expansion)
__device__ __forceinline__ void far_field(const float* __restrict__ M, float yx, float yy, float yz, float eps2,
                                          float& gx, float& gy, float& gz, float& pot) {
  const float s2 = __builtin_fmaf(yx, yx, __builtin_fmaf(yy, yy, __builtin_fmaf(yz, yz, eps2)));
  const float inv = __builtin_amdgcn_rsqf(s2), i2 = inv * inv;
  const float g0 = inv, g1 = -inv * i2, g2 = -3.f * g1 * i2, g3 = -5.f * g2 * i2, g4 = -7.f * g3 * i2,
              g5 = -9.f * g4 * i2;
  const float M0 = M[4], trM2 = M[5], trQ = M[6];
  const float* m2 = M + 8;
  const float* t3 = M + 14;
  const float* Q = M + 17;
  const float* m3 = M + 23;
  const float* m4 = M + 33;
  const float M2x = __builtin_fmaf(m2[0], yx, __builtin_fmaf(m2[1], yy, m2[2] * yz));
  const float M2y = __builtin_fmaf(m2[1], yx, __builtin_fmaf(m2[3], yy, m2[4] * yz));
  const float M2z = __builtin_fmaf(m2[2], yx, __builtin_fmaf(m2[4], yy, m2[5] * yz));
  const float yM2y = __builtin_fmaf(M2x, yx, __builtin_fmaf(M2y, yy, M2z * yz));
  const float Qx = __builtin_fmaf(Q[0], yx, __builtin_fmaf(Q[1], yy, Q[2] * yz));
  const float Qy = __builtin_fmaf(Q[1], yx, __builtin_fmaf(Q[3], yy, Q[4] * yz));
  const float Qz = __builtin_fmaf(Q[2], yx, __builtin_fmaf(Q[4], yy, Q[5] * yz));
  const float yQy = __builtin_fmaf(Qx, yx, __builtin_fmaf(Qy, yy, Qz * yz));
  const float t3y = __builtin_fmaf(t3[0], yx, __builtin_fmaf(t3[1], yy, t3[2] * yz));
  // quadratic monomials, off-diagonal ones doubled: (M3 . y y)_k
  const float qxx = yx * yx, qxy = 2.f * yx * yy, qxz = 2.f * yx * yz, qyy = yy * yy, qyz = 2.f * yy * yz,
              qzz = yz * yz;
  auto dot6 = [&](float a, float b, float c, float d, float e, float f) {
    return __builtin_fmaf(qxx, a, __builtin_fmaf(qxy, b, __builtin_fmaf(qxz, c,
           __builtin_fmaf(qyy, d, __builtin_fmaf(qyz, e, qzz * f)))));
  };
  const float M3x = dot6(m3[0], m3[1], m3[2], m3[3], m3[4], m3[5]);
  const float M3y = dot6(m3[1], m3[3], m3[4], m3[6], m3[7], m3[8]);
  const float M3z = dot6(m3[2], m3[4], m3[5], m3[7], m3[8], m3[9]);
  const float T3 = __builtin_fmaf(M3x, yx, __builtin_fmaf(M3y, yy, M3z * yz));
  // cubic monomials with their multinomial weights: (M4 . y y y)_k
  const float cxxx = qxx * yx, cxxy = 3.f * qxx * yy, cxxz = 3.f * qxx * yz, cxyy = 3.f * yx * qyy,
              cxyz = 3.f * qxy * yz, cxzz = 3.f * yx * qzz, cyyy = qyy * yy, cyyz = 3.f * qyy * yz,
              cyzz = 3.f * yy * qzz, czzz = qzz * yz;
  auto dot10 = [&](int a, int b, int c, int d, int e, int f, int g, int h, int k, int l) {
    return __builtin_fmaf(cxxx, m4[a], __builtin_fmaf(cxxy, m4[b], __builtin_fmaf(cxxz, m4[c],
           __builtin_fmaf(cxyy, m4[d], __builtin_fmaf(cxyz, m4[e], __builtin_fmaf(cxzz, m4[f],
           __builtin_fmaf(cyyy, m4[g], __builtin_fmaf(cyyz, m4[h], __builtin_fmaf(cyzz, m4[k], czzz * m4[l])))))))));
  };
  const float M4x = dot10(0, 1, 2, 3, 4, 5, 6, 7, 8, 9);
  const float M4y = dot10(1, 3, 4, 6, 7, 8, 10, 11, 12, 13);
  const float M4z = dot10(2, 4, 5, 7, 8, 9, 11, 12, 13, 14);
  const float T4 = __builtin_fmaf(M4x, yx, __builtin_fmaf(M4y, yy, M4z * yz));
  const float S = M0 * g0 + 0.5f * (g2 * yM2y + g1 * trM2) - (g3 * T3 + 3.f * g2 * t3y) * (1.f / 6.f) +
                  (g4 * T4 + 6.f * g3 * yQy + 3.f * g2 * trQ) * (1.f / 24.f);
  const float A = M0 * g1 + 0.5f * (g3 * yM2y + g2 * trM2) - (g4 * T3 + 3.f * g3 * t3y) * (1.f / 6.f) +
                  (g5 * T4 + 6.f * g4 * yQy + 3.f * g3 * trQ) * (1.f / 24.f);
  const float b2 = g2, b3 = -0.5f * g3, bt = -0.5f * g2, b4 = g4 * (1.f / 6.f), bq = 0.5f * g3;
  gx = A * yx + b2 * M2x + b3 * M3x + bt * t3[0] + b4 * M4x + bq * Qx;
  gy = A * yy + b2 * M2y + b3 * M3y + bt * t3[1] + b4 * M4y + bq * Qy;
  gz = A * yz + b2 * M2z + b3 * M3z + bt * t3[2] + b4 * M4z + bq * Qz;
  pot = -S;
}

// mixed form, Plummer: split positions, FMA pair arithmetic, blocked
// compensated sums.  hi = (x,y,z,m) rounded, lo = (x-hi, y-hi, z-hi, 0).
//
// Two ways to form r = x_i - x_j, chosen per (wavefront, sub-tile of 64 j):
//  split:  r = (hi_i - hi_j) + (lo_i - lo_j), 9 operations per pair, relative
//          error ~2^-24 for every pair;
//  tile:   the sub-tile is expressed relative to its middle record c,
//          d_j = (hi_j - c) + lo_j (staged in LDS) and d_i = (hi_i - c) + lo_i
//          (once per sub-tile per lane), r = d_i - d_j: 3 operations per pair,
//          absolute error <= 2^-24 (|d_i| + |d_j|) + rounding of r.
// The tile form is used only when every lane of the wavefront has
// |d_i| >= 3 R, R = max_j |d_j| over the sub-tile: then |r| >= |d_i| - R >= 2 R
// and (|d_i| + |d_j|) / |r| <= (|d_i| + R) / (|d_i| - R) <= 2, so r keeps a
// relative error of a few 2^-24 -- the accuracy of the split form.
// Wavefronts near a sub-tile (the pairs whose separation the split
// coordinates exist for) use the split form.  With particles in
// Peano-Hilbert order 64 consecutive records are compact and most
// (wavefront, sub-tile) combinations are far apart; with an unordered input
// every sub-tile is as large as the system and every wavefront uses the split
// form.  split_only bit 0 forces the split form everywhere; bit 1 disables
// the unit-mass body; bit 2 the expanded unit-mass tile body (measurement
// switches).
// After r: 3 FMAs for r^2 + eps^2, one rsq, 3 multiplies, 3 FMAs into the
// partial acceleration, 1 subtraction for the potential.  With every mass
// equal, tile-form sub-tiles skip r altogether (the expanded body below).
//
// Grid: `parts` j-parts per i-block, block b = i_block * parts + part.
// Workgroups are dispatched round-robin over the 8 XCDs, so with 8 parts
// part p runs on XCD p and each XCD streams only its 1/8 of the j-set
// (4 MiB of hi+lo records at N = 2^20), which its L2 holds; without the
// split every XCD streams the whole j-set and re-fetches it from HBM.  Each
// block writes its compensated (hi, lo) partial sums; k_gravity_combine adds
// the parts in order.  (gravity_mixed_parts picks fewer parts where 8 would
// fill most of the GPU's block slots in one round.)
// The far-field instantiation holds the 48 moment words of an expansion in
// VGPRs next to the pairwise state: left alone the compiler takes 127
// registers (4 wavefronts per SIMD); capped at 5 wavefronts it fits 96
// registers without spilling (at 6 it spills 15 registers and runs 4-8 %
// slower).  The pairwise instantiation is left unbounded (63 registers, 8
// wavefronts).
template <int JT, bool FAR>
__global__ __launch_bounds__(kGravBlock) __attribute__((amdgpu_waves_per_eu(FAR ? 5 : 1)))
void k_gravity_mixed_plummer(const F4* __restrict__ hi,
                                                                       const F4* __restrict__ lo,
                                                                       F4* __restrict__ part_hi,
                                                                       F4* __restrict__ part_lo, int n_i,
                                                                       int n_j, int j_chunk, float eps2,
                                                                       int split_only,
                                                                       const float* __restrict__ mass_range,
                                                                       const float* __restrict__ mom,
                                                                       FarRatios far_q2,
                                                                       const float* __restrict__ cells,
                                                                       const float* __restrict__ halves,
                                                                       int n_real, int parts) {
  // every mass equal (mass_range = {min, max} over the real particles): the
  // pair body drops its mass factor and k_gravity_combine scales by m once;
  // padding records sit at 1e30, where r^2 overflows and rsq returns 0, so
  // they add exactly nothing without their zero mass
  const bool uniform = (split_only & 2) == 0 && mass_range[0] == mass_range[1];
  const bool expand = uniform && (split_only & 4) == 0;
  // a separate instantiation: the far-field code raises the VGPR count, which
  // would lower the occupancy of the pairwise kernel
  const bool far_on = FAR && (split_only & 1) == 0;
  // split_only bit 6: 32-record halves of a sub-tile not taken whole
  const bool half_on = far_on && (split_only & 64) != 0 && halves != nullptr;
  split_only &= 1;
  constexpr int kSub = JT / kTile;   // sub-tiles of kTile records, each with its own origin
  __shared__ F4 s_hi[JT], s_lo[JT], s_rel[JT];
  __shared__ F4 s_c[kSub];
  __shared__ float s_r2[kSub];
  __shared__ __align__(16) float s_mom[kSub][FAR ? kMomWords : 1];
  static_assert(kTile == 64 && JT % kGravBlock == 0, "each wavefront stages whole sub-tiles");
  const int part = blockIdx.x % parts;
  const int i = (blockIdx.x / parts) * kGravBlock + threadIdx.x;
  const int ii = i < n_i ? i : n_i - 1;
  const int lane = threadIdx.x & 63;
  const F4 pi = hi[ii], li = lo[ii];
  // padding lanes (records at 1e30) take no far-field term: its products
  // would overflow there; they pass every acceptance test
  const bool real_i = !FAR || i < n_real;
  GAcc<float, 1> a[4];
  auto add_far = [&](const float* M, float yx, float yy, float yz) {
    float gx, gy, gz, pt;
    far_field(M, yx, yy, yz, eps2, gx, gy, gz, pt);
    if (real_i) {
      a[0].part += gx;
      a[1].part += gy;
      a[2].part += gz;
      a[3].part += pt;
    }
  };
  const int j_begin = min(part * j_chunk, n_j), j_end = min(j_begin + j_chunk, n_j);
  int skip_end = j_begin;   // this wavefront is done with the records below
  for (int t0 = j_begin; t0 < j_end; t0 += JT) {
    const int cnt = min(JT, j_end - t0);
    bool need = true;
    if constexpr (FAR) {
      // the coarsest cell starting here, inside this part (or the array's
      // last, partial cell), that every lane sees from beyond q cell radii:
      // one expansion for all of it
      if (far_on) {
        need = t0 >= skip_end;
        const float* Mc = nullptr;
        int Cc = 0;
        size_t off = 0;
        for (int L = 0; L < kFarLevels; ++L) {
          const int C = kCellSize[L];
          if (need && !Mc && t0 % C == 0 && (t0 + C <= j_end || j_end == n_j)) {
            const float* M = cells + off + (size_t)(t0 / C) * kMomWords;
            const float yx = ((pi.x - M[48]) + li.x) - M[0], yy = ((pi.y - M[49]) + li.y) - M[1],
                        yz = ((pi.z - M[50]) + li.z) - M[2];
            const float y2 = __builtin_fmaf(yx, yx, __builtin_fmaf(yy, yy, yz * yz));
            if (M[4] == 0.f || __all(!real_i || y2 >= far_q2.q2[L] * M[3])) {
              Mc = M;
              Cc = C;
            }
          }
          off += (size_t)((n_j + C - 1) / C) * kMomWords;
        }
        if (Mc) {   // one expansion, whichever level accepted (none for an empty cell)
          if (Mc[4] != 0.f)
            add_far(Mc, ((pi.x - Mc[48]) + li.x) - Mc[0], ((pi.y - Mc[49]) + li.y) - Mc[1],
                    ((pi.z - Mc[50]) + li.z) - Mc[2]);
          for (int k = 0; k < 4; ++k) a[k].tile_end();
          skip_end = min(t0 + Cc, j_end);
          need = false;
        }
      }
      // the barrier that orders the previous tile's LDS reads before this
      // tile's staging also tells whether any wavefront needs the records
      if (!__syncthreads_or(need)) continue;
    } else {
      __syncthreads();
    }
    // stage: this wavefront's sub-tiles, each relative to its middle real
    // record, and each sub-tile's squared radius about that record.  A
    // padding record (at 1e30) as origin would collapse every real offset
    // to -1e30 and make its squared radius infinite; the tile test would
    // then pass for every wavefront and the real records would add nothing.
    // A sub-tile of padding only keeps its middle record.
    for (int q = threadIdx.x; q < JT; q += kGravBlock) {
      const int sub = q / kTile, sub0 = sub * kTile;
      float r2 = 0.f;
      if (sub0 < cnt) {
        const int nj = min(kTile, cnt - sub0), nr = max(0, min(nj, n_real - (t0 + sub0)));
        const F4 c = hi[t0 + sub0 + (nr > 0 ? nr : nj) / 2];
        if (q < cnt) {
          const F4 h = hi[t0 + q], l = lo[t0 + q];
          F4 d{(h.x - c.x) + l.x, (h.y - c.y) + l.y, (h.z - c.z) + l.z, h.w};
          r2 = __builtin_fmaf(d.x, d.x, __builtin_fmaf(d.y, d.y, d.z * d.z));
          if (expand) d.w = r2;   // the expanded body needs |d_j|^2, not m
          s_hi[q] = h;
          s_lo[q] = l;
          s_rel[q] = d;
        }
        for (int o = 32; o > 0; o >>= 1) r2 = fmaxf(r2, __shfl_xor(r2, o));
        if (lane == 0) {
          s_c[sub] = c;
          s_r2[sub] = r2;
        }
      }
    }
    if constexpr (FAR) {   // the tile's sub-tile records (t0 is a multiple of kTile)
      for (int q = threadIdx.x; q < kSub * kMomWords; q += kGravBlock) {
        const int sub = q / kMomWords, w = q % kMomWords;
        if (sub * kTile < cnt) s_mom[sub][w] = mom[(size_t)(t0 / kTile + sub) * kMomWords + w];
      }
    }
    __syncthreads();
    if (FAR && !need) continue;   // a cell covered this tile for this wavefront
    auto pair = [&](float rx, float ry, float rz, float mj) {
      const float d2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, eps2)));
      const float inv = __builtin_amdgcn_rsqf(d2);
      const float mi = mj * inv;
      const float s3 = mi * (inv * inv);
      a[0].fsub(rx, s3);
      a[1].fsub(ry, s3);
      a[2].fsub(rz, s3);
      a[3].sub(mi);
    };
    auto pair_unit = [&](float rx, float ry, float rz) {   // unit mass: 12 VALU + rsq
      const float d2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, eps2)));
      const float inv = __builtin_amdgcn_rsqf(d2);
      const float s3 = inv * (inv * inv);
      a[0].fsub(rx, s3);
      a[1].fsub(ry, s3);
      a[2].fsub(rz, s3);
      a[3].sub(inv);
    };
    for (int sub = 0; sub * kTile < cnt; ++sub) {
      const int j0 = sub * kTile, nj = min(kTile, cnt - j0);
      int jb = j0, je = j0 + nj;   // the records taken pair by pair
      const F4 c = s_c[sub];
      const float dx = (pi.x - c.x) + li.x, dy = (pi.y - c.y) + li.y, dz = (pi.z - c.z) + li.z;
      // tile form only if every lane of the wavefront has |d_i| >= 3 R_sub
      // (uniform per wavefront: no divergence); the squared test carries a
      // margin of 1.1 for the rounding of the squares
      if (FAR && far_on) {
        // multipole far field when every lane is q sub-tile radii from its
        // centre of mass (far_q2 = q^2; uniform per wavefront)
        const float* M = &s_mom[sub][0];
        const float yx = ((pi.x - M[48]) + li.x) - M[0], yy = ((pi.y - M[49]) + li.y) - M[1],
                    yz = ((pi.z - M[50]) + li.z) - M[2];
        const float y2 = __builtin_fmaf(yx, yx, __builtin_fmaf(yy, yy, yz * yz));
        if (M[4] == 0.f || __all(!real_i || y2 >= far_q2.q2[kFarLevels] * M[3])) {
          if (M[4] != 0.f) add_far(M, yx, yy, yz);
          if ((sub & 1) || j0 + kTile >= cnt)
            for (int k = 0; k < 4; ++k) a[k].tile_end();
          continue;
        }
        if (half_on && nj == kTile) {
          // each 32-record half the wavefront sees from beyond q of its radii
          // takes its expansion; the pairs run over the rest
          bool took[2];
          for (int h = 0; h < 2; ++h) {
            const float* H = halves + ((size_t)(t0 / kTile + sub) * 2 + h) * kMomWords;
            const float hx = ((pi.x - H[48]) + li.x) - H[0], hy = ((pi.y - H[49]) + li.y) - H[1],
                        hz = ((pi.z - H[50]) + li.z) - H[2];
            const float h2 = __builtin_fmaf(hx, hx, __builtin_fmaf(hy, hy, hz * hz));
            took[h] = H[4] == 0.f || __all(!real_i || h2 >= far_q2.q2[kFarLevels + 1] * H[3]);
            if (took[h] && H[4] != 0.f) add_far(H, hx, hy, hz);
          }
          if (took[0] && took[1]) {
            if ((sub & 1) || j0 + kTile >= cnt)
              for (int k = 0; k < 4; ++k) a[k].tile_end();
            continue;
          }
          if (took[0]) jb = j0 + kTile / 2;
          if (took[1]) je = j0 + kTile / 2;
        }
      }
      const bool far = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)) >= 9.9f * s_r2[sub];
      const bool tile_form = !split_only && __all(far);
      auto pair_tile = [&](int j) {
        const F4 pj = s_rel[j];
        pair(dx - pj.x, dy - pj.y, dz - pj.z, pj.w);
      };
      auto pair_split = [&](int j) {
        const F4 pj = s_hi[j], lj = s_lo[j];
        pair((pi.x - pj.x) + (li.x - lj.x), (pi.y - pj.y) + (li.y - lj.y), (pi.z - pj.z) + (li.z - lj.z), pj.w);
      };
      auto unit_tile = [&](int j) {
        const F4 pj = s_rel[j];
        pair_unit(dx - pj.x, dy - pj.y, dz - pj.z);
      };
      auto unit_split = [&](int j) {
        const F4 pj = s_hi[j], lj = s_lo[j];
        pair_unit((pi.x - pj.x) + (li.x - lj.x), (pi.y - pj.y) + (li.y - lj.y), (pi.z - pj.z) + (li.z - lj.z));
      };
      // full sub-tiles (all of them when n_j is a multiple of kTile) have a
      // fixed trip count, unrolled so LDS reads are issued ahead of their use;
      // the branches are uniform per wavefront
      // the pairs of `len` consecutive staged records from `base`: a whole
      // sub-tile (len 64) or, in the far-field kernel, the half the
      // wavefront did not take by its expansion (len 32); fixed trip counts,
      // unrolled so LDS reads are issued ahead of their use
      auto run = [&](int base, auto len_c) {
        constexpr int len = decltype(len_c)::value;
        if (uniform) {
          if (tile_form && expand) {
            // expanded equal-mass body: r^2 + eps^2 = (|d_i|^2 + eps^2 + |d_j|^2)
            // - 2 d_i.d_j and sum_j s3 (d_j - d_i) = M - d_i S over the
            // records, 11 VALU + rsq per pair.  With |d_j| <= |d_i| / 3 the
            // expansion loses at most a factor (4/2)^2 of relative accuracy
            // in r^2 to cancellation, and M - d_i S none (|M| <= |d_i| S / 3).
            const float ci = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
            const float bx = -2.f * dx, by = -2.f * dy, bz = -2.f * dz;
            float sa = 0.f, mx = 0.f, my = 0.f, mz = 0.f;
#pragma unroll 16
            for (int j = 0; j < len; ++j) {
              const F4 pj = s_rel[base + j];
              const float d2 =
                  __builtin_fmaf(pj.x, bx, __builtin_fmaf(pj.y, by, __builtin_fmaf(pj.z, bz, ci + pj.w)));
              const float inv = __builtin_amdgcn_rsqf(d2);
              const float s3 = inv * (inv * inv);
              sa += s3;
              mx = __builtin_fmaf(pj.x, s3, mx);
              my = __builtin_fmaf(pj.y, s3, my);
              mz = __builtin_fmaf(pj.z, s3, mz);
              a[3].sub(inv);
            }
            a[0].part += __builtin_fmaf(-dx, sa, mx);
            a[1].part += __builtin_fmaf(-dy, sa, my);
            a[2].part += __builtin_fmaf(-dz, sa, mz);
          } else if (tile_form) {
#pragma unroll 16
            for (int j = 0; j < len; ++j) unit_tile(base + j);
          } else {
#pragma unroll 16
            for (int j = 0; j < len; ++j) unit_split(base + j);
          }
        } else if (tile_form) {
#pragma unroll 16
          for (int j = 0; j < len; ++j) pair_tile(base + j);
        } else {
#pragma unroll 16
          for (int j = 0; j < len; ++j) pair_split(base + j);
        }
      };
      // full sub-tiles (all of them when n_j is a multiple of kTile); the
      // branches are uniform per wavefront
      if (nj == kTile) {
        bool whole = true;
        if constexpr (FAR) {
          if (je - jb != kTile) {
            run(jb, std::integral_constant<int, kTile / 2>{});
            whole = false;
          }
        }
        if (whole) run(j0, std::integral_constant<int, kTile>{});
      } else {
        for (int j = j0; j < j0 + nj; ++j) {
          if (uniform) {
            if (tile_form) unit_tile(j);
            else unit_split(j);
          } else if (tile_form) {
            pair_tile(j);
          } else {
            pair_split(j);
          }
        }
      }
      // fold the FP32 partials every two sub-tiles (128 pairs) and at the end
      if ((sub & 1) || j0 + kTile >= cnt)
        for (int k = 0; k < 4; ++k) a[k].tile_end();
    }
  }
  if (i < n_i) {
    const size_t o = (size_t)part * n_i + i;
    part_hi[o] = {a[0].v.hi, a[1].v.hi, a[2].v.hi, a[3].v.hi};
    part_lo[o] = {a[0].v.lo, a[1].v.lo, a[2].v.lo, a[3].v.lo};
  }
}

// sum of the `parts` compensated partials of each i, in part order, times
// the common mass when every mass is equal
__global__ __launch_bounds__(kBlock) void k_gravity_combine(const F4* __restrict__ part_hi,
                                                            const F4* __restrict__ part_lo, F4* __restrict__ out,
                                                            int n_i, const float* __restrict__ mass_range,
                                                            int flags, int parts) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n_i) return;
  const float m = (flags & 2) == 0 && mass_range[0] == mass_range[1] ? mass_range[0] : 1.f;
  DR a[4] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
  for (int p = 0; p < parts; ++p) {
    const F4 h = part_hi[(size_t)p * n_i + i], l = part_lo[(size_t)p * n_i + i];
    dr_add(a[0], h.x);
    dr_add(a[0], l.x);
    dr_add(a[1], h.y);
    dr_add(a[1], l.y);
    dr_add(a[2], h.z);
    dr_add(a[2], l.z);
    dr_add(a[3], h.w);
    dr_add(a[3], l.w);
  }
  out[i] = {m * (a[0].hi + a[0].lo), m * (a[1].hi + a[1].lo), m * (a[2].hi + a[2].lo), m * (a[3].hi + a[3].lo)};
}

// {min, max} of the masses of the real particles: per-block partials, then
// one block (fixed tree order, no atomics)
__global__ __launch_bounds__(kBlock) void k_mass_range_partials(const F4* __restrict__ hi, int n,
                                                                float* __restrict__ partial) {
  __shared__ float smin[kBlock], smax[kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const float m = i < n ? hi[i].w : hi[0].w;
  smin[threadIdx.x] = m;
  smax[threadIdx.x] = m;
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + k]);
      smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = smin[0];
    partial[2 * blockIdx.x + 1] = smax[0];
  }
}
__global__ __launch_bounds__(kBlock) void k_mass_range_final(const float* __restrict__ partial, int nb,
                                                             float* __restrict__ out) {
  __shared__ float smin[kBlock], smax[kBlock];
  float lo = partial[0], hi = partial[1];
  for (int b = threadIdx.x; b < nb; b += kBlock) {
    lo = fminf(lo, partial[2 * b]);
    hi = fmaxf(hi, partial[2 * b + 1]);
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + k]);
      smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = smin[0];
    out[1] = smax[0];
  }
}

// mixed form, spline softening: same accumulation and split positions
__global__ __launch_bounds__(kGravBlock) void k_gravity_mixed_spline(const F4* __restrict__ hi,
                                                                      const F4* __restrict__ lo,
                                                                      const F4* __restrict__ gvel,
                                                                      F4* __restrict__ out, int n_i, int n_j) {
  __shared__ F4 s_hi[kJTile], s_lo[kJTile];
  __shared__ float s_r2[kJTile];
  const int i = blockIdx.x * kGravBlock + threadIdx.x;
  const int ii = i < n_i ? i : n_i - 1;
  const F4 pi = hi[ii], li = lo[ii];
  const float ri2 = gvel[ii].w;
  GAcc<float, 1> a[4];
  for (int t0 = 0; t0 < n_j; t0 += kJTile) {
    const int cnt = stage_tile(s_hi, hi, t0, n_j);
    for (int q = threadIdx.x; q < cnt; q += kGravBlock) {
      s_lo[q] = lo[t0 + q];
      s_r2[q] = gvel[t0 + q].w;
    }
    __syncthreads();
    for (int j0 = 0; j0 < cnt; j0 += kTile) {
      const int j1 = min(j0 + kTile, cnt);
      for (int j = j0; j < j1; ++j) {
        const F4 pj = s_hi[j], lj = s_lo[j];
        const float rx = (pi.x - pj.x) + (li.x - lj.x);
        const float ry = (pi.y - pj.y) + (li.y - lj.y);
        const float rz = (pi.z - pj.z) + (li.z - lj.z);
        pair_spline<float, false>(a, rx, ry, rz, ri2, pj.w, s_r2[j]);
      }
      for (int k = 0; k < 4; ++k) a[k].tile_end();
    }
  }
  if (i < n_i) out[i] = {a[0].get(), a[1].get(), a[2].get(), a[3].get()};
}

// ---------------------------------------------------------------------------
// FP64 form, Plummer softening.  The pair body in fused multiply-adds (the
// FP64 form has no bit contract with the reference's FP32 arithmetic):
// 3 sub, 3 FMA for s^2 = r^2 + eps^2, v_rsq_f64 and two Newton steps
// (y <- y (1.5 - (s^2/2) y^2), 3 operations each after s^2/2), 3 for
// m/s^3, 3 FMA into the acceleration and one into the potential.  With few
// i-blocks (N below 2^19) the j-range is split into parts, one block per
// (i-block, part), so that the GPU holds enough wavefronts to hide the FP64
// latencies; the parts' sums are added in part order (deterministic).
constexpr int kG64Tile = kGravBlock;
__global__ __launch_bounds__(kGravBlock) void k_gravity64_plummer(const D4* __restrict__ gpos, D4* __restrict__ out,
                                                                   int n_i, int n_j, int parts, int chunk,
                                                                   double eps2) {
  __shared__ D4 s_pos[kG64Tile];
  const int part = (int)blockIdx.x % parts, ib = (int)blockIdx.x / parts;
  const int i = ib * kGravBlock + threadIdx.x;
  const D4 pi = gpos[i < n_i ? i : n_i - 1];
  const int j_begin = part * chunk, j_end = min(n_j, j_begin + chunk);
  double ax = 0.0, ay = 0.0, az = 0.0, ph = 0.0;
  for (int t0 = j_begin; t0 < j_end; t0 += kG64Tile) {
    const int cnt = min(kG64Tile, j_end - t0);
    __syncthreads();
    if ((int)threadIdx.x < cnt) s_pos[threadIdx.x] = gpos[t0 + threadIdx.x];
    __syncthreads();
    for (int j = 0; j < cnt; ++j) {
      const D4 pj = s_pos[j];
      const double rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
      const double d2 = __builtin_fma(rx, rx, __builtin_fma(ry, ry, __builtin_fma(rz, rz, eps2)));
      const double h = 0.5 * d2;
      double y = __builtin_amdgcn_rsq(d2);
      y = y * __builtin_fma(-h, y * y, 1.5);
      y = y * __builtin_fma(-h, y * y, 1.5);
      const double my = pj.w * y;
      const double s3 = my * (y * y);
      ax = __builtin_fma(-rx, s3, ax);
      ay = __builtin_fma(-ry, s3, ay);
      az = __builtin_fma(-rz, s3, az);
      ph -= my;
    }
  }
  if (i < n_i) out[(size_t)part * n_i + i] = {ax, ay, az, ph};
}

// the parts' sums, in part order
__global__ __launch_bounds__(kBlock) void k_gravity64_combine(const D4* __restrict__ part, D4* __restrict__ out,
                                                               int n_i, int parts) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n_i) return;
  D4 a = part[i];
  for (int p = 1; p < parts; ++p) {
    const D4 b = part[(size_t)p * n_i + i];
    a = {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w};
  }
  out[i] = a;
}

int gravity64_parts(int n_i) {
  const int blocks = ceil_div(n_i, kGravBlock);
  return std::max(1, std::min(8, ceil_div(2048, blocks)));
}

// The forward pass runs beside the FP64 SPH passes.  With no more i-blocks
// than CUs it takes parts for two blocks per CU, leaving the passes the rest
// of each CU; otherwise the reverse pass's split.  Measured, FP64 step /
// gravity alone: 65 536 particles, 8 parts 5.51 / 3.54 ms, 3 parts 5.04 /
// 3.67, 2 parts 4.71 / 3.81; 262 144 particles, 2 parts 50.6 / 50.9 ms,
// 1 part 70.2 / 54.1 (there the passes beside a half-empty grid slow it).
int gravity64_forward_parts(int n_i) {
  int dev = 0, cus = 0;
  CS_HIP(hipGetDevice(&dev));
  CS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = ceil_div(n_i, kGravBlock);
  if (blocks > cus) return gravity64_parts(n_i);
  return std::max(1, std::min(8, ceil_div(2 * cus, blocks)));
}

void launch_gravity64_plummer(const D4* gpos, D4* out, D4* part, int n_i, int n_j, double eps, hipStream_t s,
                              int parts) {
  if (parts <= 0) parts = gravity64_forward_parts(n_i);
  if (parts > 8) throw std::invalid_argument("gravity: at most 8 j-parts");
  const int chunk = ceil_div(ceil_div(n_j, parts), kG64Tile) * kG64Tile;
  hipLaunchKernelGGL(k_gravity64_plummer, dim3(ceil_div(n_i, kGravBlock) * parts), dim3(kGravBlock), 0, s, gpos,
                     parts > 1 ? part : out, n_i, n_j, parts, chunk, eps * eps);
  CS_CHECK_LAUNCH();
  if (parts > 1) {
    hipLaunchKernelGGL(k_gravity64_combine, dim3(ceil_div(n_i, kBlock)), dim3(kBlock), 0, s, part, out, n_i, parts);
    CS_CHECK_LAUNCH();
  }
}

int gravity_sums_mode(int n_i, int sums) {
  if (sums >= 0) return sums;
  return n_i <= kFourSumsMaxN ? 1 : (n_i <= kTwoSumsMaxN ? 2 : 0);
}

template <class T>
void launch_gravity(const V4<T>* gpos, const V4<T>* gvel, V4<T>* out, int n_i, int n_j, int soft, T eps,
                    int exact, hipStream_t s, int sums) {
  const dim3 grid(ceil_div(n_i, kGravBlock)), block(kGravBlock);
  const T eps2 = eps * eps;
  if constexpr (sizeof(T) == 4) {
    if (soft == 1) {
      const int mode = gravity_sums_mode(n_i, sums);
      if (mode == 2) {
        const dim3 grid2(ceil_div(n_i, kTwoParticles));
        if (exact) hipLaunchKernelGGL(k_gravity_two<true>, grid2, block, 0, s, gpos, out, n_i, n_j, eps2);
        else hipLaunchKernelGGL(k_gravity_two<false>, grid2, block, 0, s, gpos, out, n_i, n_j, eps2);
      } else if (mode == 1) {
        const dim3 grid4(ceil_div(n_i, kSumParticles));
        if (exact) hipLaunchKernelGGL(k_gravity_sums<true>, grid4, block, 0, s, gpos, out, n_i, n_j, eps2);
        else hipLaunchKernelGGL(k_gravity_sums<false>, grid4, block, 0, s, gpos, out, n_i, n_j, eps2);
      } else {
        if (exact) hipLaunchKernelGGL(k_gravity_lane<true>, grid, block, 0, s, gpos, out, n_i, n_j, eps2);
        else hipLaunchKernelGGL(k_gravity_lane<false>, grid, block, 0, s, gpos, out, n_i, n_j, eps2);
      }
    } else {
      if (exact) hipLaunchKernelGGL((k_gravity<float, 0, 0, true>), grid, block, 0, s, gpos, gvel, out, n_i, n_j, eps2);
      else hipLaunchKernelGGL((k_gravity<float, 0, 0, false>), grid, block, 0, s, gpos, gvel, out, n_i, n_j, eps2);
    }
  } else {
    if (soft == 0) hipLaunchKernelGGL((k_gravity<double, 0, 2, true>), grid, block, 0, s, gpos, gvel, out, n_i, n_j, eps2);
    else hipLaunchKernelGGL((k_gravity<double, 1, 2, true>), grid, block, 0, s, gpos, gvel, out, n_i, n_j, eps2);
  }
  CS_CHECK_LAUNCH();
}

// j-parts of the pairwise mixed kernel: 8, one per XCD, unless the 8-part
// grid would take more than half of the GPU's block slots (8 blocks of
// 63-register wavefronts per CU) in a single round.  Such a grid holds the
// CUs until gravity ends and the latency-bound SPH passes queued beside it
// find no slot; with fewer parts gravity keeps about half of each SIMD's 8
// wavefront slots, which still fill its VALU, and the passes run in the rest
// (their walks at the higher issue priority).  Mixed step at 65 536
// particles (256 i-blocks): 8 parts 3.21 ms, 4 parts 2.19, 3 parts 2.39
// (gravity alone 2.03 / 2.07 / 2.22 ms).  Grids of several rounds keep 8
// (262 144 particles: 8 parts 22.6 ms/step, 2 parts 25.0, 1 part 28.0).  The
// far-field kernel keeps 8.
int gravity_mixed_parts(int n_i, int far) {
  if (far) return kGravParts;
  int dev = 0, cus = 0;
  CS_HIP(hipGetDevice(&dev));
  CS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const long slots = 8l * cus, blocks = (long)ceil_div(n_i, kGravBlock);
  if (blocks * kGravParts <= slots / 2 || blocks * kGravParts > slots) return kGravParts;
  return (int)std::max(1l, slots / 2 / blocks);
}

void launch_gravity_mixed(const F4* hi, const F4* lo, const F4* gvel, F4* out, F4* part, float* mass_range,
                          float* moments, float far_ratio, float far_grade, int n_i, int n_j, int n_real,
                          int soft, float eps, int split_only, hipStream_t s, int parts) {
  const dim3 grid(ceil_div(n_i, kGravBlock)), block(kGravBlock);
  if (soft == 0) {
    hipLaunchKernelGGL(k_gravity_mixed_spline, grid, block, 0, s, hi, lo, gvel, out, n_i, n_j);
  } else {
    // mass_range: 2 floats of result followed by 2 per block of n_real
    const int nb = ceil_div(n_real, kBlock);
    hipLaunchKernelGGL(k_mass_range_partials, dim3(nb), dim3(kBlock), 0, s, hi, n_real, mass_range + 2);
    CS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_mass_range_final, dim3(1), dim3(kBlock), 0, s, mass_range + 2, nb, mass_range);
    CS_CHECK_LAUNCH();
    const bool far = (split_only & 16) && moments;
    if (parts <= 0 || far) parts = gravity_mixed_parts(n_i, far);
    if (parts > kGravParts) throw std::invalid_argument("gravity: at most 8 j-parts");
    const int chunk = ceil_div(ceil_div(n_j, parts), kTile) * kTile;
    F4* part_hi = part;
    F4* part_lo = part + (size_t)kGravParts * n_i;
    float* cells = nullptr;
    float* halves = nullptr;
    if ((split_only & 16) && moments) {
      const int n_sub = ceil_div(n_j, kTile);
      hipLaunchKernelGGL(k_subtile_moments, dim3(ceil_div(n_sub, kGravBlock / 64)), block, 0, s, hi, lo, moments,
                         n_j, n_real, n_sub, kTile, split_only, mass_range);
      CS_CHECK_LAUNCH();
      // the cell levels follow the sub-tile records in the same buffer
      cells = moments + (size_t)n_sub * kMomWords;
      size_t off = 0;
      for (int L = 0; L < kFarLevels; ++L) {
        const int n_cells = ceil_div(n_j, kCellSize[L]);
        hipLaunchKernelGGL(k_cell_moments, dim3(n_cells), block, 0, s, hi, lo, cells + off, n_j, n_real,
                           kCellSize[L], split_only, mass_range);
        CS_CHECK_LAUNCH();
        off += (size_t)n_cells * kMomWords;
      }
      if (split_only & 64) {   // the 32-record halves of the sub-tiles, after the cells
        halves = cells + off;
        const int n_half = 2 * n_sub;
        hipLaunchKernelGGL(k_subtile_moments, dim3(ceil_div(n_half, kGravBlock / 64)), block, 0, s, hi, lo, halves,
                           n_j, n_real, n_half, kTile / 2, split_only, mass_range);
        CS_CHECK_LAUNCH();
      }
    }
    // ratio q at the coarsest level, graded by (C / C_coarsest)^far_grade:
    // an order-4 group's truncation error goes as M R^5 / y^7, and with
    // equal masses M grows with the record count C while R grows as
    // C^(1/3), so a grade near 1/15 asks every level for the same bound
    FarRatios fq{};
    for (int L = 0; L <= kFarLevels + 1; ++L) {
      const float c = (float)(L < kFarLevels ? kCellSize[L] : (L == kFarLevels ? kTile : kTile / 2)) /
                      (float)kCellSize[0];
      const float q = far_ratio * powf(c, far_grade);
      fq.q2[L] = q * q;
    }
    const dim3 ggrid(ceil_div(n_i, kGravBlock) * parts);
    if ((split_only & 16) && moments)
      hipLaunchKernelGGL((k_gravity_mixed_plummer<kJTileMixed, true>), ggrid, block, 0, s, hi, lo, part_hi, part_lo,
                         n_i, n_j, chunk, eps * eps, split_only, mass_range, moments, fq, cells, halves, n_real,
                         parts);
    else
      hipLaunchKernelGGL((k_gravity_mixed_plummer<kJTileMixed, false>), ggrid, block, 0, s, hi, lo, part_hi, part_lo,
                         n_i, n_j, chunk, eps * eps, split_only, mass_range, nullptr, fq, nullptr, nullptr, n_real,
                         parts);
    CS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_gravity_combine, dim3(ceil_div(n_i, kBlock)), dim3(kBlock), 0, s, part_hi, part_lo, out,
                       n_i, mass_range, split_only, parts);
  }
  CS_CHECK_LAUNCH();
}

template void launch_gravity<float>(const F4*, const F4*, F4*, int, int, int, float, int, hipStream_t, int);
template void launch_gravity<double>(const D4*, const D4*, D4*, int, int, int, double, int, hipStream_t, int);

}  // namespace cudasph
