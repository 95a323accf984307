// Device-wide primitives: exclusive scan, stable LSD radix sort of
// (key,value) records, the counting-sort transpose that turns the i-side
// neighbour lists into j-side reaction lists, and Peano-Hilbert keys.
//
// The radix sort replaces the reference's RadixSort (radixsort.cu:50-77).
// Each pass (8 key bits) is three launches: per-tile digit histograms,
// an exclusive scan of the digit-major histogram table, and a stable
// scatter.  The scatter ranks a tile in 16 striped rounds of 256 records; in
// each round a wavefront finds the lanes holding the same digit with eight
// ballots (one per digit bit), takes its rank from a popcount, and adds the
// counts of the lower wavefronts from LDS.  Equal keys keep their input
// order, which the reaction pass relies on.
#include <algorithm>

#include "kernels.h"

namespace cudasph {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;   // 2048

__device__ __forceinline__ int scan_load(const int* in, long idx, long n, int mode) {
  if (idx >= n) return 0;
  const int v = in[idx];
  return mode == 1 ? n_norm(v, 4) : v;
}

// block-wide exclusive scan of one value per thread; returns the block total
__device__ __forceinline__ int block_exclusive_scan(int v, int* s_tmp, int& total) {
  const int tid = threadIdx.x;
  s_tmp[tid] = v;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    const int add = tid >= off ? s_tmp[tid - off] : 0;
    __syncthreads();
    s_tmp[tid] += add;
    __syncthreads();
  }
  const int incl = s_tmp[tid];
  total = s_tmp[kScanBlock - 1];
  __syncthreads();
  return incl - v;
}

// Batched scans (blockIdx.y = batch b): array b's input and output start
// `stride` ints after array b-1's, its tile partials `pstride` ints after
// b-1's.
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const int* __restrict__ in, long n, int mode,
                                                             int* __restrict__ partial, long stride, long pstride) {
  __shared__ int s_tmp[kScanBlock];
  in += blockIdx.y * stride;
  partial += blockIdx.y * pstride;
  const long base = (long)blockIdx.x * kScanTile + (long)threadIdx.x * kScanItems;
  int sum = 0;
  for (int k = 0; k < kScanItems; ++k) sum += scan_load(in, base + k, n, mode);
  int total;
  block_exclusive_scan(sum, s_tmp, total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// one block scans the tile partials in chunks of 256, carrying the sum
__global__ __launch_bounds__(kScanBlock) void k_scan_partials(int* __restrict__ partial, int n_part, long pstride) {
  __shared__ int s_tmp[kScanBlock];
  partial += blockIdx.y * pstride;
  int carry = 0;
  for (int c0 = 0; c0 < n_part; c0 += kScanBlock) {
    const int idx = c0 + threadIdx.x;
    const int v = idx < n_part ? partial[idx] : 0;
    int total;
    const int ex = block_exclusive_scan(v, s_tmp, total);
    if (idx < n_part) partial[idx] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) partial[n_part] = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_down(const int* __restrict__ in, long n, int mode,
                                                           const int* __restrict__ partial, int* __restrict__ out,
                                                           int n_part, long stride, long pstride) {
  __shared__ int s_tmp[kScanBlock];
  in += blockIdx.y * stride;
  partial += blockIdx.y * pstride;
  out += blockIdx.y * stride;
  const long base = (long)blockIdx.x * kScanTile + (long)threadIdx.x * kScanItems;
  int v[kScanItems];
  int sum = 0;
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = scan_load(in, base + k, n, mode);
    sum += v[k];
  }
  int total;
  int run = block_exclusive_scan(sum, s_tmp, total) + partial[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = partial[n_part];
}

long scan_scratch_ints(long n) { return (n + kScanTile - 1) / kScanTile + 1; }

// `batch` scans at once: in and out of scan b at b * stride ints past the
// first's, its scratch at b * pstride
static void exclusive_scan_batch(const int* in, int* out, long n, int mode, int* scratch, hipStream_t s,
                                 int batch, long stride, long pstride) {
  const int n_part = (int)((n + kScanTile - 1) / kScanTile);
  if (n_part == 0) {
    for (int b = 0; b < batch; ++b) CS_HIP(hipMemsetAsync(out + b * stride, 0, sizeof(int), s));
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3(n_part, batch), dim3(kScanBlock), 0, s, in, n, mode, scratch, stride,
                     pstride);
  CS_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_partials, dim3(1, batch), dim3(kScanBlock), 0, s, scratch, n_part, pstride);
  CS_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_down, dim3(n_part, batch), dim3(kScanBlock), 0, s, in, n, mode, scratch, out, n_part,
                     stride, pstride);
  CS_CHECK_LAUNCH();
}

void launch_exclusive_scan(const int* in, int* out, long n, int mode, int* scratch, hipStream_t s) {
  exclusive_scan_batch(in, out, n, mode, scratch, s, 1, 0, 0);
}

void launch_exclusive_scan_pair(const int* in, int* out, long n, long stride, int* scratch, hipStream_t s) {
  exclusive_scan_batch(in, out, n, 0, scratch, s, 2, stride, scan_scratch_ints(n));
}

// ---------------------------------------------------------------------------
// radix sort
constexpr int kRadixBlock = 256;
constexpr int kRadixRounds = 16;
constexpr int kRadixTile = kRadixBlock * kRadixRounds;   // 4096 records per block

__global__ __launch_bounds__(kRadixBlock) void k_radix_hist(const KV* __restrict__ in, long n, int shift,
                                                             int* __restrict__ counts, int n_blocks,
                                                             long data_stride, long scratch_stride) {
  __shared__ int s_hist[256];
  in += blockIdx.y * data_stride;
  counts += blockIdx.y * scratch_stride;
  s_hist[threadIdx.x] = 0;
  __syncthreads();
  const long base = (long)blockIdx.x * kRadixTile;
  for (int r = 0; r < kRadixRounds; ++r) {
    const long p = base + r * kRadixBlock + threadIdx.x;
    if (p < n) atomicAdd(&s_hist[(in[p].key >> shift) & 255], 1);
  }
  __syncthreads();
  counts[(long)threadIdx.x * n_blocks + blockIdx.x] = s_hist[threadIdx.x];
}

__global__ __launch_bounds__(kRadixBlock) void k_radix_scatter(const KV* __restrict__ in, KV* __restrict__ out,
                                                                long n, int shift,
                                                                const int* __restrict__ offsets, int n_blocks,
                                                                long data_stride, long scratch_stride) {
  __shared__ int s_base[256];
  in += blockIdx.y * data_stride;
  out += blockIdx.y * data_stride;
  offsets += blockIdx.y * scratch_stride;
  __shared__ int s_wave[4][256];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  s_base[tid] = offsets[(long)tid * n_blocks + blockIdx.x];
  const long base = (long)blockIdx.x * kRadixTile;
  for (int r = 0; r < kRadixRounds; ++r) {
    s_wave[0][tid] = 0;
    s_wave[1][tid] = 0;
    s_wave[2][tid] = 0;
    s_wave[3][tid] = 0;
    __syncthreads();
    const long p = base + r * kRadixBlock + tid;
    const bool valid = p < n;
    KV rec{0u, 0u};
    int d = 0;
    if (valid) {
      rec = in[p];
      d = (int)((rec.key >> shift) & 255);
    }
    unsigned long long same = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int bit = (d >> b) & 1;
      const unsigned long long m = __ballot(valid && bit);
      same &= bit ? m : ~m;
    }
    const int rank = __popcll(same & ((1ull << lane) - 1ull));
    if (valid && rank == 0) s_wave[wave][d] = __popcll(same);
    __syncthreads();
    if (valid) {
      int before = 0;
      for (int w = 0; w < wave; ++w) before += s_wave[w][d];
      out[s_base[d] + before + rank] = rec;
    }
    __syncthreads();
    s_base[tid] += s_wave[0][tid] + s_wave[1][tid] + s_wave[2][tid] + s_wave[3][tid];
    __syncthreads();
  }
}

long radix_scratch_ints(long n) {
  const long nb = (n + kRadixTile - 1) / kRadixTile;
  const long table = 256 * nb;
  return 2 * table + 2 + scan_scratch_ints(table);
}

void launch_radix_sort_batch(KV* data, KV* alt, long n, int bits, int* scratch, hipStream_t s, int batch,
                             long data_stride) {
  if (n <= 1 || bits <= 0 || batch <= 0) return;
  const int nb = (int)((n + kRadixTile - 1) / kRadixTile);
  const long table = 256L * nb;
  const long sstride = radix_scratch_ints(n);   // each sort's scratch
  int* counts = scratch;
  int* offsets = scratch + table;
  int* scan_tmp = scratch + 2 * table + 2;
  KV* src = data;
  KV* dst = alt;
  int passes = 0;
  for (int shift = 0; shift < bits; shift += 8, ++passes) {
    hipLaunchKernelGGL(k_radix_hist, dim3(nb, batch), dim3(kRadixBlock), 0, s, src, n, shift, counts, nb,
                       data_stride, sstride);
    CS_CHECK_LAUNCH();
    // the scan's input, output and partials sit at the same offsets in
    // every sort's scratch
    exclusive_scan_batch(counts, offsets, table, 0, scan_tmp, s, batch, sstride, sstride);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb, batch), dim3(kRadixBlock), 0, s, src, dst, n, shift, offsets, nb,
                       data_stride, sstride);
    CS_CHECK_LAUNCH();
    KV* t = src;
    src = dst;
    dst = t;
  }
  if (passes & 1)
    for (int b = 0; b < batch; ++b)
      CS_HIP(hipMemcpyAsync(data + b * data_stride, alt + b * data_stride, n * sizeof(KV), hipMemcpyDeviceToDevice,
                            s));
}

void launch_radix_sort(KV* data, KV* alt, long n, int bits, int* scratch, hipStream_t s) {
  launch_radix_sort_batch(data, alt, n, bits, scratch, s, 1, 0);
}

// ---------------------------------------------------------------------------
// reaction transpose: for every j, the list slots k that name j, in
// ascending k (ascending owner i: the slots of i are contiguous and the
// offsets ascend with i), as pairs {key = owner i, value = k} in
// pairs[start[j], start[j + 1]).  A counting sort: per-j counts, their
// exclusive scan, a scatter that places each slot in its j's bucket (in
// whatever order the atomics resolve), and a sort of each bucket by k.

// one wavefront per particle i (measured faster than 16 lanes per particle:
// 0.69 against 1.28 ms for the count at N = 2^20): each real slot adds one to
// its j's count
__global__ __launch_bounds__(kBlock) void k_tr_count(const int* __restrict__ offset, const int* __restrict__ count,
                                                      const int* __restrict__ list, int* __restrict__ tcnt, int n) {
  const int i = blockIdx.x * (kBlock / 64) + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const int off = offset[i], c = count[i];
  for (int k = off + lane; k < off + c; k += 64) atomicAdd(&tcnt[list[k]], 1);
}

// each real slot takes a place in its j's bucket; tcnt counts down to zero
__global__ __launch_bounds__(kBlock) void k_tr_fill(const int* __restrict__ offset, const int* __restrict__ count,
                                                     const int* __restrict__ list, const int* __restrict__ start,
                                                     int* __restrict__ tcnt, KV* __restrict__ pairs, int n) {
  const int i = blockIdx.x * (kBlock / 64) + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const int off = offset[i], c = count[i];
  for (int k = off + lane; k < off + c; k += 64) {
    const int j = list[k];
    const int p = atomicSub(&tcnt[j], 1) - 1;
    pairs[start[j] + p] = KV{(uint32_t)i, (uint32_t)k};
  }
}

// bitonic sort of the m <= kTrSortMax pairs at p by k (value), ascending, in
// the one-wavefront block's LDS array s
__device__ __forceinline__ void bucket_sort(KV* __restrict__ p, int m, unsigned long long* s, int lane) {
  int P = 2;
  while (P < m) P <<= 1;
  for (int t = lane; t < P; t += 64) {
    if (t < m) {
      const KV e = p[t];
      s[t] = ((unsigned long long)e.value << 32) | e.key;
    } else {
      s[t] = ~0ull;
    }
  }
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1)
    for (int h = k >> 1; h > 0; h >>= 1) {
      for (int t = lane; t < P / 2; t += 64) {
        const int x = ((t & ~(h - 1)) << 1) | (t & (h - 1)), y = x + h;
        const unsigned long long u = s[x], v = s[y];
        const bool up = (x & k) == 0;
        if ((u > v) == up) {
          s[x] = v;
          s[y] = u;
        }
      }
      __syncthreads();
    }
  for (int t = lane; t < m; t += 64) p[t] = KV{(uint32_t)(s[t] & 0xffffffffu), (uint32_t)(s[t] >> 32)};
}

// one one-wavefront block per bucket (measured faster than blocks that loop
// over buckets: 1.0 against 1.17 ms at N = 2^20): a bitonic sort by k in LDS
// (the k are distinct) for buckets of up to kTrSortMax slots; a larger bucket
// is ranked instead (each slot's place is the number of smaller k), through
// `alt`
constexpr int kTrSortMax = 512;
__global__ __launch_bounds__(64) void k_tr_sort(KV* __restrict__ pairs, KV* __restrict__ alt,
                                                 const int* __restrict__ start, int n) {
  __shared__ unsigned long long s[kTrSortMax];
  const int j = blockIdx.x, lane = threadIdx.x;
  if (j >= n) return;
  const int b = start[j], m = start[j + 1] - b;
  if (m <= 1) return;
  if (m > kTrSortMax) {
    for (int t = lane; t < m; t += 64) {
      const KV e = pairs[b + t];
      int r = 0;
      for (int q = 0; q < m; ++q) r += pairs[b + q].value < e.value;
      alt[b + r] = e;
    }
    __syncthreads();
    for (int t = lane; t < m; t += 64) pairs[b + t] = alt[b + t];
    return;
  }
  bucket_sort(pairs + b, m, s, lane);
}

void launch_reaction_transpose(const int* offset, const int* count, const int* list, int n, int* tcnt,
                               int* start, int* end, KV* pairs, KV* alt, int* scan_scratch, hipStream_t s) {
  CS_HIP(hipMemsetAsync(tcnt, 0, (size_t)n * sizeof(int), s));
  const dim3 waves(ceil_div(n, kBlock / 64));
  hipLaunchKernelGGL(k_tr_count, waves, dim3(kBlock), 0, s, offset, count, list, tcnt, n);
  CS_CHECK_LAUNCH();
  launch_exclusive_scan(tcnt, start, n, 0, scan_scratch, s);
  CS_HIP(hipMemcpyAsync(end, start + 1, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_tr_fill, waves, dim3(kBlock), 0, s, offset, count, list, start, tcnt, pairs, n);
  CS_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_tr_sort, dim3(n), dim3(64), 0, s, pairs, alt, start, n);
  CS_CHECK_LAUNCH();
}

__global__ __launch_bounds__(kBlock) void k_set_keys(KV* __restrict__ kv, const uint32_t* __restrict__ keys, int n) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t < n) kv[t].key = keys[t];
}
void launch_set_keys(KV* kv, const uint32_t* keys, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_set_keys, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, s, kv, keys, n);
  CS_CHECK_LAUNCH();
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_reaction(const V4<T>* __restrict__ pos, V4<T>* __restrict__ bdots,
                                                      const T* __restrict__ ngb_acc, const KV* __restrict__ pairs,
                                                      const int* __restrict__ start, const int* __restrict__ end,
                                                      int n) {
  const int j = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (j >= n) return;
  const V4<T> pj = pos[j];
  V4<T> d = bdots[j];
  const int p0 = start[j], p1 = end[j];
  // software pipeline: the next slot's owner, coefficient and position are
  // requested before the current term is added (the sum keeps slot order)
  KV e = p0 < p1 ? pairs[p0] : KV{(uint32_t)j, 0u};
  T an = p0 < p1 ? ngb_acc[e.value] : T(0);
  V4<T> qn = pos[e.key];
  for (int p = p0; p < p1; ++p) {
    const T acc = an;
    const V4<T> pi = qn;
    if (p + 1 < p1) {
      e = pairs[p + 1];
      const int in_ = (int)e.key;
      an = ngb_acc[e.value];
      qn = pos[in_];
    }
    d.x += acc * (pi.x - pj.x);
    d.y += acc * (pi.y - pj.y);
    d.z += acc * (pi.z - pj.z);
  }
  bdots[j] = d;
}

template <class T>
void launch_reaction(const V4<T>* pos, V4<T>* bdots, const T* ngb_acc, const KV* pairs, const int* start,
                     const int* end, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_reaction<T>, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, s, pos, bdots, ngb_acc, pairs,
                     start, end, n);
  CS_CHECK_LAUNCH();
}
template void launch_reaction<float>(const F4*, F4*, const float*, const KV*, const int*, const int*, int,
                                     hipStream_t);
template void launch_reaction<double>(const D4*, D4*, const double*, const KV*, const int*, const int*, int,
                                      hipStream_t);

// mixed form: the same gather with split positions and compensated sums; a
// group of 8 lanes per j, each lane taking every 8th of j's slots, partial
// sums combined with exact two-sums
__global__ __launch_bounds__(kBlock) void k_reaction_mixed(const F4* __restrict__ pos, const F4* __restrict__ lo,
                                                            F4* __restrict__ bdots, const float* __restrict__ ngb_acc,
                                                            const KV* __restrict__ pairs,
                                                            const int* __restrict__ start,
                                                            const int* __restrict__ end, int n) {
  constexpr int kG = 8;
  const int g = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const int j = g / kG, lane = g % kG;
  if (j >= n) return;   // whole groups leave together
  const F4 pj = pos[j], lj = lo[j];
  float s[3] = {0.f, 0.f, 0.f}, c[3] = {0.f, 0.f, 0.f};
  auto kadd = [&](int q, float x) {
    const float y = x - c[q];
    const float t = s[q] + y;
    c[q] = (t - s[q]) - y;
    s[q] = t;
  };
  for (int p = start[j] + lane; p < end[j]; p += kG) {
    const KV e = pairs[p];
    const int i = (int)e.key;
    const float acc = ngb_acc[e.value];
    const F4 pi = pos[i], li = lo[i];
    kadd(0, acc * ((pi.x - pj.x) + (li.x - lj.x)));
    kadd(1, acc * ((pi.y - pj.y) + (li.y - lj.y)));
    kadd(2, acc * ((pi.z - pj.z) + (li.z - lj.z)));
  }
  for (int q = 0; q < 3; ++q)
    for (int o = kG / 2; o > 0; o >>= 1) {
      const float s2 = __shfl_xor(s[q], o, kG), c2 = __shfl_xor(c[q], o, kG);
      const float t = s[q] + s2;
      const float bp = t - s[q];
      const float err = (s[q] - (t - bp)) + (s2 - bp);
      c[q] = (c[q] + c2) - err;
      s[q] = t;
    }
  if (lane != 0) return;
  F4 d = bdots[j];
  d.x += s[0] - c[0];
  d.y += s[1] - c[1];
  d.z += s[2] - c[2];
  bdots[j] = d;
}

void launch_reaction_mixed(const F4* pos, const F4* pos_lo, F4* bdots, const float* ngb_acc, const KV* pairs,
                           const int* start, const int* end, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_reaction_mixed, dim3(ceil_div(n * 8, kBlock)), dim3(kBlock), 0, s, pos, pos_lo, bdots, ngb_acc,
                     pairs, start, end, n);
  CS_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// Peano-Hilbert keys (sort_bodies.cpp:11-105, 119-163).  The state tables are
// Gadget-2's; they fix the curve and therefore the particle order.
__constant__ int c_ph_quad[24 * 8] = {
    0, 7, 1, 6, 3, 4, 2, 5,  7, 4, 6, 5, 0, 3, 1, 2,  4, 3, 5, 2, 7, 0, 6, 1,  3, 0, 2, 1, 4, 7, 5, 6,
    1, 0, 6, 7, 2, 3, 5, 4,  0, 3, 7, 4, 1, 2, 6, 5,  3, 2, 4, 5, 0, 1, 7, 6,  2, 1, 5, 6, 3, 0, 4, 7,
    6, 1, 7, 0, 5, 2, 4, 3,  1, 2, 0, 3, 6, 5, 7, 4,  2, 5, 3, 4, 1, 6, 0, 7,  5, 6, 4, 7, 2, 1, 3, 0,
    7, 6, 0, 1, 4, 5, 3, 2,  6, 5, 1, 2, 7, 4, 0, 3,  5, 4, 2, 3, 6, 7, 1, 0,  4, 7, 3, 0, 5, 6, 2, 1,
    6, 7, 5, 4, 1, 0, 2, 3,  7, 0, 4, 3, 6, 1, 5, 2,  0, 1, 3, 2, 7, 6, 4, 5,  1, 6, 2, 5, 0, 7, 3, 4,
    2, 3, 1, 0, 5, 4, 6, 7,  3, 4, 0, 7, 2, 5, 1, 6,  4, 5, 7, 6, 3, 2, 0, 1,  5, 2, 6, 1, 4, 3, 7, 0};
__constant__ int c_ph_rotx[24] = {4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 0, 1, 2, 3, 17, 18, 19, 16, 23, 20, 21, 22};
__constant__ int c_ph_roty[24] = {1, 2, 3, 0, 16, 17, 18, 19, 11, 8, 9, 10, 22, 23, 20, 21, 14, 15, 12, 13, 4, 5, 6, 7};
__constant__ int c_ph_nx[8] = {3, 0, 0, 2, 2, 0, 0, 1};
__constant__ int c_ph_ny[8] = {0, 1, 1, 2, 2, 3, 3, 0};
__constant__ int c_ph_sense[8] = {-1, -1, -1, +1, +1, -1, -1, -1};

__device__ unsigned long long ph_key_dev(int x, int y, int z, int bits) {
  unsigned long long key = 0;
  int rot = 0, sense = 1;
  for (int b = bits - 1; b >= 0; --b) {
    const int q = c_ph_quad[rot * 8 + (((x >> b) & 1) << 2) + (((y >> b) & 1) << 1) + ((z >> b) & 1)];
    key = (key << 3) + (unsigned long long)(sense == 1 ? q : 7 - q);
    for (int k = 0; k < c_ph_nx[q]; ++k) rot = c_ph_rotx[rot];
    for (int k = 0; k < c_ph_ny[q]; ++k) rot = c_ph_roty[rot];
    sense *= c_ph_sense[q];
  }
  return key;
}

// bounding box (sort_bodies.cpp:121-134): per-block min/max of x,y,z, then
// one block folds the block results into box[6 nb .. 6 nb + 5] (min and max
// are exact in any order)
template <class T>
__global__ __launch_bounds__(kBlock) void k_bbox(const V4<T>* __restrict__ pos, int n, T* __restrict__ out) {
  __shared__ T s[6][kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const V4<T> p = pos[i < n ? i : 0];
  s[0][threadIdx.x] = p.x; s[1][threadIdx.x] = p.y; s[2][threadIdx.x] = p.z;
  s[3][threadIdx.x] = p.x; s[4][threadIdx.x] = p.y; s[5][threadIdx.x] = p.z;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      for (int c = 0; c < 3; ++c) s[c][threadIdx.x] = fmin(s[c][threadIdx.x], s[c][threadIdx.x + off]);
      for (int c = 3; c < 6; ++c) s[c][threadIdx.x] = fmax(s[c][threadIdx.x], s[c][threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_bbox_fold(T* __restrict__ box, int nb) {
  __shared__ T s[6][kBlock];
  T v[6];
  for (int c = 0; c < 6; ++c) v[c] = box[c];   // block 0 always exists
  for (int b = threadIdx.x; b < nb; b += kBlock) {
    for (int c = 0; c < 3; ++c) v[c] = fmin(v[c], box[6 * b + c]);
    for (int c = 3; c < 6; ++c) v[c] = fmax(v[c], box[6 * b + c]);
  }
  for (int c = 0; c < 6; ++c) s[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      for (int c = 0; c < 3; ++c) s[c][threadIdx.x] = fmin(s[c][threadIdx.x], s[c][threadIdx.x + off]);
      for (int c = 3; c < 6; ++c) s[c][threadIdx.x] = fmax(s[c][threadIdx.x], s[c][threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) box[6 * nb + threadIdx.x] = s[threadIdx.x][0];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_ph_keys(const V4<T>* __restrict__ pos, int n, int bits,
                                                     const T* __restrict__ box, uint32_t* key_hi,
                                                     uint32_t* key_lo) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  T lo[3], hi[3];
  for (int c = 0; c < 3; ++c) { lo[c] = box[c]; hi[c] = box[3 + c]; }
  const T size = fmax(hi[2] - lo[2], fmax(hi[1] - lo[1], hi[0] - lo[0]));
  const T fac = T(1.0 / (double)size * (double)(1LL << bits));
  const V4<T> p = pos[i];
  const int x = (int)((p.x - lo[0]) * fac), y = (int)((p.y - lo[1]) * fac), z = (int)((p.z - lo[2]) * fac);
  const unsigned long long key = ph_key_dev(x, y, z, bits);
  key_hi[i] = (uint32_t)(key >> 32);
  key_lo[i] = (uint32_t)key;
}

template <class T>
void launch_ph_keys(const V4<T>* pos, int n, int bits, uint32_t* key_hi, uint32_t* key_lo, int* scratch,
                    hipStream_t s) {
  const int nb = ceil_div(n, kBlock);
  T* box = reinterpret_cast<T*>(scratch);
  hipLaunchKernelGGL(k_bbox<T>, dim3(nb), dim3(kBlock), 0, s, pos, n, box);
  CS_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_bbox_fold<T>, dim3(1), dim3(kBlock), 0, s, box, nb);
  CS_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ph_keys<T>, dim3(nb), dim3(kBlock), 0, s, pos, n, bits, box + 6 * (size_t)nb, key_hi,
                     key_lo);
  CS_CHECK_LAUNCH();
}
template void launch_ph_keys<float>(const F4*, int, int, uint32_t*, uint32_t*, int*, hipStream_t);
template void launch_ph_keys<double>(const D4*, int, int, uint32_t*, uint32_t*, int*, hipStream_t);

}  // namespace cudasph
