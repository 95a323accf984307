// Device build of the reference's kd-tree (kd_tree.cpp:39-128, packing
// sph_accelerations.cpp:21-33).
//
// The tree is the left-balanced implicit tree: node p holds one particle, its
// children are 2p and 2p+1, the split axis of depth d is d % 3, and the node's
// particle is the one at rank size(2p) (the size of its left subtree) of its
// subtree's particles along that axis.  Every node of the nodes 1..n exists,
// so the tree's shape, every subtree's size and where it starts among the
// nodes of a level are closed forms of n (subtree_size, subtree_start).
//
// The build is level by level, on three lists of particle ids, one per axis,
// each sorted once by that coordinate (stable LSD radix sort on an
// order-preserving integer image of the coordinate) and kept grouped by
// subtree: at level L every particle not yet placed carries the heap index t
// of the level-L subtree that holds it, and each list holds the unplaced
// particles subtree by subtree (t ascending), each subtree's run still in
// its axis's order.  One level:
//  * the list of the split axis L % 3 holds each subtree's particles in split
//    order, so the particle at rank size(2t) of the run becomes node t, the
//    ones before go to subtree 2t and the ones after to 2t+1 -- both runs stay
//    contiguous and in order, shifted by the closed-form count of nodes placed
//    before them;
//  * the other two lists are partitioned stably: a particle going to 2t is
//    placed after the earlier ones of its run going to 2t (an exclusive scan
//    of "goes left" flags), one going to 2t+1 after the earlier ones going
//    right (its rank in the run minus those going left and the one placed).
// Levels: floor(log2 n) + 1.  Work per level is O(n); nothing is sorted after
// the first three sorts.
//
// Ties in a coordinate: the reference's first-element-pivot quicksort decides
// which of two equal-coordinate particles becomes the node; here the order of
// the first sort (ascending particle id) does.  Either way every particle of a
// left subtree has coordinate <= the node's and every one of a right subtree
// >= it, which is all the walks' pruning (dev_solve_range.cu:48-61) uses, so
// the neighbour sets are the same; without ties the two trees are identical.
#include "common.h"
#include "kernels.h"

namespace cudasph {
namespace {

// nodes in the subtree of heap node t of the complete tree of n nodes
__host__ __device__ __forceinline__ int subtree_size(int t, int n) {
  int size = 0;
  long long lo = t, hi = t;
  while (lo <= n) {
    size += (int)((hi < n ? hi : n) - lo + 1);
    lo <<= 1;
    hi = (hi << 1) | 1;
  }
  return size;
}

// among the nodes of levels >= L, laid out subtree by subtree, the offset of
// the subtree of level-L node t: the nodes below the level-L nodes left of t
__host__ __device__ __forceinline__ int subtree_start(int t, int L, int n) {
  int s = 0;
  long long a = 1ll << L, b = t;
  while (a <= n) {
    s += (int)((b < (long long)n + 1 ? b : (long long)n + 1) - a);
    a <<= 1;
    b <<= 1;
  }
  return s;
}

// order-preserving unsigned images of a coordinate (IEEE: flip all bits of a
// negative value, the sign bit of a positive one)
__device__ __forceinline__ uint32_t ordered_bits(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ uint64_t ordered_bits(double x) {
  const uint64_t u = (uint64_t)__double_as_longlong(x);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

template <class T>
__device__ __forceinline__ T coord(const V4<T>& p, int axis) {
  return axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
}

__device__ __forceinline__ float node_word(float, int bits) { return __int_as_float(bits); }
__device__ __forceinline__ double node_word(double, int bits) { return (double)bits; }

// key word `word` (0: low 32 bits, 1: high 32 bits of a double's image) of
// the coordinate of particle `value` (the value of the record, so a second
// pass re-keys records already ordered by the first)
template <class T>
__global__ __launch_bounds__(kBlock) void k_tree_keys(const V4<T>* __restrict__ pos, KV* __restrict__ kv, int n,
                                                       int word, int first) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int axis = blockIdx.y;   // the three axes' records one after another in kv
  kv += (size_t)axis * n;
  if (i >= n) return;
  const int id = first ? i : (int)kv[i].value;
  const auto b = ordered_bits(coord(pos[id], axis));
  kv[i] = {(uint32_t)(word ? (uint64_t)b >> 32 : (uint64_t)b & 0xffffffffu), (uint32_t)id};
}

// axis blockIdx.y's sorted ids into its list (the lists 2 n ints apart)
__global__ __launch_bounds__(kBlock) void k_tree_ids(const KV* __restrict__ kv, int* __restrict__ list,
                                                      int* __restrict__ tag, int n) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int axis = blockIdx.y;
  if (i >= n) return;
  list[(size_t)axis * 2 * n + i] = (int)kv[(size_t)axis * n + i].value;
  if (axis == 0) tag[i] = 1;
}

// fixed slots (sph_accelerations.cpp:21-33, read_state.cpp:85-88): slot 0 and
// slot n_dev+1 zero, slots n+1..n_dev far sentinels with split 3; the
// identity hash, or the padding part of the heap-order hash
template <class T>
__global__ __launch_bounds__(kBlock) void k_tree_fixed(V4<T>* __restrict__ tree, KV* __restrict__ hash, int n,
                                                        int n_dev, int heap_order) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= n_dev + 2) return;
  if (t == 0 || t == n_dev + 1) tree[t] = {T(0), T(0), T(0), T(0)};
  else if (t > n) tree[t] = {T(1e10), T(1e10), T(1e10), node_word(T(0), -1)};
  if (t < n_dev) {
    if (!heap_order) hash[t] = {(uint32_t)t, (uint32_t)t};
    else if (t >= n) hash[t] = {1u << 24, (uint32_t)t};
  }
}

// level L on the split axis's list (m unplaced particles): place the medians,
// give every other particle its child subtree, and move it to its run of the
// next level's layout
template <class T>
__global__ __launch_bounds__(kBlock) void k_tree_split(const int* __restrict__ la, int* __restrict__ la_next,
                                                        int* __restrict__ tag, const V4<T>* __restrict__ pos,
                                                        V4<T>* __restrict__ tree, KV* __restrict__ hash, int m,
                                                        int L, int n, int axis, int heap_order,
                                                        int* __restrict__ ties) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= m) return;
  const int id = la[idx];
  const int t = tag[id];
  const int p = idx - subtree_start(t, L, n);
  const int lt = 2 * t <= n ? subtree_size(2 * t, n) : 0;
  const int placed_before = t - (1 << L);   // level-L nodes left of t
  if (p < lt) {
    tag[id] = 2 * t;
    la_next[idx - placed_before] = id;
  } else if (p == lt) {
    const V4<T> q = pos[id];
    tree[t] = {q.x, q.y, q.z, node_word(T(0), id | (axis << 24))};
    if (heap_order) hash[t - 1] = {(uint32_t)(t - 1), (uint32_t)id};
    // a run neighbour with the node's coordinate: which of the two becomes
    // the node is the presort's choice, not necessarily the reference's
    if (ties) {
      const T c = coord(q, axis);
      if ((p > 0 && coord(pos[la[idx - 1]], axis) == c) ||
          (p + 1 < subtree_size(t, n) && coord(pos[la[idx + 1]], axis) == c))
        atomicOr(ties, 1);
    }
  } else {
    tag[id] = 2 * t + 1;
    la_next[idx - placed_before - 1] = id;
  }
}

// the other two lists: "goes left" flags for the scan, and where in each list
// the run's placed particle sits
__global__ __launch_bounds__(kBlock) void k_tree_flags(const int* __restrict__ lb, const int* __restrict__ lc,
                                                        const int* __restrict__ tag, int* __restrict__ fb,
                                                        int* __restrict__ fc, int* __restrict__ at_b,
                                                        int* __restrict__ at_c, int m, int L) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= m) return;
  const int next_level = 2 << L;   // tags of the next level are >= 2^(L+1); a placed particle kept t < that
  const int tb = tag[lb[idx]], tc = tag[lc[idx]];
  fb[idx] = tb >= next_level && !(tb & 1);
  fc[idx] = tc >= next_level && !(tc & 1);
  if (tb < next_level) at_b[tb] = idx;
  if (tc < next_level) at_c[tc] = idx;
}

__device__ __forceinline__ void regroup_one(const int* __restrict__ l, int* __restrict__ l_next,
                                            const int* __restrict__ sc, const int* __restrict__ at,
                                            const int* __restrict__ tag, int idx, int L, int n) {
  const int id = l[idx];
  const int nt = tag[id];
  if (nt < (2 << L)) return;   // placed at this level
  const int t = nt >> 1;
  const int s = subtree_start(t, L, n);
  const int base = s - (t - (1 << L));   // where subtree 2t starts in the next layout
  const int lefts = sc[idx] - sc[s];
  if (!(nt & 1)) {
    l_next[base + lefts] = id;
  } else {
    const int lt = subtree_size(2 * t, n);
    const int rights = (idx - s) - lefts - (at[t] < idx ? 1 : 0);
    l_next[base + lt + rights] = id;
  }
}

__global__ __launch_bounds__(kBlock) void k_tree_regroup(const int* __restrict__ lb, int* __restrict__ lb_next,
                                                          const int* __restrict__ sb, const int* __restrict__ at_b,
                                                          const int* __restrict__ lc, int* __restrict__ lc_next,
                                                          const int* __restrict__ sc, const int* __restrict__ at_c,
                                                          const int* __restrict__ tag, int m, int L, int n) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= m) return;
  regroup_one(lb, lb_next, sb, at_b, tag, idx, L, n);
  regroup_one(lc, lc_next, sc, at_c, tag, idx, L, n);
}

// ---------------------------------------------------------------------------
// The last levels in LDS.  Once every subtree of level L0 holds at most
// kLdsMax particles, one block per subtree builds the rest of it in its own
// LDS: the block reads its three runs (particle ids, in the level-L0 layout),
// names each particle by its rank in the first run (a slot), and runs the
// split / flags / scans / regroup of the level kernels above on slots, with
// per-level positions taken relative to where the subtree's runs start in
// that level's layout.  The level kernels' dependent global-memory passes
// (the build's cost at small N) become LDS passes; the results are the same
// nodes in the same places.
constexpr int kLdsMax = 2048;
// threads of a finishing block: few blocks run (one per subtree), so each
// takes a whole CU's worth of wavefronts to keep its passes short
constexpr int kLdsThreads = 1024;

struct LocalLevel {
  int L, r0, r1, r0n;        // level, the runs' range in its layout, and their start at level L+1
  int t_base;                // first level-L node of the subtree
};

// exclusive scans of two values per thread (kLdsThreads threads) at once: within
// each wavefront by shuffles, then across the block's wavefronts through
// LDS (one barrier)
__device__ __forceinline__ void block_excl_scan2(int& x, int& y, int* s_tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int ix = x, iy = y;
  for (int o = 1; o < 64; o <<= 1) {
    const int ux = __shfl_up(ix, o), uy = __shfl_up(iy, o);
    if (lane >= o) {
      ix += ux;
      iy += uy;
    }
  }
  if (lane == 63) {
    s_tot[wave] = ix;
    s_tot[kLdsThreads / 64 + wave] = iy;
  }
  __syncthreads();
  int bx = 0, by = 0;
  for (int w = 0; w < wave; ++w) {
    bx += s_tot[w];
    by += s_tot[kLdsThreads / 64 + w];
  }
  x = bx + ix - x;
  y = by + iy - y;
}

template <class T>
__global__ __launch_bounds__(kLdsThreads) void k_tree_lds(const int* __restrict__ work, int* __restrict__ slot_of, int n,
                                                      int L0, int cur0, int cur1, int cur2,
                                                      const V4<T>* __restrict__ pos, V4<T>* __restrict__ tree,
                                                      KV* __restrict__ hash, int heap_order, int* __restrict__ ties) {
  __shared__ uint16_t s_list[3][2][kLdsMax];   // slots, per axis, double-buffered
  __shared__ int s_tag[kLdsMax];               // the heap node each slot's particle belongs to
  __shared__ uint8_t s_fb[kLdsMax], s_fc[kLdsMax];
  __shared__ uint16_t s_sb[kLdsMax], s_sc[kLdsMax];
  __shared__ uint16_t s_at_b[kLdsMax], s_at_c[kLdsMax];   // local position of each node's placed particle
  __shared__ int s_tot[2 * kLdsThreads / 64];
  const int tid = threadIdx.x;
  const int T0 = (1 << L0) + (int)blockIdx.x;
  if (T0 > n) return;
  const int base0 = subtree_start(T0, L0, n);
  const int m0 = subtree_size(T0, n);
  const int* gl[3];
  const int curs[3] = {cur0, cur1, cur2};
  for (int a = 0; a < 3; ++a) gl[a] = work + (size_t)(2 * a + curs[a]) * n + base0;
  // slots: rank in the first run; slot_of (indexed by particle) carries
  // the names to the other two runs
  for (int k = tid; k < m0; k += kLdsThreads) {
    const int id = gl[0][k];
    s_tag[k] = T0;
    s_list[0][0][k] = (uint16_t)k;
    slot_of[id] = k;
  }
  __syncthreads();
  for (int k = tid; k < m0; k += kLdsThreads) {
    s_list[1][0][k] = (uint16_t)slot_of[gl[1][k]];
    s_list[2][0][k] = (uint16_t)slot_of[gl[2][k]];
  }
  __syncthreads();
  int cur[3] = {0, 0, 0};
  for (int L = L0; L < 31; ++L) {
    const long long d0 = (long long)T0 << (L - L0), d1 = (long long)(T0 + 1) << (L - L0);
    if (d0 > n) break;
    const int r0 = subtree_start((int)d0, L, n);
    const int r1 = subtree_start((int)min(d1, (long long)n + 1), L, n);
    if (r0 >= r1) break;
    const long long e0 = d0 << 1;
    const int r0n = subtree_start((int)min(e0, (long long)n + 1), L + 1, n);
    const int a = L % 3, b = (L + 1) % 3, c = (L + 2) % 3;
    const int next_level = 2 << L;
    const int t_base = (int)d0;
    // split on axis a
    for (int i = tid; i < r1 - r0; i += kLdsThreads) {
      const int g = r0 + i;
      const int slot = s_list[a][cur[a]][i];
      const int t = s_tag[slot];
      const int p = g - subtree_start(t, L, n);
      const int lt = 2 * t <= n ? subtree_size(2 * t, n) : 0;
      const int pb = t - (1 << L);
      if (p < lt) {
        s_tag[slot] = 2 * t;
        s_list[a][cur[a] ^ 1][g - pb - r0n] = (uint16_t)slot;
      } else if (p == lt) {
        const int id = gl[0][slot];   // the slot is the particle's rank in the first run
        const V4<T> q = pos[id];
        tree[t] = {q.x, q.y, q.z, node_word(T(0), id | (a << 24))};
        if (heap_order) hash[t - 1] = {(uint32_t)(t - 1), (uint32_t)id};
        if (ties) {   // as in k_tree_split
          const T cq = coord(q, a);
          if ((p > 0 && coord(pos[gl[0][s_list[a][cur[a]][i - 1]]], a) == cq) ||
              (p + 1 < subtree_size(t, n) && coord(pos[gl[0][s_list[a][cur[a]][i + 1]]], a) == cq))
            atomicOr(ties, 1);
        }
      } else {
        s_tag[slot] = 2 * t + 1;
        s_list[a][cur[a] ^ 1][g - pb - 1 - r0n] = (uint16_t)slot;
      }
    }
    __syncthreads();
    // flags of the other two lists, and where each node's placed particle sits in them
    for (int i = tid; i < r1 - r0; i += kLdsThreads) {
      const int tb = s_tag[s_list[b][cur[b]][i]], tc = s_tag[s_list[c][cur[c]][i]];
      s_fb[i] = tb >= next_level && !(tb & 1);
      s_fc[i] = tc >= next_level && !(tc & 1);
      if (tb < next_level) s_at_b[tb - t_base] = (uint16_t)i;
      if (tc < next_level) s_at_c[tc - t_base] = (uint16_t)i;
    }
    __syncthreads();
    // exclusive scans of the flags over the runs (chunks of ceil(R / kLdsThreads) per thread)
    {
      const int R = r1 - r0;
      const int chunk = (R + kLdsThreads - 1) / kLdsThreads;
      const int i0 = min(R, tid * chunk), i1 = min(R, i0 + chunk);
      int cb = 0, cc = 0;
      for (int i = i0; i < i1; ++i) {
        cb += s_fb[i];
        cc += s_fc[i];
      }
      int ob = cb, oc = cc;
      block_excl_scan2(ob, oc, s_tot);
      for (int i = i0; i < i1; ++i) {
        s_sb[i] = (uint16_t)ob;
        s_sc[i] = (uint16_t)oc;
        ob += s_fb[i];
        oc += s_fc[i];
      }
    }
    __syncthreads();
    // regroup the other two lists into the next level's runs
    for (int i = tid; i < r1 - r0; i += kLdsThreads) {
      const int g = r0 + i;
      for (int k = 0; k < 2; ++k) {
        const int ax = k ? c : b;
        const int slot = s_list[ax][cur[ax]][i];
        const int nt = s_tag[slot];
        if (nt < next_level) continue;   // placed at this level
        const int t = nt >> 1;
        const int s = subtree_start(t, L, n);
        const int base = s - (t - (1 << L));
        const uint16_t* scn = k ? s_sc : s_sb;
        const int lefts = scn[i] - scn[s - r0];
        int dst;
        if (!(nt & 1)) {
          dst = base + lefts;
        } else {
          const int at = (k ? s_at_c : s_at_b)[t - t_base];
          dst = base + subtree_size(2 * t, n) + (g - s) - lefts - (at < i ? 1 : 0);
        }
        s_list[ax][cur[ax] ^ 1][dst - r0n] = (uint16_t)slot;
      }
    }
    __syncthreads();
    for (int k = 0; k < 3; ++k) cur[k] ^= 1;
  }
}

}  // namespace

long tree_work_ints(long n) { return 13 * n + 8; }

template <class T>
void launch_tree_build(const V4<T>* pos, int n, int n_dev, V4<T>* tree, KV* hash, int heap_order, int* work,
                       KV* kv, KV* kv_alt, int* sort_scratch, int* scan_scratch, hipStream_t s, int* ties) {
  if (n <= 0 || n > n_dev) throw std::invalid_argument("tree build: need 0 < n <= n_dev");
  hipLaunchKernelGGL((k_tree_fixed<T>), dim3(ceil_div(n_dev + 2, kBlock)), dim3(kBlock), 0, s, tree, hash, n, n_dev,
                     heap_order);
  CS_CHECK_LAUNCH();
  int* list[3][2];
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < 2; ++k) list[a][k] = work + (size_t)(2 * a + k) * n;
  // the scans write m + 1 outputs (the total last), so each holds n + 1;
  // fc and sc lie n + 1 after fb and sb, so one batched scan does both
  int* tag = work + 6 * (size_t)n;
  int* fb = work + 7 * (size_t)n;
  int* fc = fb + (size_t)n + 1;
  int* sb = work + 9 * (size_t)n + 2;
  int* sc = sb + (size_t)n + 1;
  int* at_b = work + 11 * (size_t)n + 4 - 1;   // indexed by heap node 1..n
  int* at_c = work + 12 * (size_t)n + 4 - 1;
  // the three presorted lists, sorted side by side (FP64 coordinates: low
  // key word, then high)
  const dim3 g3(ceil_div(n, kBlock), 3);
  hipLaunchKernelGGL((k_tree_keys<T>), g3, dim3(kBlock), 0, s, pos, kv, n, 0, 1);
  CS_CHECK_LAUNCH();
  launch_radix_sort_batch(kv, kv_alt, n, 32, sort_scratch, s, 3, n);
  if (sizeof(T) == 8) {
    hipLaunchKernelGGL((k_tree_keys<T>), g3, dim3(kBlock), 0, s, pos, kv, n, 1, 0);
    CS_CHECK_LAUNCH();
    launch_radix_sort_batch(kv, kv_alt, n, 32, sort_scratch, s, 3, n);
  }
  hipLaunchKernelGGL(k_tree_ids, g3, dim3(kBlock), 0, s, kv, list[0][0], tag, n);
  CS_CHECK_LAUNCH();
  int cur[3] = {0, 0, 0};
  int m = n;
  // level kernels while the level's subtrees hold more than kLdsMax
  // particles (the leftmost is the largest), then one launch in LDS
  int L_lds = 0;
  while (subtree_size(1 << L_lds, n) > kLdsMax) ++L_lds;
  for (int L = 0; m > 0 && L < L_lds; ++L) {
    const int a = L % 3, b = (L + 1) % 3, c = (L + 2) % 3;
    const dim3 gm(ceil_div(m, kBlock));
    hipLaunchKernelGGL((k_tree_split<T>), gm, dim3(kBlock), 0, s, list[a][cur[a]], list[a][cur[a] ^ 1], tag, pos,
                       tree, hash, m, L, n, a, heap_order, ties);
    CS_CHECK_LAUNCH();
    const int placed = (int)std::min<long>(1l << L, (long)n - ((1l << L) - 1));
    if (m - placed > 0) {
      hipLaunchKernelGGL(k_tree_flags, gm, dim3(kBlock), 0, s, list[b][cur[b]], list[c][cur[c]], tag, fb, fc, at_b,
                         at_c, m, L);
      CS_CHECK_LAUNCH();
      launch_exclusive_scan_pair(fb, sb, m, (long)n + 1, scan_scratch, s);
      hipLaunchKernelGGL(k_tree_regroup, gm, dim3(kBlock), 0, s, list[b][cur[b]], list[b][cur[b] ^ 1], sb, at_b,
                         list[c][cur[c]], list[c][cur[c] ^ 1], sc, at_c, tag, m, L, n);
      CS_CHECK_LAUNCH();
    }
    for (int k = 0; k < 3; ++k) cur[k] ^= 1;
    m -= placed;
  }
  if (m > 0) {
    const int roots = (int)std::min<long>(1l << L_lds, (long)n - (1l << L_lds) + 1);
    hipLaunchKernelGGL((k_tree_lds<T>), dim3(roots), dim3(kLdsThreads), 0, s, work, tag, n, L_lds, cur[0], cur[1], cur[2],
                       pos, tree, hash, heap_order, ties);
    CS_CHECK_LAUNCH();
  }
}

namespace {

// a coordinate as a float bound: rounded down (lo) or up (hi) when it is not
// a float already
__device__ __forceinline__ float lo_bound(float x) { return x; }
__device__ __forceinline__ float hi_bound(float x) { return x; }
// the neighbouring float below / above a finite f (its bit pattern moved by
// one unit away from or towards zero)
__device__ __forceinline__ float float_below(float f) {
  const int b = __float_as_int(f);
  return f > 0.f ? __int_as_float(b - 1) : (f < 0.f ? __int_as_float(b + 1) : -__int_as_float(1));
}
__device__ __forceinline__ float float_above(float f) {
  const int b = __float_as_int(f);
  return f > 0.f ? __int_as_float(b + 1) : (f < 0.f ? __int_as_float(b - 1) : __int_as_float(1));
}
__device__ __forceinline__ float lo_bound(double x) {
  const float f = (float)x;
  return (double)f > x ? float_below(f) : f;
}
__device__ __forceinline__ float hi_bound(double x) {
  const float f = (float)x;
  return (double)f < x ? float_above(f) : f;
}

// box of node k: the node's own particle and the boxes of the children a
// walk can enter from it (their boxes written before)
template <class T>
__device__ __forceinline__ void box_node(const V4<T>* __restrict__ tree, F4* box, int n_nodes, int fix_phantom,
                                         int k) {
  const V4<T> nd = tree[k];
  F4 lo = {lo_bound(nd.x), lo_bound(nd.y), lo_bound(nd.z), 0.f};
  F4 hi = {hi_bound(nd.x), hi_bound(nd.y), hi_bound(nd.z), 0.f};
  const int c = 2 * k;
  for (int ch = c; ch <= c + 1; ++ch) {
    // children 2k, 2k+1 <= n_nodes; node n_nodes+1 (the phantom) is entered
    // from 2k = n_nodes unless fix_phantom
    if (ch > n_nodes && !(ch == n_nodes + 1 && c == n_nodes && !fix_phantom)) continue;
    const F4 a = box[2 * ch], b = box[2 * ch + 1];
    lo = {fminf(lo.x, a.x), fminf(lo.y, a.y), fminf(lo.z, a.z), 0.f};
    hi = {fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z), 0.f};
  }
  box[2 * k] = lo;
  box[2 * k + 1] = hi;
}

// One block per subtree rooted at level S (node 2^S + blockIdx.x): its levels
// deepest first, with a barrier between levels, so that a node's children
// were written by the same block before it reads them.
template <class T>
__global__ __launch_bounds__(kBlock) void k_tree_boxes_sub(const V4<T>* __restrict__ tree, F4* box, int n_nodes,
                                                            int fix_phantom, int S, int levels, int last_node) {
  const long r = (1l << S) + blockIdx.x;
  for (int L = levels; L >= S; --L) {
    const long first = r << (L - S), cnt = 1l << (L - S);
    for (long q = threadIdx.x; q < cnt; q += kBlock)
      if (first + q <= last_node) box_node<T>(tree, box, n_nodes, fix_phantom, (int)(first + q));
    __syncthreads();
  }
}

// the levels above S, in one block
template <class T>
__global__ __launch_bounds__(kBlock) void k_tree_boxes_top(const V4<T>* __restrict__ tree, F4* box, int n_nodes,
                                                            int fix_phantom, int S, int last_node) {
  for (int L = S - 1; L >= 0; --L) {
    const int last = (int)std::min<long>((2l << L) - 1, (long)last_node);
    for (int k = (1 << L) + (int)threadIdx.x; k <= last; k += kBlock) box_node<T>(tree, box, n_nodes, fix_phantom, k);
    __syncthreads();
  }
}

}  // namespace

constexpr int kBoxSubtreeLevel = 10;

template <class T>
void launch_tree_boxes(const V4<T>* tree, F4* box, int n_nodes, int fix_phantom, hipStream_t s) {
  const int last_node = n_nodes + 1;   // the phantom's box is its record's point
  int levels = 0;
  while ((2l << levels) <= (long)last_node) ++levels;   // deepest level: floor(log2(last_node))
  // subtrees rooted at level S, up to 2^10 of them, then the levels above
  const int S = std::min(levels, kBoxSubtreeLevel);
  hipLaunchKernelGGL(k_tree_boxes_sub<T>, dim3(1 << S), dim3(kBlock), 0, s, tree, box, n_nodes, fix_phantom, S, levels,
                     last_node);
  CS_CHECK_LAUNCH();
  if (S > 0) {
    hipLaunchKernelGGL(k_tree_boxes_top<T>, dim3(1), dim3(kBlock), 0, s, tree, box, n_nodes, fix_phantom, S,
                       last_node);
    CS_CHECK_LAUNCH();
  }
}

template void launch_tree_build<float>(const F4*, int, int, F4*, KV*, int, int*, KV*, KV*, int*, int*, hipStream_t,
                                       int*);
template void launch_tree_build<double>(const D4*, int, int, D4*, KV*, int, int*, KV*, KV*, int*, int*, hipStream_t,
                                        int*);
template void launch_tree_boxes<float>(const F4*, F4*, int, int, hipStream_t);
template void launch_tree_boxes<double>(const D4*, F4*, int, int, hipStream_t);

}  // namespace cudasph
