// SPH passes on the GPU: kd-tree walk, smoothing-length solve, neighbour
// lists, density, EOS and hydro accelerations.
//
// One lane per particle.  Lanes take particles in the order of the hash
// array (thread t -> bodies_hash[t].value, as in the reference).  The engine
// fills it with kd-tree heap order, or with memory order when the particles
// are already in Peano-Hilbert order; either way the 64 lanes of a wavefront
// hold spatially adjacent particles, walk nearly the same nodes, and their
// neighbour gathers hit the same cache lines.  The per-lane walk stack lives
// in LDS, one column per lane (bank-conflict free).  Per-particle sums run in
// list order, so in the plain form each particle's result is the same bit
// pattern as the CPU restatement's.  The mixed form has its own density and
// hydro kernels (FP32 pair arithmetic, compensated sums).
#include "kernels.h"
#include "sph_math.h"

namespace cudasph {

// ---------------------------------------------------------------------------
// kd-tree walk (dev_solve_range.cu:17-75, dev_find_ngb.cu:13-73).
// Pop a node; push its left child if the query sphere reaches below the split
// plane and its right child if it reaches above (so the right subtree is
// visited first); then test the node's own particle.  Children are pushed for
// every node with 2p <= n_nodes, which for even n_nodes includes the
// nonexistent node n+1; the tree buffer holds a zero record there, which is
// what the reference's out-of-range texture fetch returns.  fix_phantom drops
// that push instead.  Returns the number of nodes visited.  With `phantom`
// given, bit 0 of *phantom is set when the walk reads node n+1 and bit 1 when
// that read passes the distance test (a spurious visit of particle 0 when the
// zero record sits there, n = n_dev).
//
// With `boxes` (subtree bounding boxes, launch_tree_boxes) a child is pushed
// only if its box also comes within the range: a subtree whose box lies
// beyond it holds no neighbour, so the neighbours, their order and every sum
// over them stay the same, and only the walk length shrinks.  The children's
// boxes (64 adjacent bytes) are fetched together with the node, so the chain
// of dependent fetches does not grow.  The test keeps a margin of 2^-16 of
// r^2 over the rounding of the box distance and of d^2.
template <class T>
__device__ __forceinline__ T box_dist2(const F4 lo, const F4 hi, T qx, T qy, T qz) {
  const T dx = fmax(fmax(T(lo.x) - qx, qx - T(hi.x)), T(0));
  const T dy = fmax(fmax(T(lo.y) - qy, qy - T(hi.y)), T(0));
  const T dz = fmax(fmax(T(lo.z) - qz, qz - T(hi.z)), T(0));
  return dx * dx + dy * dy + dz * dz;
}
template <class T>
__device__ __forceinline__ bool box_within(const F4 lo, const F4 hi, T qx, T qy, T qz, T r2m) {
  return box_dist2<T>(lo, hi, qx, qy, qz) <= r2m;
}
template <class T>
__device__ __forceinline__ T range2_margin(T r2) {   // the box test's r^2 with its margin
  return r2 * T(1.0 + 1.0 / 65536.0);
}

template <class T, class Visit>
__device__ __forceinline__ int kd_walk(const V4<T> q, const V4<T>* __restrict__ tree, int n_nodes,
                                       int fix_phantom, int* __restrict__ stk, Visit&& visit,
                                       int* phantom = nullptr, const F4* __restrict__ boxes = nullptr) {
  int top = 0;
  stk[0] = 1;
  int n_walk = 0;
  const int stride = (int)blockDim.x;   // the stack: one LDS column per lane of the block
  // the query by value, in scalars: selecting a component of a reference
  // becomes a select of addresses, which keeps the caller's copy in scratch
  const T qx = q.x, qy = q.y, qz = q.z, qw = q.w;
  const T r2 = qw * qw;
  const T r2m = range2_margin(r2);
  while (top >= 0) {
    ++n_walk;
    int id = stk[top * stride];
    const bool ph = phantom && id == n_nodes + 1;
    const V4<T> node = tree[id];
    F4 b0, b1, b2, b3;
    if (boxes) {   // children's boxes, beside the node fetch (a leaf fetches the root's: unused)
      const int cb = 2 * id <= n_nodes ? 4 * id : 4;
      b0 = boxes[cb];
      b1 = boxes[cb + 1];
      b2 = boxes[cb + 2];
      b3 = boxes[cb + 3];
    }
    const int bits = node_bits(node.w);
    const int split = (bits >> 24) & 3;
    id <<= 1;
    if (id <= n_nodes && split < 3) {
      const T qc = (split & 2) ? qz : ((split & 1) ? qy : qx);
      const T nc = (split & 2) ? node.z : ((split & 1) ? node.y : node.x);
      if (qc - qw < nc && (!boxes || box_within<T>(b0, b1, qx, qy, qz, r2m))) stk[(top++) * stride] = id;
      if (qc + qw > nc && (!fix_phantom || id < n_nodes) && (!boxes || box_within<T>(b2, b3, qx, qy, qz, r2m)))
        stk[(top++) * stride] = id + 1;
    }
    --top;
    const T dx = qx - node.x, dy = qy - node.y, dz = qz - node.z;
    const T ds2 = dx * dx + dy * dy + dz * dz;
    if (ds2 < r2) visit(ds2, bits & ~(3 << 24));
    if (ph) *phantom |= ds2 < r2 ? 3 : 1;
  }
  return n_walk;
}

template <class T>
struct NgbEval {
  T f, step, count;
};

// the sums of one evaluation, neighbour by neighbour in walk order
template <class T>
struct NgbSum {
  T f = T(0), df = T(0), cnt = T(0);
  __device__ __forceinline__ void add(T ds2, T range) {
    const T u = T(2) * dsqrt(ds2) / range;
    f += kern_ig<T>(u);
    df -= kern_digdh<T>(u);
    cnt += T(1);
  }
  __device__ __forceinline__ NgbEval<T> finish(T range, float desired_ngb) {
    f -= T(desired_ngb);
    df = T((double)df * (1.0 / (double)range));
    if (df != T(0)) df = f / df;
    else df = f < T(0) ? -range : range;
    return {f, df, cnt};
  }
};

// ---------------------------------------------------------------------------
// h-solve candidate cache (few particles, where a wavefront's walks are
// latency-bound and its slowest lane's several evaluations set its time).
// The first evaluation walks at 9/8 of the range and keeps every particle it
// accepts there, in walk order, with what decides whether a walk at a
// smaller range r would reach it: the walk reaches a node iff every push on
// its path passes at r.  A push's split-plane test, fl(qc - r) < nc (or
// fl(qc + r) > nc), holds for r above the margin qc - nc (nc - qc) up to
// rounding, so each record carries the path's largest margin m and a bound
// d >= 8 * eps * max(|qc|, |nc|, range) on the rounding: r >= m + 2d passes
// every split test, r <= m' - 2d (m' the largest lower bound) fails one.
// The box tests compare a path's largest box distance^2 b, which does not
// depend on r, with r^2 (1 + 2^-16): exactly the walk's comparison.  A later
// evaluation at r <= the cached range filters the records in order: a
// record inside r whose reach is decided adds its term, and one whose reach
// falls between the bounds sends that evaluation back to a walk.  So every
// evaluation sums the terms of the walk at r, in its order: the same bits.
template <class T>
struct SolveCache {
  V4<T>* rec = nullptr;   // this lane's first record, record k at rec[k * stride]; null: no cache
  int stride = 0;
  int count = -1;         // -1: not filled; > kSolveCacheK: overflowed (walks only)
  T radius = T(0);
};

template <class T>
__device__ __forceinline__ T round_bound(T a, T b, T c) {
  constexpr T eps = sizeof(T) == 4 ? T(1.0 / 8388608.0) : T(1.0 / 4503599627370496.0);
  return T(8) * eps * fmax(fmax(dabs(a), dabs(b)), c);
}

// kd_walk with the path bounds: the LDS stack holds, per entry, the node and
// the largest upper (hi) and lower (lo) margin bound and box distance^2 (b)
// of the pushes on its path.  visit(ds2, hi, lo, b) for each accepted node.
template <class T, class Visit>
__device__ __forceinline__ void kd_walk_paths(const V4<T> q, const V4<T>* __restrict__ tree, int n_nodes,
                                              int fix_phantom, int* __restrict__ stk, T* __restrict__ s_hi,
                                              T* __restrict__ s_lo, T* __restrict__ s_b, Visit&& visit,
                                              const F4* __restrict__ boxes) {
  const int stride = (int)blockDim.x;
  const T qx = q.x, qy = q.y, qz = q.z, qw = q.w;
  const T r2 = qw * qw;
  const T r2m = range2_margin(r2);
  const T lowest = -qw;   // the root's path: no push, reached at any r > 0
  int top = 0;
  stk[0] = 1;
  s_hi[0] = lowest;
  s_lo[0] = lowest;
  s_b[0] = T(0);
  while (top >= 0) {
    int id = stk[top * stride];
    const T hi = s_hi[top * stride], lo = s_lo[top * stride], pb = s_b[top * stride];
    const V4<T> node = tree[id];
    F4 b0, b1, b2, b3;
    if (boxes) {
      const int cb = 2 * id <= n_nodes ? 4 * id : 4;
      b0 = boxes[cb];
      b1 = boxes[cb + 1];
      b2 = boxes[cb + 2];
      b3 = boxes[cb + 3];
    }
    const int bits = node_bits(node.w);
    const int split = (bits >> 24) & 3;
    id <<= 1;
    if (id <= n_nodes && split < 3) {
      const T qc = (split & 2) ? qz : ((split & 1) ? qy : qx);
      const T nc = (split & 2) ? node.z : ((split & 1) ? node.y : node.x);
      const T d = round_bound<T>(qc, nc, qw);
      const T bl = boxes ? box_dist2<T>(b0, b1, qx, qy, qz) : T(0);
      const T br = boxes ? box_dist2<T>(b2, b3, qx, qy, qz) : T(0);
      const bool push_l = qc - qw < nc && (!boxes || bl <= r2m);
      const bool push_r = qc + qw > nc && (!fix_phantom || id < n_nodes) && (!boxes || br <= r2m);
      if (push_l) {
        stk[top * stride] = id;
        s_hi[top * stride] = fmax(hi, (qc - nc) + T(2) * d);
        s_lo[top * stride] = fmax(lo, (qc - nc) - T(2) * d);
        s_b[top * stride] = fmax(pb, bl);
        ++top;
      }
      if (push_r) {
        stk[top * stride] = id + 1;
        s_hi[top * stride] = fmax(hi, (nc - qc) + T(2) * d);
        s_lo[top * stride] = fmax(lo, (nc - qc) - T(2) * d);
        s_b[top * stride] = fmax(pb, br);
        ++top;
      }
    }
    --top;
    const T dx = qx - node.x, dy = qy - node.y, dz = qz - node.z;
    const T ds2 = dx * dx + dy * dy + dz * dz;
    if (ds2 < r2) visit(ds2, hi, lo, pb);
  }
}

// the cached records filtered at range r: false when some record's reach is
// not decided (the caller walks instead)
template <class T>
__device__ __forceinline__ bool filter_cache(const SolveCache<T>& c, T r, NgbSum<T>& s) {
  const T r2 = r * r;
  const T r2m = range2_margin(r2);
  for (int k = 0; k < c.count; ++k) {
    const V4<T> e = c.rec[(size_t)k * c.stride];   // {ds2, hi, lo, b}
    if (!(e.x < r2)) continue;
    const bool box_ok = e.w <= r2m;
    if (box_ok && r >= e.y) s.add(e.x, r);
    else if (box_ok && r > e.z) return false;
  }
  return true;
}

// smoothed count minus target, Newton step and integer count at range q.w
// (dev_solve_range.cu:17-89); with a cache, from its records where it can
template <class T, bool CACHE>
__device__ __forceinline__ NgbEval<T> eval_ngb(const V4<T>& q, const V4<T>* __restrict__ tree,
                                               const WalkParams& wp, int* stk, SolveCache<T>& sc) {
  if (CACHE && sc.count < 0) {   // first evaluation: fill the cache from a walk at 9/8 of the range
    const int stride = (int)blockDim.x, depth = walk_stack_depth(wp.n_nodes);
    T* s_hi = reinterpret_cast<T*>(stk - threadIdx.x + depth * stride) + threadIdx.x;
    T* s_lo = s_hi + depth * stride;
    T* s_b = s_lo + depth * stride;
    V4<T> qc = q;
    qc.w = q.w * T(1.125);
    int n = 0;
    kd_walk_paths<T>(qc, tree, wp.n_nodes, wp.fix_phantom, stk, s_hi, s_lo, s_b, [&](T ds2, T hi, T lo, T b) {
      if (n < kSolveCacheK) sc.rec[(size_t)n * sc.stride] = {ds2, hi, lo, b};
      ++n;
    }, wp.boxes);
    sc.count = n;
    sc.radius = qc.w;
  }
  if (CACHE && sc.count <= kSolveCacheK && q.w <= sc.radius) {
    NgbSum<T> s;
    if (filter_cache<T>(sc, q.w, s)) return s.finish(q.w, wp.desired_ngb);
  }
  NgbSum<T> s;
  kd_walk<T>(q, tree, wp.n_nodes, wp.fix_phantom, stk, [&](T ds2, int) { s.add(ds2, q.w); }, nullptr, wp.boxes);
  return s.finish(q.w, wp.desired_ngb);
}

template <class T>
__device__ __forceinline__ T clamp_step(T dx, T range) {   // f_scale = 0.1, dev_solve_range.cu:92,100-103
  if ((double)dabs(dx) > 0.1 * (double)range) dx = T((dx < T(0) ? -0.1 : 0.1) * (double)range);
  return dx;
}

// bracket [lo, hi] around the target (dev_solve_range.cu:93-135); z = count
template <class T, bool CACHE>
__device__ __forceinline__ V4<T> bracket(V4<T> q, const V4<T>* __restrict__ tree, const WalkParams& wp,
                                         int* stk, SolveCache<T>& sc) {
  const T rel = T(wp.rel_err);
  NgbEval<T> f = eval_ngb<T, CACHE>(q, tree, wp, stk, sc);
  T dx = f.step;
  if (dabs(dx) < rel * q.w)
    return {T((double)q.w - 0.5 * (double)dabs(dx)), T((double)q.w + 0.5 * (double)dabs(dx)), f.count, T(0)};
  dx = clamp_step<T>(dx, q.w);
  T x_prev = q.w, f_prev = f.f;
  int iter = 0;
  while ((f_prev * f.f >= T(0)) && (iter++ < wp.max_iter)) {
    x_prev = q.w;
    f_prev = f.f;
    q.w -= dx;
    f = eval_ngb<T, CACHE>(q, tree, wp, stk, sc);
    dx = f.step;
    if (dabs(dx) < rel * q.w)
      return {T((double)q.w - 0.5 * (double)dabs(dx)), T((double)q.w + 0.5 * (double)dabs(dx)), f.count, T(0)};
    dx = clamp_step<T>(dx, q.w);
  }
  if (x_prev < q.w) return {x_prev, q.w, f.count, T(0)};
  return {q.w, x_prev, f.count, T(0)};
}

// safeguarded Newton inside the bracket (dev_solve_range.cu:137-172)
template <class T, bool CACHE>
__device__ __forceinline__ void solve_range(V4<T> q, const V4<T>* __restrict__ tree, const WalkParams& wp,
                                            int* stk, T& range, T& count, SolveCache<T>& sc) {
  const T rel = T(wp.rel_err);
  V4<T> br = bracket<T, CACHE>(q, tree, wp, stk, sc);
  T lo = br.x, hi = br.y;
  count = br.z;
  q.w = T(0.5 * (double)(lo + hi));
  T dx = T(0);
  if (dabs(lo - hi) > rel * q.w) {
    bool keep = true;
    int iter = 0;
    while (keep && (iter++ < wp.max_iter)) {
      q.w -= dx;
      if ((q.w - hi) * (q.w - lo) >= T(0)) {          // left the bracket: bisect
        dx = T(0.5 * (double)(hi - lo));
        q.w = T(0.5 * (double)(lo + hi));
      }
      const NgbEval<T> f = eval_ngb<T, CACHE>(q, tree, wp, stk, sc);
      count = f.count;
      if (f.f < T(0)) lo = q.w;
      else hi = q.w;
      dx = f.step;
      keep = dabs(dx) > rel * q.w;
    }
  }
  range = q.w;
}

// The walk kernels are latency-bound: a wavefront waits on one dependent
// tree fetch after another and then needs only a few instructions.  Beside
// gravity, whose wavefronts always have VALU work ready, they would wait for
// issue behind them too (an h-solve of 3.9 ms alone took 23 ms beside the
// plain form's gravity at 262 144 particles).  Their wavefronts take the
// highest issue priority instead, so they issue as soon as their data is
// back and gravity fills the remaining cycles: plain step at 262 144
// particles 67.8 -> 62.9 ms, mixed 22.1 -> 21.4 ms.
__device__ __forceinline__ void walk_priority() { __builtin_amdgcn_s_setprio(3); }

// dev_solve_range.cu:174-189
template <class T, bool CACHE>
__global__ __launch_bounds__(kBlock) void k_solve_range(V4<T>* __restrict__ pos, int* __restrict__ n_ngb,
                                                         const KV* __restrict__ hash,
                                                         const V4<T>* __restrict__ tree, int n_dev,
                                                         WalkParams wp) {
  walk_priority();
  // walk_stack_depth(n_nodes) entries per lane; with the cache, three
  // columns of T per entry after them (the path bounds)
  extern __shared__ int s_stack[];
  const int t = xcd_block(blockIdx.x, gridDim.x) * (int)blockDim.x + threadIdx.x;
  if (t >= n_dev) return;
  const int i = (int)hash[t].value;
  const V4<T> q = pos[i];
  T range = T(0), count = T(0);
  SolveCache<T> sc;
  if (CACHE) {
    sc.rec = static_cast<V4<T>*>(wp.solve_cache) + t;
    sc.stride = n_dev;
  }
  if (q.w > T(0)) solve_range<T, CACHE>(q, tree, wp, s_stack + threadIdx.x, range, count, sc);
  pos[i].w = range;
  n_ngb[i] = (int)count;
}

// dev_find_ngb.cu:75-87.  Capacity of particle i's slot is
// n_norm(cap_count[i], 4) (the layout of sph_accelerations.cpp:58-64); at most
// that many ids are written, the true count goes to count_out[i], and
// flags[0] is raised when a count exceeds its slot; flags[1] and flags[2]
// count the walks that read node n+1 and that took it as a neighbour.  With
// clamp set, count_out is clamped to the slot (for callers that own the
// offsets).  flags holds at least 3 ints, zeroed by the caller.
template <class T>
__global__ __launch_bounds__(kBlock) void k_find_ngb(const V4<T>* __restrict__ pos, const V4<T>* __restrict__ tree,
                                                      KV* __restrict__ hash, const int* __restrict__ offset,
                                                      const int* cap_count, int* count_out,
                                                      int* __restrict__ list, int* __restrict__ flags,
                                                      int n_dev, WalkParams wp, int clamp) {
  walk_priority();
  extern __shared__ int s_stack[];   // walk_stack_depth(n_nodes) entries per lane
  const int t = xcd_block(blockIdx.x, gridDim.x) * (int)blockDim.x + threadIdx.x;
  if (t >= n_dev) return;
  const int i = (int)hash[t].value;
  const V4<T> q = pos[i];
  const int off = offset[i];
  const int cap = n_norm(cap_count[i], 4);
  // ids go out four at a time as one 16-byte store when the slot run is
  // 16-byte aligned (offsets are multiples of 4 in the reference layout): a
  // wavefront's lanes write 64 different lists, so a 4-byte store per id
  // would dirty a cache line per lane per id
  const bool vec = (off & 3) == 0 && (reinterpret_cast<uintptr_t>(list) & 15) == 0;
  int c = 0, ph = 0;
  int b0 = 0, b1 = 0, b2 = 0;   // the current group's first three ids
  const int nw = kd_walk<T>(q, tree, wp.n_nodes, wp.fix_phantom, s_stack + threadIdx.x, [&](T, int id) {
    const int r = c & 3;
    if (vec) {
      if (r == 3) {
        if (c < cap) *reinterpret_cast<int4*>(list + off + c - 3) = make_int4(b0, b1, b2, id);
      } else {
        b0 = r == 0 ? id : b0;
        b1 = r == 1 ? id : b1;
        b2 = r == 2 ? id : b2;
      }
    } else if (c < cap) {
      list[off + c] = id;
    }
    ++c;
  }, &ph, wp.boxes);
  if (vec) {   // the last, partial group (slots past the count stay untouched)
    const int g = c & ~3, r = c & 3;
    if (r > 0 && g < cap) {
      list[off + g] = b0;
      if (r > 1) list[off + g + 1] = b1;
      if (r > 2) list[off + g + 2] = b2;
    }
  }
  hash[t].key = (uint32_t)nw;
  if (ph & 1) atomicAdd(flags + 1, 1);   // walks that read node n+1 (rare: a few per tree)
  if (ph & 2) atomicAdd(flags + 2, 1);   // ... and took it as a neighbour
  if (c > cap) {
    atomicOr(flags, 1);
    if (clamp) c = cap;
  }
  count_out[i] = c;
}

// ---------------------------------------------------------------------------
// compensated (two-sum) accumulator for the mixed form
struct KSum {
  float s = 0.f, c = 0.f;   // the compensated value is s - c
  __device__ __forceinline__ void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ float get() const { return s; }
  __device__ __forceinline__ float value() const { return s - c; }
};

// lanes per particle in the cooperative mixed-form kernels
constexpr int kGroup = 8;

// sum of the group's KSums into every lane of the group: a butterfly over
// the kGroup lanes, each step an exact two-sum of the partial values with the
// rounding error carried into the compensation
__device__ __forceinline__ void group_sum(KSum& a) {
  for (int o = kGroup / 2; o > 0; o >>= 1) {
    const float s2 = __shfl_xor(a.s, o, kGroup), c2 = __shfl_xor(a.c, o, kGroup);
    const float t = a.s + s2;
    const float bp = t - a.s;
    const float err = (a.s - (t - bp)) + (s2 - bp);   // a.s + s2 = t + err exactly
    a.c = (a.c + c2) - err;
    a.s = t;
  }
}
__device__ __forceinline__ float group_max(float v) {
  for (int o = kGroup / 2; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kGroup));
  return v;
}
// plain running sum in list order (the reference's accumulation)
template <class T>
struct Sum {
  T s = T(0);
  __device__ __forceinline__ void add(T x) { s += x; }
  __device__ __forceinline__ void sub(T x) { s -= x; }
  __device__ __forceinline__ T get() const { return s; }
};

// density, grad-h terms, div v and |curl v| (dev_density_loop.cu:29-111);
// plain and FP64 forms
template <class T>
__global__ __launch_bounds__(kBlock) void k_density(const V4<T>* __restrict__ pos, const V4<T>* __restrict__ vel,
                                                     V4<T>* __restrict__ hydro,
                                                     V4<T>* __restrict__ dots, const KV* __restrict__ hash,
                                                     const int* __restrict__ offset, const int* __restrict__ count,
                                                     const int* __restrict__ list, int n_dev) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (t >= n_dev) return;
  const int i = (int)hash[t].value;
  const V4<T> bp = pos[i], bv = vel[i];
  Sum<T> rho, om, psi, den, divv, cx, cy, cz;
  const int off = offset[i], n = count[i];
  // software pipeline: the next neighbour's id and records are requested
  // before the current pair is computed (the sums keep list order)
  int jn = n > 0 ? list[off] : i;
  V4<T> pn = pos[jn], vn = vel[jn];
  for (int k = off; k < off + n; ++k) {
    const V4<T> pj = pn, vj = vn;
    if (k + 1 < off + n) {
      jn = list[k + 1];
      pn = pos[jn];
      vn = vel[jn];
    }
    const T dx = bp.x - pj.x, dy = bp.y - pj.y, dz = bp.z - pj.z;
    const T u = T(2) * dsqrt(dx * dx + dy * dy + dz * dz) / bp.w;
    rho.add(vj.w * kern_w<T>(u));
    om.add(vj.w * kern_dwdh<T>(u));
    psi.add(vj.w * kern_dphidh<T>(u));
    den.sub(kern_digdh<T>(u));
    const T dvx = bv.x - vj.x, dvy = bv.y - vj.y, dvz = bv.z - vj.z;
    const T mdw = bv.w * kern_dw<T>(u);
    divv.add(mdw * (dx * dvx + dy * dvy + dz * dvz));
    cx.add(mdw * (dy * dvz - dz * dvy));
    cy.add(mdw * (dz * dvx - dx * dvz));
    cz.add(mdw * (dx * dvy - dy * dvx));
  }
  const T h = T(0.5 * (double)bp.w);
  const T h3 = h * h * h, h5 = h * h * h * h * h;
  T o = om.get(), p = psi.get();
  const T d = den.get();
  if (d != T(0)) {
    o = T((double)o * (1.0 / (double)h3 / (double)d));
    p = T((double)p * (1.0 / (double)h / (double)d));
  } else {
    o = p = T(0);
  }
  const T r = T((double)rho.get() * (1.0 / (double)h3));
  const T dv = T((double)divv.get() * (1.0 / (double)r / (double)h5));
  const T x = cx.get(), y = cy.get(), z = cz.get();
  const T curl = dsqrt(x * x + y * y + z * z) / r / h5;
  hydro[i].x = r;
  dots[i] = {dv, curl, o, p};
}

// dev_compute_pressure.cu:99-111 (every record, no hash, as the reference)
template <class T>
__global__ __launch_bounds__(kBlock) void k_eos(V4<T>* __restrict__ hydro, int n_dev, EosConst c) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n_dev) return;
  hydro[i].y = eos_record<T>(hydro[i], c);
}

// mixed form: P from the FP64 internal energy with the FP64 Newton root, and
// the per-particle factors the viscosity needs (dev_art_visc.cu:12-26), once
// per particle instead of once per pair: hydro2 = {v_sig (hydro pass),
// sound speed c = sqrt(5/3 P/rho), P/rho^2, Balsara |div v| / (|div v| +
// |curl v| + 1e-4 c/h)}
__global__ __launch_bounds__(kBlock) void k_eos_mixed(F4* __restrict__ hydro, const double* __restrict__ u64,
                                                       const F4* __restrict__ pos, const F4* __restrict__ dots,
                                                       F4* __restrict__ hydro2, int n, EosConst c) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const F4 hd = hydro[i];
  const double P = eos_pressure_fp64((double)hd.x, u64[i], (double)hd.w, c);
  hydro[i].y = (float)P;
  const double rho = hd.x;
  const double cs = sqrt(5.0 / 3.0 * P / rho);
  const F4 dd = dots[i];
  const double ad = fabs((double)dd.x);
  const double bal = ad / ((ad + (double)dd.y) + 0.0001 * cs / (0.5 * (double)pos[i].w));
  F4 h2 = hydro2[i];
  h2.y = (float)cs;
  h2.z = (float)(P / (rho * rho));
  h2.w = (float)bal;
  hydro2[i] = h2;
}

// hydro accelerations, du/dt and max signal speed (dev_hydro_dots.cu:25-113);
// writes the per-pair reaction coefficient m_i*acc to ngb_acc[k]; plain and
// FP64 forms
template <class T>
__global__ __launch_bounds__(kBlock) void k_hydro(const V4<T>* __restrict__ pos, const V4<T>* __restrict__ vel,
                                                   const V4<T>* __restrict__ hydro,
                                                   const V4<T>* __restrict__ dots, V4<T>* __restrict__ bdots,
                                                   V4<T>* __restrict__ hydro2, T* __restrict__ ngb_acc,
                                                   const KV* __restrict__ hash, const int* __restrict__ offset,
                                                   const int* __restrict__ count, const int* __restrict__ list,
                                                   int n_dev, T alpha, T beta) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (t >= n_dev) return;
  const int i = (int)hash[t].value;
  const V4<T> bp = pos[i], bv = vel[i], hd = hydro[i], dd = dots[i];
  const T prho2 = hd.y / (hd.x * hd.x);
  const double hw = 0.5 * (double)bp.w;
  // POW5(0.5*range) and POW2(0.5*range) expand left to right in the reference
  const double pow5 = hw * 0.5 * (double)bp.w * 0.5 * (double)bp.w * 0.5 * (double)bp.w * 0.5 * (double)bp.w;
  const double pow2 = hw * 0.5 * (double)bp.w;
  Sum<T> ax, ay, az, ud;
  T vsig = T(0);
  const int off = offset[i], n = count[i];
  // software pipeline, as in k_density
  int jn = n > 0 ? list[off] : i;
  V4<T> pn = pos[jn], vn = vel[jn], hn = hydro[jn], dn = dots[jn];
  for (int k = off; k < off + n; ++k) {
    const V4<T> pj = pn, vj = vn, hj = hn, dj = dn;
    if (k + 1 < off + n) {
      jn = list[k + 1];
      pn = pos[jn];
      vn = vel[jn];
      hn = hydro[jn];
      dn = dots[jn];
    }
    const T dx = bp.x - pj.x, dy = bp.y - pj.y, dz = bp.z - pj.z;
    const T dvx = bv.x - vj.x, dvy = bv.y - vj.y, dvz = bv.z - vj.z;
    const T ds = dsqrt(dx * dx + dy * dy + dz * dz);
    const T u = T((double)ds / hw);
    const T drdv = dx * dvx + dy * dvy + dz * dvz;
    const Visc<T> av = art_visc<T>(bp.w, pj.w, ds, drdv, hd, dd, hj, dj, alpha, beta);
    vsig = dmax(vsig, av.vsig);
    const double w1 = (double)kern_dw<T>(u);
    const T g = kern_dig<T>(u);
    const double pv = (double)prho2 + 0.5 * (double)av.pi;
    T acc = T(pv * w1 / pow5);
    acc = T((double)acc - ((double)(prho2 * dd.z) + 0.5 * (double)dd.w) / (double)vj.w * (double)g / pow2);
    ax.sub(vj.w * acc * dx);
    ay.sub(vj.w * acc * dy);
    az.sub(vj.w * acc * dz);
    ngb_acc[k] = bv.w * acc;
    T udot = T((double)vj.w * pv * w1 / pow5);
    udot = T((double)udot - (double)(prho2 * dd.z * g) / pow2);
    ud.add(udot * drdv);
  }
  bdots[i] = {ax.get(), ay.get(), az.get(), ud.get()};
  hydro2[i].x = vsig;
}

// ---------------------------------------------------------------------------
// Mixed form: FP32 pair arithmetic with FMA polynomials (f32k), hardware
// square root and reciprocals, split hi/lo position differences and Kahan
// sums.  The per-particle quantities of the viscosity come precomputed from
// k_eos_mixed.  Same sums as k_density / k_hydro, same list order.

// density, grad-h terms, div v, |curl v| (dev_density_loop.cu:29-111); a
// group of kGroup lanes per particle, each lane taking every kGroup-th
// neighbour, partial sums combined by group_sum
__global__ __launch_bounds__(kBlock) void k_density_mixed(const F4* __restrict__ pos, const F4* __restrict__ vel,
                                                          const F4* __restrict__ pos_lo, F4* __restrict__ hydro,
                                                          F4* __restrict__ dots, const KV* __restrict__ hash,
                                                          const int* __restrict__ offset,
                                                          const int* __restrict__ count,
                                                          const int* __restrict__ list, int n_dev) {
  const int g = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const int t = g / kGroup, lane = g % kGroup;
  if (t >= n_dev) return;   // whole groups leave together
  const int i = (int)hash[t].value;
  const F4 bp = pos[i], bv = vel[i], loi = pos_lo[i];
  const int off = offset[i], n = count[i];
  const float two_over_r = n > 0 ? 2.f / bp.w : 0.f;
  KSum rho, om, psi, den, divv, cx, cy, cz;
  for (int k = off + lane; k < off + n; k += kGroup) {
    const int j = list[k];
    const F4 pj = pos[j], vj = vel[j], lj = pos_lo[j];
    const float dx = (bp.x - pj.x) + (loi.x - lj.x);
    const float dy = (bp.y - pj.y) + (loi.y - lj.y);
    const float dz = (bp.z - pj.z) + (loi.z - lj.z);
    const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    const float u = __builtin_amdgcn_sqrtf(d2) * two_over_r;
    const float iu = __builtin_amdgcn_rcpf(u);
    rho.add(vj.w * f32k::w(u));
    om.add(vj.w * f32k::dwdh(u));
    psi.add(vj.w * f32k::dphidh(u));
    den.add(-f32k::digdh(u));
    const float dvx = bv.x - vj.x, dvy = bv.y - vj.y, dvz = bv.z - vj.z;
    const float mdw = bv.w * f32k::dw(u, iu);
    divv.add(mdw * __builtin_fmaf(dx, dvx, __builtin_fmaf(dy, dvy, dz * dvz)));
    cx.add(mdw * __builtin_fmaf(dy, dvz, -dz * dvy));
    cy.add(mdw * __builtin_fmaf(dz, dvx, -dx * dvz));
    cz.add(mdw * __builtin_fmaf(dx, dvy, -dy * dvx));
  }
  group_sum(rho);
  group_sum(om);
  group_sum(psi);
  group_sum(den);
  group_sum(divv);
  group_sum(cx);
  group_sum(cy);
  group_sum(cz);
  if (lane != 0) return;
  if (n == 0) {
    hydro[i].x = 0.f;
    dots[i] = {0.f, 0.f, 0.f, 0.f};
    return;
  }
  const double h = 0.5 * (double)bp.w, h3 = h * h * h, h5 = h3 * h * h;
  const double d = den.value();
  double o = 0.0, ps = 0.0;
  if (d != 0.0) {
    o = om.value() / (h3 * d);
    ps = psi.value() / (h * d);
  }
  const double r = rho.value() / h3;
  const double x = cx.value(), y = cy.value(), z = cz.value();
  hydro[i].x = (float)r;
  dots[i] = {(float)(divv.value() / (r * h5)), (float)(sqrt(x * x + y * y + z * z) / (r * h5)), (float)o,
             (float)ps};
}

// hydro accelerations, du/dt, v_sig and the reaction coefficients
// (dev_hydro_dots.cu:25-113, dev_art_visc.cu:8-32); a group of kGroup lanes
// per particle, as in k_density_mixed
__global__ __launch_bounds__(kBlock) void k_hydro_mixed(const F4* __restrict__ pos, const F4* __restrict__ vel,
                                                        const F4* __restrict__ pos_lo, const F4* __restrict__ hydro,
                                                        const F4* __restrict__ dots, F4* __restrict__ bdots,
                                                        F4* __restrict__ hydro2, float* __restrict__ ngb_acc,
                                                        const KV* __restrict__ hash, const int* __restrict__ offset,
                                                        const int* __restrict__ count, const int* __restrict__ list,
                                                        int n_dev, float alpha, float beta) {
  const int gl = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const int t = gl / kGroup, lane = gl % kGroup;
  if (t >= n_dev) return;   // whole groups leave together
  const int i = (int)hash[t].value;
  const int off = offset[i], n = count[i];
  const F4 bp = pos[i], bv = vel[i], loi = pos_lo[i], dd = dots[i], qi = hydro2[i];
  const float ci = qi.y, prho2 = qi.z, bal_i = qi.w;
  const float inv_h = n > 0 ? 2.f / bp.w : 0.f, inv_h2 = inv_h * inv_h, inv_h5 = inv_h2 * inv_h2 * inv_h;
  const float k1 = (prho2 * dd.z + 0.5f * dd.w) * inv_h2;   // (P/rho^2 Omega + Psi/2) / h^2
  const float k2 = prho2 * dd.z * inv_h2;                   // P/rho^2 Omega / h^2
  KSum ax, ay, az, ud;
  float vsig = 0.f;
  for (int k = off + lane; k < off + n; k += kGroup) {
    const int j = list[k];
    const F4 pj = pos[j], vj = vel[j], lj = pos_lo[j], qj = hydro2[j];
    const float dx = (bp.x - pj.x) + (loi.x - lj.x);
    const float dy = (bp.y - pj.y) + (loi.y - lj.y);
    const float dz = (bp.z - pj.z) + (loi.z - lj.z);
    const float dvx = bv.x - vj.x, dvy = bv.y - vj.y, dvz = bv.z - vj.z;
    const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    const float ds = __builtin_amdgcn_sqrtf(d2);
    const float u = ds * inv_h;
    const float iu = __builtin_amdgcn_rcpf(u);
    const float drdv = __builtin_fmaf(dx, dvx, __builtin_fmaf(dy, dvy, dz * dvz));
    const float cs = ci + qj.y;
    float pi_ = 0.f, vs = cs;
    if (drdv < 0.f) {
      const float r_ds = __builtin_amdgcn_rcpf(ds);
      vs = __builtin_fmaf(-beta * drdv, r_ds, cs);
      const float mu = drdv * r_ds * __builtin_amdgcn_rcpf(cs) * (bal_i + qj.w);
      pi_ = (prho2 + qj.z) * mu * __builtin_fmaf(beta, mu, -alpha);
    }
    vsig = fmaxf(vsig, vs);
    const float w1 = f32k::dw(u, iu), gk = f32k::dig(u, iu);
    const float t1 = __builtin_fmaf(0.5f, pi_, prho2) * w1 * inv_h5;
    const float mjA = __builtin_fmaf(vj.w, t1, -k1 * gk);      // m_j * acc
    ax.add(-mjA * dx);
    ay.add(-mjA * dy);
    az.add(-mjA * dz);
    ngb_acc[k] = bv.w * mjA * __builtin_amdgcn_rcpf(vj.w);   // m_i * acc
    ud.add(__builtin_fmaf(vj.w, t1, -k2 * gk) * drdv);
  }
  group_sum(ax);
  group_sum(ay);
  group_sum(az);
  group_sum(ud);
  vsig = group_max(vsig);
  if (lane != 0) return;
  bdots[i] = {ax.value(), ay.value(), az.value(), ud.value()};
  hydro2[i].x = vsig;
}

// ---------------------------------------------------------------------------
// launchers

// lanes per block of the tree walks: 256, or 64 when 256-lane blocks would
// leave CUs empty (below 256 blocks), so that the walks' wavefronts spread
// over every CU instead of sharing the memory pipelines of a quarter of them
static int walk_block(int n_dev) { return n_dev < 256 * kBlock ? 64 : kBlock; }

// the candidate cache where the walks run one wavefront per CU
bool solve_cache_pays(int n_dev) { return walk_block(n_dev) == 64; }

template <class T>
void launch_solve_range(V4<T>* pos, int* n_ngb, const KV* hash, const V4<T>* tree, int n_dev,
                        const WalkParams& wp, hipStream_t s) {
  const int bs = walk_block(n_dev);
  // the cache's path bounds take three more LDS columns; past 64 KB a block
  // walks without it
  const size_t stack = (size_t)walk_stack_depth(wp.n_nodes) * bs;
  const bool cache = wp.solve_cache && stack * (sizeof(int) + 3 * sizeof(T)) <= 65536;
  const size_t lds = stack * (sizeof(int) + (cache ? 3 * sizeof(T) : 0));
  if (cache)
    hipLaunchKernelGGL((k_solve_range<T, true>), dim3(ceil_div(n_dev, bs)), dim3(bs), lds, s, pos, n_ngb, hash, tree,
                       n_dev, wp);
  else
    hipLaunchKernelGGL((k_solve_range<T, false>), dim3(ceil_div(n_dev, bs)), dim3(bs), lds, s, pos, n_ngb, hash,
                       tree, n_dev, wp);
  CS_CHECK_LAUNCH();
}

template <class T>
void launch_find_ngb(const V4<T>* pos, const V4<T>* tree, KV* hash, const int* offset, const int* cap_count,
                     int* count_out, int* list, int* flags, int n_dev, const WalkParams& wp, int clamp,
                     hipStream_t s) {
  const int bs = walk_block(n_dev);
  const size_t lds = (size_t)walk_stack_depth(wp.n_nodes) * bs * sizeof(int);
  hipLaunchKernelGGL(k_find_ngb<T>, dim3(ceil_div(n_dev, bs)), dim3(bs), lds, s, pos, tree, hash, offset, cap_count,
                     count_out, list, flags, n_dev, wp, clamp);
  CS_CHECK_LAUNCH();
}

template <class T>
void launch_density(const V4<T>* pos, const V4<T>* vel, V4<T>* hydro, V4<T>* dots, const KV* hash,
                    const int* offset, const int* count, const int* list, int n_dev, hipStream_t s) {
  hipLaunchKernelGGL(k_density<T>, dim3(ceil_div(n_dev, kBlock)), dim3(kBlock), 0, s, pos, vel, hydro, dots, hash,
                     offset, count, list, n_dev);
  CS_CHECK_LAUNCH();
}

template <class T>
void launch_eos(V4<T>* hydro, int n_dev, const EosConst& c, hipStream_t s) {
  hipLaunchKernelGGL(k_eos<T>, dim3(ceil_div(n_dev, kBlock)), dim3(kBlock), 0, s, hydro, n_dev, c);
  CS_CHECK_LAUNCH();
}

void launch_eos_mixed(F4* hydro, const double* u64, const F4* pos, const F4* dots, F4* hydro2, int n,
                      const EosConst& c, hipStream_t s) {
  hipLaunchKernelGGL(k_eos_mixed, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, s, hydro, u64, pos, dots, hydro2, n, c);
  CS_CHECK_LAUNCH();
}

void launch_density_mixed(const F4* pos, const F4* vel, const F4* pos_lo, F4* hydro, F4* dots, const KV* hash,
                          const int* offset, const int* count, const int* list, int n_dev, hipStream_t s) {
  hipLaunchKernelGGL(k_density_mixed, dim3(ceil_div(n_dev * kGroup, kBlock)), dim3(kBlock), 0, s, pos, vel, pos_lo,
                     hydro, dots, hash, offset, count, list, n_dev);
  CS_CHECK_LAUNCH();
}

void launch_hydro_mixed(const F4* pos, const F4* vel, const F4* pos_lo, const F4* hydro, const F4* dots, F4* bdots,
                        F4* hydro2, float* ngb_acc, const KV* hash, const int* offset, const int* count,
                        const int* list, int n_dev, float alpha, float beta, hipStream_t s) {
  hipLaunchKernelGGL(k_hydro_mixed, dim3(ceil_div(n_dev * kGroup, kBlock)), dim3(kBlock), 0, s, pos, vel, pos_lo,
                     hydro, dots, bdots, hydro2, ngb_acc, hash, offset, count, list, n_dev, alpha, beta);
  CS_CHECK_LAUNCH();
}

template <class T>
void launch_hydro(const V4<T>* pos, const V4<T>* vel, const V4<T>* hydro, const V4<T>* dots, V4<T>* bdots,
                  V4<T>* hydro2, T* ngb_acc, const KV* hash, const int* offset, const int* count, const int* list,
                  int n_dev, T alpha, T beta, hipStream_t s) {
  hipLaunchKernelGGL(k_hydro<T>, dim3(ceil_div(n_dev, kBlock)), dim3(kBlock), 0, s, pos, vel, hydro, dots, bdots,
                     hydro2, ngb_acc, hash, offset, count, list, n_dev, alpha, beta);
  CS_CHECK_LAUNCH();
}

#define CS_INST(T)                                                                                        \
  template void launch_solve_range<T>(V4<T>*, int*, const KV*, const V4<T>*, int, const WalkParams&,      \
                                      hipStream_t);                                                       \
  template void launch_find_ngb<T>(const V4<T>*, const V4<T>*, KV*, const int*, const int*, int*, int*,   \
                                   int*, int, const WalkParams&, int, hipStream_t);                       \
  template void launch_eos<T>(V4<T>*, int, const EosConst&, hipStream_t);
CS_INST(float)
CS_INST(double)
#undef CS_INST
template void launch_density<float>(const F4*, const F4*, F4*, F4*, const KV*, const int*, const int*, const int*,
                                    int, hipStream_t);
template void launch_density<double>(const D4*, const D4*, D4*, D4*, const KV*, const int*, const int*, const int*,
                                     int, hipStream_t);
template void launch_hydro<float>(const F4*, const F4*, const F4*, const F4*, F4*, F4*, float*, const KV*, const int*,
                                  const int*, const int*, int, float, float, hipStream_t);
template void launch_hydro<double>(const D4*, const D4*, const D4*, const D4*, D4*, D4*, double*, const KV*,
                                   const int*, const int*, const int*, int, double, double, hipStream_t);

}  // namespace cudasph
