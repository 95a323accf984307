// Host kd-tree builder (kd_tree.cpp:4-128, packing sph_accelerations.cpp:21-33).
//
// The tree is the reference's left-balanced implicit tree: node p holds one
// particle, its children are 2p and 2p+1, the split axis is depth % 3, and
// each node's particle is the element at the left-subtree size after a
// first-element-pivot quicksort of its subarray.  The quicksort's handling of
// equal coordinates decides which of two tied particles becomes a node, so
// the same quicksort is run here; the output is bit-identical to the
// reference's.  Disjoint subtrees are built concurrently by a small pool of
// host threads.
#include "kdtree.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

namespace cudasph {
namespace {

template <class T>
struct Item {
  int id;
  T c[3];
};

// size of the left subtree of a left-balanced tree of n nodes (kd_tree.cpp:63-79)
inline int left_size(int n) {
  int m = 1;
  while (m <= n) m <<= 1;
  m >>= 1;
  const int last = n - (m - 1);
  return last <= m / 2 ? (m - 2) / 2 + last : (m - 2) / 2 + m / 2;
}

constexpr int kParallelGrain = 4096;   // elements below which work stays on the calling thread

template <class T>
struct Builder {
  V4<T>* tree;
  std::atomic<int> spare_threads;

  bool take_thread() {
    int v = spare_threads.load();
    while (v > 0)
      if (spare_threads.compare_exchange_weak(v, v - 1)) return true;
    return false;
  }

  // First-element-pivot quicksort, kd_tree.cpp:4-31.  Each partition step is
  // the reference's; the two resulting ranges are disjoint, so sorting them
  // in either order, or concurrently, gives the same array.  The smaller
  // range is recursed into (on a spare thread when it is large) and the
  // larger one is iterated on, which keeps the stack logarithmic.
  void quicksort(Item<T>* a, int axis, int left, int right) {
    std::vector<std::thread> spawned;
    while (left < right) {
      const int lo = left, hi = right;
      const Item<T> pivot = a[left];
      const T key = pivot.c[axis];
      while (left < right) {
        while (a[right].c[axis] >= key && left < right) --right;
        if (left != right) a[left++] = a[right];
        while (a[left].c[axis] <= key && left < right) ++left;
        if (left != right) a[right--] = a[left];
      }
      a[left] = pivot;
      const int mid = left;
      int s_lo, s_hi;   // smaller side
      if (mid - lo < hi - mid) {
        s_lo = lo; s_hi = mid - 1; left = mid + 1; right = hi;
      } else {
        s_lo = mid + 1; s_hi = hi; left = lo; right = mid - 1;
      }
      if (s_lo < s_hi) {
        if (s_hi - s_lo > kParallelGrain && take_thread())
          spawned.emplace_back([=] { quicksort(a, axis, s_lo, s_hi); spare_threads.fetch_add(1); });
        else
          quicksort(a, axis, s_lo, s_hi);
      }
    }
    for (auto& t : spawned) t.join();
  }

  void build(int node, Item<T>* a, int depth, int n) {
    const int axis = depth % 3;
    quicksort(a, axis, 0, n - 1);
    const int lt = left_size(n);
    const Item<T>& e = a[lt];
    tree[node] = {e.c[0], e.c[1], e.c[2], host_node_word(T(0), e.id | (axis << 24))};
    const int nl = lt, nr = n - lt - 1;
    std::thread worker;
    const bool spawned = nl > kParallelGrain && nr > 0 && take_thread();
    if (spawned) worker = std::thread([=] { build(2 * node, a, depth + 1, nl); });
    else if (nl > 0) build(2 * node, a, depth + 1, nl);
    if (nr > 0) build(2 * node + 1, a + lt + 1, depth + 1, nr);
    if (spawned) {
      worker.join();
      spare_threads.fetch_add(1);
    }
  }
};

}  // namespace

template <class T>
void build_kd_tree(const V4<T>* pos, int n, int n_dev, int n_threads, V4<T>* tree, KV* hash) {
  std::vector<Item<T>> items(n);
  for (int i = 0; i < n; ++i) items[i] = {i, {pos[i].x, pos[i].y, pos[i].z}};
  // slots past the last node: far sentinels with split 3 (read_state.cpp:85-88);
  // slot n_dev+1 is zero, the value of the reference's out-of-range fetch
  const T far = T(1e10);
  for (int i = 0; i < n_dev + 2; ++i) tree[i] = {far, far, far, host_node_word(T(0), -1)};
  tree[0] = {T(0), T(0), T(0), T(0)};
  tree[n_dev + 1] = {T(0), T(0), T(0), T(0)};
  if (n > 0) {
    Builder<T> b;
    b.tree = tree;
    int hw = n_threads > 0 ? n_threads : (int)std::thread::hardware_concurrency();
    b.spare_threads = std::max(0, hw - 1);
    b.build(1, items.data(), 0, n);
  }
  // particle_map (kd_tree.cpp:117-128): heap order of the nodes; padding
  // particles map to themselves
  for (int t = 0; t < n; ++t) {
    const int bits = node_word_bits(tree[t + 1].w);
    hash[t] = {(uint32_t)t, (uint32_t)(bits & ~(3 << 24))};
  }
  for (int t = n; t < n_dev; ++t) hash[t] = {1u << 24, (uint32_t)t};
}

template void build_kd_tree<float>(const F4*, int, int, int, F4*, KV*);
template void build_kd_tree<double>(const D4*, int, int, int, D4*, KV*);

}  // namespace cudasph
