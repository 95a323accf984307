// Host kd-tree builder, reference-exact (see kdtree.cpp).
#pragma once

#include "types.h"

namespace cudasph {

// tree: n_dev+2 records (index 0 unused, 1..n nodes, sentinels after, zero
// record at n_dev+1).  hash[t] = {t, particle at heap node t+1} for t < n and
// {1<<24, t} for the padding.
template <class T>
void build_kd_tree(const V4<T>* pos, int n, int n_dev, int n_threads, V4<T>* tree, KV* hash);

}  // namespace cudasph
