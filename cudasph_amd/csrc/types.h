// Plain C++ record types shared by the host code (no HIP headers needed).
#pragma once

#include <cstdint>
#include <cstring>

namespace cudasph {

// 4-component record, 16 B for float and 32 B for double, aligned so that a
// record is one vector load on the GPU.
template <class T>
struct alignas(4 * sizeof(T)) V4 {
  T x, y, z, w;
};
using F4 = V4<float>;
using D4 = V4<double>;

// KeyValuePair (radixsort.cuh:48-57)
struct alignas(8) KV {
  uint32_t key, value;
};

// kd-tree node word: the int bits (body_id | split << 24) stored in w
// (sph_accelerations.cpp:30-31).  The FP64 form stores the same int as a
// double value.
inline float host_node_word(float, int bits) {
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}
inline double host_node_word(double, int bits) { return (double)bits; }
inline int node_word_bits(float w) {
  int b;
  std::memcpy(&b, &w, 4);
  return b;
}
inline int node_word_bits(double w) { return (int)w; }

}  // namespace cudasph
