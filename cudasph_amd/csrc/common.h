// Shared types and helpers for the HIP kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "types.h"

namespace cudasph {

constexpr int kBlock = 256;        // threads per block for per-particle kernels
// kd-walk stack entries per lane.  Popping a node at depth d leaves at most
// one pending sibling per level 1..d on the stack and pushes at most two
// children, so a tree of H levels (H = floor(log2 n) + 1) needs at most H
// entries; one more is kept spare.  Sizing the LDS stack to the tree
// (22 entries at n = 2^20 instead of a fixed 28) lets more blocks share a CU.
__host__ __device__ __forceinline__ int walk_stack_depth(int n_nodes) {
  int h = 1;
  while (h < 31 && (n_nodes >> h) > 0) ++h;   // h = floor(log2 n) + 1 for n >= 1
  return h + 1;
}

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

[[noreturn]] inline void throw_hip(const char* file, int line, const char* expr, hipError_t e) {
  char buf[512];
  std::snprintf(buf, sizeof buf, "%s:%d: %s failed: %s", file, line, expr, hipGetErrorString(e));
  throw HipError(buf);
}

#define CS_HIP(expr)                                                       \
  do {                                                                     \
    hipError_t cs_e_ = (expr);                                             \
    if (cs_e_ != hipSuccess) cudasph::throw_hip(__FILE__, __LINE__, #expr, cs_e_); \
  } while (0)

#define CS_CHECK_LAUNCH() CS_HIP(hipGetLastError())

// kd-tree node word, device side (see types.h)
__device__ __forceinline__ int node_bits(float w) { return __float_as_int(w); }
__device__ __forceinline__ int node_bits(double w) { return (int)w; }

// Workgroups are handed to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8, each XCD with its own L2).  xcd_block renumbers them so that XCD p
// takes the p-th contiguous share of the grid: particles that are neighbours
// in memory (Peano-Hilbert) order then run on the same XCD, and their
// neighbour gathers and tree walks share its L2 instead of fetching the same
// lines into all eight.  A bijection on [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  constexpr int kXcd = 8;
  const int q = nb / kXcd, r = nb % kXcd, x = b % kXcd;
  return x * q + (x < r ? x : r) + b / kXcd;
}

// n_norm (CUDAsph.h:16-20): round up to a multiple of j, at least j
__host__ __device__ __forceinline__ int n_norm(int n, int j) {
  n = ((n - 1) / j) * j + j;
  if (n == 0) n = j;
  return n;
}

}  // namespace cudasph
