// The reference's C boundary (CUDAsph.h:22-48, host_calls.cu:27-376,
// radixsort.cuh:59-61) over this library's kernels.
//
// Same names, argument order, units and call-order contract: host_solve_range
// binds the arrays that host_sph_accelerations then uses, as the reference's
// texture bindings do (host_calls.cu:107-146).  Errors print and exit, as
// CUDA_SAFE_CALL does (host_calls.cu:14-22).
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "../../include/cudasph.h"
#include "common.h"
#include "kernels.h"
#include "sph_math.h"

using namespace cudasph;

namespace {

#define ABI_CALL(expr)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      std::fprintf(stderr, "\nerror: " #expr " failed with error %s\n", hipGetErrorName(e_)); \
      std::exit(1);                                                                     \
    }                                                                                   \
  } while (0)

double wall() {
  timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + 1e-6 * t.tv_usec;
}

// compile-time constants of the reference, settable through sph_abi_set_params
struct AbiState {
  sph_params p;
  int exact_rsqrt = 0;
  // bindings made by host_solve_range (the reference's textures)
  F4* pos = nullptr;
  F4* vel = nullptr;
  int* n_ngb = nullptr;
  int n_nodes = 0;
  int n_dev = 0;
  F4* tree = nullptr;     // private copy with the zero record at n_dev+1
  size_t tree_cap = 0;
  int* flags = nullptr;
  AbiState() { sph_params_default(&p, 1); }
};
AbiState& abi() {
  static AbiState s;
  return s;
}

template <class F>
void guarded(F&& f) {
  try {
    f();
  } catch (const std::exception& e) {
    std::fprintf(stderr, "\nerror: %s\n", e.what());
    std::exit(1);
  }
}

WalkParams walk(const AbiState& a, int n_nodes, float rel_err) {
  return {n_nodes, a.p.fix_phantom, rel_err, a.p.desired_ngb, a.p.max_iter};
}

}  // namespace

extern "C" {

// options with no counterpart in the reference: its DESIRED_NGB, MAX_ITER,
// EOS unit constants and softening are compile-time constants
void sph_abi_set_params(const sph_params* p, int exact_rsqrt) {
  abi().p = *p;
  abi().exact_rsqrt = exact_rsqrt;
}

void initCUDA(void) {}

int cuda_get_device(void) {
  int d = 0;
  ABI_CALL(hipGetDevice(&d));
  return d;
}
void cuda_set_device0(int device_id) {
  if (device_id > 0) ABI_CALL(hipSetDevice(device_id - 1));
}
void cuda_set_device(int) {}
int get_device_count(void) {
  int n = 0;
  ABI_CALL(hipGetDeviceCount(&n));
  return n;
}

void allocateCUDAarray(void** pos, int n, int) { ABI_CALL(hipMalloc(pos, (size_t)n)); }
void deleteCUDAarray(void* pos, int) { ABI_CALL(hipFree(pos)); }
void copyArrayToDevice(void* device, const void* host, int n, int) {
  ABI_CALL(hipMemcpy(device, host, (size_t)n, hipMemcpyHostToDevice));
}
void copyArrayFromDevice(void* host, const void* device, int n, int) {
  ABI_CALL(hipMemcpy(host, device, (size_t)n, hipMemcpyDeviceToHost));
}
void zallocateNBodyArrays(float** pos, int numBodies, int) {
  ABI_CALL(hipMalloc((void**)pos, sizeof(float) * 4 * (size_t)numBodies));
}
void zdeleteNBodyArrays(float* pos, int) { ABI_CALL(hipFree(pos)); }
void zcopyArrayFromDevice(float* host, const float* device, int numBodies, int) {
  ABI_CALL(hipMemcpy(host, device, (size_t)numBodies * 4 * sizeof(float), hipMemcpyDeviceToHost));
}
void zcopyArrayToDevice(float* device, const float* host, int numBodies, int) {
  ABI_CALL(hipMemcpy(device, host, (size_t)numBodies * 4 * sizeof(float), hipMemcpyHostToDevice));
}

// host_calls.cu:319-361; the reference softens with the M4 spline at range
void compute_CUDA_gravity(int numBodies, sph_float4* accel, float* pos, float* vel, int p, int q, int) {
  if (p <= 0 || numBodies % p != 0 || q != 1) {
    std::fprintf(stderr, "compute_CUDA_gravity: need q == 1 and p dividing numBodies\n");
    std::exit(1);
  }
  AbiState& a = abi();
  guarded([&] {
    launch_gravity<float>((const F4*)pos, (const F4*)vel, (F4*)accel, numBodies, numBodies, a.p.grav_mode,
                          a.p.grav_eps, a.exact_rsqrt, nullptr);
    CS_HIP(hipDeviceSynchronize());
  });
}

// host_calls.cu:77-153
float host_solve_range(int n_bodies, int n_nodes, float rel_err, sph_kv* bodies_hash, int* /*bodies_map*/,
                       sph_float4* bodies_pos, sph_float4* bodies_vel, sph_float4* kd_tree, int* n_ngb, int) {
  const double t0 = wall();
  AbiState& a = abi();
  if (n_nodes < 0 || n_nodes > n_bodies || n_bodies >= (1 << 24) - 2) {
    std::fprintf(stderr, "host_solve_range: need 0 <= n_nodes <= n_bodies < 2^24\n");
    std::exit(1);
  }
  guarded([&] {
    const size_t need = (size_t)n_bodies + 2;
    if (a.tree_cap < need) {
      if (a.tree) CS_HIP(hipFree(a.tree));
      CS_HIP(hipMalloc(&a.tree, need * sizeof(F4)));
      a.tree_cap = need;
    }
    if (!a.flags) CS_HIP(hipMalloc(&a.flags, 16));
    // the caller's tree has n_bodies+1 records; the walk may read record
    // n_nodes+1, which the reference's texture returns as zeros
    CS_HIP(hipMemcpy(a.tree, kd_tree, ((size_t)n_bodies + 1) * sizeof(F4), hipMemcpyDeviceToDevice));
    CS_HIP(hipMemset(a.tree + n_bodies + 1, 0, sizeof(F4)));
    launch_solve_range<float>((F4*)bodies_pos, n_ngb, (const KV*)bodies_hash, a.tree, n_bodies,
                              walk(a, n_nodes, rel_err), nullptr);
    CS_HIP(hipDeviceSynchronize());
    a.pos = (F4*)bodies_pos;
    a.vel = (F4*)bodies_vel;
    a.n_ngb = n_ngb;
    a.n_nodes = n_nodes;
    a.n_dev = n_bodies;
    a.p.rel_err = rel_err;
  });
  return (float)(wall() - t0);
}

// host_calls.cu:155-287: find_ngb, density, EOS, hydro.  n_ngb is overwritten
// with the number of ids written per particle (see the header).
float host_sph_accelerations(int n_bodies, int /*ngb_tot*/, float av_alpha, float av_beta, sph_kv* bodies_hash,
                             sph_kv* /*bodies_hash_extra*/, int* /*bodies_map*/, int* ngb_list, int* ngb_offset,
                             sph_float4* hydro_data, sph_float4* dots_data, sph_float4* bodies_dots,
                             sph_float4* hydro_data2, float* ngb_acc, int) {
  const double t0 = wall();
  AbiState& a = abi();
  if (!a.pos || a.n_dev != n_bodies) {
    std::fprintf(stderr, "host_sph_accelerations: call host_solve_range first with the same n\n");
    std::exit(1);
  }
  guarded([&] {
    CS_HIP(hipMemset(a.flags, 0, 16));
    launch_find_ngb<float>(a.pos, a.tree, (KV*)bodies_hash, ngb_offset, a.n_ngb, a.n_ngb, ngb_list, a.flags,
                           n_bodies, walk(a, a.n_nodes, a.p.rel_err), 1, nullptr);
    int fl = 0;
    CS_HIP(hipMemcpy(&fl, a.flags, 4, hipMemcpyDeviceToHost));
    if (fl)
      std::fprintf(stderr,
                   "host_sph_accelerations: a neighbour list outgrew its slot; the list was truncated to the "
                   "slot (the reference overwrites the next slot)\n");
    EosConst c;
    c.density_unit = a.p.eos_density_unit;
    c.energy_unit = a.p.eos_energy_unit;
    c.pressure_unit = a.p.eos_pressure_unit;
    c.amu = a.p.eos_amu;
    c.boltzmann = a.p.eos_boltzmann;
    c.rad_const = a.p.eos_rad_const;
    c.gamma = a.p.eos_gamma;
    c.temp_base = a.p.eos_temp_base;
    c.newton = a.p.eos_newton;
    launch_density<float>(a.pos, a.vel, (F4*)hydro_data, (F4*)dots_data, (const KV*)bodies_hash,
                                 ngb_offset, a.n_ngb, ngb_list, n_bodies, nullptr);
    launch_eos<float>((F4*)hydro_data, n_bodies, c, nullptr);
    launch_hydro<float>(a.pos, a.vel, (const F4*)hydro_data, (const F4*)dots_data,
                               (F4*)bodies_dots, (F4*)hydro_data2, ngb_acc, (const KV*)bodies_hash, ngb_offset,
                               a.n_ngb, ngb_list, n_bodies, av_alpha, av_beta, nullptr);
    CS_HIP(hipDeviceSynchronize());
  });
  return (float)(wall() - t0);
}

// radixsort.cu:50-77
void RadixSort(sph_kv* pData0, sph_kv* pData1, unsigned int elements, unsigned int bits) {
  guarded([&] {
    int* scratch = nullptr;
    CS_HIP(hipMalloc(&scratch, (size_t)(radix_scratch_ints(elements) + 16) * sizeof(int)));
    launch_radix_sort((KV*)pData0, (KV*)pData1, elements, (int)bits, scratch, nullptr);
    CS_HIP(hipDeviceSynchronize());
    CS_HIP(hipFree(scratch));
  });
}

}  // extern "C"
