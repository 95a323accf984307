// Engine C API (group 2 of include/cudasph.h): exception-safe wrappers over
// the engine, the reference state-file format, and the last-error string.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cudasph.h"
#include "engine.h"
#include "kdtree.h"
#include "kernels.h"

using namespace cudasph;

namespace {

thread_local std::string g_last_error;

template <class F>
auto guard(F&& f, decltype(f()) fail) -> decltype(f()) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_last_error = e.what();
  } catch (...) {
    g_last_error = "unknown error";
  }
  return fail;
}

EngineBase* E(void* e) {
  if (!e) throw std::invalid_argument("null engine");
  return static_cast<EngineBase*>(e);
}

// reference state file: header (read_state.cpp:5-23) + 23 doubles per body
// (read_state.cpp:34-67)
struct StateHeader {
  int n_bodies, nnopt;
  double hmin, hmax, sep0, t_end, dt_out;
  int snapshot_number, iteration_number;
  double global_time;
  int nav;
  double alpha, beta, eta2;
  int ngr, nrelax;
  double trelax, dt, omega2;
};

static_assert(sizeof(StateHeader) <= sizeof(EngineBase::state_header), "state header storage");

bool read_header(FILE* f, StateHeader& h) {
  bool ok = true;
  auto rd = [&](void* p, size_t n) { ok = ok && std::fread(p, n, 1, f) == 1; };
  rd(&h.n_bodies, 4); rd(&h.nnopt, 4);
  rd(&h.hmin, 8); rd(&h.hmax, 8); rd(&h.sep0, 8); rd(&h.t_end, 8); rd(&h.dt_out, 8);
  rd(&h.snapshot_number, 4); rd(&h.iteration_number, 4);
  rd(&h.global_time, 8);
  rd(&h.nav, 4);
  rd(&h.alpha, 8); rd(&h.beta, 8); rd(&h.eta2, 8);
  rd(&h.ngr, 4); rd(&h.nrelax, 4);
  rd(&h.trelax, 8); rd(&h.dt, 8); rd(&h.omega2, 8);
  return ok;
}
void write_header(FILE* f, const StateHeader& h) {
  auto wr = [&](const void* p, size_t n) { std::fwrite(p, n, 1, f); };
  wr(&h.n_bodies, 4); wr(&h.nnopt, 4);
  wr(&h.hmin, 8); wr(&h.hmax, 8); wr(&h.sep0, 8); wr(&h.t_end, 8); wr(&h.dt_out, 8);
  wr(&h.snapshot_number, 4); wr(&h.iteration_number, 4);
  wr(&h.global_time, 8);
  wr(&h.nav, 4);
  wr(&h.alpha, 8); wr(&h.beta, 8); wr(&h.eta2, 8);
  wr(&h.ngr, 4); wr(&h.nrelax, 4);
  wr(&h.trelax, 8); wr(&h.dt, 8); wr(&h.omega2, 8);
}

constexpr double kAmuCgs = 1.66054e-24;   // read_state.cpp:61, write_state.cpp:53

template <class T>
void read_state_into(EngineBase* e, FILE* f, const StateHeader& h, int nd) {
  const int n = h.n_bodies;
  std::vector<V4<T>> pos(nd), vel(nd), hyd(nd), bd(nd), gr(nd), dd(nd);
  e->get("pos", pos.data(), (long)(nd * sizeof(V4<T>)));
  e->get("vel", vel.data(), (long)(nd * sizeof(V4<T>)));
  e->get("hydro", hyd.data(), (long)(nd * sizeof(V4<T>)));
  e->get("dots", dd.data(), (long)(nd * sizeof(V4<T>)));
  std::vector<double> rec(23);
  for (int i = 0; i < n; ++i) {
    if (std::fread(rec.data(), 8, 23, f) != 23) throw std::runtime_error("state file truncated");
    pos[i] = {T(rec[0]), T(rec[1]), T(rec[2]), T(2 * rec[4])};
    vel[i] = {T(rec[6]), T(rec[7]), T(rec[8]), T(rec[3])};
    hyd[i] = {T(rec[5]), T(0), T(rec[12]), T(rec[18] / kAmuCgs)};
    bd[i] = {T(rec[9]), T(rec[10]), T(rec[11]), T(rec[13])};
    gr[i] = {T(rec[14]), T(rec[15]), T(rec[16]), T(rec[17])};
    dd[i].x = T(rec[22]);
  }
  e->set("pos", pos.data(), (long)(nd * sizeof(V4<T>)));
  e->set("vel", vel.data(), (long)(nd * sizeof(V4<T>)));
  e->set("hydro", hyd.data(), (long)(nd * sizeof(V4<T>)));
  e->set("bdots", bd.data(), (long)(nd * sizeof(V4<T>)));
  e->set("grav", gr.data(), (long)(nd * sizeof(V4<T>)));
  e->set("dots", dd.data(), (long)(nd * sizeof(V4<T>)));
  if (e->get_scalar("form") == SPH_FORM_MIXED) {
    std::vector<D4> p64(nd), v64(nd);
    std::vector<double> u64(nd);
    std::fseek(f, -(long)n * 23 * 8, SEEK_CUR);
    for (int i = 0; i < n; ++i) {
      if (std::fread(rec.data(), 8, 23, f) != 23) throw std::runtime_error("state file truncated");
      p64[i] = {rec[0], rec[1], rec[2], 0.0};
      v64[i] = {rec[6], rec[7], rec[8], rec[3]};
      u64[i] = rec[12];
    }
    e->set("pos64", p64.data(), (long)(nd * sizeof(D4)));
    e->set("vel64", v64.data(), (long)(nd * sizeof(D4)));
    e->set("u64", u64.data(), (long)(nd * sizeof(double)));
  }
}

template <class T>
void write_state_from(EngineBase* e, FILE* f, int n, int nd) {
  std::vector<V4<T>> pos(nd), vel(nd), hyd(nd), bd(nd), gr(nd), dd(nd);
  e->get("pos", pos.data(), (long)(nd * sizeof(V4<T>)));
  e->get("vel", vel.data(), (long)(nd * sizeof(V4<T>)));
  e->get("hydro", hyd.data(), (long)(nd * sizeof(V4<T>)));
  e->get("bdots", bd.data(), (long)(nd * sizeof(V4<T>)));
  e->get("grav", gr.data(), (long)(nd * sizeof(V4<T>)));
  e->get("dots", dd.data(), (long)(nd * sizeof(V4<T>)));
  // the mixed form integrates positions, velocities and u in FP64: the file
  // (23 doubles per body) keeps that state, so a restart loses nothing
  const bool mixed = e->get_scalar("form") == SPH_FORM_MIXED;
  std::vector<D4> p64(mixed ? nd : 0), v64(mixed ? nd : 0);
  std::vector<double> u64(mixed ? nd : 0);
  if (mixed) {
    e->get("pos64", p64.data(), (long)(nd * sizeof(D4)));
    e->get("vel64", v64.data(), (long)(nd * sizeof(D4)));
    e->get("u64", u64.data(), (long)(nd * sizeof(double)));
  }
  for (int i = 0; i < n; ++i) {
    double rec[23] = {(double)pos[i].x, (double)pos[i].y, (double)pos[i].z, (double)vel[i].w,
                      (double)pos[i].w * 0.5, (double)hyd[i].x, (double)vel[i].x, (double)vel[i].y,
                      (double)vel[i].z, (double)bd[i].x, (double)bd[i].y, (double)bd[i].z,
                      (double)hyd[i].z, (double)bd[i].w, (double)gr[i].x, (double)gr[i].y,
                      (double)gr[i].z, (double)gr[i].w, (double)hyd[i].w * kAmuCgs, -1.0, -1.0, -1.0,
                      (double)dd[i].x};
    if (mixed) {
      rec[0] = p64[i].x;
      rec[1] = p64[i].y;
      rec[2] = p64[i].z;
      rec[3] = v64[i].w;
      rec[6] = v64[i].x;
      rec[7] = v64[i].y;
      rec[8] = v64[i].z;
      rec[12] = u64[i];
    }
    std::fwrite(rec, 8, 23, f);
  }
}

// Scratch of sph_radix_sort_noalloc: the digit-count and offset tables and
// the scan partials of launch_radix_sort for up to SPH_NOALLOC_SORT_MAX
// records (4096 records per tile, 256 digits; radix_scratch_ints).  Each GPU
// the library runs on holds its own copy until the process ends.  Host code
// never uses this name as an address: sort_scratch_ptr() looks the device
// address up once per device.
constexpr long kNoAllocSortMax = SPH_NOALLOC_SORT_MAX;
constexpr long kNoAllocTable = 256L * ((kNoAllocSortMax + 4095) / 4096);
constexpr long kNoAllocScratchInts = 2 * kNoAllocTable + 2 + (kNoAllocTable + 2047) / 2048 + 1;
static __device__ __align__(256) int g_sort_scratch[kNoAllocScratchInts];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_ptr[kMaxDevices] = {};

// the current device's address of g_sort_scratch (hipGetSymbolAddress on the
// first call on each device, with that device current)
int* sort_scratch_ptr() {
  int dev = 0;
  CS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("sph_radix_sort_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    if (radix_scratch_ints(kNoAllocSortMax) > kNoAllocScratchInts)
      throw std::logic_error("sph_radix_sort_noalloc: scratch array smaller than the sort needs");
    void* p = nullptr;
    CS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(g_sort_scratch)));
    g_scratch_ptr[dev] = static_cast<int*>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

// RadixSort (radixsort.cu:50-77) without any allocation: the scratch comes
// from the library's fixed device array
int sph_radix_sort_noalloc(sph_kv* pData0, sph_kv* pData1, unsigned int elements, unsigned int bits,
                           void* stream) {
  return guard([&] {
    if (elements > (unsigned)kNoAllocSortMax)
      throw std::invalid_argument("sph_radix_sort_noalloc: " + std::to_string(elements) +
                                  " elements exceed the maximum of " + std::to_string(kNoAllocSortMax));
    if (bits > 32) throw std::invalid_argument("sph_radix_sort_noalloc: bits must be at most 32");
    if (elements > 1 && bits > 0 && (!pData0 || !pData1))
      throw std::invalid_argument("sph_radix_sort_noalloc: null data pointer");
    int* scratch = sort_scratch_ptr();
    launch_radix_sort(reinterpret_cast<KV*>(pData0), reinterpret_cast<KV*>(pData1), (long)elements, (int)bits,
                      scratch, static_cast<hipStream_t>(stream));
    return 0;
  }, -1);
}


const char* sph_last_error(void) { return g_last_error.c_str(); }

void* sph_engine_create(int form, const sph_params* p, int device_id) {
  return guard([&]() -> void* { return make_engine(form, *p, device_id); }, (void*)nullptr);
}

void sph_engine_destroy(void* e) {
  guard([&] { delete static_cast<EngineBase*>(e); return 0; }, 0);
}

long sph_engine_set(void* e, const char* name, const void* host, long n_bytes) {
  return guard([&] { return E(e)->set(name, host, n_bytes); }, -2L);
}
long sph_engine_get(void* e, const char* name, void* host, long max_bytes) {
  return guard([&] { return E(e)->get(name, host, max_bytes); }, -2L);
}
double sph_engine_get_scalar(void* e, const char* name) {
  return guard([&] { return E(e)->get_scalar(name); }, (double)NAN);
}
int sph_engine_set_scalar(void* e, const char* name, double v) {
  return guard([&] { E(e)->set_scalar(name, v); return 0; }, -1);
}
int sph_engine_pass(void* e, int pass) {
  return guard([&] { E(e)->pass(pass); return 0; }, -1);
}
int sph_engine_run(void* e, int steps) {
  return guard([&] { E(e)->run(steps); return 0; }, -1);
}
int sph_engine_sync(void* e) {
  return guard([&] { E(e)->sync(); return 0; }, -1);
}
int sph_engine_statistics(void* e, double* out) {
  return guard([&] { E(e)->statistics(out); return (int)SPH_STAT_COUNT; }, -1);
}

int sph_engine_stats(void* e, double* out11) {
  return guard([&] { E(e)->stats(out11); return 0; }, -1);
}
int sph_engine_timers(void* e, const char** names, double* ms, long* calls, int max_n) {
  return guard([&] { return E(e)->timers(names, ms, calls, max_n); }, -1);
}
void sph_engine_reset_timers(void* e) {
  guard([&] { E(e)->reset_timers(); return 0; }, 0);
}

int sph_build_kd_tree(int n, int n_dev, int n_threads, const sph_float4* pos, sph_float4* tree,
                      sph_kv* particle_map) {
  return guard([&] {
    build_kd_tree<float>((const F4*)pos, n, n_dev, n_threads, (F4*)tree, (KV*)particle_map);
    return 0;
  }, -1);
}

int sph_build_kd_tree64(int n, int n_dev, int n_threads, const double* pos4, double* tree4, sph_kv* particle_map) {
  return guard([&] {
    build_kd_tree<double>((const D4*)pos4, n, n_dev, n_threads, (D4*)tree4, (KV*)particle_map);
    return 0;
  }, -1);
}

int sph_adjoint_forward(void* e, int steps, int ckpt_every, double* loss) {
  return guard([&] {
    const double l = E(e)->adjoint_forward(steps, ckpt_every);
    if (loss) *loss = l;
    return 0;
  }, -1);
}

int sph_adjoint_backward(void* e, double* grad_x, double* grad_v, double* grad_u) {
  return guard([&] {
    E(e)->adjoint_backward(grad_x, grad_v, grad_u);
    return 0;
  }, -1);
}

int sph_adjoint_replay(void* e, const double* x0, const double* v0, const double* u0, double* loss) {
  return guard([&] {
    const double l = E(e)->adjoint_replay(x0, v0, u0);
    if (loss) *loss = l;
    return 0;
  }, -1);
}

double sph_copy_calibration(long bytes) {
  return guard([&] {
    const long n = bytes / 16;
    F4 *a = nullptr, *b = nullptr;
    CS_HIP(hipMalloc(&a, n * 16));
    CS_HIP(hipMalloc(&b, n * 16));
    CS_HIP(hipMemset(a, 0, n * 16));
    hipEvent_t e0, e1;
    CS_HIP(hipEventCreate(&e0));
    CS_HIP(hipEventCreate(&e1));
    launch_copy16(a, b, n, nullptr);
    CS_HIP(hipEventRecord(e0, nullptr));
    launch_copy16(a, b, n, nullptr);
    CS_HIP(hipEventRecord(e1, nullptr));
    CS_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CS_HIP(hipEventElapsedTime(&ms, e0, e1));
    CS_HIP(hipEventDestroy(e0));
    CS_HIP(hipEventDestroy(e1));
    CS_HIP(hipFree(a));
    CS_HIP(hipFree(b));
    return (double)ms;
  }, -1.0);
}

double sph_fma_peak_form(int form) {
  return guard([&] {
    int dev = 0, cus = 0;
    CS_HIP(hipGetDevice(&dev));
    CS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // 8 wavefronts per SIMD on every CU: 4 SIMDs x 8 wavefronts / 4 per block
    const int blocks = cus * 8, iters = 16384;
    float* out = nullptr;
    CS_HIP(hipMalloc(&out, (size_t)blocks * kBlock * sizeof(float)));
    hipEvent_t e0, e1;
    CS_HIP(hipEventCreate(&e0));
    CS_HIP(hipEventCreate(&e1));
    launch_fma_peak(out, blocks, iters, form, nullptr);
    CS_HIP(hipEventRecord(e0, nullptr));
    launch_fma_peak(out, blocks, iters, form, nullptr);
    CS_HIP(hipEventRecord(e1, nullptr));
    CS_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CS_HIP(hipEventElapsedTime(&ms, e0, e1));
    CS_HIP(hipEventDestroy(e0));
    CS_HIP(hipEventDestroy(e1));
    CS_HIP(hipFree(out));
    const double flops = 2.0 * 8.0 * iters * (double)blocks * kBlock;
    return flops / (ms * 1e-3) / 1e12;
  }, -1.0);
}

double sph_pair_body_probe(int variant) {
  return guard([&] {
    int dev = 0, cus = 0;
    CS_HIP(hipGetDevice(&dev));
    CS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = cus * 8, rounds = 2048;
    float* out = nullptr;
    CS_HIP(hipMalloc(&out, (size_t)blocks * kBlock * sizeof(float)));
    hipEvent_t e0, e1;
    CS_HIP(hipEventCreate(&e0));
    CS_HIP(hipEventCreate(&e1));
    launch_pair_body(out, blocks, rounds, variant, nullptr);
    CS_HIP(hipEventRecord(e0, nullptr));
    launch_pair_body(out, blocks, rounds, variant, nullptr);
    CS_HIP(hipEventRecord(e1, nullptr));
    CS_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CS_HIP(hipEventElapsedTime(&ms, e0, e1));
    CS_HIP(hipEventDestroy(e0));
    CS_HIP(hipEventDestroy(e1));
    CS_HIP(hipFree(out));
    return 16.0 * rounds * (double)blocks * kBlock / (ms * 1e-3);
  }, -1.0);
}

int sph_state_file_count(const char* path) {
  return guard([&] {
    FILE* f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open file ") + path);
    StateHeader h;
    const bool ok = read_header(f, h);
    std::fclose(f);
    if (!ok) throw std::runtime_error("state header truncated");
    return h.n_bodies;
  }, -1);
}

int sph_engine_read_state(void* e, const char* path) {
  return guard([&] {
    EngineBase* en = E(e);
    FILE* f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open file ") + path);
    StateHeader h;
    if (!read_header(f, h)) { std::fclose(f); throw std::runtime_error("state header truncated"); }
    const sph_params& p = en->params();
    if (h.n_bodies != p.n_bodies) { std::fclose(f); throw std::runtime_error("state file N differs from the engine's"); }
    if (en->get_scalar("real_bytes") == 8) read_state_into<double>(en, f, h, p.n_dev);
    else read_state_into<float>(en, f, h, p.n_dev);
    std::fclose(f);
    en->set_scalar("dt", h.dt);
    en->set_scalar("global_time", h.global_time);
    en->set_scalar("iteration", h.iteration_number);
    std::memcpy(en->state_header, &h, sizeof(h));
    en->has_state_header = true;
    return 0;
  }, -1);
}

// write_state.cpp:3-62: the header read last (or, for an engine started
// from initial conditions, one carrying the parameters) with the fields the
// run advances taken from the engine
static void write_state_file(EngineBase* en, const char* path) {
  const sph_params& p = en->params();
  StateHeader h{};
  if (en->has_state_header) {
    std::memcpy(&h, en->state_header, sizeof(h));
  } else {
    h.alpha = p.av_alpha;
    h.beta = p.av_beta;
  }
  h.n_bodies = p.n_bodies;
  h.iteration_number = (int)en->get_scalar("iteration");
  h.global_time = en->get_scalar("global_time");
  h.dt = en->get_scalar("dt");
  FILE* f = std::fopen(path, "wb");
  if (!f) throw std::runtime_error(std::string("cannot open file ") + path);
  write_header(f, h);
  if (en->get_scalar("real_bytes") == 8) write_state_from<double>(en, f, p.n_bodies, p.n_dev);
  else write_state_from<float>(en, f, p.n_bodies, p.n_dev);
  const bool ok = std::ferror(f) == 0;
  std::fclose(f);
  if (!ok) throw std::runtime_error(std::string("write failed: ") + path);
}

int sph_engine_write_state(void* e, const char* path) {
  return guard([&] {
    write_state_file(E(e), path);
    return 0;
  }, -1);
}

// output.cpp:3-15: a numbered snapshot each time global_time passes the next
// multiple of dt_out (dt_out > 0, from the state file's header), and the
// restart file every call
int sph_engine_output(void* e, const char* dir) {
  return guard([&] {
    EngineBase* en = E(e);
    StateHeader h{};
    if (en->has_state_header) std::memcpy(&h, en->state_header, sizeof(h));
    const std::string base = std::string(dir && *dir ? dir : ".") + "/";
    int written = 0;
    if (en->get_scalar("global_time") > h.snapshot_number * h.dt_out && h.dt_out > 0) {
      ++h.snapshot_number;
      std::memcpy(en->state_header, &h, sizeof(h));
      char name[32];
      std::snprintf(name, sizeof(name), "out%04d.bin", h.snapshot_number);
      write_state_file(en, (base + name).c_str());
      written = h.snapshot_number;
    }
    write_state_file(en, (base + "restart.bin").c_str());
    return written;
  }, -1);
}

}  // extern "C"
