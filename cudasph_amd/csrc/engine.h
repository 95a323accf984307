// Device-resident SPH simulation (host side).
//
// Engine mirrors the reference's host class CUDAsph (CUDAsph.h:72-163): the
// same step() / sph_accelerations() / gravity() / compute_timestep() /
// system_statistics() / sort_bodies() structure and the same float4 state
// arrays, but the state stays in HBM for the whole run.  Per step the only
// host round trips are the positions for the kd-tree build (the reference
// builds the tree on the host, kd_tree.cpp) and one small read of the
// neighbour-list total.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "../../include/cudasph.h"
#include "common.h"

namespace cudasph {

class EngineBase {
 public:
  virtual ~EngineBase() = default;
  virtual long set(const std::string& name, const void* host, long n_bytes) = 0;
  virtual long get(const std::string& name, void* host, long max_bytes) = 0;
  virtual double get_scalar(const std::string& name) = 0;
  virtual void set_scalar(const std::string& name, double v) = 0;
  virtual void pass(int id) = 0;
  virtual void run(int steps) = 0;
  virtual void sync() = 0;
  virtual void stats(double* out11) = 0;
  // everything system_statistics.cpp:4-82 reports, SPH_STAT_COUNT values
  virtual void statistics(double* out) = 0;
  virtual int timers(const char** names, double* ms, long* calls, int max_n) = 0;
  virtual void reset_timers() = 0;
  virtual const sph_params& params() const = 0;
  // reverse mode (FP64 form): record K steps, gradient of the loss with
  // respect to the initial x, v, u, and re-runs of the recorded map
  virtual double adjoint_forward(int steps, int ckpt_every) = 0;
  virtual void adjoint_backward(double* gx, double* gv, double* gu) = 0;
  virtual double adjoint_replay(const double* x0, const double* v0, const double* u0) = 0;
  // the header of the last state file read (read_state.cpp:5-23), kept so
  // that write_state writes back the fields this engine does not evolve
  // (t_end, dt_out, snapshot_number, ...); opaque here, laid out by api.hip
  alignas(8) unsigned char state_header[192] = {};
  bool has_state_header = false;
};

EngineBase* make_engine(int form, const sph_params& p, int device);

// device buffer owned by the engine
template <class X>
struct DBuf {
  X* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t count) {
    if (p) CS_HIP(hipFree(p));
    p = nullptr;
    n = count;
    if (count) CS_HIP(hipMalloc(&p, count * sizeof(X)));
  }
  size_t bytes() const { return n * sizeof(X); }
  operator X*() const { return p; }
};

// pinned host buffer
template <class X>
struct HBuf {
  X* p = nullptr;
  size_t n = 0;
  HBuf() = default;
  HBuf(const HBuf&) = delete;
  HBuf& operator=(const HBuf&) = delete;
  ~HBuf() {
    if (p) (void)hipHostFree(p);
  }
  void alloc(size_t count) {
    if (p) CS_HIP(hipHostFree(p));
    p = nullptr;
    n = count;
    if (count) CS_HIP(hipHostMalloc(&p, count * sizeof(X), hipHostMallocDefault));
  }
  X& operator[](size_t i) { return p[i]; }
};

// named GPU timers: HIP events around each launch group on the engine stream
class Timers {
 public:
  void begin(const char* name, hipStream_t s);
  void add_host(const char* name, double ms);
  void end(hipStream_t s);
  void resolve();
  int list(const char** names, double* ms, long* calls, int max_n);
  void reset();
  ~Timers();
  bool enabled = false;

 private:
  struct Pending {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> pool_;
  const char* open_ = nullptr;
  hipEvent_t open_ev_ = nullptr;
  std::map<std::string, std::pair<double, long>> acc_;
  hipEvent_t take();
};

}  // namespace cudasph
