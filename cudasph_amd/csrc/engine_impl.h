// Engine<T>: the device-resident simulation's class definition, shared by
// engine.hip (forward passes) and adjoint.hip (reverse passes).
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "engine.h"
#include "kdtree.h"
#include "kernels.h"
#include "sph_math.h"

namespace cudasph {

struct TimerScope {
  Timers& t;
  hipStream_t s;
  TimerScope(Timers& t_, const char* name, hipStream_t s_) : t(t_), s(s_) { t.begin(name, s); }
  ~TimerScope() { t.end(s); }
};

inline int bits_for(long v) {
  int b = 0;
  while ((1L << b) <= v) ++b;
  return b;
}

// ---------------------------------------------------------------------------
template <class T>
class Engine final : public EngineBase {
 public:
  Engine(int form, const sph_params& p, int device);
  ~Engine() override {
    if (s_) (void)hipStreamSynchronize(s_);
    if (gs_) (void)hipStreamSynchronize(gs_);
    if (fetched_) (void)hipEventDestroy(fetched_);
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    if (ev_join_) (void)hipEventDestroy(ev_join_);
    if (gs_) (void)hipStreamDestroy(gs_);
    if (s_) (void)hipStreamDestroy(s_);
  }
  long set(const std::string& name, const void* host, long n_bytes) override;
  long get(const std::string& name, void* host, long max_bytes) override;
  double get_scalar(const std::string& name) override;
  void set_scalar(const std::string& name, double v) override;
  void pass(int id) override;
  void run(int steps) override {
    for (int k = 0; k < steps; ++k) step();
  }
  void sync() override { CS_HIP(hipStreamSynchronize(s_)); }
  void stats(double* out11) override;
  void statistics(double* out) override;
  int timers(const char** names, double* ms, long* calls, int max_n) override {
    sync();
    return timers_.list(names, ms, calls, max_n);
  }
  void reset_timers() override {
    sync();
    timers_.reset();
  }
  const sph_params& params() const override { return p_; }

 private:
  static constexpr bool kF32 = sizeof(T) == 4;
  bool mixed() const { return form_ == SPH_FORM_MIXED; }
  void ensure_boxes();
  WalkParams walk_params() const {
    return {p_.n_bodies, p_.fix_phantom, p_.rel_err, p_.desired_ngb, p_.max_iter,
            walk_boxes_ ? boxes_.p : nullptr};
  }
  EosConst eos_const() const {
    EosConst c;
    c.density_unit = p_.eos_density_unit;
    c.energy_unit = p_.eos_energy_unit;
    c.pressure_unit = p_.eos_pressure_unit;
    c.amu = p_.eos_amu;
    c.boltzmann = p_.eos_boltzmann;
    c.rad_const = p_.eos_rad_const;
    c.gamma = p_.eos_gamma;
    c.temp_base = p_.eos_temp_base;
    c.newton = p_.eos_newton;
    return c;
  }
  void refresh_if_dirty() {
    if (mixed() && dirty64_) {
      launch_refresh_mixed(pos64_, vel64_, u64_, (F4*)pos_.p, pos_lo_, (F4*)vel_.p, (F4*)hydro_.p, n_, s_);
      dirty64_ = false;
    }
  }

  void tree();
  void tree_fetch();
  void tree_build();
  void sph_after_tree();
  void solve();
  void offsets();
  long check_total();
  void find();
  void density();
  void eos();
  void hydro();
  void reaction();
  void sph_accelerations();
  void gravity(bool async = false);
  void gravity_join();
  void gravity_pack();
  void gravity_kernels(hipStream_t gs);
  void add_gravity();
  void timestep(double* out);
  void timestep_compute(double* out);
  void predict();
  void correct(const double* dt_new, const double* dt_old);
  void stats_pass();
  void sort_bodies();
  void init();
  void step();
  void build_transpose();

 public:
  // reverse mode of the FP64 form (adjoint.hip)
  double adjoint_forward(int steps, int ckpt_every) override;
  void adjoint_backward(double* gx, double* gv, double* gu) override;
  double adjoint_replay(const double* x0, const double* v0, const double* u0) override;

 private:
  enum { kAdjOff = 0, kAdjRecord = 1, kAdjReplay = 2, kAdjReplayStoredGravity = 3 };
  void adj_forces_backward(bool seed_rho);
  void adj_save(DBuf<V4<T>>* dst[4], long slot);   // pos, vel, hydro, bdots
  void adj_load(DBuf<V4<T>>* src[4], long slot);
  double adj_loss();
  int adj_mode_ = kAdjOff, adj_K_ = 0, adj_k_ = 1, adj_f_ = 0, adj_iter0_ = 0;
  int adj_free_ranges_ = 0;   // sph_adjoint_replay re-solves the ranges (timesteps stay recorded)
  int adj_implicit_h_ = 0;    // sph_adjoint_backward adds the h-solve's derivative
  DBuf<double> adj_rbar_;     // per-particle range adjoint of one evaluation
  bool adj_ready_ = false;
  std::vector<double> adj_dt_;                       // timestep set by force evaluation f = 0..K
  DBuf<double> adj_range_;                           // ranges of evaluation f, (K+1) x n_dev
  DBuf<int> adj_nngb_;                               // h-solve counts of evaluation f
  DBuf<V4<T>> adj_grav_;                             // gravity of evaluation f
  DBuf<V4<T>> adj_ck_pos_, adj_ck_vel_, adj_ck_hyd_, adj_ck_bd_;     // checkpoints, every k steps
  // force-evaluation data of each step of one segment, captured during its
  // recomputation: state at the evaluation, lists and their transpose
  struct EvalSlot {
    DBuf<V4<T>> pos, vel, hyd, dots;
    DBuf<int> off, cnt, list, tr_start, tr_end;
    DBuf<KV> pairs;
    long tot = 0;
  };
  std::vector<std::unique_ptr<EvalSlot>> adj_eval_;
  int adj_capture_ = -1;           // slot to capture into during step(), or -1
  void adj_capture(EvalSlot& e);
  void adj_restore(const EvalSlot& e);
  DBuf<V4<T>> adj_init_pos_, adj_init_vel_, adj_init_hyd_, adj_final_vel_;
  // the state adjoint_forward leaves (pos vel hydro dots bdots hydro2 grav;
  // dt, the stats, force and diagnostics sums; step count and time), put
  // back when adjoint_backward ends so that step() continues the recorded run
  DBuf<V4<T>> adj_end_[7];
  DBuf<double> adj_end_small_;
  int adj_end_iter_ = 0;
  double adj_end_time_ = 0, adj_t0_ = 0;
  void adj_keep_end();
  void adj_put_back_end();
  DBuf<V4<T>> ax_, av_, aa_;       // carried adjoints of the state: x, v, (a, udot)
  DBuf<double> au_;                // carried adjoint of u
  DBuf<V4<T>> xs_, vs_, hb1_, hb2_;  // one evaluation: x, v, (rho, P, u), (divv, curl, Omega, Psi)
  DBuf<double> adj_out_, adj_partial_;
  DBuf<double> pairbuf_;           // per-slot j-side adjoint contributions, 10 doubles each

  int form_;
  sph_params p_;
  int n_, nd_;
  long cap_;
  hipStream_t s_ = nullptr;
  hipEvent_t fetched_ = nullptr, ev_fork_ = nullptr, ev_join_ = nullptr;
  hipStream_t gs_ = nullptr;       // gravity stream (overlap of gravity with the SPH passes)
  bool grav_pending_ = false;
  Timers timers_;

  DBuf<V4<T>> pos_, vel_, hydro_, dots_, bdots_, hydro2_, grav_, tree_, gpos_, gvel_, old_vu_, tmp4_;
  DBuf<KV> hash_, pairs_, pairs_alt_, perm_;
  DBuf<int> n_ngb_, ngb_count_, ngb_offset_, list_, owner_, tr_start_, tr_end_, flags_, scan_tmp_, sort_tmp_;
  DBuf<T> acc_;
  DBuf<double> dt_cur_, dt_next_, partials_, force6_, stats6_;
  DBuf<double> diag_part_, diag_;   // system statistics beyond the energies (launch_diagnostics)
  DBuf<uint32_t> key_hi_, key_lo_;
  // mixed form
  DBuf<D4> pos64_, vel64_, old64_;
  DBuf<D4> gpart64_;            // FP64 form: per-j-part gravity sums (launch_gravity64_plummer)
  DBuf<double> u64_;
  DBuf<F4> pos_lo_, ghi_, glo_, gvel32_, gpart_;
  DBuf<F4> boxes_;              // subtree bounding boxes of tree_ (launch_tree_boxes), 2 per node
  bool boxes_stale_ = true;     // tree_ changed since they were built
  DBuf<float> gmass_;   // {min, max} of the masses + reduction partials (mixed gravity)
  DBuf<float> gmom_;    // per-sub-tile multipole records (mixed gravity far field)
  bool dirty64_ = false;

  HBuf<V4<T>> h_pos_, h_tree_;
  HBuf<KV> h_hash_;
  HBuf<long long> h_small_;

  double dt_ = 0, global_time_ = 0;
  int iteration_ = 0, exact_rsqrt_ = 0;
  // 1 (default): gravity on gs_ beside the SPH passes (Plummer softening),
  // 0: in order (measured in Engine::step)
  int overlap_ = 1;
  int grav_split_only_ = 0;
  int grav_unit_mass_ = 1;    // mixed gravity: unit-mass pair body when every mass is equal
  int grav_expand_ = 1;       // mixed gravity: expanded r^2 in the equal-mass tile body
  int grav_parts_ = 0;        // mixed pairwise gravity: j-parts, 0 automatic
  int solve_cache_ = -1;      // h-solve candidate cache: 1 on, 0 off, -1 where it pays
  DBuf<V4<T>> solve_rec_;     // its records, kSolveCacheK per particle
  int grav_sums_ = -1;        // plain Plummer gravity: four sums in four wavefronts (1), one lane (0), auto (-1)
  int grav_far_ = 0;          // mixed gravity: order-4 multipole far field for distant sub-tiles
  double grav_far_ratio_ = 8;  // ... taken beyond this many radii of a 16384-record cell
  int grav_far_half_ = 1;      // ... also the 32-record halves of a sub-tile not taken whole
  double grav_far_grade_ = 0;  // ... the ratio at a level of C records scaled by (C / 16384)^grade
  int i_order_ = 1;   // 1: lanes take particles in kd-tree heap order, 0: in memory order
  // 1: the kd-tree is built on the device (same tree without coordinate
  // ties; with ties the same neighbour sets; the mixed form's default);
  // 0: on the host, bit-identical to the reference's quicksort build
  int tree_gpu_ = 0;            // 0 host build, 1 device build, 2 device build checked for ties (host on a tie)
  long tree_fallbacks_ = 0;     // steps whose verified device build fell back to the host build
  DBuf<int> ties_;
  int walk_boxes_ = 0;   // walks prune subtrees by their bounding boxes (mixed and FP64 forms)
  DBuf<int> tree_work_;   // allocated on first use of the device build
  long ngb_tot_ = 0;
  bool tot_stale_ = true;   // ngb_tot_ not yet read for the current offsets (a host round trip)
  int overflow_reruns_ = 0;
  int phantom_reads_ = 0, phantom_ngb_ = 0;   // walks of the last find_ngb that read node n+1 / took it
};

}  // namespace cudasph
