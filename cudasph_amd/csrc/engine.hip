// Device-resident SPH engine: the passes of step.cpp:13-147 over HBM state.
#include "engine_impl.h"

namespace cudasph {

// ---------------------------------------------------------------------------
hipEvent_t Timers::take() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  CS_HIP(hipEventCreate(&e));
  return e;
}
void Timers::begin(const char* name, hipStream_t s) {
  if (!enabled) return;
  open_ = name;
  open_ev_ = take();
  CS_HIP(hipEventRecord(open_ev_, s));
}
void Timers::end(hipStream_t s) {
  if (!enabled || !open_) return;
  hipEvent_t b = take();
  CS_HIP(hipEventRecord(b, s));
  pending_.push_back({open_, open_ev_, b});
  open_ = nullptr;
}
void Timers::add_host(const char* name, double ms) {
  if (!enabled) return;
  auto& a = acc_[name];
  a.first += ms;
  a.second += 1;
}
void Timers::resolve() {
  for (auto& p : pending_) {
    CS_HIP(hipEventSynchronize(p.b));
    float ms = 0.f;
    CS_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    auto& a = acc_[p.name];
    a.first += ms;
    a.second += 1;
    pool_.push_back(p.a);
    pool_.push_back(p.b);
  }
  pending_.clear();
}
int Timers::list(const char** names, double* ms, long* calls, int max_n) {
  resolve();
  int k = 0;
  for (auto& kv : acc_) {
    if (k >= max_n) break;
    // the map key owns the string; hand out a stable pointer
    names[k] = kv.first.c_str();
    ms[k] = kv.second.first;
    calls[k] = kv.second.second;
    ++k;
  }
  return (int)acc_.size();
}
void Timers::reset() {
  resolve();
  acc_.clear();
}
Timers::~Timers() {
  for (auto& p : pending_) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : pool_) (void)hipEventDestroy(e);
}

template <class T>
Engine<T>::Engine(int form, const sph_params& p, int device) : form_(form), p_(p) {
  n_ = p.n_bodies;
  nd_ = p.n_dev;
  if (n_ <= 0 || nd_ < n_) throw std::invalid_argument("sph_engine: need 0 < n_bodies <= n_dev");
  if (nd_ >= (1 << 24) - 2) throw std::invalid_argument("sph_engine: n_dev must be below 2^24 (tree node word)");
  if (form == SPH_FORM_MIXED && !kF32) throw std::invalid_argument("sph_engine: mixed form is FP32");
  CS_HIP(hipSetDevice(device));
  CS_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
  CS_HIP(hipStreamCreateWithFlags(&gs_, hipStreamNonBlocking));   // default priority, as s_
  CS_HIP(hipEventCreateWithFlags(&fetched_, hipEventDisableTiming));
  CS_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
  CS_HIP(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming));
  cap_ = 100L * nd_;  // ngb_tot_max (read_state.cpp:29)
  // the plain form's lists follow the reference's tie-breaking: its device
  // build is checked for ties and falls back to the host build when one
  // decides a node; the other forms build on the device
  tree_gpu_ = form != SPH_FORM_PLAIN ? 1 : 2;
  // the plain form walks exactly as the reference (its walk lengths are an
  // output, bodies_hash.key); the others prune by subtree boxes
  walk_boxes_ = form != SPH_FORM_PLAIN;

  for (auto* b : {&pos_, &vel_, &hydro_, &dots_, &bdots_, &hydro2_, &grav_, &gpos_, &gvel_, &old_vu_, &tmp4_})
    b->alloc(nd_);
  tree_.alloc(nd_ + 2);
  ties_.alloc(1);
  boxes_.alloc(2 * ((size_t)nd_ + 2));
  if (!kF32) {   // the larger of the forward and reverse passes' splits
    const int parts = std::max(gravity64_parts(nd_), gravity64_forward_parts(nd_));
    if (parts > 1) gpart64_.alloc((size_t)parts * nd_);
  }
  hash_.alloc(nd_);
  perm_.alloc(nd_);
  n_ngb_.alloc(nd_);
  ngb_count_.alloc(nd_);
  ngb_offset_.alloc(nd_ + 1);
  list_.alloc(cap_);
  owner_.alloc(nd_);   // per-j counts of the reaction transpose
  acc_.alloc(cap_);
  pairs_.alloc(cap_);
  pairs_alt_.alloc(cap_);
  tr_start_.alloc(nd_ + 1);
  tr_end_.alloc(nd_ + 1);
  flags_.alloc(4);
  scan_tmp_.alloc(scan_scratch_ints(std::max<long>(cap_, nd_ + 1)) + 16);
  sort_tmp_.alloc(radix_scratch_ints(cap_) + 16);
  dt_cur_.alloc(1);
  dt_next_.alloc(1);
  partials_.alloc(reduce_partials(nd_) + 64);
  force6_.alloc(6);
  stats6_.alloc(6);
  diag_part_.alloc(diag_partials(n_));
  diag_.alloc(19);
  key_hi_.alloc(nd_);
  key_lo_.alloc(nd_);
  if (mixed()) {
    pos64_.alloc(nd_);
    vel64_.alloc(nd_);
    old64_.alloc(nd_);
    u64_.alloc(nd_);
    pos_lo_.alloc(nd_);
    ghi_.alloc(nd_);
    glo_.alloc(nd_);
    gvel32_.alloc(nd_);
    gpart_.alloc(2 * 8 * (size_t)nd_);
    gmass_.alloc(2 + 2 * (size_t)ceil_div(n_, kBlock));
    // far-field records: 64-record sub-tiles, then cells of 16384, 4096, 1024, 256,
    // then the sub-tiles' 32-record halves
    gmom_.alloc(52 * ((size_t)ceil_div(nd_, 64) + ceil_div(nd_, 16384) + ceil_div(nd_, 4096) + ceil_div(nd_, 1024) +
                      ceil_div(nd_, 256) + 2 * (size_t)ceil_div(nd_, 64)));
    CS_HIP(hipMemsetAsync(pos_lo_, 0, pos_lo_.bytes(), s_));
    CS_HIP(hipMemsetAsync(pos64_, 0, pos64_.bytes(), s_));
    CS_HIP(hipMemsetAsync(vel64_, 0, vel64_.bytes(), s_));
    CS_HIP(hipMemsetAsync(u64_, 0, u64_.bytes(), s_));
  }
  h_pos_.alloc(nd_);
  h_tree_.alloc(nd_ + 2);
  h_hash_.alloc(nd_);
  h_small_.alloc(8);

  for (auto* b : {&pos_, &vel_, &hydro_, &dots_, &bdots_, &hydro2_, &grav_, &gpos_, &gvel_, &old_vu_})
    CS_HIP(hipMemsetAsync(b->p, 0, b->bytes(), s_));
  for (auto* b : {&n_ngb_, &ngb_count_, &ngb_offset_, &tr_start_, &tr_end_, &flags_})
    CS_HIP(hipMemsetAsync(b->p, 0, b->bytes(), s_));
  CS_HIP(hipMemsetAsync(dt_cur_, 0, 8, s_));
  CS_HIP(hipMemsetAsync(dt_next_, 0, 8, s_));
  CS_HIP(hipMemsetAsync(stats6_, 0, 48, s_));
  CS_HIP(hipMemsetAsync(diag_, 0, diag_.bytes(), s_));
  CS_HIP(hipMemsetAsync(force6_, 0, 48, s_));

  // identity hash and far padding records (read_state.cpp:70-89)
  std::vector<KV> hh(nd_);
  for (int i = 0; i < nd_; ++i) hh[i] = {i < n_ ? (uint32_t)i : (1u << 24), (uint32_t)i};
  CS_HIP(hipMemcpyAsync(hash_, hh.data(), nd_ * sizeof(KV), hipMemcpyHostToDevice, s_));
  if (nd_ > n_) {
    std::vector<V4<T>> far(nd_ - n_, V4<T>{T(1e10), T(1e10), T(1e10), T(0)});
    CS_HIP(hipMemcpyAsync(pos_.p + n_, far.data(), far.size() * sizeof(V4<T>), hipMemcpyHostToDevice, s_));
    CS_HIP(hipMemcpyAsync(vel_.p + n_, far.data(), far.size() * sizeof(V4<T>), hipMemcpyHostToDevice, s_));
  }
  sync();
}

template <class T>
long Engine<T>::set(const std::string& name, const void* host, long nb) {
  auto put = [&](void* dev, size_t have) {
    const size_t c = std::min<size_t>(have, (size_t)nb);
    CS_HIP(hipMemcpyAsync(dev, host, c, hipMemcpyHostToDevice, s_));
    sync();
    return (long)c;
  };
  if (name == "pos") return put(pos_, pos_.bytes());
  if (name == "vel") return put(vel_, vel_.bytes());
  if (name == "hydro") return put(hydro_, hydro_.bytes());
  if (name == "dots") return put(dots_, dots_.bytes());
  if (name == "bdots") return put(bdots_, bdots_.bytes());
  if (name == "hydro2") return put(hydro2_, hydro2_.bytes());
  if (name == "grav") return put(grav_, grav_.bytes());
  if (name == "tree") {
    boxes_stale_ = true;
    return put(tree_, tree_.bytes());
  }
  if (name == "hash") return put(hash_, hash_.bytes());
  if (name == "n_ngb") return put(n_ngb_, n_ngb_.bytes());
  if (name == "ngb_count") return put(ngb_count_, ngb_count_.bytes());
  if (name == "ngb_offset") {
    tot_stale_ = true;
    return put(ngb_offset_, ngb_offset_.bytes());
  }
  if (name == "ngb_list") return put(list_, list_.bytes());
  if (name == "ngb_acc") return put(acc_, acc_.bytes());
  if (mixed()) {
    if (name == "pos64") { dirty64_ = true; return put(pos64_, pos64_.bytes()); }
    if (name == "vel64") { dirty64_ = true; return put(vel64_, vel64_.bytes()); }
    if (name == "u64") { dirty64_ = true; return put(u64_, u64_.bytes()); }
  }
  return -1;
}

template <class T>
long Engine<T>::get(const std::string& name, void* host, long mb) {
  refresh_if_dirty();
  auto take = [&](const void* dev, size_t have) {
    const size_t c = std::min<size_t>(have, (size_t)mb);
    sync();
    CS_HIP(hipMemcpy(host, dev, c, hipMemcpyDeviceToHost));
    return (long)c;
  };
  if (name == "pos") return take(pos_, pos_.bytes());
  if (name == "vel") return take(vel_, vel_.bytes());
  if (name == "hydro") return take(hydro_, hydro_.bytes());
  if (name == "dots") return take(dots_, dots_.bytes());
  if (name == "bdots") return take(bdots_, bdots_.bytes());
  if (name == "hydro2") return take(hydro2_, hydro2_.bytes());
  if (name == "grav") return take(grav_, grav_.bytes());
  if (name == "tree") return take(tree_, tree_.bytes());
  if (name == "boxes") return take(boxes_, boxes_.bytes());
  if (name == "hash") return take(hash_, hash_.bytes());
  if (name == "n_ngb") return take(n_ngb_, n_ngb_.bytes());
  if (name == "ngb_count") return take(ngb_count_, ngb_count_.bytes());
  if (name == "ngb_offset") return take(ngb_offset_, ngb_offset_.bytes());
  if (name == "ngb_list") return take(list_, list_.bytes());
  if (name == "ngb_acc") return take(acc_, acc_.bytes());
  if (name == "adj_range") return take(adj_range_, adj_range_.bytes());
  if (name == "adj_dt") {
    const size_t c = std::min<size_t>(adj_dt_.size() * 8, (size_t)mb);
    std::memcpy(host, adj_dt_.data(), c);
    return (long)c;
  }
  if (mixed()) {
    if (name == "pos64") return take(pos64_, pos64_.bytes());
    if (name == "vel64") return take(vel64_, vel64_.bytes());
    if (name == "u64") return take(u64_, u64_.bytes());
    if (name == "pos_lo") return take(pos_lo_, pos_lo_.bytes());
  }
  return -1;
}

template <class T>
double Engine<T>::get_scalar(const std::string& name) {
  if (name == "dt") {
    sync();
    CS_HIP(hipMemcpy(&dt_, dt_cur_, 8, hipMemcpyDeviceToHost));
    return dt_;
  }
  if (name == "global_time") return global_time_;
  if (name == "iteration") return iteration_;
  if (name == "ngb_tot") return (double)ngb_tot_;
  if (name == "overflow_reruns") return overflow_reruns_;
  if (name == "phantom_reads") return phantom_reads_;
  if (name == "phantom_ngb") return phantom_ngb_;
  if (name == "exact_rsqrt") return exact_rsqrt_;
  if (name == "i_order") return i_order_;
  if (name == "grav_split_only") return grav_split_only_;
  if (name == "grav_unit_mass") return grav_unit_mass_;
  if (name == "grav_expand") return grav_expand_;
  if (name == "grav_sums") return grav_sums_;
  if (name == "grav_parts") return grav_parts_;
  if (name == "grav_parts_auto") {   // the j-parts the automatic choice takes at this size
    if (!kF32) return gravity64_forward_parts(nd_);
    return mixed() ? gravity_mixed_parts(nd_, grav_far_) : 1;
  }
  if (name == "solve_cache") return solve_cache_;
  if (name == "adj_free_ranges") return adj_free_ranges_;
  if (name == "adj_implicit_h") return adj_implicit_h_;
  if (name == "grav_far") return grav_far_;
  if (name == "grav_far_ratio") return grav_far_ratio_;
  if (name == "grav_far_grade") return grav_far_grade_;
  if (name == "grav_far_half") return grav_far_half_;
  if (name == "walk_boxes") return walk_boxes_;
  if (name == "overlap") return overlap_;
  if (name == "tree_gpu") return tree_gpu_;
  if (name == "tree_fallbacks") return (double)tree_fallbacks_;
  if (name == "timing") return timers_.enabled;
  if (name == "cap") return (double)cap_;
  if (name == "real_bytes") return (double)sizeof(T);
  if (name == "form") return form_;
  return NAN;
}

template <class T>
void Engine<T>::set_scalar(const std::string& name, double v) {
  if (name == "dt") {
    sync();
    dt_ = v;
    CS_HIP(hipMemcpy(dt_cur_, &dt_, 8, hipMemcpyHostToDevice));
  } else if (name == "global_time") {
    global_time_ = v;
  } else if (name == "iteration") {
    iteration_ = (int)v;
  } else if (name == "exact_rsqrt") {
    exact_rsqrt_ = (int)v;
  } else if (name == "i_order") {
    i_order_ = (int)v;
  } else if (name == "grav_split_only") {
    grav_split_only_ = (int)v;
  } else if (name == "grav_unit_mass") {
    grav_unit_mass_ = (int)v;
  } else if (name == "grav_expand") {
    grav_expand_ = (int)v;
  } else if (name == "solve_cache") {
    if (v != -1 && v != 0 && v != 1) throw std::invalid_argument("solve_cache must be -1, 0 or 1");
    solve_cache_ = (int)v;
  } else if (name == "grav_parts") {
    if (!(v >= 0 && v <= 8) || v != (int)v) throw std::invalid_argument("grav_parts must be 0 (automatic) or 1..8");
    grav_parts_ = (int)v;
  } else if (name == "grav_sums") {
    if (v != -1 && v != 0 && v != 1 && v != 2) throw std::invalid_argument("grav_sums must be -1, 0, 1 or 2");
    grav_sums_ = (int)v;
  } else if (name == "adj_free_ranges") {
    adj_free_ranges_ = (int)v;
  } else if (name == "adj_implicit_h") {
    adj_implicit_h_ = (int)v;
  } else if (name == "grav_far") {
    grav_far_ = (int)v;
  } else if (name == "grav_far_ratio") {
    if (!(v >= 4)) throw std::invalid_argument("grav_far_ratio must be at least 4");
    grav_far_ratio_ = v;
  } else if (name == "grav_far_half") {
    grav_far_half_ = (int)v;
  } else if (name == "grav_far_grade") {
    if (!(v >= 0 && v <= 0.2)) throw std::invalid_argument("grav_far_grade must lie in [0, 0.2]");
    grav_far_grade_ = v;
  } else if (name == "walk_boxes") {
    walk_boxes_ = v != 0;
  } else if (name == "overlap") {
    if (v != 0 && v != 1) throw std::invalid_argument("overlap must be 0 or 1");
    overlap_ = (int)v;
  } else if (name == "tree_gpu") {
    if (v != 0 && v != 1 && v != 2) throw std::invalid_argument("tree_gpu must be 0, 1 or 2");
    tree_gpu_ = (int)v;
  } else if (name == "timing") {
    sync();
    timers_.enabled = v != 0;
  } else {
    throw std::invalid_argument("unknown scalar " + name);
  }
}

template <class T>
void Engine<T>::stats(double* o) {
  sync();
  double s6[6], f6[6];
  CS_HIP(hipMemcpy(s6, stats6_, 48, hipMemcpyDeviceToHost));
  CS_HIP(hipMemcpy(f6, force6_, 48, hipMemcpyDeviceToHost));
  o[0] = s6[0];
  o[1] = s6[1];
  o[2] = s6[2];
  o[3] = s6[0] + s6[1] + s6[2];
  o[4] = std::sqrt(s6[3] * s6[3] + s6[4] * s6[4] + s6[5] * s6[5]);
  for (int k = 0; k < 6; ++k) o[5 + k] = f6[k];
}

// system_statistics.cpp:4-82 (values of the last STATS pass, which runs where
// the reference's does: after add_gravity, before the corrector; the list
// counters of the last find_ngb)
template <class T>
void Engine<T>::statistics(double* o) {
  stats(o);
  double s6[6], d[19];
  CS_HIP(hipMemcpy(s6, stats6_, 48, hipMemcpyDeviceToHost));
  CS_HIP(hipMemcpy(d, diag_, sizeof d, hipMemcpyDeviceToHost));
  for (int c = 0; c < 3; ++c) {
    o[SPH_STAT_JX + c] = s6[3 + c];
    o[SPH_STAT_XMIN + c] = d[c];
    o[SPH_STAT_XMAX + c] = d[3 + c];
  }
  const double m = d[6];
  o[SPH_STAT_MTOT] = m;
  double r2 = 0, v2 = 0;
  for (int c = 0; c < 3; ++c) {
    o[SPH_STAT_XCOM + c] = m != 0 ? d[7 + c] / m : 0.0;
    o[SPH_STAT_VXCOM + c] = m != 0 ? d[10 + c] / m : 0.0;
    r2 += o[SPH_STAT_XCOM + c] * o[SPH_STAT_XCOM + c];
    v2 += o[SPH_STAT_VXCOM + c] * o[SPH_STAT_VXCOM + c];
  }
  o[SPH_STAT_RCM] = std::sqrt(r2);
  o[SPH_STAT_VCM] = std::sqrt(v2);
  const double mean = d[13] / n_, msq = d[14] / n_;
  o[SPH_STAT_NGB_MEAN] = mean;
  o[SPH_STAT_NGB_SIGMA] = std::sqrt(std::max(0.0, msq - mean * mean));
  o[SPH_STAT_NGB_MIN] = d[16];
  o[SPH_STAT_NGB_MAX] = d[17];
  o[SPH_STAT_COUNT_MISMATCH] = d[15];
  o[SPH_STAT_MAX_COUNT] = d[18];
  o[SPH_STAT_PHANTOM_READS] = phantom_reads_;
  o[SPH_STAT_PHANTOM_NGB] = phantom_ngb_;
  o[SPH_STAT_OVERFLOW_RERUNS] = overflow_reruns_;
  o[SPH_STAT_ITERATION] = iteration_;
  o[SPH_STAT_TIME] = global_time_;
}

// ---------------------------------------------------------------------------
// passes

// kd-tree build (sph_accelerations.cpp:14-33).
// Host build (tree_gpu 0), reference-exact: tree_fetch enqueues the position
// download; tree_build waits for it, builds the tree and enqueues the upload,
// so GPU work queued in between (gravity under Plummer softening) runs while
// the host builds.  Device build (tree_gpu 1): tree_fetch enqueues it, ahead
// of that GPU work, and tree_build only collects dt.  Verified device build
// (tree_gpu 2, the plain form's default): the device build, which is the
// reference's tree unless some node's particle ties its split coordinate
// with another particle of its subtree (then the reference's quicksort may
// have picked the other one); the build flags that case, and the host build
// runs for that step instead, from positions copied beside the device build.
template <class T>
void Engine<T>::tree_fetch() {
  if (tree_gpu_) {
    const long need = tree_work_ints(n_);
    if ((long)tree_work_.n < need) tree_work_.alloc(need);
    TimerScope ts(timers_, "tree_build", s_);
    int* ties = nullptr;
    if (tree_gpu_ == 2) {
      CS_HIP(hipMemsetAsync(ties_, 0, sizeof(int), s_));
      ties = ties_.p;
    }
    launch_tree_build<T>(pos_, n_, nd_, tree_, hash_, i_order_, tree_work_, pairs_, pairs_alt_, sort_tmp_, scan_tmp_,
                         s_, ties);
    boxes_stale_ = true;
    if (tree_gpu_ == 2) {
      CS_HIP(hipMemcpyAsync(h_pos_.p, pos_, (size_t)nd_ * sizeof(V4<T>), hipMemcpyDeviceToHost, s_));
      CS_HIP(hipMemcpyAsync(&h_small_[5], ties_, sizeof(int), hipMemcpyDeviceToHost, s_));
    }
  } else {
    CS_HIP(hipMemcpyAsync(h_pos_.p, pos_, (size_t)nd_ * sizeof(V4<T>), hipMemcpyDeviceToHost, s_));
  }
  CS_HIP(hipMemcpyAsync(&h_small_[0], dt_cur_, 8, hipMemcpyDeviceToHost, s_));
  CS_HIP(hipEventRecord(fetched_, s_));
}

template <class T>
void Engine<T>::tree_build() {
  CS_HIP(hipEventSynchronize(fetched_));
  std::memcpy(&dt_, &h_small_[0], 8);
  if (tree_gpu_ == 1) return;
  if (tree_gpu_ == 2) {
    int tied = 0;
    std::memcpy(&tied, &h_small_[5], sizeof(int));
    if (!tied) return;
    ++tree_fallbacks_;
  }
  const auto t0 = std::chrono::steady_clock::now();
  build_kd_tree<T>(h_pos_.p, n_, nd_, p_.n_threads, h_tree_.p, h_hash_.p);
  if (i_order_ == 0)   // lanes take particles in memory order (the reference's identity hash)
    for (int t = 0; t < nd_; ++t) h_hash_.p[t] = {(uint32_t)t, (uint32_t)t};
  timers_.add_host("tree_build_host",
                   std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  CS_HIP(hipMemcpyAsync(tree_, h_tree_.p, (size_t)(nd_ + 2) * sizeof(V4<T>), hipMemcpyHostToDevice, s_));
  boxes_stale_ = true;
  CS_HIP(hipMemcpyAsync(hash_, h_hash_.p, (size_t)nd_ * sizeof(KV), hipMemcpyHostToDevice, s_));
}

template <class T>
void Engine<T>::tree() {
  tree_fetch();
  tree_build();
}

// the subtree boxes of the current tree, when the walks use them
template <class T>
void Engine<T>::ensure_boxes() {
  if (!walk_boxes_ || !boxes_stale_) return;
  TimerScope ts(timers_, "tree_boxes", s_);
  launch_tree_boxes<T>(tree_, boxes_, p_.n_bodies, p_.fix_phantom, s_);
  boxes_stale_ = false;
}

template <class T>
void Engine<T>::solve() {
  const size_t rec = (size_t)adj_f_ * nd_;
  // the recorded ranges and counts of this evaluation (a replay with
  // adj_free_ranges set solves them afresh instead, for measuring what the
  // frozen ranges leave out of the gradient)
  if (adj_mode_ >= kAdjReplay && !(adj_mode_ == kAdjReplay && adj_free_ranges_)) {
    launch_set_range<T>(pos_, adj_range_.p + rec, nd_, s_);
    CS_HIP(hipMemcpyAsync(n_ngb_, adj_nngb_.p + rec, (size_t)nd_ * 4, hipMemcpyDeviceToDevice, s_));
    return;
  }
  ensure_boxes();
  {
    TimerScope ts(timers_, "solve_range", s_);
    WalkParams wp = walk_params();
    if (solve_cache_ > 0 || (solve_cache_ < 0 && solve_cache_pays(nd_))) {
      if (solve_rec_.n < (size_t)kSolveCacheK * nd_) solve_rec_.alloc((size_t)kSolveCacheK * nd_);
      wp.solve_cache = solve_rec_.p;
    }
    launch_solve_range<T>(pos_, n_ngb_, hash_, tree_, nd_, wp, s_);
  }
  if (adj_mode_ == kAdjRecord) {
    launch_get_range<T>(pos_, adj_range_.p + rec, nd_, s_);
    CS_HIP(hipMemcpyAsync(adj_nngb_.p + rec, n_ngb_, (size_t)nd_ * 4, hipMemcpyDeviceToDevice, s_));
  }
}

// list slot layout (sph_accelerations.cpp:58-72); the total is checked
// against the allocation before any list is written
template <class T>
void Engine<T>::offsets() {
  {
    TimerScope ts(timers_, "offsets_scan", s_);
    launch_exclusive_scan(n_ngb_, ngb_offset_, nd_, 1, scan_tmp_, s_);
  }
  check_total();
}

template <class T>
long Engine<T>::check_total() {
  int tot = 0;
  CS_HIP(hipMemcpyAsync(&h_small_[1], ngb_offset_.p + nd_, 4, hipMemcpyDeviceToHost, s_));
  sync();
  std::memcpy(&tot, &h_small_[1], 4);
  if (tot < 0 || tot > cap_)
    throw std::runtime_error("neighbour list overflow: ngb_tot exceeds 100 n_dev (sph_accelerations.cpp:66-72)");
  ngb_tot_ = tot;
  tot_stale_ = false;
  return tot;
}

// dev_find_ngb; if a particle finds more ids than its slot holds (the
// reference's count/list mismatch), rebuild the offsets from the true
// counts and walk again
template <class T>
void Engine<T>::find() {
  ensure_boxes();
  const WalkParams wp = walk_params();
  if (tot_stale_) check_total();   // offsets set from outside since the last scan
  for (int attempt = 0; attempt < 2; ++attempt) {
    {
      TimerScope ts(timers_, "find_ngb", s_);
      CS_HIP(hipMemsetAsync(flags_, 0, 16, s_));
      launch_find_ngb<T>(pos_, tree_, hash_, ngb_offset_, attempt == 0 ? n_ngb_.p : ngb_count_.p, ngb_count_,
                         list_, flags_, nd_, wp, 0, s_);
    }
    int fl[3] = {0, 0, 0};
    CS_HIP(hipMemcpyAsync(&h_small_[2], flags_, 12, hipMemcpyDeviceToHost, s_));
    sync();
    std::memcpy(fl, &h_small_[2], 12);
    phantom_reads_ = fl[1];
    phantom_ngb_ = fl[2];
    if (!fl[0]) return;
    ++overflow_reruns_;
    launch_exclusive_scan(ngb_count_, ngb_offset_, nd_, 1, scan_tmp_, s_);
    check_total();
  }
  throw std::runtime_error("find_ngb: slot overflow persisted after recount");
}

template <class T>
void Engine<T>::density() {
  refresh_if_dirty();
  TimerScope ts(timers_, "density", s_);
  if constexpr (kF32) {
    if (mixed())
      launch_density_mixed(pos_, vel_, pos_lo_, hydro_, dots_, hash_, ngb_offset_, ngb_count_, list_, nd_, s_);
    else
      launch_density<float>(pos_, vel_, hydro_, dots_, hash_, ngb_offset_, ngb_count_, list_, nd_, s_);
  } else {
    launch_density<double>(pos_, vel_, hydro_, dots_, hash_, ngb_offset_, ngb_count_, list_, nd_, s_);
  }
}

template <class T>
void Engine<T>::eos() {
  TimerScope ts(timers_, "eos", s_);
  const EosConst c = eos_const();
  if constexpr (kF32) {
    if (mixed()) {
      launch_eos_mixed((F4*)hydro_.p, u64_, (const F4*)pos_.p, (const F4*)dots_.p, (F4*)hydro2_.p, n_, c, s_);
      return;
    }
  }
  launch_eos<T>(hydro_, nd_, c, s_);
}

template <class T>
void Engine<T>::hydro() {
  TimerScope ts(timers_, "hydro", s_);
  const T a = T(p_.av_alpha), b = T(p_.av_beta);
  if constexpr (kF32) {
    if (mixed())
      launch_hydro_mixed(pos_, vel_, pos_lo_, hydro_, dots_, bdots_, hydro2_, acc_, hash_, ngb_offset_, ngb_count_,
                         list_, nd_, a, b, s_);
    else
      launch_hydro<float>(pos_, vel_, hydro_, dots_, bdots_, hydro2_, acc_, hash_, ngb_offset_,
                                 ngb_count_, list_, nd_, a, b, s_);
  } else {
    launch_hydro<double>(pos_, vel_, hydro_, dots_, bdots_, hydro2_, acc_, hash_, ngb_offset_,
                                ngb_count_, list_, nd_, a, b, s_);
  }
}

// reaction scatter (sph_accelerations.cpp:104-115) as a gather: the list
// slots grouped by target j, each group in ascending slot = ascending i order
// (the reference's serial order), and each j sums its own reactions
template <class T>
void Engine<T>::build_transpose() {
  if (tot_stale_) check_total();
  launch_reaction_transpose(ngb_offset_, ngb_count_, list_, nd_, owner_, tr_start_, tr_end_, pairs_, pairs_alt_,
                            scan_tmp_, s_);
}

template <class T>
void Engine<T>::reaction() {
  TimerScope ts(timers_, "reaction", s_);
  build_transpose();
  if constexpr (kF32) {
    if (mixed()) {
      launch_reaction_mixed(pos_, pos_lo_, bdots_, acc_, pairs_, tr_start_, tr_end_, n_, s_);
      return;
    }
  }
  launch_reaction<T>(pos_, bdots_, acc_, pairs_, tr_start_, tr_end_, n_, s_);
}

template <class T>
void Engine<T>::sph_accelerations() {
  tree();
  sph_after_tree();
}

template <class T>
void Engine<T>::sph_after_tree() {
  solve();
  offsets();
  find();
  density();
  eos();
  hydro();
  reaction();
}

// CUDA_gravity.cpp:4-71 + sph_grav_kernel.cu
// Gravity.  async: the pack runs on the main stream (it reads the predicted
// positions before the h-solve rewrites their range words), the kernels on the
// gravity stream gs_, so the SPH passes queued next on the main stream run
// beside them; gravity_join() makes the main stream wait for the result.
template <class T>
void Engine<T>::gravity(bool async) {
  refresh_if_dirty();
  if (adj_mode_ == kAdjReplayStoredGravity) {
    CS_HIP(hipMemcpyAsync(grav_, adj_grav_.p + (size_t)adj_f_ * nd_, (size_t)nd_ * sizeof(V4<T>),
                          hipMemcpyDeviceToDevice, s_));
    return;
  }
  gravity_pack();
  hipStream_t gs = s_;
  if (async) {
    CS_HIP(hipEventRecord(ev_fork_, s_));
    CS_HIP(hipStreamWaitEvent(gs_, ev_fork_, 0));
    gs = gs_;
  }
  gravity_kernels(gs);
  if (adj_mode_ == kAdjRecord)
    CS_HIP(hipMemcpyAsync(adj_grav_.p + (size_t)adj_f_ * nd_, grav_, (size_t)nd_ * sizeof(V4<T>),
                          hipMemcpyDeviceToDevice, gs));
  if (async) {
    CS_HIP(hipEventRecord(ev_join_, gs_));
    grav_pending_ = true;
  }
}

template <class T>
void Engine<T>::gravity_join() {
  if (!grav_pending_) return;
  CS_HIP(hipStreamWaitEvent(s_, ev_join_, 0));
  grav_pending_ = false;
}

template <class T>
void Engine<T>::gravity_pack() {
  if constexpr (kF32) {
    if (mixed()) {
      launch_gravity_pack_mixed(pos64_, pos_, vel_, ghi_, glo_, gvel32_, n_, nd_, s_);
      return;
    }
  }
  launch_gravity_pack<T>(pos_, vel_, gpos_, gvel_, n_, nd_, s_);
}

template <class T>
void Engine<T>::gravity_kernels(hipStream_t gs) {
  TimerScope ts(timers_, "gravity", gs);
  if constexpr (kF32) {
    if (mixed()) {
      launch_gravity_mixed(ghi_, glo_, gvel32_, (F4*)grav_.p, gpart_, gmass_, gmom_, (float)grav_far_ratio_, (float)grav_far_grade_, nd_, nd_,
                           n_, p_.grav_mode,
                           p_.grav_eps,
                           grav_split_only_ | (grav_unit_mass_ ? 0 : 2) | (grav_expand_ ? 0 : 4) |
                               (grav_far_ ? 16 : 0) | (grav_far_ && grav_far_half_ ? 64 : 0),
                           gs, grav_parts_);
      return;
    }
  }
  if constexpr (!kF32) {
    if (p_.grav_mode == 1) {
      const int parts = grav_parts_ ? grav_parts_ : gravity64_forward_parts(nd_);
      if (parts > 1 && gpart64_.n < (size_t)parts * nd_) gpart64_.alloc((size_t)parts * nd_);
      launch_gravity64_plummer(gpos_, grav_, gpart64_.p, nd_, nd_, (double)p_.grav_eps, gs, parts);
      return;
    }
  }
  launch_gravity<T>(gpos_, gvel_, grav_, nd_, nd_, p_.grav_mode, T(p_.grav_eps), exact_rsqrt_, gs, grav_sums_);
}

template <class T>
void Engine<T>::add_gravity() {
  launch_add_gravity<T>(bdots_, grav_, pos_, vel_, n_, partials_, force6_, s_);
}

template <class T>
void Engine<T>::timestep(double* out) {
  if (adj_mode_ >= kAdjReplay) {   // the recorded timestep of this evaluation
    CS_HIP(hipMemcpyAsync(out, &adj_dt_[adj_f_], 8, hipMemcpyHostToDevice, s_));
    sync();
    return;
  }
  timestep_compute(out);
  if (adj_mode_ == kAdjRecord) {
    sync();
    CS_HIP(hipMemcpy(&adj_dt_[adj_f_], out, 8, hipMemcpyDeviceToHost));
  }
}

template <class T>
void Engine<T>::timestep_compute(double* out) {
  if constexpr (kF32) {
    if (mixed()) {
      launch_timestep_mixed(pos_, u64_, hydro2_, bdots_, n_, p_.cn_vel, p_.cn_eth, partials_, out, s_);
      return;
    }
  }
  launch_timestep<T>(pos_, hydro_, hydro2_, bdots_, n_, p_.cn_vel, p_.cn_eth, kF32 ? 1 : 0, partials_, out, s_);
}

template <class T>
void Engine<T>::predict() {
  refresh_if_dirty();
  if constexpr (kF32) {
    if (mixed()) {
      launch_predict_mixed(pos64_, vel64_, u64_, old64_, pos_, pos_lo_, vel_, hydro_, bdots_, grav_, dt_cur_, n_, s_);
      return;
    }
  }
  launch_predict<T>(pos_, vel_, hydro_, bdots_, grav_, old_vu_, dt_cur_, n_, s_);
}

template <class T>
void Engine<T>::correct(const double* dt_new, const double* dt_old) {
  if constexpr (kF32) {
    if (mixed()) {
      launch_correct_mixed(vel64_, u64_, old64_, vel_, hydro_, bdots_, dt_new, dt_old, n_, s_);
      return;
    }
  }
  launch_correct<T>(vel_, hydro_, bdots_, old_vu_, dt_new, dt_old, n_, s_);
}

template <class T>
void Engine<T>::stats_pass() {
  if constexpr (kF32) {
    if (mixed()) {
      launch_stats_mixed(pos64_, vel64_, u64_, grav_, n_, partials_, stats6_, s_);
      launch_diagnostics<double>(pos64_, vel64_, n_ngb_, ngb_count_, n_, diag_part_, diag_, s_);
      return;
    }
  }
  launch_stats<T>(pos_, vel_, hydro_, grav_, n_, partials_, stats6_, s_);
  launch_diagnostics<T>(pos_, vel_, n_ngb_, ngb_count_, n_, diag_part_, diag_, s_);
}

// sort_bodies.cpp:119-186 on the GPU: 3*bits-bit Peano-Hilbert keys sorted
// stably (low 32 bits, then the high bits), then the seven state arrays
// gathered through the permutation
template <class T>
void Engine<T>::sort_bodies() {
  refresh_if_dirty();
  TimerScope ts(timers_, "ph_sort", s_);
  launch_ph_keys<T>(pos_, n_, p_.ph_bits, key_hi_, key_lo_, (int*)partials_.p, s_);
  // (key_lo, i) -> sort -> (key_hi[perm], perm) -> sort
  KV* kv = pairs_.p;
  KV* alt = pairs_alt_.p;
  launch_iota_kv(kv, n_, s_);
  launch_set_keys(kv, key_lo_, n_, s_);
  launch_radix_sort(kv, alt, n_, 32, sort_tmp_, s_);
  // second pass keys: key_hi of the permuted order
  launch_gather<uint32_t>(key_hi_, key_lo_, kv, n_, s_);
  launch_set_keys(kv, key_lo_, n_, s_);
  const int hi_bits = std::max(0, 3 * p_.ph_bits - 32);
  launch_radix_sort(kv, alt, n_, hi_bits, sort_tmp_, s_);
  auto permute = [&](DBuf<V4<T>>& a) {
    launch_gather<V4<T>>(a, tmp4_, kv, n_, s_);
    CS_HIP(hipMemcpyAsync(a, tmp4_, (size_t)n_ * sizeof(V4<T>), hipMemcpyDeviceToDevice, s_));
  };
  permute(pos_);
  permute(vel_);
  permute(bdots_);
  permute(hydro_);
  permute(dots_);
  permute(hydro2_);
  permute(grav_);
  if constexpr (kF32) {
    if (mixed()) {
      for (DBuf<D4>* a : {&pos64_, &vel64_}) {
        launch_gather<D4>(*a, old64_, kv, n_, s_);
        CS_HIP(hipMemcpyAsync(a->p, old64_, (size_t)n_ * sizeof(D4), hipMemcpyDeviceToDevice, s_));
      }
      launch_gather<double>(u64_, (double*)tmp4_.p, kv, n_, s_);
      CS_HIP(hipMemcpyAsync(u64_, tmp4_, (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
      launch_refresh_mixed(pos64_, vel64_, u64_, (F4*)pos_.p, pos_lo_, (F4*)vel_.p, (F4*)hydro_.p, n_, s_);
    }
  }
}

template <class T>
void Engine<T>::init() {
  adj_f_ = 0;
  sph_accelerations();
  gravity();
  add_gravity();
  stats_pass();
  timestep(dt_cur_);
}

// step.cpp:13-147
template <class T>
void Engine<T>::step() {
  ++iteration_;
  adj_f_ = iteration_ - adj_iter0_;
  // no re-sort while recording or replaying an adjoint trajectory (the
  // records are indexed by particle)
  if (p_.ph_period > 0 && iteration_ % p_.ph_period == 0 && adj_mode_ == kAdjOff) sort_bodies();
  predict();
  // Under Plummer softening gravity reads only the predicted positions and
  // masses, so it is queued before the SPH passes and runs while the host
  // builds the kd-tree; the results are the same as in the reference's order
  // (sph_accelerations, then CUDA_gravity, step.cpp:65,94).  The spline
  // softening needs the new ranges and keeps that order.
  const bool grav_first = p_.grav_mode == 1;
  // overlap < 0 (the default) is on: the mixed form's SPH passes take ~12 ms
  // at N = 2^20 and slow its gravity by ~4 ms when they run beside it
  // (288.5 -> 282.9 ms/step; with the far field 106.2 -> 101.6); FP64 827 ->
  // 814 ms/step at 2^20, 58.2 -> 53.4 at 262 144; plain, with its checked
  // device tree build, 3.63 -> 2.22 ms/step at 16 384, 12.8 -> 8.5 at 65 536,
  // 81.5 -> 71.3 at 262 144, 1056 -> 1036 at 2^20
  const bool overlap = overlap_ != 0;
  // on its own stream, gravity starts before the device tree build, which
  // then runs beside it (the build is on the SPH passes' path, not gravity's)
  const bool grav_now = grav_first && overlap && tree_gpu_;
  if (grav_now) gravity(true);
  tree_fetch();            // also reads back dt_
  if (grav_first && !grav_now) gravity(overlap);
  tree_build();
  sph_after_tree();
  if (adj_capture_ >= 0) adj_capture(*adj_eval_[adj_capture_]);
  if (!grav_first) gravity(false);
  gravity_join();
  add_gravity();
  global_time_ += dt_;
  stats_pass();
  timestep(dt_next_);
  correct(dt_next_, dt_cur_);
  CS_HIP(hipMemcpyAsync(dt_cur_, dt_next_, 8, hipMemcpyDeviceToDevice, s_));
}

template <class T>
void Engine<T>::pass(int id) {
  switch (id) {
    case SPH_TREE: tree(); break;
    case SPH_SOLVE: solve(); break;
    case SPH_OFFSETS: offsets(); break;
    case SPH_FIND: find(); break;
    case SPH_DENSITY: density(); break;
    case SPH_EOS: eos(); break;
    case SPH_HYDRO: hydro(); break;
    case SPH_REACTION: reaction(); break;
    case SPH_GRAVITY: gravity(); break;
    case SPH_ADD_GRAVITY: add_gravity(); break;
    case SPH_TIMESTEP: timestep(dt_cur_); break;
    case SPH_PREDICT: predict(); break;
    case SPH_CORRECT: correct(dt_cur_, dt_cur_); break;
    case SPH_STATS: stats_pass(); break;
    case SPH_PH_SORT: sort_bodies(); break;
    case SPH_SPH_ACC: sph_accelerations(); break;
    case SPH_INIT: init(); break;
    case SPH_STEP: step(); break;
    default: throw std::invalid_argument("unknown pass id");
  }
}

template class Engine<float>;
template class Engine<double>;

EngineBase* make_engine(int form, const sph_params& p, int device) {
  if (form == SPH_FORM_PLAIN || form == SPH_FORM_MIXED) return new Engine<float>(form, p, device);
  if (form == SPH_FORM_FP64) return new Engine<double>(form, p, device);
  throw std::invalid_argument("unknown form");
}

}  // namespace cudasph
