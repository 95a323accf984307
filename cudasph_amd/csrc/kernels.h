// Host-side declarations of the kernel launchers (implemented in the .hip
// files).  Every launcher enqueues on the given stream and checks the launch.
#pragma once

#include "common.h"

namespace cudasph {

struct WalkParams {
  int n_nodes;        // real particles (the reference passes n_bodies, CUDA_calls.cpp:5)
  int fix_phantom;
  float rel_err;
  float desired_ngb;
  int max_iter;
  // subtree bounding boxes (launch_tree_boxes), or null: with them the walk
  // also skips a child whose subtree box lies beyond the range.  The same
  // neighbours in the same order; fewer nodes visited (the walk length)
  const F4* boxes = nullptr;
  // h-solve candidate cache, or null: kSolveCacheK records of V4<T> per
  // lane, record k of lane t at k * n_dev + t (launch_solve_range)
  void* solve_cache = nullptr;
};
constexpr int kSolveCacheK = 128;

struct EosConst;  // sph_math.h

// ---- sph_passes.hip ----
template <class T>
void launch_solve_range(V4<T>* pos, int* n_ngb, const KV* hash, const V4<T>* tree, int n_dev,
                        const WalkParams& wp, hipStream_t s);
bool solve_cache_pays(int n_dev);
template <class T>
void launch_find_ngb(const V4<T>* pos, const V4<T>* tree, KV* hash, const int* offset, const int* cap_count,
                     int* count_out, int* list, int* flags, int n_dev, const WalkParams& wp, int clamp,
                     hipStream_t s);
// plain and FP64 forms
template <class T>
void launch_density(const V4<T>* pos, const V4<T>* vel, V4<T>* hydro, V4<T>* dots, const KV* hash,
                    const int* offset, const int* count, const int* list, int n_dev, hipStream_t s);
template <class T>
void launch_eos(V4<T>* hydro, int n_dev, const EosConst& c, hipStream_t s);
// mixed form: FP64 pressure and the per-particle viscosity factors into hydro2.yzw
void launch_eos_mixed(F4* hydro, const double* u64, const F4* pos, const F4* dots, F4* hydro2, int n,
                      const EosConst& c, hipStream_t s);
// mixed form: FP32 pair arithmetic (f32k), Kahan sums, split positions
void launch_density_mixed(const F4* pos, const F4* vel, const F4* pos_lo, F4* hydro, F4* dots, const KV* hash,
                          const int* offset, const int* count, const int* list, int n_dev, hipStream_t s);
void launch_hydro_mixed(const F4* pos, const F4* vel, const F4* pos_lo, const F4* hydro, const F4* dots, F4* bdots,
                        F4* hydro2, float* ngb_acc, const KV* hash, const int* offset, const int* count,
                        const int* list, int n_dev, float alpha, float beta, hipStream_t s);
template <class T>
void launch_hydro(const V4<T>* pos, const V4<T>* vel, const V4<T>* hydro, const V4<T>* dots, V4<T>* bdots,
                  V4<T>* hydro2, T* ngb_acc, const KV* hash, const int* offset, const int* count, const int* list,
                  int n_dev, T alpha, T beta, hipStream_t s);

// ---- gravity.hip ----
// sums: for the FP32 Plummer kernel, 1 keeps each particle's four sums in
// four wavefronts (k_gravity_sums), 2 in two (k_gravity_two), 0 in one lane
// (k_gravity_lane), -1 (auto) four up to kFourSumsMaxN particles, two up to
// kTwoSumsMaxN, one lane above; the same bits either way
constexpr int kFourSumsMaxN = 16384;
constexpr int kTwoSumsMaxN = 65536;
int gravity_sums_mode(int n_i, int sums);
template <class T>
void launch_gravity(const V4<T>* gpos, const V4<T>* gvel, V4<T>* out, int n_i, int n_j, int soft, T eps,
                    int exact, hipStream_t s, int sums = -1);
// FP64 form, Plummer softening (fused pair body; the j-range split into
// `parts` parts, 0 for gravity64_forward_parts(n_i), `part` then holding
// parts * n_i records when parts > 1).  gravity64_parts(n_i) is the
// reverse pass's split: enough parts to fill the GPU when the i-blocks alone
// would not.
int gravity64_parts(int n_i);
int gravity64_forward_parts(int n_i);
void launch_gravity64_plummer(const D4* gpos, D4* out, D4* part, int n_i, int n_j, double eps, hipStream_t s,
                              int parts = 0);
// part: scratch of 2 * 8 * n_i records (per-j-part partial sums, Plummer only);
// mass_range: scratch of 2 + 2 * ceil(n_real / 256) floats ({min, max} of the
// real particles' masses, then per-block partials)
// moments: scratch of 52 floats per 64-record j sub-tile followed by 52 per
// cell of 16384, 4096, 1024 and 256 records and, with split_only bit 6, 52
// per 32-record half sub-tile (the optional multipole far
// field, split_only bit 4, taken where every lane of a wavefront is
// far_ratio radii away, the ratio at a level of C records scaled by
// (C / 16384)^far_grade; may be null otherwise)
// parts: the number of j-parts (1..8), 0 for gravity_mixed_parts(n_i, far)
int gravity_mixed_parts(int n_i, int far);
void launch_gravity_mixed(const F4* hi, const F4* lo, const F4* gvel, F4* out, F4* part, float* mass_range,
                          float* moments, float far_ratio, float far_grade, int n_i, int n_j, int n_real,
                          int soft, float eps, int split_only, hipStream_t s, int parts = 0);

// ---- sort.hip ----
// exclusive scan of in[0..n) into out[0..n]; out[n] = total.  mode 1 scans
// n_norm(in[i], 4) (the neighbour-list slot layout).  scratch holds
// scan_scratch_ints(n) ints.
long scan_scratch_ints(long n);
void launch_exclusive_scan(const int* in, int* out, long n, int mode, int* scratch, hipStream_t s);
// two such scans (mode 0) in one launch sequence: the second's input and
// output `stride` ints after the first's; scratch 2 scan_scratch_ints(n)
void launch_exclusive_scan_pair(const int* in, int* out, long n, long stride, int* scratch, hipStream_t s);
// stable LSD radix sort of KV records by key bits [0, bits); the result is
// left in `data` (alt is scratch of the same length).  scratch holds
// radix_scratch_ints(n) ints.
long radix_scratch_ints(long n);
void launch_radix_sort(KV* data, KV* alt, long n, int bits, int* scratch, hipStream_t s);
// `batch` independent sorts of n records in one sequence of launches: sort b
// on data + b * data_stride (alt likewise) with scratch + b *
// radix_scratch_ints(n)
void launch_radix_sort_batch(KV* data, KV* alt, long n, int bits, int* scratch, hipStream_t s, int batch,
                             long data_stride);
// the reaction lists: for each of the n particles j, the list slots k
// naming j in ascending k, as pairs {key = owner i, value = k} in
// pairs[start[j], end[j]) (end[j] = start[j + 1]; start holds n + 1 ints).
// tcnt holds n ints of scratch, alt as many pairs as pairs, scan_scratch
// scan_scratch_ints(n) ints.
void launch_reaction_transpose(const int* offset, const int* count, const int* list, int n, int* tcnt,
                               int* start, int* end, KV* pairs, KV* alt, int* scan_scratch, hipStream_t s);
// kv[t].key = keys[t]
void launch_set_keys(KV* kv, const uint32_t* keys, int n, hipStream_t s);
// bdots[j].xyz += ngb_acc[k] (r_i - r_j) over the reaction pairs of j, in
// ascending k (sph_accelerations.cpp:104-115)
template <class T>
void launch_reaction(const V4<T>* pos, V4<T>* bdots, const T* ngb_acc, const KV* pairs, const int* start,
                     const int* end, int n, hipStream_t s);
void launch_reaction_mixed(const F4* pos, const F4* pos_lo, F4* bdots, const float* ngb_acc, const KV* pairs,
                           const int* start, const int* end, int n, hipStream_t s);
// Peano-Hilbert keys on the bounding cube (sort_bodies.cpp:121-163); keys
// are 3*bits wide, written as (hi,lo) 32-bit halves; scratch holds
// 6 * (ceil(n / 256) + 1) reals of type T
template <class T>
void launch_ph_keys(const V4<T>* pos, int n, int bits, uint32_t* key_hi, uint32_t* key_lo, int* scratch,
                    hipStream_t s);

// ---- tree.hip ----
// The kd-tree of the first n positions (kd_tree.cpp:39-128) built on the
// device: tree[0..n_dev+1] and hash[0..n_dev) as the host builder writes them
// (heap_order 1: hash in the nodes' heap order, 0: identity).  work holds
// tree_work_ints(n) ints; kv and kv_alt 3 n records each; sort_scratch
// 3 radix_scratch_ints(n) and scan_scratch 2 scan_scratch_ints(n) ints.
long tree_work_ints(long n);
// ties (may be null): set to 1 when some node's particle shares its split
// coordinate with another particle of its subtree -- the one case in which
// the reference's quicksort may have picked the other particle as the node;
// otherwise the tree is the reference's, bit for bit.
template <class T>
void launch_tree_build(const V4<T>* pos, int n, int n_dev, V4<T>* tree, KV* hash, int heap_order, int* work,
                       KV* kv, KV* kv_alt, int* sort_scratch, int* scan_scratch, hipStream_t s,
                       int* ties = nullptr);
// Subtree bounding boxes of a kd-tree over n_nodes particles: box[2k] (low
// corner) and box[2k+1] (high corner) of node k's subtree, for k up to
// n_nodes+1; FP32, rounded outward.  Two launches: one block per subtree
// below level 10, then one block for the levels above.
template <class T>
void launch_tree_boxes(const V4<T>* tree, F4* box, int n_nodes, int fix_phantom, hipStream_t s);

// ---- integrate.hip ----
template <class T>
void launch_gravity_pack(const V4<T>* pos, const V4<T>* vel, V4<T>* gpos, V4<T>* gvel, int n, int n_dev,
                         hipStream_t s);
void launch_gravity_pack_mixed(const D4* pos64, const F4* pos, const F4* vel, F4* hi, F4* lo, F4* gvel, int n,
                               int n_dev, hipStream_t s);
template <class T>
void launch_predict(V4<T>* pos, V4<T>* vel, V4<T>* hydro, V4<T>* bdots, V4<T>* grav, V4<T>* old_vu,
                    const double* dt, int n, hipStream_t s);
template <class T>
void launch_correct(V4<T>* vel, V4<T>* hydro, const V4<T>* bdots, const V4<T>* old_vu, const double* dt_new,
                    const double* dt_old, int n, hipStream_t s);
void launch_predict_mixed(D4* pos64, D4* vel64, double* u64, D4* old64, F4* pos, F4* pos_lo, F4* vel, F4* hydro,
                          F4* bdots, F4* grav, const double* dt, int n, hipStream_t s);
void launch_correct_mixed(D4* vel64, double* u64, const D4* old64, F4* vel, F4* hydro, const F4* bdots,
                          const double* dt_new, const double* dt_old, int n, hipStream_t s);
void launch_refresh_mixed(const D4* pos64, const D4* vel64, const double* u64, F4* pos, F4* pos_lo, F4* vel,
                          F4* hydro, int n, hipStream_t s);
// dots += grav, and per-block partial sums of force and torque
template <class T>
void launch_add_gravity(V4<T>* bdots, const V4<T>* grav, const V4<T>* pos, const V4<T>* vel, int n,
                        double* partials, double* out6, hipStream_t s);
// dt = min_i 1/(1/dt_vel + 1/dt_eth) (compute_timestep.cpp:3-25); round_float
// stores the float-rounded value as the reference returns float
template <class T>
void launch_timestep(const V4<T>* pos, const V4<T>* hydro, const V4<T>* hydro2, const V4<T>* bdots, int n,
                     float cn_vel, float cn_eth, int round_float, double* partials, double* dt_out,
                     hipStream_t s);
void launch_timestep_mixed(const F4* pos, const double* u64, const F4* hydro2, const F4* bdots, int n,
                           float cn_vel, float cn_eth, double* partials, double* dt_out, hipStream_t s);
// T U W Jx Jy Jz sums
template <class T>
void launch_stats(const V4<T>* pos, const V4<T>* vel, const V4<T>* hydro, const V4<T>* grav, int n,
                  double* partials, double* out6, hipStream_t s);
void launch_stats_mixed(const D4* pos64, const D4* vel64, const double* u64, const F4* grav, int n,
                        double* partials, double* out6, hipStream_t s);
long reduce_partials(int n);   // doubles needed for partial sums of n items (x6 components)
// bounding box, mass and mass-weighted position and velocity sums, and the
// h-solve counts n_ngb (min, max, sum, sum of squares) of the n real
// particles, the number whose list length ngb_count differs from n_ngb, and
// the longest list: 19 doubles into out (layout: integrate.hip, kDg*);
// partials holds diag_partials(n) doubles
long diag_partials(int n);
template <class P>
void launch_diagnostics(const V4<P>* pos, const V4<P>* vel, const int* n_ngb, const int* ngb_count, int n,
                        double* partials, double* out, hipStream_t s);
// permutation gather of a 4-record array: dst[t] = src[perm[t]]
template <class R>
void launch_gather(const R* src, R* dst, const KV* perm, int n, hipStream_t s);
void launch_iota_kv(KV* kv, int n, hipStream_t s);
// pos[i].w <-> range[i]
template <class T>
void launch_set_range(V4<T>* pos, const double* range, int n, hipStream_t s);
template <class T>
void launch_get_range(const V4<T>* pos, double* range, int n, hipStream_t s);
void launch_copy16(const F4* src, F4* dst, long n, hipStream_t s);
// FP32 FMA probe: blocks x kBlock lanes, 8 * iters FMAs per lane (form 0:
// 8 chains of v_fma_f32 with scalar coefficients; form 1: 16 chains of the
// accumulate form v_fmac_f32 with vector operands); out holds blocks * kBlock
// floats
void launch_fma_peak(float* out, int blocks, int iters, int form, hipStream_t s);
// the expanded gravity pair body alone: blocks x kBlock lanes, 16 * rounds
// pairs per lane (variant 1: the rsq replaced by a multiply)
void launch_pair_body(float* out, int blocks, int rounds, int variant, hipStream_t s);

}  // namespace cudasph
