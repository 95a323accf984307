/*
 * cudasph.h -- C ABI of the MI355X-native SPH engine (libcudasph.so).
 *
 * Two groups of entry points:
 *
 *  1. The reference's own boundary, with the same names, argument order and
 *     meaning, so that the reference's host code (CUDAsph class) links
 *     against this library unchanged.  Each declaration cites the reference
 *     line it replaces.  float4 arguments are 16-byte {x,y,z,w} float records
 *     and KeyValuePair is {uint32 key, uint32 value}, 8-byte aligned; both are
 *     declared here as plain structs so this header needs no GPU headers.
 *     Errors behave as in the reference: a failing HIP call prints the error
 *     and exits the process (host_calls.cu:14-22).
 *
 *  2. The engine API (sph_engine_*): a device-resident simulation in one of
 *     three forms (plain FP32 drop-in, mixed precision, FP64), its passes,
 *     state transfer and diagnostics, plus the adjoint.  These return 0 / a
 *     byte count on success and a negative value on failure; the message is
 *     available from sph_last_error().  They never exit the process.
 *
 * All pointers in group 1 are device pointers unless the name says host.
 */
#ifndef CUDASPH_H
#define CUDASPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sph_float4 { float x, y, z, w; } sph_float4;
typedef struct sph_kv { uint32_t key, value; } sph_kv; /* KeyValuePair, radixsort.cuh:48-57 */

/* ---------------------------------------------------------------------- */
/* 1. reference boundary (CUDAsph.h:22-48, radixsort.cuh:59-61)            */

/* CUDAsph.h:24, host_calls.cu:30-32 -- no-op */
void initCUDA(void);
/* CUDAsph.h:25, host_calls.cu:55-59 -- device allocation of n bytes */
void allocateCUDAarray(void** pos, int n, int device_id);
/* CUDAsph.h:26, host_calls.cu:60-64 */
void deleteCUDAarray(void* pos, int device_id);
/* CUDAsph.h:27, host_calls.cu:65-69 -- n bytes host -> device */
void copyArrayToDevice(void* device, const void* host, int n, int device_id);
/* CUDAsph.h:28, host_calls.cu:70-74 -- n bytes device -> host */
void copyArrayFromDevice(void* host, const void* device, int n, int device_id);
/* CUDAsph.h:30-31, host_calls.cu:319-361 -- softened direct-N gravity.
 * pos = n x {x,y,z,m}, vel = n x {vx,vy,vz,range^2}; accel = n x {gx,gy,gz,phi}.
 * p and q are the reference's launch shape; this library chooses its own
 * (p must still divide n, as in the reference). */
void compute_CUDA_gravity(int numBodies, sph_float4* accel, float* pos, float* vel,
                          int p, int q, int device_id);
/* CUDAsph.h:35-39, host_calls.cu:292-317 -- 16-byte-per-body arrays */
void zallocateNBodyArrays(float** pos, int numBodies, int device_id);
void zdeleteNBodyArrays(float* pos, int device_id);
void zcopyArrayFromDevice(float* host, const float* device, int numBodies, int device_id);
void zcopyArrayToDevice(float* device, const float* host, int numBodies, int device_id);
/* CUDAsph.h:41, host_calls.cu:77-153 -- smoothing-length solve.  In place:
 * bodies_pos[i].w <- range (= 2h), n_ngb[i] <- neighbour count at that range.
 * Binds pos/vel/n_ngb/kd_tree for the following host_sph_accelerations call,
 * as the reference's texture bindings do.  kd_tree has n_bodies+1 records
 * (index 0 unused).  Returns elapsed wall seconds. */
float host_solve_range(int n_bodies, int n_nodes, float rel_err, sph_kv* bodies_hash,
                       int* bodies_map, sph_float4* bodies_pos, sph_float4* bodies_vel,
                       sph_float4* kd_tree, int* n_ngb, int device_id);
/* CUDAsph.h:42-43, host_calls.cu:155-287 -- neighbour lists, density, EOS and
 * hydro accelerations over the arrays bound by host_solve_range.  On return
 * n_ngb holds the number of ids find_ngb wrote for each particle (the
 * h-solve count except in the reference's bracket edge case), and
 * bodies_hash[t].key holds the walk length of thread t.  Returns wall seconds. */
float host_sph_accelerations(int n_bodies, int ngb_tot, float av_alpha, float av_beta,
                             sph_kv* bodies_hash, sph_kv* bodies_hash_extra, int* bodies_map,
                             int* ngb_list, int* ngb_offset, sph_float4* hydro_data,
                             sph_float4* dots_data, sph_float4* bodies_dots,
                             sph_float4* hydro_data2, float* ngb_acc, int device_id);
/* CUDAsph.h:44, host_calls.cu:364-373 */
int get_device_count(void);
/* CUDAsph.h:45, host_calls.cu:47-53 -- no-op, as in the reference */
void cuda_set_device(int device_id);
/* CUDAsph.h:46, host_calls.cu:39-45 -- 1-based device selection */
void cuda_set_device0(int device_id);
/* CUDAsph.h:47, host_calls.cu:33-38 */
int cuda_get_device(void);
/* radixsort.cuh:59-61, radixsort.cu:50-77 -- stable LSD sort of
 * (key,value) pairs by the low `bits` bits of key; pData1 is scratch of the
 * same size.  Sorted output ends in pData0. */
void RadixSort(sph_kv* pData0, sph_kv* pData1, unsigned int elements, unsigned int bits);

/* No counterpart in the reference: RadixSort for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The same stable LSD sort of the low
 * `bits` bits of key, result in pData0, pData1 scratch of the same size; the
 * digit tables live in a fixed device array of this library (about 8.4 MB
 * per GPU, held until the process ends) instead of a hipMalloc per call, so
 * at most SPH_NOALLOC_SORT_MAX elements.  Enqueued on `stream` (a
 * hipStream_t; NULL is the null stream) and not waited for.  Every call on a
 * GPU uses that GPU's one array: calls that overlap on one GPU race.
 * Returns 0, or -1 with sph_last_error() (more than SPH_NOALLOC_SORT_MAX
 * elements, bits above 32, a null pointer, a HIP error). */
#define SPH_NOALLOC_SORT_MAX (1 << 24)
int sph_radix_sort_noalloc(sph_kv* pData0, sph_kv* pData1, unsigned int elements, unsigned int bits,
                           void* stream);

/* kd_tree.cpp:97-128 with the packing of sph_accelerations.cpp:21-33 (host
 * code, no GPU): the reference's left-balanced kd-tree of pos[0..n) (x,y,z
 * used), built with the reference's quicksort so ties resolve identically,
 * on up to n_threads host threads (0 = all).  tree receives n_dev+2 records
 * in the device layout (w = int bits of id | split<<24; far sentinels past n;
 * a zero record at n_dev+1); particle_map[t] = {t, particle at heap node t+1}
 * (kd_tree.cpp:117-128), padding entries map to themselves.  Returns 0. */
int sph_build_kd_tree(int n, int n_dev, int n_threads, const sph_float4* pos, sph_float4* tree,
                      sph_kv* particle_map);
/* the same on FP64 positions (the FP64 form's tree; w holds the int as a double) */
int sph_build_kd_tree64(int n, int n_dev, int n_threads, const double* pos4, double* tree4,
                        sph_kv* particle_map);

/* No counterpart in the reference (its DESIRED_NGB, MAX_ITER, EOS unit
 * constants and M4 softening are compile-time constants, dev_solve_range.cu:14-15,
 * dev_compute_pressure.cu:55-81): sets the values the group-1 entry points use.
 * exact_rsqrt = 1 makes compute_CUDA_gravity use a correctly rounded 1/sqrt. */
struct sph_params;
void sph_abi_set_params(const struct sph_params* p, int exact_rsqrt);

/* ---------------------------------------------------------------------- */
/* 2. engine API                                                           */

/* Parameters.  Defaults (sph_params_default) are the reference's values;
 * sph_params_preset gives the five benchmark configurations. */
typedef struct sph_params {
  int    n_bodies;       /* real particles N                                  */
  int    n_dev;          /* padded count, n_norm(N,256) (read_state.cpp:27)    */
  float  rel_err;        /* h-solve tolerance, 1e-3 (CUDA_calls.cpp:4)         */
  float  desired_ngb;    /* smoothed neighbour target, 24 (dev_solve_range.cu:14) */
  int    max_iter;       /* 20 (dev_solve_range.cu:15)                         */
  int    fix_phantom;    /* 0: the walk reads the zero node n+1 like the reference */
  float  av_alpha;       /* 1 (CUDA_calls.cpp:13)                              */
  float  av_beta;        /* 2 (CUDA_calls.cpp:14)                              */
  float  cn_vel;         /* 0.3 (compute_timestep.cpp:4)                       */
  float  cn_eth;         /* 0.2 (compute_timestep.cpp:5)                       */
  int    grav_mode;      /* 0: M4 spline at range (reference), 1: Plummer eps  */
  float  grav_eps;       /* Plummer softening length                           */
  double eos_density_unit, eos_energy_unit, eos_pressure_unit; /* dev_compute_pressure.cu:55-81 */
  double eos_amu, eos_boltzmann, eos_rad_const, eos_gamma, eos_temp_base;
  int    ph_period;      /* Peano-Hilbert re-sort period, 100 (step.cpp:26)    */
  int    ph_bits;        /* bits per dimension, 18 (sort_bodies.cpp:3)         */
  int    eos_newton;     /* 0: Ferrari closed form (reference), 1: FP64 Newton */
  int    n_threads;      /* host threads for the kd-tree build (0 = all)       */
} sph_params;

enum {
  SPH_FORM_PLAIN = 0,    /* FP32 drop-in with the reference's arithmetic       */
  SPH_FORM_MIXED = 1,    /* FP32 pair math, compensated sums, hi/lo positions,
                            FP64 integration, EOS Newton and diagnostics      */
  SPH_FORM_FP64 = 2      /* every pass in FP64                                 */
};

/* pass ids (same numbering as the oracle) */
enum {
  SPH_TREE = 0, SPH_SOLVE = 1, SPH_OFFSETS = 2, SPH_FIND = 3, SPH_DENSITY = 4,
  SPH_EOS = 5, SPH_HYDRO = 6, SPH_REACTION = 7, SPH_GRAVITY = 8,
  SPH_ADD_GRAVITY = 9, SPH_TIMESTEP = 10, SPH_PREDICT = 11, SPH_CORRECT = 12,
  SPH_STATS = 13, SPH_PH_SORT = 14, SPH_SPH_ACC = 15, SPH_INIT = 16, SPH_STEP = 17
};

void        sph_params_default(sph_params* p, int n_bodies);
/* config 1..5 of the benchmark set; n_bodies <= 0 takes the config's own N */
int         sph_params_preset(sph_params* p, int config, int n_bodies);
/* Seeded initial conditions for config 1..5 in FP64: pos4 = {x,y,z,range},
 * vel4 = {vx,vy,vz,m}, hydro4 = {rho,P,u,mu}; n records each.  Identical on
 * every x86-64 host with the same libm. */
int         sph_ic_generate(int config, int n, uint64_t seed, double* pos4, double* vel4,
                            double* hydro4);
/* The velocity grid behind config 5's particle velocities, for the same n
 * and seed: three 128^3 fields (component-major, index (x*128 + y)*128 + z
 * over [-1,1)^3, periodic) of white noise filtered by k^-2, so P(k) ~ k^-4,
 * before the trilinear interpolation and the scaling to rms Mach 2.
 * field3 receives 3*128^3 doubles.  Returns 0, or -1 on bad arguments. */
int sph_ic_turbulent_field(int n, uint64_t seed, double* field3);

const char* sph_last_error(void);

void*  sph_engine_create(int form, const sph_params* p, int device_id);
void   sph_engine_destroy(void* e);
/* State arrays by name, in the engine's real type (float for PLAIN, double
 * for FP64, and for MIXED: float for the SPH arrays plus "pos64", "vel64",
 * "u64" in double).  Names: pos vel hydro dots bdots hydro2 grav (4 reals per
 * particle), tree (n_dev+2 records), hash, n_ngb, ngb_count, ngb_offset
 * (n_dev+1), ngb_list, ngb_acc.  Returns bytes copied or < 0. */
long   sph_engine_set(void* e, const char* name, const void* host, long n_bytes);
long   sph_engine_get(void* e, const char* name, void* host, long max_bytes);
/* Scalars.  Read and write: dt, global_time, iteration; exact_rsqrt (plain
 * form: correctly rounded 1/sqrt in gravity, the CPU restatement's bits; 0
 * is the hardware rsq the reference uses); i_order (per-particle kernels in
 * memory order 0 or kd-tree heap order 1); overlap (1, the default: gravity
 * on a second stream beside the SPH passes, 0: in order; the results are the
 * same bits either way);
 * timing (per-kernel HIP-event timers);
 * tree_gpu (1: the kd-tree built on the device, the mixed and FP64 forms'
 * default -- the same tree as the host build when no two particles share a
 * coordinate, otherwise the same neighbour sets; 0: the host build,
 * bit-identical to kd_tree.cpp:39-128; 2, the plain form's default: the
 * device build, checked at every node for a particle that shares the node's
 * split coordinate, the only case where the two builds can differ, and the
 * host build for that step when one does -- the reference's tree either way;
 * read-only tree_fallbacks counts those steps);
 * walk_boxes (1, the mixed and FP64 forms' default: the tree walks also skip
 * a child whose subtree bounding box lies beyond the range -- the same
 * neighbours in the same order, so the same results, with fewer nodes
 * visited; 0, the plain form's default: the reference's walk, whose node
 * counts it leaves in bodies_hash.key).
 * Mixed-form gravity measurement switches, all results within the form's
 * tolerance: grav_split_only (1: split hi/lo differences for every pair),
 * grav_unit_mass (0: the general-mass body even when every mass is equal),
 * grav_expand (0: the equal-mass tile body on r instead of the expanded one),
 * grav_parts (the pairwise kernel's j-parts, 1..8; 0, the default: 8, or
 * fewer where 8 would take most of the GPU's block slots in one round -- 4
 * at 65 536 particles, leaving the SPH passes beside it room on every CU;
 * the sums' grouping follows it; also the FP64 form's forward gravity;
 * read-only grav_parts_auto: the count the default takes at this size).
 * solve_cache (the h-solve's candidate cache: 1 on, 0 off, -1, the default:
 * on below 65 536 particles; the same bits either way).
 * Plain-form Plummer gravity, the same bits either way: grav_sums (1: each
 * particle's four sums in four wavefronts, 2: in two, 0: in one lane; -1,
 * the default: four wavefronts up to 16 384 particles, two up to 65 536).
 * Mixed-form gravity options: grav_far (1: an order-4 multipole far field for
 * each cell of 16384, 4096, 1024 or 256 records, or 64-record sub-tile, that a
 * whole wavefront sees from beyond grav_far_ratio of its radii, default 8,
 * coarsest first; not pairwise, so off by default); grav_far_half (1, the
 * default: a sub-tile not taken whole is tried as two 32-record halves);
 * grav_far_grade (the ratio at a level of C records scaled by
 * (C / 16384)^grade, default 0).
 * Read only: ngb_tot, overflow_reruns, cap, real_bytes, form.  Unknown names
 * return NaN (get) or -1 with sph_last_error() (set). */
double sph_engine_get_scalar(void* e, const char* name);
int    sph_engine_set_scalar(void* e, const char* name, double v);
int    sph_engine_pass(void* e, int pass);
/* run `steps` full steps (step.cpp:13-147) */
int    sph_engine_run(void* e, int steps);
int    sph_engine_sync(void* e);
/* T U W E |J| Fx Fy Fz Tx Ty Tz of the last STATS / ADD_GRAVITY passes */
int    sph_engine_stats(void* e, double* out11);

/* Indices of sph_engine_statistics' output. */
enum {
  SPH_STAT_T = 0, SPH_STAT_U, SPH_STAT_W, SPH_STAT_E, SPH_STAT_J,    /* as sph_engine_stats */
  SPH_STAT_FX, SPH_STAT_FY, SPH_STAT_FZ, SPH_STAT_TX, SPH_STAT_TY, SPH_STAT_TZ,
  SPH_STAT_JX, SPH_STAT_JY, SPH_STAT_JZ,                              /* angular momentum      */
  SPH_STAT_XMIN, SPH_STAT_YMIN, SPH_STAT_ZMIN,                        /* computational box     */
  SPH_STAT_XMAX, SPH_STAT_YMAX, SPH_STAT_ZMAX,
  SPH_STAT_MTOT,                                                      /* total mass            */
  SPH_STAT_XCOM, SPH_STAT_YCOM, SPH_STAT_ZCOM,                        /* centre of mass        */
  SPH_STAT_VXCOM, SPH_STAT_VYCOM, SPH_STAT_VZCOM,                     /* its velocity          */
  SPH_STAT_RCM, SPH_STAT_VCM,                                         /* |Rcm|, |Vcm|          */
  SPH_STAT_NGB_MEAN, SPH_STAT_NGB_SIGMA, SPH_STAT_NGB_MIN, SPH_STAT_NGB_MAX,  /* h-solve counts */
  SPH_STAT_COUNT_MISMATCH,   /* particles whose list length differs from the h-solve count */
  SPH_STAT_MAX_COUNT,        /* longest neighbour list                                      */
  SPH_STAT_PHANTOM_READS,    /* walks of the last find_ngb that read node n+1 (fewer with walk_boxes) */
  SPH_STAT_PHANTOM_NGB,      /* ... and took its record as a neighbour (only when n = n_dev) */
  SPH_STAT_OVERFLOW_RERUNS,  /* find_ngb re-walks after a slot overflow, since creation     */
  SPH_STAT_ITERATION, SPH_STAT_TIME,
  SPH_STAT_COUNT
};
/* system_statistics.cpp:4-82: what the reference prints every step --
 * energies, box, centre of mass, |J| and the neighbour-count mean, sigma, min
 * and max (over the real particles' h-solve counts n_ngb, as there) -- plus
 * the list diagnostics above, from the last STATS pass (in step(): after the
 * gravity is added, before the corrector, as step.cpp:125).  Sums are in
 * double (the reference sums the energies in float).  out receives
 * SPH_STAT_COUNT doubles.  Returns SPH_STAT_COUNT, or -1. */
int    sph_engine_statistics(void* e, double* out);
/* Per-kernel GPU milliseconds accumulated since the last reset, measured with
 * HIP events on the engine's stream; names[i] is a static string. */
int    sph_engine_timers(void* e, const char** names, double* ms, long* calls, int max_n);
void   sph_engine_reset_timers(void* e);

/* Reverse mode of the FP64 form (form SPH_FORM_FP64, Plummer gravity).  No
 * counterpart in the reference.  The differentiated map runs init() and
 * `steps` step()s from the engine's current state and returns
 *     L = sum_i m_i |v_i|^2 / 2 + sum_i rho_i^2      (final state, real particles)
 * with the smoothing ranges, h-solve counts and timesteps of every force
 * evaluation held at the values of the recorded run.
 *  sph_adjoint_forward:  runs and records the map (state checkpointed every
 *                        `ckpt_every` steps, per-evaluation gravity stored);
 *                        *loss receives L.
 *  sph_adjoint_backward: dL/dx0, dL/dv0 (n x 3 doubles each) and dL/du0 (n
 *                        doubles) of the recorded run into host arrays (any may
 *                        be NULL); recomputes each checkpoint segment, then
 *                        puts back the state, dt, step count and time that
 *                        sph_adjoint_forward left, so a following step()
 *                        continues the recorded run.
 *  sph_adjoint_replay:   L of the recorded (frozen) map from another initial
 *                        state: x0, v0 (n x 3) and u0 (n) host arrays, NULL =
 *                        the recorded run's own.  Leaves the engine in the
 *                        replayed final state (time and step count as at the
 *                        end of the recorded run).  With the engine scalar
 *                        adj_free_ranges = 1 the replay solves the ranges and
 *                        counts afresh at every evaluation (timesteps still
 *                        recorded): the map whose derivative the frozen-range
 *                        gradient approximates.
 * With the engine scalar adj_implicit_h = 1, sph_adjoint_backward adds each
 * evaluation's h-solve derivative (dr/dx through the smoothed-count equation)
 * and returns the gradient of that re-solved map instead of the frozen one
 * (meaningful where the h-solve is converged tightly, e.g. rel_err 1e-10).
 * All return 0, or -1 with sph_last_error() set. */
int    sph_adjoint_forward(void* e, int steps, int ckpt_every, double* loss);
int    sph_adjoint_backward(void* e, double* grad_x, double* grad_v, double* grad_u);
int    sph_adjoint_replay(void* e, const double* x0, const double* v0, const double* u0, double* loss);

/* Reference state file (read_state.cpp:5-67 / write_state.cpp:6-61): 124-byte
 * header + 23 doubles per particle.  The engine keeps the header it reads and
 * writes it back with n_bodies, iteration_number, global_time and dt from the
 * run (the other fields, t_end, dt_out, snapshot_number, ..., as read). */
int    sph_state_file_count(const char* path);   /* N from the header, or < 0 */
int    sph_engine_read_state(void* e, const char* path);
int    sph_engine_write_state(void* e, const char* path);
/* CUDAsph::output (output.cpp:3-15) into directory `dir`: when global_time >
 * snapshot_number * dt_out and dt_out > 0, snapshot_number += 1 and
 * out%04d.bin is written; restart.bin is written every call.  Returns the
 * snapshot number written, 0 if only restart.bin, -1 on error. */
int    sph_engine_output(void* e, const char* dir);

/* Copies `bytes` between two fresh device buffers twice with a plain copy
 * kernel (k_copy16) and returns the milliseconds of the second copy: the
 * reference point for scaling the FETCH_SIZE / WRITE_SIZE counters of
 * rocprofv3.  No counterpart in the reference. */
double sph_copy_calibration(long bytes);
/* Measured FP32 FMA rate in TFLOP/s (2 flops per FMA) on 8 wavefronts per
 * SIMD of every CU, single-lane (unpacked) FMAs, the second of two launches
 * timed: form 0, 8 chains of v_fma_f32 with scalar coefficients; form 1, 16
 * chains of the accumulate form v_fmac_f32 with vector operands.  The
 * achievable peak beside the published one.  No counterpart in the reference. */
double sph_fma_peak_form(int form);
/* Pairs per second of the mixed gravity kernel's expanded equal-mass pair
 * body alone (11 VALU + rsq per pair) on records held in registers, 8
 * wavefronts per SIMD of every CU, the second of two launches timed; variant 1
 * replaces the rsq by a multiply.  No counterpart in the reference. */
double sph_pair_body_probe(int variant);

#ifdef __cplusplus
}
#endif
#endif
