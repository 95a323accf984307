/*
 * CPU oracle for the SPH hot path of 3gx/CUDAsph -- C API.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load liboracle.so, and only as the reference
 * the results are compared with or as the timed CPU baseline.  The product library (cudasph_amd/) never links,
 * loads or calls anything in oracle/.
 *
 * The parameter block has the same field layout as sph_params in
 * include/cudasph.h so that one ctypes structure drives both, but it is
 * declared here independently: the oracle shares no code with the product.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  int    n_bodies;       /* real particles N                                  */
  int    n_dev;          /* padded count, n_norm(N,256) (read_state.cpp:27)    */
  float  rel_err;        /* h-solve tolerance, 1e-3 (CUDA_calls.cpp:4)         */
  float  desired_ngb;    /* smoothed count target, 24 (dev_solve_range.cu:14)  */
  int    max_iter;       /* 20 (dev_solve_range.cu:15)                         */
  int    fix_phantom;    /* 0: walk reads the zero node n+1 like the reference */
  float  av_alpha;       /* 1 (CUDA_calls.cpp:13)                              */
  float  av_beta;        /* 2 (CUDA_calls.cpp:14)                              */
  float  cn_vel;         /* 0.3 (compute_timestep.cpp:4)                       */
  float  cn_eth;         /* 0.2 (compute_timestep.cpp:5)                       */
  int    grav_mode;      /* 0: M4 spline at range (reference), 1: Plummer eps  */
  float  grav_eps;       /* Plummer softening length                           */
  double eos_density_unit, eos_energy_unit, eos_pressure_unit;
  double eos_amu, eos_boltzmann, eos_rad_const, eos_gamma, eos_temp_base;
  int    ph_period;      /* Peano-Hilbert re-sort period, 100 (step.cpp:26)    */
  int    ph_bits;        /* bits per dimension, 18 (sort_bodies.cpp:3)         */
  int    eos_newton;     /* 0: Ferrari closed form (reference), 1: FP64 Newton */
  int    n_threads;      /* OpenMP threads for the oracle (0 = runtime default) */
} orc_params;

/* passes, in step() order */
enum {
  ORC_TREE = 0, ORC_SOLVE = 1, ORC_OFFSETS = 2, ORC_FIND = 3, ORC_DENSITY = 4,
  ORC_EOS = 5, ORC_HYDRO = 6, ORC_REACTION = 7, ORC_GRAVITY = 8,
  ORC_ADD_GRAVITY = 9, ORC_TIMESTEP = 10, ORC_PREDICT = 11, ORC_CORRECT = 12,
  ORC_STATS = 13, ORC_PH_SORT = 14, ORC_SPH_ACC = 15, ORC_INIT = 16,
  ORC_STEP = 17
};

/* real_bits: 32 = reference FP32 semantics, 64 = FP64 oracle */
void*  orc_create(int real_bits, const orc_params* p);
void   orc_destroy(void* h);
/* arrays by name: pos vel hydro dots bdots hydro2 grav (4 reals each),
 * tree (n_dev+2 x 4 reals), hash (2 x u32), n_ngb ngb_count ngb_offset
 * (ints; ngb_offset has n_dev+1), ngb_list (int), ngb_acc (real).
 * Returns the byte count copied, or -1 for an unknown name. */
long   orc_get(void* h, const char* name, void* dst, long max_bytes);
long   orc_set(void* h, const char* name, const void* src, long n_bytes);
double orc_get_scalar(void* h, const char* name);
void   orc_set_scalar(void* h, const char* name, double v);
int    orc_pass(void* h, int pass);
/* gravity for i in [i0, i1) only (bounded CPU-baseline samples) */
void   orc_gravity_range(void* h, int i0, int i1);
/* statistics of the last ORC_STATS pass: T U W E |J| Fx Fy Fz Tx Ty Tz
 * (force and torque from the last ORC_ADD_GRAVITY) */
void   orc_stats(void* h, double* out11);
/* everything system_statistics.cpp:4-82 reports plus the list counters, in
 * the layout of the engine's sph_engine_statistics (40 doubles) */
void   orc_statistics(void* h, double* out40);

/* stateless helpers for unit tests */
void   orc_kernel_fns(int real_bits, int n, const double* u, double* out9);
double orc_pressure(int real_bits, const orc_params* p, double rho, double u, double mu);
double orc_find_root(int real_bits, int newton, double q, double r);
void   orc_art_visc(int real_bits, float alpha, float beta, double range_i, double range_j,
                    double ds, double drdv, const double* hd_i, const double* dd_i,
                    const double* hd_j, const double* dd_j, double* out2);
void   orc_dreal_sub(const float* x2, float y, float* out2);
long long orc_ph_key(int x, int y, int z, int bits);
/* direct O(N^2) neighbour count at each particle's own range (the brute-force
 * check of stride_setup.cpp:35-54) */
void   orc_brute_ngb(int n, const float* pos4, int* count);

#ifdef __cplusplus
}
#endif
#endif
