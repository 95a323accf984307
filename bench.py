#!/usr/bin/env python3
"""Benchmark: particle-updates/s of the SPH step at N = 1,048,576 (two n=1.5
polytropes on a parabolic encounter, 64 +- 8 neighbours, Plummer eps = 0.002),
mixed-precision form vs the plain FP32 drop-in, with energy drift measured
against an FP64 run of the same steps.

`value` is the mixed form's whole-step throughput, N * K / wall(K steps), with
the state resident in HBM (the per-step kd-tree build is inside the timed
region, as in the reference's step(): on the device in the mixed form, on the
host with its transfers in the plain form).

Extra objects on the JSON line:
  roofline      gravity kernel (the dominant kernel) of the mixed form:
                20 flop per pair-interaction x N^2 per launch / mean launch
                time from HIP events on the engine's stream, against MI355X's
                157.3 TFLOP/s FP32 peak; traffic = HBM bytes per launch from
                the committed rocprofv3 counter summary when present
  cpu_baseline  the CPU oracle (reference arithmetic, OpenMP over particles)
                on the host cores, on a bounded sample of the same workload,
                and on one core
  scale         the scale workload (turbulent uniform sphere, N = 4,194,304):
                5 mixed steps with per-pass SPH GB/s and the gravity FP32
                fraction, the same steps with the far field, and one
                reverse-mode step of the FP64 form
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# MI355X's published FP32 vector peak: 256 CUs x 4 SIMDs x 2.4 GHz x 64 flops
# per SIMD per cycle, i.e. a wave64 FMA every two cycles.  This build emits no
# packed FP32 (v_pk_*); the measured single-lane FMA loop (k_fma_peak) reaches
# 0.8 of this figure, so it is the unpacked rate, not a packed one.
PEAK_FP32_TFLOPS = 157.3
FLOP_PER_PAIR = 20
WORKLOAD = "two n=1.5 polytropes (R=1, M=0.5 each), parabolic orbit q=0.5 from separation 5; " \
           "gas+radiation EOS mu=0.6 a=1e-3; 64+-8 neighbours; Plummer eps=0.002"
# the seeded workloads of `--config` (3, the default, is the headline)
WORKLOADS = {
    1: "uniform cube [0,1)^3, m=1/N, v=0, u=1; 32+-4 neighbours; Plummer eps=0.01",
    2: "Plummer sphere (M=1, a=1, r<10), velocities from its distribution function, u for 10% of the "
       "virial support; 50+-5 neighbours; Plummer eps=0.005",
    3: WORKLOAD,
    4: "Plummer sphere as config 2 at N=65,536; 50+-5 neighbours; Plummer eps=0.005",
    5: "uniform sphere R=1, M=1, Gaussian velocity field P(k)~k^-4 on 128^3 at rms Mach 2, u=1; "
       "64+-8 neighbours; Plummer eps=0.001",
}
CONFIG = 3


def progress(msg):
    """one line on stderr per finished leg (stdout carries only the JSON line)"""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_form(cs, form, n, seed, warmup, steps, timing=True, i_order=0, scalars=None):
    # the initial conditions are put in Peano-Hilbert order (the reference's
    # periodic sort_bodies) so neighbour gathers read nearby memory
    e = cs.make_engine(form, CONFIG, n, seed, ph_sort=True, i_order=i_order)
    e.set_scalar("timing", 1 if timing else 0)
    for k, v in (scalars or {}).items():
        e.set_scalar(k, v)
    e.init()
    energies = [e.stats()["E"]]
    for _ in range(warmup):
        e.step()
        energies.append(e.stats()["E"])
    e.reset_timers()
    e.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        e.step()
        energies.append(e.stats()["E"])
    e.sync()
    wall = time.perf_counter() - t0
    timers = e.timers()
    st = e.stats()
    kbar = float(e.get("ngb_count")[:n].mean())
    import numpy as np
    # the list pass leaves each lane's count of visited tree nodes in the hash
    # keys (dev_find_ngb.cu:84-85)
    walks = float(e.get("hash")[:, 0].astype(np.int64).sum())
    v = (e.get("vel64") if form == "mixed" else e.get("vel"))[:n].astype(np.float64)
    mom = (v[:, :3] * v[:, 3:4]).sum(0)
    out = {"wall_s": wall, "kbar": kbar, "walks": walks, "energies": energies, "timers": timers, "dt": e.scalar("dt"),
           "momentum": mom, "momentum_scale": float((np.linalg.norm(v[:, :3], axis=1) * v[:, 3]).sum()),
           "angular_momentum": st["J"]}
    if timing and form == "mixed":
        # gravity runs beside the SPH passes, so the timed steps' per-pass
        # times include the co-running gravity; two more steps in order,
        # after everything above was taken, time each mixed-form pass alone
        # (the SPH rates are reported for that form)
        out["timers_in_order"] = timers_in_order(e, 2)
    e.close()
    progress(f"{form}{' ' + str(scalars) if scalars else ''}: {steps} steps in {wall:.2f} s")
    return out


def timers_in_order(e, steps):
    e.set_scalar("overlap", 0)
    e.sync()
    e.reset_timers()
    e.step(steps)
    e.sync()
    t = e.timers()
    e.set_scalar("overlap", 1)
    return {k: (v[0] * 1.0, v[1]) for k, v in t.items()}, steps


# compulsory HBM bytes per particle of each SPH pass (k = mean list length
# including self): each state array element read or written once per pass,
# lists read once per pass
SPH_BYTES = {
    "solve_range": lambda k: 60.0,        # hash 8 + pos R/W 32 + tree 16 + n_ngb 4
    "find_ngb": lambda k: 48.0 + 4 * k,   # hash R/W 12 + pos 16 + offset 4 + tree 16 + list writes
    "density": lambda k: 96.0 + 4 * k,    # hash 8 + pos, vel 32 + n, offset 8 + list + hydro R/W 32 + dots 16
    "eos": lambda k: 32.0,                # hydro R/W
    "hydro": lambda k: 128.0 + 8 * k,     # hash 8 + pos, vel, hydro, dots 64 + n, offset 8 + list + acc + 48
    "reaction": lambda k: 56.0 + 8 * k,   # n, offset 8 + list + acc + pos 16 + bodies_dots R/W 32
}


def sph_bandwidth(timers, steps, n, kbar, walks=None):
    out = {}
    for name, f in SPH_BYTES.items():
        if name not in timers:
            continue
        ms, calls = timers[name]
        per_launch_s = ms / max(calls, 1) / 1e3
        gbps = f(kbar) * n / per_launch_s / 1e9
        out[name] = {"bytes_per_particle": f(kbar), "ms": per_launch_s * 1e3, "GBps": gbps,
                     "frac_of_8TBps": gbps / 8000.0}
        if name == "find_ngb" and walks:
            # the walk is latency-bound: tree nodes visited per second
            out[name]["node_visits_per_particle"] = walks / n
            out[name]["node_visits_per_s"] = walks / per_launch_s
    return out


def traffic_from_profile(kernel_prefix):
    """HBM bytes per launch of the gravity kernel from the committed counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    path = ROOT / "profiles" / "gravity_mixed_c3_1M_pmc.txt"
    if not path.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9.eE+-]+)", path.read_text())
    return float(m.group(1)) if m else None


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo) and logical CPU count."""
    name = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                name = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return name, os.cpu_count()


def cpu_step(cs, n, seed, threads, gravity_sample):
    """One oracle FP32 step on `threads` host threads: SPH passes and
    integration on all N, the N^2 gravity on `gravity_sample` i-particles
    against all N j, scaled to N.  Returns (seconds per step, parts)."""
    import oracle
    p = cs.preset_params(CONFIG, n)
    p.n_threads = threads
    pos, vel, hyd = cs.generate_ic(CONFIG, n, seed)
    gravity_sample = min(gravity_sample, p.n_dev)
    o = oracle.Oracle(32, p)
    o.load_ic(pos, vel, hyd)
    t0 = time.perf_counter()
    o.run_pass(oracle.SPH_ACC)
    t_sph = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.gravity_range(0, gravity_sample)
    t_grav = (time.perf_counter() - t0) * (p.n_dev / gravity_sample)
    o.set_scalar("dt", 1e-4)
    t0 = time.perf_counter()
    o.run_pass(oracle.ADD_GRAVITY)
    o.run_pass(oracle.TIMESTEP)
    o.run_pass(oracle.PREDICT)
    o.run_pass(oracle.CORRECT)
    t_int = time.perf_counter() - t0
    return t_sph + t_grav + t_int, (t_sph, t_grav, t_int, p.n_dev / gravity_sample)


def cpu_share():
    """Host CPUs this process may use: the CPUs in its affinity mask, capped by
    the cgroup's CPU quota (cpu.max) and by OMP_NUM_THREADS when they are set.
    Returns (threads, affinity count, quota or None)."""
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    threads = affinity
    if quota:
        threads = min(threads, max(1, int(quota)))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    return threads, affinity, quota


def cpu_baseline(cs, n, seed, gravity_sample, gravity_sample_1core):
    """The CPU oracle (the reference's FP32 arithmetic, OpenMP over
    particles) on every host core this process may use, and on one core."""
    threads, affinity, quota = cpu_share()
    step, (t_sph, t_grav, t_int, scale) = cpu_step(cs, n, seed, threads, gravity_sample)
    progress(f"cpu baseline on {threads} threads: {step:.1f} s per step")
    step1, (s1, g1, i1, scale1) = cpu_step(cs, n, seed, 1, gravity_sample_1core)
    progress(f"cpu baseline on 1 thread: {step1:.1f} s per step")
    model, logical = cpu_model()
    return {
        "value": n / step, "unit": "particle-updates/s", "cores": threads, "kind": "port",
        "cpu": model, "logical_cpus": logical, "affinity_cpus": affinity, "cgroup_cpu_quota": quota,
        "sample": (f"one plain-FP32 step of the same {n}-particle workload on {threads} threads (the "
                   f"CPUs this process may use: affinity mask, cgroup quota, OMP_NUM_THREADS): tree+h-solve+lists+density+EOS+hydro+reaction on all "
                   f"particles ({t_sph:.2f} s), direct gravity for {gravity_sample} i-particles against all j "
                   f"scaled by {scale:.0f} ({t_grav:.1f} s), integration ({t_int:.3f} s)"),
        "one_core": {"value": n / step1, "unit": "particle-updates/s", "cores": 1,
                     "sample": (f"the same step on 1 thread: SPH passes on all particles ({s1:.2f} s), gravity "
                                f"for {gravity_sample_1core} i-particles scaled by {scale1:.0f} ({g1:.1f} s), "
                                f"integration ({i1:.3f} s)")},
    }


ADJ_WORKLOAD = "Plummer sphere (M=1, a=1, truncated at 10), 50+-5 neighbours, Plummer eps=0.005; " \
               "loss = kinetic energy + sum rho^2 of the final state; gradient w.r.t. x0, v0, u0"


def adjoint_leg(cs, n, steps, ckpt, fd_dirs, seed):
    """Reverse mode of the FP64 form: wall time of the backward pass (with
    checkpoint recomputation) against the plain FP64 forward of the same run,
    and the gradient against central differences of the recorded map."""
    import numpy as np
    p = cs.preset_params(4, n)
    pos, vel, hyd = cs.generate_ic(4, n, seed)

    def engine():
        e = cs.Engine("fp64", p)
        e.load_ic(pos, vel, hyd)
        e.sync()
        return e

    e0 = engine()
    t0 = time.perf_counter()
    e0.init()
    e0.step(steps)
    e0.sync()
    t_fwd = time.perf_counter() - t0
    e0.close()
    e = engine()
    t0 = time.perf_counter()
    loss = e.adjoint_forward(steps, ckpt)
    t_rec = time.perf_counter() - t0
    e.set_scalar("timing", 1)
    t0 = time.perf_counter()
    gx, gv, gu = e.adjoint_backward()
    t_bwd = time.perf_counter() - t0
    timers = {k: v[0] for k, v in e.timers().items() if k.startswith("adjoint")}
    e.set_scalar("timing", 0)
    g = np.concatenate([gx.ravel(), gv.ravel(), gu])
    z = np.concatenate([pos[:n, :3].ravel(), vel[:n, :3].ravel(), hyd[:n, 2]])

    def L(zz):
        return e.adjoint_replay(zz[: 3 * n].reshape(n, 3), zz[3 * n: 6 * n].reshape(n, 3), zz[6 * n:])

    rs = np.random.default_rng(2024)
    errs = []
    for _ in range(fd_dirs):
        d = rs.standard_normal(z.size)
        d /= np.linalg.norm(d)
        fd = (L(z + 1e-4 * d) - L(z - 1e-4 * d)) / 2e-4
        errs.append(abs(g @ d - fd) / abs(fd))
    e.close()
    return {
        "workload": ADJ_WORKLOAD, "n": n, "steps": steps, "ckpt_every": ckpt, "dtype": "f64",
        "loss": loss, "forward_s": t_fwd, "recording_forward_s": t_rec, "backward_s": t_bwd,
        "backward_kernel_ms": timers,
        "adjoint_over_forward": t_bwd / t_fwd,
        "fd_check": {"directions": fd_dirs, "step": 1e-4,
                     "max_rel_err": max(errs) if errs else None,
                     "median_rel_err": float(np.median(errs)) if errs else None},
    }


SCALE_WORKLOAD = "uniform-density sphere (R=1, M=1) with a Gaussian k^-4 velocity field from a 128^3 grid, " \
                 "rms Mach 2; u=1, gas+radiation EOS; 64+-8 neighbours; Plummer eps=0.001"


def scale_leg(cs, n, steps, seed):
    """The scale workload: `steps` mixed-form steps (SPH GB/s, gravity FP32
    fraction) and one reverse-mode step of the FP64 form at the same N."""
    import numpy as np
    e = cs.make_engine("mixed", 5, n, seed, ph_sort=True, i_order=0)
    e.set_scalar("timing", 1)
    e.init()
    e0 = e.stats()["E"]
    e.reset_timers()
    e.sync()
    t0 = time.perf_counter()
    e.step(steps)
    e.sync()
    wall = time.perf_counter() - t0
    timers = e.timers()
    e_pair = e.stats()["E"]
    kbar = float(e.get("ngb_count")[:n].mean())
    walks = float(e.get("hash")[:, 0].astype(np.int64).sum())
    nd = e.nd
    t_in_order, s_in_order = timers_in_order(e, 1)   # each SPH pass alone
    e.close()
    # the same steps with the multipole far field (grav_far = 1)
    e = cs.make_engine("mixed", 5, n, seed, ph_sort=True, i_order=0)
    e.set_scalar("timing", 1)
    e.set_scalar("grav_far", 1)
    e.init()
    e.reset_timers()
    e.sync()
    t0 = time.perf_counter()
    e.step(steps)
    e.sync()
    wall_far = time.perf_counter() - t0
    t_far = e.timers()
    e_far = e.stats()["E"]
    e.close()
    g_ms, g_calls = timers.get("gravity", (float("nan"), 1))
    g_avg = g_ms / max(g_calls, 1) / 1e3
    achieved = FLOP_PER_PAIR * float(nd) * float(nd) / g_avg / 1e12
    out = {"workload": SCALE_WORKLOAD, "n": n, "steps": steps, "form": "mixed",
           "value": n * steps / wall, "unit": "particle-updates/s", "ms_per_step": wall / steps * 1e3,
           "kernel_ms_per_step": {k: v[0] / steps for k, v in sorted(timers.items())},
           "sph_hbm": {"mean_list_length": kbar,
                       "measured": "1 step after the timed region with gravity in order, so each pass runs alone",
                       "passes": sph_bandwidth(t_in_order, s_in_order, n, kbar, walks)},
           "gravity": {"kernel": "k_gravity_mixed_plummer", "launch_ms": g_avg * 1e3, "achieved": achieved,
                       "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS},
           "far_field": {"value": n * steps / wall_far, "ms_per_step": wall_far / steps * 1e3,
                         "gravity_ms": t_far["gravity"][0] / max(t_far["gravity"][1], 1),
                         "energy_vs_pairwise": {"definition": "|E_far - E_pairwise| / |E(0)| / steps after the run",
                                                "value": abs(e_far - e_pair) / abs(e0) / steps}}}
    progress(f"scale: {steps} mixed steps at N={n}, {out['ms_per_step']:.0f} ms/step; far field "
             f"{out['far_field']['ms_per_step']:.0f} ms/step")
    p = cs.preset_params(5, n)
    pos, vel, hyd = cs.generate_ic(5, n, seed)
    f = cs.Engine("fp64", p)
    f.load_ic(pos, vel, hyd)
    f.sync()
    t0 = time.perf_counter()
    loss = f.adjoint_forward(1, 1)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    f.adjoint_backward()
    t_bwd = time.perf_counter() - t0
    f.close()
    progress(f"scale: one FP64 adjoint step, forward {t_fwd:.1f} s, backward {t_bwd:.1f} s")
    out["adjoint_step"] = {"form": "fp64", "steps": 1, "loss": loss, "forward_s": t_fwd, "backward_s": t_bwd,
                           "adjoint_over_forward": t_bwd / t_fwd}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=0, help="particles (default: the config's own N)")
    ap.add_argument("--config", type=int, default=3, choices=(1, 2, 3, 4, 5),
                    help="the seeded workload; 3 (the default) is the headline")
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-gravity-sample", type=int, default=4096)
    ap.add_argument("--cpu-gravity-sample-1core", type=int, default=512)
    ap.add_argument("--i-order", type=int, default=0, help="0: memory order, 1: kd-tree heap order")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-far", action="store_true",
                    help="skip the mixed form with the multipole far field (reported beside the headline)")
    ap.add_argument("--adjoint-n", type=int, default=65536)
    ap.add_argument("--adjoint-steps", type=int, default=20)
    ap.add_argument("--adjoint-ckpt", type=int, default=5)
    ap.add_argument("--fd-dirs", type=int, default=64,
                    help="seeded unit directions for the finite-difference check of the gradient")
    ap.add_argument("--adjoint-only", action="store_true", help="run only the reverse-mode leg")
    ap.add_argument("--scale", action="store_true",
                    help="run only the scale workload (N = 4,194,304: 5 mixed steps + one adjoint step)")
    ap.add_argument("--no-scale", action="store_true",
                    help="skip the scale workload, which the default run reports in a `scale` object")
    ap.add_argument("--scale-n", type=int, default=4194304)
    ap.add_argument("--scale-steps", type=int, default=5)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel over BYTES (for scaling memory counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import cudasph_amd as cs
        ms = cs.lib().sph_copy_calibration(args.copy_calibration)
        print(json.dumps({"copy_bytes": args.copy_calibration, "ms": ms,
                          "GBps": 2 * args.copy_calibration / ms / 1e6}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import cudasph_amd as cs
    global CONFIG
    CONFIG = args.config
    n, K, W = args.n or cs.preset_params(CONFIG, 0).n_bodies, args.steps, args.warmup
    if args.scale:
        print(json.dumps(scale_leg(cs, args.scale_n, args.scale_steps, args.seed)))
        return
    if args.adjoint_only:
        print(json.dumps(adjoint_leg(cs, args.adjoint_n, args.adjoint_steps, args.adjoint_ckpt, args.fd_dirs,
                                     args.seed)))
        return
    mixed = run_form(cs, "mixed", n, args.seed, W, K, i_order=args.i_order)
    plain = None if args.no_plain else run_form(cs, "plain", n, args.seed, W, K, i_order=args.i_order)
    fp64 = None if args.no_fp64 else run_form(cs, "fp64", n, args.seed, W, K, timing=False)
    # the optional multipole far field (not pairwise, so not the headline)
    far = None if args.no_far else run_form(cs, "mixed", n, args.seed, W, K, i_order=args.i_order,
                                            scalars={"grav_far": 1})

    # achievable peaks on this GPU beside the published ones: a 1 GiB copy
    # kernel and an FMA loop (both the second of two launches)
    copy_ms = cs.lib().sph_copy_calibration(1 << 30)
    peaks = {"hbm_copy_GBps": 2 * (1 << 30) / copy_ms / 1e6,
             "fp32_fma_TFLOPs": cs.lib().sph_fma_peak_form(1),
             "fp32_fma_scalar_coeff_TFLOPs": cs.lib().sph_fma_peak_form(0),
             "gravity_pair_body_TFLOPs": cs.lib().sph_pair_body_probe(0) * FLOP_PER_PAIR / 1e12,
             "what": "k_copy16 over 1 GiB (read + written bytes / time); k_fma_peak, 8 wavefronts per SIMD "
                     "on every CU, 2 flops per single-lane FMA: 16 chains of the accumulate form v_fmac_f32 "
                     "with vector operands (fp32_fma_TFLOPs) and 8 chains of v_fma_f32 with scalar "
                     "coefficients; k_pair_body: the gravity kernel's expanded equal-mass pair body alone "
                     "(11 VALU + rsq) on register-held records, 8 wavefronts per SIMD, in the 20-flop "
                     "convention: the ceiling of that pair body"}
    value = n * K / mixed["wall_s"]
    g_ms, g_calls = mixed["timers"].get("gravity", (float("nan"), 0))
    g_avg = g_ms / max(g_calls, 1) / 1e3
    nd = cs.preset_params(CONFIG, n).n_dev
    achieved = FLOP_PER_PAIR * float(nd) * float(nd) / g_avg / 1e12

    def drift(run):
        if run is None or fp64 is None:
            return None
        e0 = abs(fp64["energies"][0])
        per = [abs(a - b) / e0 / max(k, 1) for k, (a, b) in enumerate(zip(run["energies"], fp64["energies"]))]
        return {"max_per_step": max(per[1:]) if len(per) > 1 else 0.0,
                "final": abs(run["energies"][-1] - fp64["energies"][-1]) / e0}

    def mom_drift(run):
        if run is None or fp64 is None:
            return None
        import numpy as np
        return {"linear": float(np.linalg.norm(run["momentum"] - fp64["momentum"]) / fp64["momentum_scale"]),
                "angular": abs(run["angular_momentum"] - fp64["angular_momentum"]) / abs(fp64["angular_momentum"])}

    line = {
        "metric": ("particle-updates/s @1M (mixed vs plain FP32, rel err <=1e-6); adjoint/fwd ratio" if CONFIG == 3
                   else f"particle-updates/s, config {CONFIG} (mixed vs plain FP32)"),
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": mixed["wall_s"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded initial conditions generated in-process)",
        "config": {"workload": WORKLOADS[CONFIG], "n": n, "form": "mixed", "seed": args.seed,
                   "initial_order": "Peano-Hilbert", "i_order": ["memory", "kd-tree heap"][args.i_order]},
        "plain_fp32": None if plain is None else {
            "value": n * K / plain["wall_s"], "ms_per_step": plain["wall_s"] / K * 1e3,
            "gravity_ms": plain["timers"].get("gravity", (0, 1))[0] / max(plain["timers"].get("gravity", (0, 1))[1], 1)},
        "mixed_over_plain": None if plain is None else plain["wall_s"] / mixed["wall_s"],
        "energy_drift_vs_fp64": {"definition": "|E_form(k) - E_fp64(k)| / |E_fp64(0)| / k over the run",
                                 "mixed": drift(mixed), "plain": drift(plain)},
        "momentum_drift_vs_fp64": {
            "definition": "|P_form - P_fp64| / sum(m |v|) and |J_form - J_fp64| / |J_fp64| after the run",
            "mixed": mom_drift(mixed), "plain": mom_drift(plain)},
        "kernel_ms_per_step": {k: v[0] / max(K, 1) for k, v in sorted(mixed["timers"].items())},
        "kernel_ms_per_step_in_order": {k: v[0] / sp for (t, sp) in [mixed["timers_in_order"]]
                                        for k, v in sorted(t.items())},
        "measured_peaks": peaks,
        "sph_hbm": {"mean_list_length": mixed["kbar"],
                    "measured": "2 steps after the timed region with gravity in order, so each pass runs alone",
                    "passes": sph_bandwidth(mixed["timers_in_order"][0], mixed["timers_in_order"][1], n,
                                            mixed["kbar"], mixed["walks"])},
        "roofline": {"bound": "valu", "kernel": "k_gravity_mixed_plummer",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic_from_profile("k_gravity") if CONFIG == 3 else None,
                     "frac_of_measured_fma": achieved / peaks["fp32_fma_TFLOPs"],
                     "frac_of_pair_body": achieved / peaks["gravity_pair_body_TFLOPs"],
                     "per_launch": f"{FLOP_PER_PAIR} flop x {nd}^2 pairs",
                     "launch_ms": g_avg * 1e3},
        "far_field": None if far is None else {
            "what": "mixed form with the order-4 multipole far field (grav_far = 1): per wavefront, "
                    "the coarsest cell of 16384/4096/1024/256 records, 64-record sub-tile or 32-record "
                    "half seen from beyond 8 of its radii takes one expansion; the rest unchanged",
            "value": n * K / far["wall_s"], "unit": "particle-updates/s", "ms_per_step": far["wall_s"] / K * 1e3,
            "gravity_ms": far["timers"].get("gravity", (0, 1))[0] / max(far["timers"].get("gravity", (0, 1))[1], 1),
            "over_pairwise_mixed": mixed["wall_s"] / far["wall_s"],
            "energy_drift_vs_fp64": drift(far), "momentum_drift_vs_fp64": mom_drift(far)},
        "adjoint_over_forward": None,
    }
    if not args.no_adjoint:
        adj = adjoint_leg(cs, args.adjoint_n, args.adjoint_steps, args.adjoint_ckpt, args.fd_dirs, args.seed)
        line["adjoint_over_forward"] = adj["adjoint_over_forward"]
        line["adjoint"] = adj
    if not args.no_scale:
        line["scale"] = scale_leg(cs, args.scale_n, args.scale_steps, args.seed)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cs, n, args.seed, args.cpu_gravity_sample,
                                            args.cpu_gravity_sample_1core)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
