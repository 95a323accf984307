"""MI355X-native SPH engine: Python bindings over libcudasph.so.

The compute path is the HIP library built from ``cudasph_amd/csrc`` (see
``cudasph_amd.build``).  This module only marshals arrays through the C ABI of
``include/cudasph.h``; it has no numerical fallback and raises if the library
is missing or the GPU call fails.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libcudasph.so"

FORM_PLAIN, FORM_MIXED, FORM_FP64 = 0, 1, 2
FORMS = {"plain": FORM_PLAIN, "mixed": FORM_MIXED, "fp64": FORM_FP64}

# pass ids (include/cudasph.h)
(TREE, SOLVE, OFFSETS, FIND, DENSITY, EOS, HYDRO, REACTION, GRAVITY, ADD_GRAVITY, TIMESTEP,
 PREDICT, CORRECT, STATS, PH_SORT, SPH_ACC, INIT, STEP) = range(18)


class Params(C.Structure):
    """sph_params (include/cudasph.h); same layout as the oracle's orc_params."""

    _fields_ = [
        ("n_bodies", C.c_int), ("n_dev", C.c_int), ("rel_err", C.c_float),
        ("desired_ngb", C.c_float), ("max_iter", C.c_int), ("fix_phantom", C.c_int),
        ("av_alpha", C.c_float), ("av_beta", C.c_float), ("cn_vel", C.c_float),
        ("cn_eth", C.c_float), ("grav_mode", C.c_int), ("grav_eps", C.c_float),
        ("eos_density_unit", C.c_double), ("eos_energy_unit", C.c_double),
        ("eos_pressure_unit", C.c_double), ("eos_amu", C.c_double),
        ("eos_boltzmann", C.c_double), ("eos_rad_const", C.c_double),
        ("eos_gamma", C.c_double), ("eos_temp_base", C.c_double),
        ("ph_period", C.c_int), ("ph_bits", C.c_int), ("eos_newton", C.c_int),
        ("n_threads", C.c_int),
    ]

    def copy(self) -> "Params":
        p = Params()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(Params))
        return p

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib() -> C.CDLL:
    """Load libcudasph.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m cudasph_amd.build` first")
    from . import build as _build
    if _build.is_stale():
        import sys
        sys.stderr.write(f"warning: {LIB_PATH} was built from other sources than csrc/ holds now; "
                         "run `python -m cudasph_amd.build`\n")
    L = C.CDLL(str(LIB_PATH))
    vp, cp, i, l, d = C.c_void_p, C.c_char_p, C.c_int, C.c_long, C.c_double
    L.sph_params_default.argtypes = [C.POINTER(Params), i]
    L.sph_params_preset.argtypes = [C.POINTER(Params), i, i]
    L.sph_params_preset.restype = i
    L.sph_ic_generate.argtypes = [i, i, C.c_uint64, vp, vp, vp]
    L.sph_ic_generate.restype = i
    L.sph_last_error.restype = cp
    L.sph_engine_create.argtypes = [i, C.POINTER(Params), i]
    L.sph_engine_create.restype = vp
    L.sph_engine_destroy.argtypes = [vp]
    L.sph_engine_set.argtypes = [vp, cp, vp, l]
    L.sph_engine_set.restype = l
    L.sph_engine_get.argtypes = [vp, cp, vp, l]
    L.sph_engine_get.restype = l
    L.sph_engine_get_scalar.argtypes = [vp, cp]
    L.sph_engine_get_scalar.restype = d
    L.sph_engine_set_scalar.argtypes = [vp, cp, d]
    L.sph_engine_set_scalar.restype = i
    for fn in ("sph_engine_pass", "sph_engine_run"):
        getattr(L, fn).argtypes = [vp, i]
        getattr(L, fn).restype = i
    L.sph_engine_sync.argtypes = [vp]
    L.sph_engine_sync.restype = i
    L.sph_engine_stats.argtypes = [vp, C.POINTER(C.c_double)]
    L.sph_engine_stats.restype = i
    L.sph_engine_statistics.argtypes = [vp, C.POINTER(C.c_double)]
    L.sph_engine_statistics.restype = i
    L.sph_engine_timers.argtypes = [vp, C.POINTER(C.c_char_p), C.POINTER(C.c_double), C.POINTER(C.c_long), i]
    L.sph_engine_timers.restype = i
    L.sph_engine_reset_timers.argtypes = [vp]
    L.sph_engine_read_state.argtypes = [vp, cp]
    L.sph_engine_read_state.restype = i
    L.sph_engine_write_state.argtypes = [vp, cp]
    L.sph_engine_write_state.restype = i
    L.sph_engine_output.argtypes = [vp, cp]
    L.sph_engine_output.restype = i
    L.sph_state_file_count.argtypes = [cp]
    L.sph_state_file_count.restype = i
    L.sph_abi_set_params.argtypes = [C.POINTER(Params), i]
    L.sph_copy_calibration.argtypes = [l]
    L.sph_copy_calibration.restype = d
    L.sph_fma_peak_form.argtypes = [i]
    L.sph_fma_peak_form.restype = d
    L.sph_pair_body_probe.argtypes = [i]
    L.sph_pair_body_probe.restype = d
    L.sph_adjoint_forward.argtypes = [vp, i, i, C.POINTER(C.c_double)]
    L.sph_adjoint_forward.restype = i
    L.sph_adjoint_backward.argtypes = [vp, vp, vp, vp]
    L.sph_adjoint_backward.restype = i
    L.sph_adjoint_replay.argtypes = [vp, vp, vp, vp, C.POINTER(C.c_double)]
    L.sph_adjoint_replay.restype = i
    L.sph_build_kd_tree.argtypes = [i, i, i, vp, vp, vp]
    L.sph_build_kd_tree.restype = i
    L.sph_build_kd_tree64.argtypes = [i, i, i, vp, vp, vp]
    L.sph_build_kd_tree64.restype = i
    _lib = L
    return L


def default_params(n_bodies: int) -> Params:
    p = Params()
    lib().sph_params_default(C.byref(p), int(n_bodies))
    return p


def preset_params(config: int, n_bodies: int = 0) -> Params:
    p = Params()
    if lib().sph_params_preset(C.byref(p), int(config), int(n_bodies)) != 0:
        raise ValueError(f"unknown config {config}")
    return p


def generate_ic(config: int, n: int, seed: int = 12345):
    """Seeded initial conditions in FP64: pos4 {x,y,z,range}, vel4 {v,m}, hydro4 {rho,P,u,mu}."""
    pos = np.zeros((n, 4), np.float64)
    vel = np.zeros((n, 4), np.float64)
    hyd = np.zeros((n, 4), np.float64)
    if lib().sph_ic_generate(int(config), int(n), int(seed), pos.ctypes.data, vel.ctypes.data,
                             hyd.ctypes.data) != 0:
        raise ValueError(f"ic generation failed for config {config}")
    return pos, vel, hyd


def build_kd_tree(pos4, n_dev: int | None = None, n_threads: int = 0):
    """Host kd-tree (reference-exact) of pos4[:, :3]; returns (tree, particle_map)."""
    pos4 = np.ascontiguousarray(pos4)
    n = len(pos4)
    nd = n if n_dev is None else int(n_dev)
    hm = np.zeros((nd, 2), np.uint32)
    if pos4.dtype == np.float64:
        tree = np.zeros((nd + 2, 4), np.float64)
        rc = lib().sph_build_kd_tree64(n, nd, n_threads, pos4.ctypes.data, tree.ctypes.data, hm.ctypes.data)
    else:
        pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
        tree = np.zeros((nd + 2, 4), np.float32)
        rc = lib().sph_build_kd_tree(n, nd, n_threads, pos4.ctypes.data, tree.ctypes.data, hm.ctypes.data)
    if rc != 0:
        raise _err("sph_build_kd_tree")
    return tree, hm


def _err(what: str):
    return RuntimeError(f"{what}: {lib().sph_last_error().decode()}")


STATS_KEYS = ("T", "U", "W", "E", "J", "Fx", "Fy", "Fz", "Tx", "Ty", "Tz")
# sph_engine_statistics' layout (SPH_STAT_* in include/cudasph.h)
STATISTICS_KEYS = STATS_KEYS + (
    "Jx", "Jy", "Jz", "xmin", "ymin", "zmin", "xmax", "ymax", "zmax", "m_tot",
    "xcom", "ycom", "zcom", "vxcom", "vycom", "vzcom", "Rcm", "Vcm",
    "ngb_mean", "ngb_sigma", "ngb_min", "ngb_max", "count_mismatch", "max_count",
    "phantom_reads", "phantom_ngb", "overflow_reruns", "iteration", "time")


class Engine:
    """A device-resident simulation (one HIP stream on one GPU)."""

    def __init__(self, form, params: Params, device: int = 0):
        self.form = FORMS[form] if isinstance(form, str) else int(form)
        self.params = params.copy()
        self.n = params.n_bodies
        self.nd = params.n_dev
        self.real = np.float64 if self.form == FORM_FP64 else np.float32
        h = lib().sph_engine_create(self.form, C.byref(self.params), int(device))
        if not h:
            raise _err("sph_engine_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().sph_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # array layout by name
    def _spec(self, name: str):
        r, nd = self.real, self.nd
        cap = int(100 * nd)
        table = {
            "pos": (r, (nd, 4)), "vel": (r, (nd, 4)), "hydro": (r, (nd, 4)), "dots": (r, (nd, 4)),
            "bdots": (r, (nd, 4)), "hydro2": (r, (nd, 4)), "grav": (r, (nd, 4)),
            "tree": (r, (nd + 2, 4)), "hash": (np.uint32, (nd, 2)), "n_ngb": (np.int32, (nd,)),
            "ngb_count": (np.int32, (nd,)), "ngb_offset": (np.int32, (nd + 1,)),
            "ngb_list": (np.int32, (cap,)), "ngb_acc": (r, (cap,)),
            "pos64": (np.float64, (nd, 4)), "vel64": (np.float64, (nd, 4)), "u64": (np.float64, (nd,)),
            "pos_lo": (np.float32, (nd, 4)), "boxes": (np.float32, (2 * (nd + 2), 4)),
        }
        return table[name]

    def get(self, name: str, count: int | None = None) -> np.ndarray:
        dt, shape = self._spec(name)
        out = np.zeros(shape, dt)
        nb = out.nbytes if count is None else int(count) * out.itemsize * (out.shape[1] if out.ndim > 1 else 1)
        got = lib().sph_engine_get(self._h, name.encode(), out.ctypes.data, nb)
        if got < 0:
            raise _err(f"get {name}")
        return out if count is None else out[:count]

    def set(self, name: str, arr) -> None:
        dt, _ = self._spec(name)
        a = np.ascontiguousarray(arr, dtype=dt)
        if lib().sph_engine_set(self._h, name.encode(), a.ctypes.data, a.nbytes) < 0:
            raise _err(f"set {name}")

    def load_ic(self, pos, vel, hyd) -> None:
        """Load FP64 initial conditions (n records); padding records keep their defaults."""
        n = self.n
        for name, src in (("pos", pos), ("vel", vel), ("hydro", hyd)):
            cur = self.get(name)
            cur[:n] = np.asarray(src)[:n]
            self.set(name, cur)
        if self.form == FORM_MIXED:
            p64 = np.zeros((self.nd, 4))
            p64[:n, :3] = pos[:n, :3]
            v64 = np.zeros((self.nd, 4))
            v64[:n] = vel[:n]
            u64 = np.zeros(self.nd)
            u64[:n] = hyd[:n, 2]
            self.set("pos64", p64)
            self.set("vel64", v64)
            self.set("u64", u64)

    def run_pass(self, pid: int) -> None:
        if lib().sph_engine_pass(self._h, int(pid)) != 0:
            raise _err(f"pass {pid}")

    def init(self) -> None:
        self.run_pass(INIT)

    def step(self, n: int = 1) -> None:
        if lib().sph_engine_run(self._h, int(n)) != 0:
            raise _err("run")

    def sync(self) -> None:
        if lib().sph_engine_sync(self._h) != 0:
            raise _err("sync")

    def scalar(self, name: str) -> float:
        return lib().sph_engine_get_scalar(self._h, name.encode())

    def set_scalar(self, name: str, v: float) -> None:
        if lib().sph_engine_set_scalar(self._h, name.encode(), float(v)) != 0:
            raise _err(f"set_scalar {name}")

    def stats(self) -> dict:
        out = (C.c_double * 11)()
        if lib().sph_engine_stats(self._h, out) != 0:
            raise _err("stats")
        return dict(zip(STATS_KEYS, list(out)))

    def statistics(self) -> dict:
        """What system_statistics.cpp:4-82 reports (energies, box, centre of
        mass, |J|, neighbour-count mean/sigma/min/max) plus the list counters
        of the last find_ngb, from the last STATS pass."""
        out = (C.c_double * len(STATISTICS_KEYS))()
        if lib().sph_engine_statistics(self._h, out) != len(STATISTICS_KEYS):
            raise _err("statistics")
        return dict(zip(STATISTICS_KEYS, list(out)))

    def timers(self) -> dict:
        k = 64
        names = (C.c_char_p * k)()
        ms = (C.c_double * k)()
        calls = (C.c_long * k)()
        m = lib().sph_engine_timers(self._h, names, ms, calls, k)
        if m < 0:
            raise _err("timers")
        return {names[i].decode(): (ms[i], calls[i]) for i in range(min(m, k))}

    def reset_timers(self) -> None:
        lib().sph_engine_reset_timers(self._h)

    # ---- reverse mode (FP64 form) ----------------------------------------
    def adjoint_forward(self, steps: int, ckpt_every: int) -> float:
        """Run init() + `steps` steps recording the map; returns the loss
        sum m|v|^2/2 + sum rho^2 of the final state."""
        loss = C.c_double()
        if lib().sph_adjoint_forward(self._h, int(steps), int(ckpt_every), C.byref(loss)) != 0:
            raise _err("adjoint_forward")
        return loss.value

    def adjoint_backward(self):
        """Gradient of the recorded run's loss: (dL/dx0 [n,3], dL/dv0 [n,3], dL/du0 [n])."""
        n = self.n
        gx, gv, gu = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
        if lib().sph_adjoint_backward(self._h, gx.ctypes.data, gv.ctypes.data, gu.ctypes.data) != 0:
            raise _err("adjoint_backward")
        return gx, gv, gu

    def adjoint_replay(self, x0=None, v0=None, u0=None) -> float:
        """Loss of the recorded (frozen) map from another initial state."""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (x0, v0, u0)]
        ptrs = [None if a is None else a.ctypes.data for a in arrs]
        loss = C.c_double()
        if lib().sph_adjoint_replay(self._h, *ptrs, C.byref(loss)) != 0:
            raise _err("adjoint_replay")
        return loss.value

    def adjoint_records(self):
        """(ranges [K+1, n], timesteps [K+1]) of the recorded force evaluations."""
        dts = np.zeros(4096)
        got = lib().sph_engine_get(self._h, b"adj_dt", dts.ctypes.data, dts.nbytes)
        if got < 0:
            raise _err("get adj_dt")
        dts = dts[: got // 8].copy()
        rng = np.zeros((len(dts), self.nd))
        if lib().sph_engine_get(self._h, b"adj_range", rng.ctypes.data, rng.nbytes) < 0:
            raise _err("get adj_range")
        return rng[:, : self.n], dts

    def write_state(self, path: str) -> None:
        if lib().sph_engine_write_state(self._h, str(path).encode()) != 0:
            raise _err("write_state")

    def read_state(self, path: str) -> None:
        if lib().sph_engine_read_state(self._h, str(path).encode()) != 0:
            raise _err("read_state")

    def output(self, directory: str) -> int:
        """output.cpp:3-15: the numbered snapshot when due, and restart.bin;
        returns the snapshot number written (0: restart.bin only)"""
        r = lib().sph_engine_output(self._h, str(directory).encode())
        if r < 0:
            raise _err("output")
        return r


def make_engine(form, config: int, n: int | None = None, seed: int = 12345, params: Params | None = None,
                device: int = 0, ph_sort: bool = False, i_order: int | None = None) -> Engine:
    """Engine loaded with the seeded initial conditions of a benchmark config.
    ph_sort puts the particles in Peano-Hilbert order first (the reference's
    sort_bodies, which it otherwise applies every 100 steps); i_order picks
    the particle order of the per-particle kernels (0 memory, 1 kd-tree heap)."""
    p = params if params is not None else preset_params(config, n or 0)
    if (FORMS[form] if isinstance(form, str) else form) != FORM_PLAIN:
        p = p.copy()
        p.eos_newton = 1
    e = Engine(form, p, device)
    pos, vel, hyd = generate_ic(config, p.n_bodies, seed)
    e.load_ic(pos, vel, hyd)
    if ph_sort:
        e.run_pass(PH_SORT)
    if i_order is not None:
        e.set_scalar("i_order", i_order)
    return e
