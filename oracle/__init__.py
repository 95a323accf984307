"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the reference the results are
compared with or as the timed CPU baseline.  The product package (cudasph_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

(TREE, SOLVE, OFFSETS, FIND, DENSITY, EOS, HYDRO, REACTION, GRAVITY, ADD_GRAVITY, TIMESTEP,
 PREDICT, CORRECT, STATS, PH_SORT, SPH_ACC, INIT, STEP) = range(18)
STATS_KEYS = ("T", "U", "W", "E", "J", "Fx", "Fy", "Fz", "Tx", "Ty", "Tz")
STATISTICS_KEYS = STATS_KEYS + (
    "Jx", "Jy", "Jz", "xmin", "ymin", "zmin", "xmax", "ymax", "zmax", "m_tot",
    "xcom", "ycom", "zcom", "vxcom", "vycom", "vzcom", "Rcm", "Vcm",
    "ngb_mean", "ngb_sigma", "ngb_min", "ngb_max", "count_mismatch", "max_count",
    "phantom_reads", "phantom_ngb", "overflow_reruns", "iteration", "time")


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        vp, cp, i, l, d = C.c_void_p, C.c_char_p, C.c_int, C.c_long, C.c_double
        L.orc_create.argtypes = [i, vp]
        L.orc_create.restype = vp
        L.orc_destroy.argtypes = [vp]
        L.orc_get.argtypes = [vp, cp, vp, l]
        L.orc_get.restype = l
        L.orc_set.argtypes = [vp, cp, vp, l]
        L.orc_set.restype = l
        L.orc_get_scalar.argtypes = [vp, cp]
        L.orc_get_scalar.restype = d
        L.orc_set_scalar.argtypes = [vp, cp, d]
        L.orc_pass.argtypes = [vp, i]
        L.orc_pass.restype = i
        L.orc_stats.argtypes = [vp, C.POINTER(C.c_double)]
        L.orc_statistics.argtypes = [vp, C.POINTER(C.c_double)]
        L.orc_gravity_range.argtypes = [vp, i, i]
        L.orc_kernel_fns.argtypes = [i, i, vp, vp]
        L.orc_pressure.argtypes = [i, vp, d, d, d]
        L.orc_pressure.restype = d
        L.orc_find_root.argtypes = [i, i, d, d]
        L.orc_find_root.restype = d
        L.orc_art_visc.argtypes = [i, C.c_float, C.c_float, d, d, d, d, vp, vp, vp, vp, vp]
        L.orc_dreal_sub.argtypes = [vp, C.c_float, vp]
        L.orc_ph_key.argtypes = [i, i, i, i]
        L.orc_ph_key.restype = C.c_longlong
        L.orc_brute_ngb.argtypes = [i, vp, vp]
        _lib = L
    return _lib


class Oracle:
    """CPU restatement of the reference's step, in FP32 (reference arithmetic) or FP64."""

    def __init__(self, bits: int, params):
        self.bits = bits
        self.params = params.copy()
        self.n = params.n_bodies
        self.nd = params.n_dev
        self.real = np.float32 if bits == 32 else np.float64
        self._h = lib().orc_create(bits, C.byref(self.params))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def _spec(self, name):
        r, nd = self.real, self.nd
        return {
            "pos": (r, (nd, 4)), "vel": (r, (nd, 4)), "hydro": (r, (nd, 4)), "dots": (r, (nd, 4)),
            "bdots": (r, (nd, 4)), "hydro2": (r, (nd, 4)), "grav": (r, (nd, 4)),
            "tree": (r, (nd + 2, 4)), "hash": (np.uint32, (nd, 2)), "n_ngb": (np.int32, (nd,)),
            "ngb_count": (np.int32, (nd,)), "ngb_offset": (np.int32, (nd + 1,)),
            "ngb_list": (np.int32, (100 * nd,)), "ngb_acc": (r, (100 * nd,)),
        }[name]

    def get(self, name):
        dt, shape = self._spec(name)
        out = np.zeros(shape, dt)
        lib().orc_get(self._h, name.encode(), out.ctypes.data, out.nbytes)
        return out

    def set(self, name, arr):
        dt, _ = self._spec(name)
        a = np.ascontiguousarray(arr, dtype=dt)
        lib().orc_set(self._h, name.encode(), a.ctypes.data, a.nbytes)

    def load_ic(self, pos, vel, hyd):
        n = self.n
        for name, src in (("pos", pos), ("vel", vel), ("hydro", hyd)):
            cur = self.get(name)
            cur[:n] = np.asarray(src)[:n]
            self.set(name, cur)

    def run_pass(self, pid):
        if lib().orc_pass(self._h, int(pid)) != 0:
            raise RuntimeError(f"oracle pass {pid} failed")

    def init(self):
        self.run_pass(INIT)

    def step(self, n=1):
        for _ in range(n):
            self.run_pass(STEP)

    def gravity_range(self, i0, i1):
        lib().orc_gravity_range(self._h, int(i0), int(i1))

    def scalar(self, name):
        return lib().orc_get_scalar(self._h, name.encode())

    def set_scalar(self, name, v):
        lib().orc_set_scalar(self._h, name.encode(), float(v))

    def stats(self):
        out = (C.c_double * 11)()
        lib().orc_stats(self._h, out)
        return dict(zip(STATS_KEYS, list(out)))

    def statistics(self):
        out = (C.c_double * len(STATISTICS_KEYS))()
        lib().orc_statistics(self._h, out)
        return dict(zip(STATISTICS_KEYS, list(out)))


def kernel_fns(bits, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros((u.size, 9))
    lib().orc_kernel_fns(bits, u.size, u.ctypes.data, out.ctypes.data)
    return out   # w dw dwdh dphidh iw diw ig digdh dig


def pressure(bits, params, rho, u, mu):
    return lib().orc_pressure(bits, C.byref(params), rho, u, mu)


def find_root(bits, newton, q, r):
    return lib().orc_find_root(bits, newton, q, r)


def art_visc(bits, alpha, beta, range_i, range_j, ds, drdv, hi, di, hj, dj):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (hi, di, hj, dj)]
    out = np.zeros(2)
    lib().orc_art_visc(bits, alpha, beta, range_i, range_j, ds, drdv, *[a.ctypes.data for a in arrs],
                       out.ctypes.data)
    return out


def dreal_sub(x2, y):
    x = np.ascontiguousarray(x2, dtype=np.float32)
    out = np.zeros(2, np.float32)
    lib().orc_dreal_sub(x.ctypes.data, np.float32(y), out.ctypes.data)
    return out


def ph_key(x, y, z, bits=18):
    return lib().orc_ph_key(int(x), int(y), int(z), int(bits))


def brute_ngb(pos4):
    p = np.ascontiguousarray(pos4, dtype=np.float32)
    out = np.zeros(len(p), np.int32)
    lib().orc_brute_ngb(len(p), p.ctypes.data, out.ctypes.data)
    return out
