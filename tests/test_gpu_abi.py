"""GPU tests of the reference's C boundary (include/cudasph.h group 1), called
through ctypes exactly as the reference's host code calls it
(sph_accelerations.cpp, CUDA_gravity.cpp), against the CPU oracle."""
import ctypes as C

import numpy as np
import pytest

from helpers import assert_bits

pytestmark = pytest.mark.gpu


class Dev:
    """device buffers through the reference ABI (allocateCUDAarray & co.)"""

    def __init__(self, L):
        self.L = L
        self.bufs = []
        L.allocateCUDAarray.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int]
        L.deleteCUDAarray.argtypes = [C.c_void_p, C.c_int]
        L.copyArrayToDevice.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        L.copyArrayFromDevice.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int]

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.L.allocateCUDAarray(C.byref(p), a.nbytes, 0)
        self.L.copyArrayToDevice(p, a.ctypes.data, a.nbytes, 0)
        self.bufs.append(p)
        return p

    def get(self, p, like):
        out = np.zeros_like(like)
        self.L.copyArrayFromDevice(out.ctypes.data, p, out.nbytes, 0)
        return out

    def free(self):
        for p in self.bufs:
            self.L.deleteCUDAarray(p, 0)


def n_norm(n, j):
    n = ((n - 1) // j) * j + j
    return j if n == 0 else n


@pytest.mark.parametrize("preset", [None, 1])
def test_solve_range_and_sph_accelerations_drop_in(cs, orc, preset):
    """The reference's call sequence (sph_accelerations.cpp:16-100) over our ABI
    reproduces the oracle's h, lists, density, EOS and hydro bit for bit."""
    n = 16384
    p = cs.default_params(n) if preset is None else cs.preset_params(preset, n)
    L = cs.lib()
    L.sph_abi_set_params(C.byref(p), 1)
    pos, vel, hyd = cs.generate_ic(1, n, 21)
    o = orc.Oracle(32, p)
    o.load_ic(pos, vel, hyd)
    o.run_pass(orc.TREE)
    nd = p.n_dev
    tree = o.get("tree")[: nd + 1]                     # the reference allocates n_dev+1 records
    d = Dev(L)
    try:
        hash_ = np.zeros((nd, 2), np.uint32)
        hash_[:, 0] = np.arange(nd)
        hash_[:, 1] = np.arange(nd)
        dpos, dvel = d.put(o.get("pos")), d.put(o.get("vel"))
        dtree, dhash = d.put(tree), d.put(hash_)
        dmap = d.put(np.zeros(nd, np.int32))
        dn = d.put(np.zeros(nd, np.int32))
        L.host_solve_range.argtypes = [C.c_int, C.c_int, C.c_float] + [C.c_void_p] * 6 + [C.c_int]
        L.host_solve_range.restype = C.c_float
        L.host_solve_range(nd, n, p.rel_err, dhash, dmap, dpos, dvel, dtree, dn, 0)
        o.run_pass(orc.SOLVE)
        assert_bits(d.get(dpos, o.get("pos")), o.get("pos"), "range")
        n_ngb = d.get(dn, o.get("n_ngb"))
        assert np.array_equal(n_ngb, o.get("n_ngb"))
        # host prefix sum of n_norm(n_ngb, 4) (sph_accelerations.cpp:58-64)
        off = np.zeros(nd + 1, np.int32)
        off[1:] = np.cumsum([n_norm(int(c), 4) for c in n_ngb])
        tot = int(off[-1])
        dlist = d.put(np.zeros(100 * nd, np.int32))
        dacc = d.put(np.zeros(100 * nd, np.float32))
        doff = d.put(off[:nd])
        dhyd, ddots = d.put(o.get("hydro")), d.put(np.zeros((nd, 4), np.float32))
        dbd, dh2 = d.put(np.zeros((nd, 4), np.float32)), d.put(np.zeros((nd, 4), np.float32))
        L.host_sph_accelerations.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float] + [C.c_void_p] * 10 + [C.c_int]
        L.host_sph_accelerations.restype = C.c_float
        L.host_sph_accelerations(nd, tot, p.av_alpha, p.av_beta, dhash, None, dmap, dlist, doff, dhyd, ddots, dbd,
                                 dh2, dacc, 0)
        for pid in (orc.OFFSETS, orc.FIND, orc.DENSITY, orc.EOS, orc.HYDRO):
            o.run_pass(pid)
        assert np.array_equal(d.get(dn, n_ngb), o.get("ngb_count"))
        lst = d.get(dlist, np.zeros(100 * nd, np.int32))[:tot]
        olst, ocnt = o.get("ngb_list"), o.get("ngb_count")
        for i in range(0, n, 97):
            assert np.array_equal(lst[off[i]:off[i] + ocnt[i]], olst[off[i]:off[i] + ocnt[i]])
        assert_bits(d.get(dhyd, o.get("hydro")), o.get("hydro"), "hydro")
        assert_bits(d.get(ddots, o.get("dots")), o.get("dots"), "dots")
        assert_bits(d.get(dbd, o.get("bdots")), o.get("bdots"), "bodies_dots")
        assert_bits(d.get(dh2, o.get("hydro2"))[:, 0], o.get("hydro2")[:, 0], "v_sig")
        walk = d.get(dhash, hash_)[:, 0]
        assert np.all(walk[:n] > 0)
    finally:
        d.free()


def test_sph_accelerations_truncates_an_outgrown_slot(cs, orc, capfd):
    """Through the reference ABI the caller owns the offsets, so a list that
    outgrows its slot cannot be re-laid out: it is truncated to the slot, the
    actual length goes back into n_ngb and a message is printed (the reference
    would overwrite the next particle's slot).  Forced by shrinking every
    n_ngb to 1 (4-slot runs) after host_solve_range."""
    n = 16384
    p = cs.preset_params(1, n)
    L = cs.lib()
    L.sph_abi_set_params(C.byref(p), 1)
    pos, vel, hyd = cs.generate_ic(1, n, 21)
    o = orc.Oracle(32, p)
    o.load_ic(pos, vel, hyd)
    o.run_pass(orc.TREE)
    nd = p.n_dev
    d = Dev(L)
    try:
        hash_ = np.zeros((nd, 2), np.uint32)
        hash_[:, 0] = np.arange(nd)
        hash_[:, 1] = np.arange(nd)
        dpos, dvel = d.put(o.get("pos")), d.put(o.get("vel"))
        dtree, dhash = d.put(o.get("tree")[: nd + 1]), d.put(hash_)
        dmap = d.put(np.zeros(nd, np.int32))
        dn = d.put(np.zeros(nd, np.int32))
        L.host_solve_range.argtypes = [C.c_int, C.c_int, C.c_float] + [C.c_void_p] * 6 + [C.c_int]
        L.host_solve_range.restype = C.c_float
        L.host_solve_range(nd, n, p.rel_err, dhash, dmap, dpos, dvel, dtree, dn, 0)
        ones = np.ones(nd, np.int32)
        L.copyArrayToDevice(dn, ones.ctypes.data, ones.nbytes, 0)
        off = (4 * np.arange(nd + 1)).astype(np.int32)
        dlist, dacc = d.put(np.full(4 * nd, -7, np.int32)), d.put(np.zeros(4 * nd, np.float32))
        doff = d.put(off[:nd])
        dhyd, ddots = d.put(o.get("hydro")), d.put(np.zeros((nd, 4), np.float32))
        dbd, dh2 = d.put(np.zeros((nd, 4), np.float32)), d.put(np.zeros((nd, 4), np.float32))
        L.host_sph_accelerations.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float] + [C.c_void_p] * 10 + [C.c_int]
        L.host_sph_accelerations.restype = C.c_float
        capfd.readouterr()
        L.host_sph_accelerations(nd, 4 * nd, p.av_alpha, p.av_beta, dhash, None, dmap, dlist, doff, dhyd, ddots,
                                 dbd, dh2, dacc, 0)
        err = capfd.readouterr().err
        assert "truncated" in err
        # the oracle's untruncated lists at the same ranges (its h-solve from
        # the same start gives the ABI's ranges bit for bit)
        for pid in (orc.SOLVE, orc.OFFSETS, orc.FIND):
            o.run_pass(pid)
        assert_bits(d.get(dpos, o.get("pos")), o.get("pos"), "range")
        ocnt, ooff, olst = o.get("ngb_count"), o.get("ngb_offset"), o.get("ngb_list")
        cnt = d.get(dn, ones)
        assert np.array_equal(cnt[:n], np.minimum(ocnt[:n], 4))
        assert (ocnt[:n] > 4).mean() > 0.99, "the slots did not overflow"
        lst = d.get(dlist, np.zeros(4 * nd, np.int32))
        for i in range(0, n, 37):
            k = int(cnt[i])
            assert np.array_equal(lst[4 * i: 4 * i + k], olst[ooff[i]: ooff[i] + k]), i
        assert np.isfinite(d.get(dbd, o.get("bdots"))[:n]).all()
    finally:
        d.free()


@pytest.mark.parametrize("mode", [0, 1])
def test_compute_cuda_gravity_drop_in(cs, orc, mode):
    n = 8192
    p = cs.preset_params(2, n)
    p.grav_mode = mode
    L = cs.lib()
    L.sph_abi_set_params(C.byref(p), 1)
    pos, vel, hyd = cs.generate_ic(2, n, 8)
    o = orc.Oracle(32, p)
    o.load_ic(pos, vel, hyd)
    o.init()
    P, V = o.get("pos"), o.get("vel")
    # CUDA_gravity.cpp:36-46 packing
    gpos = np.c_[P[:, :3], V[:, 3]].astype(np.float32)
    gvel = np.c_[V[:, :3], P[:, 3] * P[:, 3]].astype(np.float32)
    d = Dev(L)
    try:
        dp, dv = d.put(gpos), d.put(gvel)
        da = d.put(np.zeros((n, 4), np.float32))
        L.compute_CUDA_gravity.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.compute_CUDA_gravity(n, da, dp, dv, 128, 1, 0)
        o.run_pass(orc.GRAVITY)
        assert_bits(d.get(da, gpos), o.get("grav"), "grav")
    finally:
        d.free()


def test_radix_sort_is_stable(cs):
    L = cs.lib()
    rng = np.random.default_rng(3)
    for n, bits in ((1, 8), (5, 32), (1000, 8), (20000, 32), (32768, 32), (32769, 24), (70001, 20), (300000, 32)):
        keys = rng.integers(0, 1 << min(bits, 16), n, dtype=np.uint64).astype(np.uint32)
        if bits == 32:
            keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
            keys[::7] = keys[0]
        kv = np.zeros((n, 2), np.uint32)
        kv[:, 0], kv[:, 1] = keys, np.arange(n)
        d = Dev(L)
        try:
            a, b = d.put(kv), d.put(np.zeros_like(kv))
            L.RadixSort.argtypes = [C.c_void_p, C.c_void_p, C.c_uint, C.c_uint]
            L.RadixSort(a, b, n, bits)
            got = d.get(a, kv)
        finally:
            d.free()
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(got[:, 0], keys[order]) and np.array_equal(got[:, 1], order), (n, bits)


def test_measurement_probes(cs):
    """The achievable-peak probes the bench reports: a 1 GiB copy and the FMA
    loop in both forms return finite, plausible rates (a failed launch returns
    a negative number)."""
    L = cs.lib()
    ms = L.sph_copy_calibration(1 << 30)
    assert 0 < ms < 1e3
    gbps = 2 * (1 << 30) / ms / 1e6
    assert 500 < gbps < 20000, gbps
    for form in (0, 1):
        t = L.sph_fma_peak_form(form)
        assert 5 < t < 500, (form, t)
    for variant in (0, 1):
        pairs = L.sph_pair_body_probe(variant)
        assert 1e11 < pairs < 1e14, (variant, pairs)
