"""Pins the oracle's system statistics (system_statistics.cpp:4-82) and its
list counters: the count/list mismatch of dev_solve_range.cu:141-171 against
dev_find_ngb.cu:67-69, and the reads of the nonexistent node n+1
(dev_solve_range.cu:33-40), which the walk takes from the zero record past
the tree when n = n_dev.  The checks are independent: numpy sums over the
oracle's own arrays, a pure-Python walk of the oracle's tree, and the
shift/boost behaviour of the centre of mass (CPU only)."""
import numpy as np
import pytest


def run_init(cs, orc, cfg, n, seed=5, shift=(0.0, 0.0, 0.0), boost=(0.0, 0.0, 0.0)):
    p = cs.preset_params(cfg, n)
    pos, vel, hyd = cs.generate_ic(cfg, n, seed)
    pos = pos.copy()
    vel = vel.copy()
    pos[:, :3] += np.asarray(shift)
    vel[:, :3] += np.asarray(boost)
    o = orc.Oracle(32, p)
    o.load_ic(pos, vel, hyd)
    o.init()
    return o


@pytest.mark.parametrize("cfg", [1, 2, 3])
def test_statistics_equal_numpy_sums_of_the_oracle_state(cs, orc, cfg):
    n = 3000
    o = run_init(cs, orc, cfg, n)
    s = o.statistics()
    pos = o.get("pos")[:n].astype(np.float64)
    vel = o.get("vel")[:n].astype(np.float64)
    k = o.get("n_ngb")[:n].astype(np.float64)
    m = vel[:, 3]
    assert s["m_tot"] == pytest.approx(m.sum(), rel=1e-13)
    for c, ax in enumerate("xyz"):
        assert s[ax + "min"] == pos[:, c].min() and s[ax + "max"] == pos[:, c].max()
        assert s[ax + "com"] == pytest.approx((m * pos[:, c]).sum() / m.sum(), rel=1e-10, abs=1e-14)
        assert s["v" + ax + "com"] == pytest.approx((m * vel[:, c]).sum() / m.sum(), rel=1e-10, abs=1e-14)
    J = (m[:, None] * np.cross(pos[:, :3], vel[:, :3])).sum(0)
    assert np.allclose([s["Jx"], s["Jy"], s["Jz"]], J, rtol=1e-10, atol=1e-14)
    assert s["J"] == pytest.approx(np.linalg.norm(J), rel=1e-10, abs=1e-14)
    assert s["ngb_mean"] == pytest.approx(k.mean(), rel=1e-14)
    assert s["ngb_sigma"] == pytest.approx(k.std(), rel=1e-9, abs=1e-12)
    assert (s["ngb_min"], s["ngb_max"]) == (k.min(), k.max())
    cnt = o.get("ngb_count")[:n]
    assert s["count_mismatch"] == np.count_nonzero(cnt != o.get("n_ngb")[:n])
    assert s["max_count"] == cnt.max()
    # the configurations' neighbour targets hold on average
    target = {1: 32, 2: 50, 3: 64}[cfg]
    assert abs(s["ngb_mean"] - target) < 0.15 * target


def test_centre_of_mass_follows_a_shift_and_a_boost(cs, orc):
    """Translating every particle by d and boosting by w moves Rcm by d and
    Vcm by w and leaves the neighbour counts and the box size unchanged."""
    n = 2048
    a = run_init(cs, orc, 2, n).statistics()
    d, w = (0.25, -0.5, 0.125), (0.5, 0.25, -0.75)
    b = run_init(cs, orc, 2, n, shift=d, boost=w).statistics()
    for c, ax in enumerate("xyz"):
        assert b[ax + "com"] - a[ax + "com"] == pytest.approx(d[c], abs=1e-6)
        assert b["v" + ax + "com"] - a["v" + ax + "com"] == pytest.approx(w[c], abs=1e-6)
        assert (b[ax + "max"] - b[ax + "min"]) == pytest.approx(a[ax + "max"] - a[ax + "min"], rel=1e-6)
    assert abs(b["ngb_mean"] - a["ngb_mean"]) < 0.05


def python_walk_phantom(tree, n, q):
    """dev_find_ngb.cu:13-73 in plain Python on the oracle's packed tree:
    (read node n+1, took it as a neighbour)."""
    bits = np.ascontiguousarray(tree[:, 3]).view(np.int32)
    stack = [1]
    read = took = False
    x, y, z, r = (float(v) for v in q)
    qc3 = (np.float32(x), np.float32(y), np.float32(z))
    r32 = np.float32(r)
    while stack:
        node_id = stack.pop()
        node = tree[node_id]
        split = (int(bits[node_id]) >> 24) & 3
        child = 2 * node_id
        pushed = []
        if child <= n and split < 3:
            qc, nc = qc3[split], node[split]
            if qc - r32 < nc:
                pushed.append(child)
            if qc + r32 > nc:
                pushed.append(child + 1)
        stack.extend(pushed)
        d = (qc3[0] - node[0], qc3[1] - node[1], qc3[2] - node[2])
        ds2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
        if node_id == n + 1:
            read = True
            took = took or bool(ds2 < r32 * r32)
    return read, took


@pytest.mark.parametrize("cfg,n", [(1, 256), (1, 1024), (2, 512), (1, 300)])
def test_phantom_node_counts_match_a_python_walk(cs, orc, cfg, n):
    """n = n_dev puts the zero record at node n+1 (origin, particle 0), so a
    walk that reads it may take particle 0 a second time; n < n_dev puts a far
    sentinel there (read, never taken)."""
    o = run_init(cs, orc, cfg, n, seed=11)
    tree = o.get("tree")
    pos = o.get("pos")
    hits = [python_walk_phantom(tree, n, pos[i]) for i in range(n)]
    reads = sum(h[0] for h in hits)
    took = sum(h[1] for h in hits)
    s = o.statistics()
    assert (s["phantom_reads"], s["phantom_ngb"]) == (reads, took)
    if n % 2 == 0:
        assert reads > 0, "an even n leaves node n/2 with a right child n+1 that some walk reads"
    if n == o.nd:
        # a read that passes the distance test lists particle 0 once more than
        # its true distance does
        lst, off, cnt = o.get("ngb_list"), o.get("ngb_offset"), o.get("ngb_count")
        p = pos[:n].astype(np.float32)
        d = p[:, :3] - p[0, :3]
        genuine = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2] < p[:, 3] * p[:, 3]
        zeros = np.array([np.count_nonzero(lst[off[i]:off[i] + cnt[i]] == 0) for i in range(n)])
        assert np.array_equal(zeros - genuine, [int(h[1]) for h in hits])
    else:
        assert took == 0
