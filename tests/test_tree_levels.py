"""The device kd-tree build (cudasph_amd/csrc/tree.hip) places one level of
the left-balanced tree per step, on three per-axis lists kept grouped by
subtree, using closed forms of n for every subtree's size and offset.  This
mirrors its four kernels in numpy, index for index (k_tree_split,
k_tree_flags, the exclusive scans, k_tree_regroup), asserting that every
write lands inside its list and that every slot of the next level's layout is
written once, and compares the result with a recursive build that follows
kd_tree.cpp:39-95 (median at rank left_size(n) of a sort along depth % 3).
Without ties the two trees must be equal; with ties the tree must keep every
left subtree <= and every right subtree >= its node along the node's axis."""
import numpy as np
import pytest


def subtree_size(t, n):
    size = np.zeros_like(t)
    lo = t.astype(np.int64)
    hi = lo.copy()
    while True:
        act = lo <= n
        if not act.any():
            return size
        size[act] += np.minimum(hi[act], n) - lo[act] + 1
        lo = lo << 1
        hi = (hi << 1) | 1


def subtree_start(t, L, n):
    s = np.zeros_like(t)
    a, b = 1 << L, t.astype(np.int64)
    while a <= n:
        s += np.minimum(b, n + 1) - a
        a <<= 1
        b = b << 1
    return s


def left_size(n):   # kd_tree.cpp:63-79, with C's truncating division
    m = 1
    while m <= n:
        m <<= 1
    m >>= 1
    last = n - (m - 1)
    half = int((m - 2) / 2)
    return half + last if last <= m // 2 else half + m // 2


def recursive_build(xyz):
    n = len(xyz)
    node_id = np.full(n + 1, -1, np.int64)
    stack = [(1, np.arange(n), 0)]
    while stack:
        node, ids, depth = stack.pop()
        order = ids[np.argsort(xyz[ids, depth % 3], kind="stable")]
        lt = left_size(len(order))
        node_id[node] = order[lt]
        if lt > 0:
            stack.append((2 * node, order[:lt], depth + 1))
        if len(order) - lt - 1 > 0:
            stack.append((2 * node + 1, order[lt + 1:], depth + 1))
    return node_id


def level_build(xyz):
    n = len(xyz)
    node_id = np.full(n + 1, -1, np.int64)
    lists = [[np.argsort(xyz[:, a], kind="stable"), np.zeros(n, np.int64)] for a in range(3)]
    tag = np.ones(n, np.int64)
    at = {1: np.full(n + 1, -1, np.int64), 2: np.full(n + 1, -1, np.int64)}
    cur = [0, 0, 0]
    m, L = n, 0
    while m > 0:
        a = L % 3
        idx = np.arange(m)
        # k_tree_split
        ids = lists[a][cur[a]][:m]
        t = tag[ids]
        p = idx - subtree_start(t, L, n)
        lt = subtree_size(2 * t, n)
        before = t - (1 << L)
        left, med, right = p < lt, p == lt, p > lt
        placed = min(1 << L, n - ((1 << L) - 1))
        assert (p >= 0).all() and med.sum() == placed
        nxt = lists[a][cur[a] ^ 1]
        nxt[:] = -1
        dst = np.concatenate([idx[left] - before[left], idx[right] - before[right] - 1])
        assert ((dst >= 0) & (dst < m - placed)).all() and len(np.unique(dst)) == len(dst)
        nxt[dst] = np.concatenate([ids[left], ids[right]])
        assert (node_id[t[med]] == -1).all()
        node_id[t[med]] = ids[med]
        tag[ids[left]] = 2 * t[left]
        tag[ids[right]] = 2 * t[right] + 1
        if m - placed > 0:
            nl = 2 << L
            for k in (1, 2):
                X = (L + k) % 3
                lx = lists[X][cur[X]][:m]
                tx = tag[lx]
                # k_tree_flags
                flag = ((tx >= nl) & (tx % 2 == 0)).astype(np.int64)
                pl = tx < nl
                at[k][tx[pl]] = idx[pl]
                scan = np.concatenate([[0], np.cumsum(flag)[:-1]])
                # k_tree_regroup
                go = ~pl
                tt, ii = tx[go] >> 1, idx[go]
                s = subtree_start(tt, L, n)
                assert (s < m).all()
                base = s - (tt - (1 << L))
                lefts = scan[ii] - scan[s]
                rights = (ii - s) - lefts - (at[k][tt] < ii)
                d = np.where(tx[go] % 2 == 0, base + lefts, base + subtree_size(2 * tt, n) + rights)
                assert ((d >= 0) & (d < m - placed)).all() and len(np.unique(d)) == len(d)
                nxt = lists[X][cur[X] ^ 1]
                nxt[:] = -1
                nxt[d] = lx[go]
        cur = [c ^ 1 for c in cur]
        m -= placed
        L += 1
    return node_id


def assert_kd_order(xyz, node_id):
    n = len(xyz)
    assert sorted(node_id[1:].tolist()) == list(range(n))
    node = np.arange(1, n + 1)
    depth = np.floor(np.log2(node)).astype(int)
    k = 1
    while (node >> k).max() >= 1:
        has = (node >> k) >= 1
        anc = (node >> k)[has]
        axis = depth[anc - 1] % 3
        mine = xyz[node_id[node[has]], axis]
        theirs = xyz[node_id[anc], axis]
        left = ((node >> (k - 1))[has] & 1) == 0
        assert (mine[left] <= theirs[left]).all() and (mine[~left] >= theirs[~left]).all()
        k += 1


SIZES = list(range(1, 40)) + [(1 << k) + d for k in (6, 9, 12) for d in (-1, 0, 1)] + [3000, 20011]


@pytest.mark.parametrize("n", SIZES)
def test_level_build_is_the_recursive_build(n):
    xyz = np.random.default_rng(n).random((n, 3)).astype(np.float32)
    ties = any(len(np.unique(xyz[:, k])) < n for k in range(3))
    got = level_build(xyz)
    assert_kd_order(xyz, got)
    if not ties:
        assert np.array_equal(got, recursive_build(xyz))


@pytest.mark.parametrize("n,cells", [(50, 2), (1000, 3), (5000, 8)])
def test_level_build_with_ties_is_a_kd_tree(n, cells):
    xyz = np.random.default_rng(n).integers(0, cells, (n, 3)).astype(np.float32)
    assert_kd_order(xyz, level_build(xyz))
