"""The h-solve candidate cache's reach bounds, mirrored in numpy float32
(csrc/sph_passes.hip: kd_walk, kd_walk_paths, filter_cache).

A walk at range R records every particle it accepts together with its path's
largest split-plane margin plus and minus a rounding bound (hi, lo).  A later
evaluation at r <= R takes a record inside r when r >= hi, skips it when
r <= lo, and walks instead when r falls between.  These tests run the same
float32 operations as the kernels (walk order, push tests, distance sums) on
small trees and check that every decided filter returns exactly the walk's
accepted distances in the walk's order, including a constructed case where
rounding makes the walk prune a particle that lies inside the range.
"""
import numpy as np

f32 = np.float32
EPS = f32(1.0 / 8388608.0)


def build(pts):
    """left-balanced kd-tree in heap layout (node t: children 2t, 2t+1),
    split axis = depth % 3; tree[t] = (point, axis)"""
    n = len(pts)
    size = [0] * (2 * n + 2)
    for t in range(n, 0, -1):
        size[t] = 1 + (size[2 * t] if 2 * t <= n else 0) + (size[2 * t + 1] if 2 * t + 1 <= n else 0)
    tree = [None] * (n + 1)

    def rec(idx, t, depth):
        if t > n:
            return
        ax = depth % 3
        idx = sorted(idx, key=lambda i: (pts[i][ax], i))
        nl = size[2 * t] if 2 * t <= n else 0
        tree[t] = (pts[idx[nl]], ax)
        rec(idx[:nl], 2 * t, depth + 1)
        rec(idx[nl + 1:], 2 * t + 1, depth + 1)

    rec(list(range(n)), 1, 0)
    return tree


def dist2(q, p):
    d = [f32(q[k] - p[k]) for k in range(3)]
    return f32(f32(f32(d[0] * d[0]) + f32(d[1] * d[1])) + f32(d[2] * d[2]))


def walk(tree, q, r):
    """kd_walk: pop, push left if fl(qc - r) < nc, right if fl(qc + r) > nc
    (right popped first), then test the node"""
    n = len(tree) - 1
    r = f32(r)
    r2 = f32(r * r)
    stk, out = [1], []
    while stk:
        t = stk.pop()
        p, ax = tree[t]
        c = 2 * t
        if c <= n:
            qc, nc = f32(q[ax]), f32(p[ax])
            if f32(qc - r) < nc:
                stk.append(c)
            if f32(qc + r) > nc and c + 1 <= n:
                stk.append(c + 1)
        ds2 = dist2(q, p)
        if ds2 < r2:
            out.append(ds2)
    return out


def walk_paths(tree, q, R):
    """kd_walk_paths: the same walk at R, each accepted node with its path's
    largest margin bounds hi = m + 2d and lo = m - 2d"""
    n = len(tree) - 1
    R = f32(R)
    r2 = f32(R * R)
    stk, recs = [(1, -R, -R)], []
    while stk:
        t, hi, lo = stk.pop()
        p, ax = tree[t]
        c = 2 * t
        if c <= n:
            qc, nc = f32(q[ax]), f32(p[ax])
            d = f32(f32(8) * EPS * max(abs(qc), abs(nc), R))
            if f32(qc - R) < nc:
                m = f32(qc - nc)
                stk.append((c, max(hi, f32(m + f32(2 * d))), max(lo, f32(m - f32(2 * d)))))
            if f32(qc + R) > nc and c + 1 <= n:
                m = f32(nc - qc)
                stk.append((c + 1, max(hi, f32(m + f32(2 * d))), max(lo, f32(m - f32(2 * d)))))
        ds2 = dist2(q, p)
        if ds2 < r2:
            recs.append((ds2, hi, lo))
    return recs


def filt(recs, r, bounds=True):
    """filter_cache: the records inside r whose reach is decided, or None
    when one is not (the kernel then walks)"""
    r = f32(r)
    r2 = f32(r * r)
    out = []
    for ds2, hi, lo in recs:
        if not ds2 < r2:
            continue
        if not bounds or r >= hi:
            out.append(ds2)
        elif r > lo:
            return None
    return out


def test_filter_is_the_walk_on_random_and_grid_points():
    rng = np.random.default_rng(7)
    decided = undecided = 0
    for grid in (0, 64):
        raw = rng.random((3000, 3))
        if grid:
            raw = np.round(raw * grid) / grid   # many equal coordinates and distances
        pts = [tuple(f32(v) for v in row) for row in raw.astype(np.float32)]
        tree = build(pts)
        for trial in range(60):
            q = pts[rng.integers(len(pts))] if trial % 2 else tuple(f32(v) for v in rng.random(3))
            r0 = f32(rng.choice([0.0625, 0.09375, 0.1, 0.125]))
            R = f32(r0 * f32(1.125))
            recs = walk_paths(tree, q, R)
            for r in (r0, f32(r0 * f32(0.9)), R, f32(0.0625), f32(r0 * f32(1.01))):
                if r > R:
                    continue
                got = filt(recs, r)
                if got is None:
                    undecided += 1
                    continue
                decided += 1
                assert got == walk(tree, q, r), (grid, trial, r)
    assert decided > 10 * max(undecided, 1)


def test_boundary_particle_pruned_by_rounding():
    """A node splits x at nc = fl(qx - r); a particle in its left subtree sits
    at x = nc with the query's y and z.  fl(qx - r) < nc is false, so the walk
    at r prunes the subtree, yet fl(qx - nc) < r puts the particle inside the
    range: a filter without the reach bounds would take it."""
    qx, r = f32(0.3), f32(0.1)
    nc = f32(qx - r)
    assert f32(qx - nc) < r            # the rounding that makes this case
    q = (qx, f32(0.5), f32(0.5))
    # root splits x at nc (its particle far up in y); its left child holds the
    # boundary particle; a right child keeps the tree left-balanced
    root = (nc, f32(0.9), f32(0.5))
    inside = (nc, f32(0.5), f32(0.5))
    right = (f32(0.6), f32(0.5), f32(0.5))
    tree = [None, (root, 0), (inside, 1), (right, 1)]
    want = walk(tree, q, r)
    assert dist2(q, inside) not in want                      # pruned by the walk at r
    recs = walk_paths(tree, q, f32(r * f32(1.125)))
    assert filt(recs, r, bounds=False) != want                # a bare filter is wrong here
    got = filt(recs, r)
    assert got is None or got == want                         # the bounds send it to a walk
    assert got is None
