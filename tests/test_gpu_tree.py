"""Device kd-tree build (engine scalar tree_gpu = 1) against the host build,
which is the reference's quicksort build bit for bit (kd_tree.cpp:39-128,
checked against the oracle in test_gpu_parity.py).

Without ties in any coordinate the left-balanced tree is unique, so the two
builds must agree in every bit: node records, split words, fixed slots and
the particle map.  With ties the builds may pick a different one of two
equal-coordinate particles for a node; the tree must then still be a
permutation of the particles in heap layout whose left subtrees lie <= and
right subtrees >= each node along its split axis, and the neighbour sets that
the walks find must be the host tree's.
"""
import numpy as np
import pytest

from helpers import assert_bits, lists_of

pytestmark = pytest.mark.gpu


def engine(cs, form, cfg, n, seed=5):
    p = cs.preset_params(cfg, n)
    e = cs.Engine(form, p)
    pos, vel, hyd = cs.generate_ic(cfg, n, seed)
    e.load_ic(pos, vel, hyd)
    return e, p


def set_xyz(e, n, xyz):
    pos = e.get("pos")
    pos[:n, :3] = xyz
    e.set("pos", pos)


def distinct_xyz(n, seed):
    """n points whose x, y and z values are each pairwise distinct (three
    independent permutations of n distinct values in [-1, 1))"""
    rng = np.random.default_rng(seed)
    vals = np.linspace(-1.0, 1.0, n, endpoint=False) + 0.5 / n
    return np.stack([vals[rng.permutation(n)] for _ in range(3)], axis=1)


def tree_and_hash(cs, e, gpu):
    e.set_scalar("tree_gpu", gpu)
    e.run_pass(cs.TREE)
    return e.get("tree"), e.get("hash")


def check_kd_valid(tree, n):
    """heap-layout permutation with the split-axis order of every subtree"""
    words = tree[1:n + 1, 3]
    bits = words.view(np.int32) if tree.dtype == np.float32 else words.astype(np.int64)
    ids = bits & ~(3 << 24)
    axis = (bits >> 24) & 3
    assert sorted(ids.tolist()) == list(range(n)), "tree nodes are not a permutation of the particles"
    assert (axis == (np.floor(np.log2(np.arange(1, n + 1))).astype(int) % 3)).all(), "split axis is not depth % 3"
    node = np.arange(1, n + 1)
    xyz = tree[1:n + 1, :3]
    k = 1
    while (node >> k).max() >= 1:
        anc = node >> k
        has = anc >= 1
        child = node >> (k - 1)
        a = axis[anc[has] - 1]
        mine = xyz[node[has] - 1, a]
        theirs = xyz[anc[has] - 1, a]
        left = (child[has] & 1) == 0
        assert (mine[left] <= theirs[left]).all(), "a left subtree reaches past its node"
        assert (mine[~left] >= theirs[~left]).all(), "a right subtree reaches below its node"
        k += 1


@pytest.mark.parametrize("form", ["plain", "fp64"])
@pytest.mark.parametrize("n", [1, 2, 3, 7, 1000, 50000, 65536])
@pytest.mark.parametrize("i_order", [0, 1])
def test_device_tree_is_the_host_tree_without_ties(cs, form, n, i_order):
    print(f"tree form={form} n={n} i_order={i_order}", flush=True)
    e, p = engine(cs, form, 1, n)
    e.set_scalar("i_order", i_order)
    set_xyz(e, n, distinct_xyz(n, seed=n))
    th, hh = tree_and_hash(cs, e, 0)
    tg, hg = tree_and_hash(cs, e, 1)
    assert_bits(tg, th, "tree")
    assert np.array_equal(hg, hh), "particle map differs"
    check_kd_valid(tg, n)
    e.close()


@pytest.mark.parametrize("cfg,n,grid", [(1, 16384, 32), (3, 30000, 64), (2, 20000, 0)])
def test_device_tree_with_ties_finds_the_same_neighbours(cs, cfg, n, grid):
    """grid > 0 snaps every coordinate to a grid of that many cells per unit,
    so most coordinates are tied; grid 0 keeps the generated positions (ties
    only from the 24-bit draws)"""
    print(f"tree ties cfg={cfg} n={n} grid={grid}", flush=True)
    e, p = engine(cs, "plain", cfg, n)
    e.init()
    if grid:
        pos = e.get("pos")
        pos[:n, :3] = np.round(pos[:n, :3] * grid) / grid
        e.set("pos", pos)
    # host tree, ranges and lists
    e.set_scalar("tree_gpu", 0)
    e.run_pass(cs.TREE)
    e.run_pass(cs.SOLVE)
    e.run_pass(cs.OFFSETS)
    e.run_pass(cs.FIND)
    host = [np.sort(x) for x in lists_of(e, n)]
    # device tree on the same positions and ranges, the same slots
    e.set_scalar("tree_gpu", 1)
    e.run_pass(cs.TREE)
    tg = e.get("tree")
    check_kd_valid(tg, n)
    e.run_pass(cs.OFFSETS)
    e.run_pass(cs.FIND)
    dev = [np.sort(x) for x in lists_of(e, n)]
    bad = [i for i in range(n) if not np.array_equal(host[i], dev[i])]
    assert not bad, f"{len(bad)} neighbour sets differ, first particle {bad[0]}"
    e.close()


def test_mixed_steps_with_device_tree(cs):
    """three mixed-form steps with either tree: the same physics to the
    rounding of the sums' order (the lists are the same sets)"""
    n = 16384
    print(f"tree mixed steps n={n}", flush=True)
    out = []
    for gpu in (0, 1):
        e = cs.make_engine("mixed", 3, n, seed=7, ph_sort=True)
        e.set_scalar("tree_gpu", gpu)
        e.init()
        e.step(3)
        out.append((e.stats(), e.get("pos64")[:n].copy(), e.get("grav")[:n].copy()))
        e.close()
    (s0, p0, g0), (s1, p1, g1) = out
    assert abs(s1["E"] - s0["E"]) <= 1e-6 * abs(s0["E"])
    assert np.abs(p1[:, :3] - p0[:, :3]).max() < 1e-6
    assert np.allclose(g1, g0, rtol=1e-5, atol=1e-6 * np.abs(g0).max())


@pytest.mark.parametrize("n,grid", [(16384, 0), (50000, 0), (16384, 32), (30000, 64)])
def test_checked_device_tree_is_the_host_tree(cs, n, grid):
    """tree_gpu = 2 (the plain form's default): the device build, flagged when
    a node's particle ties its split coordinate within its subtree, with the
    host build for that pass when flagged -- the host tree in every bit,
    ties or not.  grid 0 gives distinct coordinates (no fallback); grid > 0
    snaps them to a grid, so the build must fall back."""
    print(f"checked tree n={n} grid={grid}", flush=True)
    e, p = engine(cs, "plain", 1, n)
    if grid:
        xyz = np.round(distinct_xyz(n, seed=n) * grid) / grid
    else:
        xyz = distinct_xyz(n, seed=n)
    set_xyz(e, n, xyz)
    th, hh = tree_and_hash(cs, e, 0)
    before = e.scalar("tree_fallbacks")
    tg, hg = tree_and_hash(cs, e, 2)
    fell_back = e.scalar("tree_fallbacks") - before
    assert fell_back == (1 if grid else 0), fell_back
    assert_bits(tg, th, "tree")
    assert np.array_equal(hg, hh), "particle map differs"
    e.close()


def expected_boxes(tree, n, fix_phantom):
    """Subtree boxes restated in numpy: node k's box bounds its record's point
    (FP64 coordinates rounded outward to float) and the boxes of the children
    a walk can enter from it -- 2k and 2k+1 up to n, and node n+1 from 2k = n
    unless fix_phantom.  Children have larger indices, so one pass from the
    last node down sees them first."""
    last = n + 1
    pts = tree[:last + 1, :3]
    if pts.dtype == np.float64:
        f = pts.astype(np.float32)
        lo = np.where(f.astype(np.float64) > pts, np.nextafter(f, np.float32(-np.inf)), f)
        hi = np.where(f.astype(np.float64) < pts, np.nextafter(f, np.float32(np.inf)), f)
    else:
        lo, hi = pts.copy(), pts.copy()
    for k in range(last, 0, -1):
        for ch in (2 * k, 2 * k + 1):
            if ch > n and not (ch == n + 1 and 2 * k == n and not fix_phantom):
                continue
            lo[k] = np.minimum(lo[k], lo[ch])
            hi[k] = np.maximum(hi[k], hi[ch])
    return lo, hi


@pytest.mark.parametrize("form,n", [("mixed", 1), ("mixed", 2), ("mixed", 7), ("mixed", 1000), ("mixed", 16000),
                                    ("mixed", 16001), ("mixed", 16384), ("mixed", 70001), ("fp64", 5),
                                    ("fp64", 16384)])
@pytest.mark.parametrize("fix_phantom", [0, 1])
def test_subtree_boxes(cs, form, n, fix_phantom):
    """The walks' subtree boxes (one block per subtree below level 10, one
    block above) against the numpy restatement, every node up to n+1."""
    print(f"boxes form={form} n={n} fix_phantom={fix_phantom}", flush=True)
    cfg = 3 if n >= 5 else 1
    p = cs.preset_params(cfg, n)
    p.fix_phantom = fix_phantom
    e = cs.Engine(form, p)
    pos, vel, hyd = cs.generate_ic(cfg, n, 11)
    e.load_ic(pos, vel, hyd)
    e.run_pass(cs.TREE)
    e.run_pass(cs.SOLVE)
    tree, box = e.get("tree"), e.get("boxes")
    lo, hi = expected_boxes(tree, n, fix_phantom)
    k = np.arange(1, n + 2)
    assert np.array_equal(box[2 * k, :3], lo[k]), "low corners differ"
    assert np.array_equal(box[2 * k + 1, :3], hi[k]), "high corners differ"
    e.close()
