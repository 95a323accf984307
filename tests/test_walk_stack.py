"""The kd-walk's LDS stack is sized per launch to floor(log2 n) + 2 entries
per lane (walk_stack_depth in cudasph_amd/csrc/common.h).  This checks the
bound on the CPU by running the walk's exact stack discipline
(dev_solve_range.cu:17-75 as restated in kd_walk, sph_passes.hip): the
current node is read at slot top, its left child overwrites that slot, the
right child goes above it, then top drops by one.  The worst case pushes
every child that exists (an unbounded search radius), including the
reference's phantom node n+1 for even n."""
import pytest


def walk_stack_depth(n_nodes: int) -> int:
    h = 1
    while h < 31 and (n_nodes >> h) > 0:
        h += 1
    return h + 1


def max_slot_written(n: int, fix_phantom: bool) -> int:
    stk = {0: 1}
    top = 0
    hi = 0
    while top >= 0:
        node = stk[top]
        child = node << 1
        if child <= n and node <= n:   # the phantom (node n+1) is a leaf: its child index exceeds n
            stk[top] = child
            hi = max(hi, top)
            top += 1
            if not fix_phantom or child < n:
                stk[top] = child + 1
                hi = max(hi, top)
                top += 1
        top -= 1
    return hi


SIZES = sorted(set(list(range(1, 300)) + [(1 << k) + d for k in range(9, 17) for d in (-2, -1, 0, 1, 2)]))


@pytest.mark.parametrize("fix_phantom", [False, True])
def test_walk_stack_fits_the_tree_sized_allocation(fix_phantom):
    for n in SIZES:
        assert max_slot_written(n, fix_phantom) < walk_stack_depth(n), n


def test_walk_stack_depth_is_tree_height_plus_one():
    # a tree of n nodes in heap layout has floor(log2 n) + 1 levels
    for n in (1, 2, 3, 4, 1023, 1024, 1048576, (1 << 24) - 3):
        assert walk_stack_depth(n) == n.bit_length() + 1
    # every size the engine accepts (n < 2^24 - 2) fits the former fixed 28
    assert walk_stack_depth((1 << 24) - 3) <= 28
