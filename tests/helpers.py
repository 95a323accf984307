"""Shared helpers for the parity tests."""
import numpy as np

# arrays that describe the whole simulation state, in engine/oracle naming
STATE = ("pos", "vel", "hydro", "dots", "bdots", "hydro2", "grav", "tree", "hash", "n_ngb",
         "ngb_count", "ngb_offset", "ngb_list", "ngb_acc")


def copy_state(dst, src):
    """Copy every state array and dt from one simulation object to another."""
    for name in STATE:
        dst.set(name, src.get(name))
    dst.set_scalar("dt", src.scalar("dt"))


def lists_of(sim, n):
    """Per-particle neighbour lists (python lists) of the first n particles."""
    off, cnt, lst = sim.get("ngb_offset"), sim.get("ngb_count"), sim.get("ngb_list")
    return [lst[off[i]:off[i] + cnt[i]] for i in range(n)]


def assert_bits(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{what}: shapes {a.shape} vs {b.shape}"
    ai = a.view(np.uint32 if a.itemsize == 4 else np.uint64)
    bi = b.view(np.uint32 if b.itemsize == 4 else np.uint64)
    bad = np.nonzero(ai != bi)
    n_bad = bad[0].size
    if n_bad:
        idx = tuple(x[0] for x in bad)
        raise AssertionError(f"{what}: {n_bad} of {a.size} values differ in bits; first at {idx}: "
                             f"{a[idx]!r} vs {b[idx]!r}")


def rel_err_vec(a, b, floor):
    """Per-particle |a-b| / max(|b|, floor) for (n,3) vectors."""
    d = np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64), axis=1)
    m = np.linalg.norm(np.asarray(b, np.float64), axis=1)
    return d / np.maximum(m, floor)
