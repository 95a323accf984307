"""GPU tests of the engine beyond the per-pass parity: padding (N not a
multiple of 256), the Peano-Hilbert re-sort, the mixed-precision form against
the FP64 oracle, and reference state-file round trips."""
import numpy as np
import pytest

from helpers import assert_bits, copy_state

pytestmark = pytest.mark.gpu


def pair(cs, orc, cfg, n, form="plain", bits=32, seed=5, **over):
    p = cs.preset_params(cfg, n)
    for k, v in over.items():
        setattr(p, k, v)
    pos, vel, hyd = cs.generate_ic(cfg, n, seed)
    e = cs.Engine(form, p)
    e.load_ic(pos, vel, hyd)
    o = orc.Oracle(bits, p)
    o.load_ic(pos, vel, hyd)
    return e, o, p


@pytest.mark.parametrize("cfg,n", [(1, 3000), (2, 5000), (3, 6001)])
def test_padded_sizes_bit_exact(cs, orc, cfg, n):
    print(f"padded cfg={cfg} n={n}", flush=True)
    e, o, p = pair(cs, orc, cfg, n)
    assert p.n_dev > n
    e.set_scalar("exact_rsqrt", 1)
    e.init()
    o.init()
    for _ in range(2):
        e.step()
        o.step()
    for name in ("pos", "vel", "hydro", "bdots", "grav"):
        assert_bits(e.get(name)[:n], o.get(name)[:n], name)
    assert e.scalar("dt") == o.scalar("dt")


def test_reference_defaults_spline_gravity_bit_exact(cs, orc):
    """The drop-in defaults: CGS EOS with the Ferrari root, M4 spline softening."""
    n = 8192
    p = cs.default_params(n)
    pos, vel, hyd = cs.generate_ic(1, n, 17)
    e = cs.Engine("plain", p)
    e.set_scalar("exact_rsqrt", 1)
    e.load_ic(pos, vel, hyd)
    o = orc.Oracle(32, p)
    o.load_ic(pos, vel, hyd)
    e.init()
    o.init()
    e.step()
    o.step()
    for name in ("pos", "vel", "hydro", "dots", "bdots", "grav"):
        assert_bits(e.get(name)[:n], o.get(name)[:n], name)


def test_peano_hilbert_resort_bit_exact(cs, orc):
    n = 20000
    e, o, p = pair(cs, orc, 2, n)
    o.init()
    copy_state(e, o)
    before = e.get("pos")[:n].copy()
    e.run_pass(cs.PH_SORT)
    o.run_pass(orc.PH_SORT)
    for name in ("pos", "vel", "hydro", "dots", "bdots", "hydro2", "grav"):
        assert_bits(e.get(name)[:n], o.get(name)[:n], name)
    assert not np.array_equal(e.get("pos")[:n], before), "the seeded order is not Peano-Hilbert order"


def test_step_100_triggers_the_resort(cs, orc):
    n = 4096
    e, o, p = pair(cs, orc, 1, n, ph_period=2)
    e.set_scalar("exact_rsqrt", 1)
    e.init()
    o.init()
    for _ in range(3):
        e.step()
        o.step()
    for name in ("pos", "vel", "hydro", "bdots"):
        assert_bits(e.get(name)[:n], o.get(name)[:n], name)


@pytest.mark.parametrize("cfg", [1, 2, 3, 5])
def test_mixed_form_trajectory_against_fp64(cs, orc, cfg):
    """Three steps of the mixed form stay on the FP64 oracle's trajectory:
    energy within 1e-6 per step, positions and velocities to FP32 level."""
    print(f"mixed trajectory cfg={cfg}", flush=True)
    n = 16384
    p = cs.preset_params(cfg, n)
    pos, vel, hyd = cs.generate_ic(cfg, n, 31)
    e = cs.make_engine("mixed", cfg, n, 31)
    o = orc.Oracle(64, p)
    o.load_ic(pos, vel, hyd)
    e.init()
    o.init()
    e0, o0 = e.stats()["E"], o.stats()["E"]
    assert abs(e0 - o0) <= 1e-6 * abs(o0)
    for _ in range(3):
        e.step()
        o.step()
    x64, xo = e.get("pos64")[:n, :3], o.get("pos")[:n, :3]
    assert np.abs(x64 - xo).max() <= 1e-6 * np.abs(xo).max()
    # the two runs converge h independently to the h-solve's 1e-3 tolerance,
    # so hydro accelerations differ at ~1e-3 of their size; over three steps
    # of dt ~ 0.02 that bounds the velocity agreement at ~1e-4 of max |v|
    v64, vo = e.get("vel64")[:n, :3], o.get("vel")[:n, :3]
    assert np.abs(v64 - vo).max() <= 1e-4 * np.abs(vo).max()
    drift = abs(e.stats()["E"] - o.stats()["E"]) / abs(o0) / 3
    assert drift <= 1e-6, drift


@pytest.mark.parametrize("cfg,ph,grav_mode", [(2, False, 1), (3, False, 1), (2, True, 1), (3, True, 1),
                                              (2, False, 0), (3, True, 0)])
def test_mixed_form_passes_against_fp64_oracle(cs, orc, cfg, ph, grav_mode):
    """Density, EOS, hydro + reaction and gravity of the mixed form against the
    FP64 oracle run on the mixed form's own state (same ranges): rho and P to
    2e-6, accelerations to 1e-5 of max(|a|, 1e-3 rms).  In Peano-Hilbert order
    the Plummer kernel takes its tile-relative form for distant sub-tiles;
    grav_mode 0 is the reference's M4-spline softening at the SPH range."""
    print(f"mixed passes cfg={cfg} ph_sort={ph} grav_mode={grav_mode}", flush=True)
    n = 16384
    p = cs.preset_params(cfg, n)
    p.grav_mode = grav_mode
    e = cs.make_engine("mixed", cfg, n, 41, params=p, ph_sort=ph)
    e.init()
    e.step()
    for pid in (cs.SPH_ACC, cs.GRAVITY):
        e.run_pass(pid)
    p = e.params.copy()
    o = orc.Oracle(64, p)
    P = e.get("pos").astype(np.float64)
    P[:n, :3] = e.get("pos64")[:n, :3]
    V = e.get("vel64").copy()
    V[n:] = e.get("vel")[n:]
    H = e.get("hydro").astype(np.float64)
    H[:n, 2] = e.get("u64")[:n]
    o.set("pos", P)
    o.set("vel", V)
    o.set("hydro", H)
    o.set("n_ngb", e.get("n_ngb"))
    for pid in (orc.TREE, orc.OFFSETS, orc.FIND, orc.DENSITY, orc.EOS, orc.HYDRO, orc.REACTION, orc.GRAVITY):
        o.run_pass(pid)
    same = e.get("ngb_count")[:n] == o.get("ngb_count")[:n]
    assert same.mean() > 0.999
    he, ho = e.get("hydro")[:n].astype(np.float64), o.get("hydro")[:n]
    assert (np.abs(he[:, 0] - ho[:, 0]) / ho[:, 0])[same].max() < 2e-6
    assert (np.abs(he[:, 1] - ho[:, 1]) / ho[:, 1])[same].max() < 2e-6
    # a list that differs (a neighbour right at the range boundary) changes
    # the reactions of its members too; compare the unaffected particles
    off, cnt, lst = o.get("ngb_offset"), o.get("ngb_count"), o.get("ngb_list")
    ok = same.copy()
    for i in np.nonzero(~same)[0]:
        ok[lst[off[i]:off[i] + cnt[i]]] = False
    assert ok.mean() > 0.99
    for name, k in (("bdots", 3), ("grav", 3)):
        a, b = e.get(name)[:n, :k].astype(np.float64), o.get(name)[:n, :k]
        if name == "bdots":
            a, b = a[ok], b[ok]
        floor = 1e-3 * np.sqrt((b**2).sum(1).mean())
        err = np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), floor)
        assert err.max() < 1e-5, (name, err.max())


def test_mixed_gravity_forms_agree(cs):
    """The tile-relative (expanded equal-mass body, and the body on r), and
    split forms of the mixed gravity kernel give the same accelerations to
    FP32 rounding level on Peano-Hilbert-ordered input."""
    n = 65536
    out = {}
    for name, split, expand in (("expanded", 0, 1), ("tile", 0, 0), ("split", 1, 1)):
        e = cs.make_engine("mixed", 3, n, 7, ph_sort=True)
        e.set_scalar("grav_split_only", split)
        e.set_scalar("grav_expand", expand)
        e.run_pass(cs.GRAVITY)
        out[name] = e.get("grav")[:n].astype(np.float64)
        e.close()
    b = out["split"][:, :3]
    rms = np.sqrt((b ** 2).sum(1).mean())
    for name in ("expanded", "tile"):
        a = out[name][:, :3]
        err = np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-3 * rms)
        assert err.max() < 2e-6, (name, err.max())
        assert not np.array_equal(a, b), f"the {name} form was not taken"
        np.testing.assert_allclose(out[name][:, 3], out["split"][:, 3], rtol=2e-6)
    assert not np.array_equal(out["expanded"], out["tile"]), "the expanded body was not taken"


@pytest.mark.parametrize("n", [2, 5, 33, 100, 300, 1000])
def test_mixed_form_small_sizes(cs, orc, n):
    """Small N: most of the 8 j-parts of the gravity grid are empty, sub-tiles
    are partial and padding is a large share of the records (down to two
    particles among 256 records, where every walk takes the whole tree)."""
    p = cs.preset_params(2, n)
    pos, vel, hyd = cs.generate_ic(2, n, 3)
    e = cs.make_engine("mixed", 2, n, 3, ph_sort=True)
    o = orc.Oracle(64, e.params.copy())
    P = e.get("pos").astype(np.float64)
    P[:n, :3] = e.get("pos64")[:n, :3]
    V = e.get("vel64")
    H = e.get("hydro").astype(np.float64)
    H[:n, 2] = e.get("u64")[:n]
    o.set("pos", P)
    o.set("vel", V)
    o.set("hydro", H)
    e.init()
    o.init()
    for _ in range(2):
        e.step()
        o.step()
    e0 = abs(o.stats()["E"])
    assert abs(e.stats()["E"] - o.stats()["E"]) <= 1e-6 * e0
    g, go = e.get("grav")[:n, :3].astype(np.float64), o.get("grav")[:n, :3]
    rms = np.sqrt((go ** 2).sum(1).mean())
    err = np.linalg.norm(g - go, axis=1) / np.maximum(np.linalg.norm(go, axis=1), 1e-3 * rms)
    assert err.max() < 1e-4, err.max()
    assert np.isfinite(e.get("bdots")[:n]).all()


@pytest.mark.parametrize("n", [16384, 5000, 16404])
@pytest.mark.parametrize("far", [0, 1])
@pytest.mark.parametrize("equal", [True, False])
def test_mixed_gravity_against_fp64(cs, orc, equal, far, n):
    """The mixed gravity pass against FP64 gravity on the same positions, with
    equal masses (the unit-mass pair body) and with masses spread over a
    factor of 3 (the general body); both Peano-Hilbert ordered, so both the
    tile-relative and the split forms are taken.  N = 5000 and 16404 leave 8
    and 20 real records in the last 64-record sub-tile, whose middle record
    is padding: the sub-tile's origin must be a real record.  Pairwise and
    with the multipole far field; the total force must vanish to rounding."""
    print(f"mixed gravity vs fp64 n={n} equal={equal} far={far}", flush=True)
    e = cs.make_engine("mixed", 2, n, 23, ph_sort=True)
    e.set_scalar("grav_far", far)
    v64 = e.get("vel64")
    if not equal:
        rng = np.random.default_rng(5)
        v64[:n, 3] *= rng.uniform(0.5, 1.5, n)
        e.set("vel64", v64)
        vel = e.get("vel")
        vel[:n, 3] = v64[:n, 3]
        e.set("vel", vel)
    e.run_pass(cs.GRAVITY)
    g = e.get("grav")[:n].astype(np.float64)
    p = e.params.copy()
    o = orc.Oracle(64, p)
    P = e.get("pos").astype(np.float64)
    P[:n, :3] = e.get("pos64")[:n, :3]
    V = e.get("vel").astype(np.float64)
    V[:n] = v64[:n]
    o.set("pos", P)
    o.set("vel", V)
    o.run_pass(orc.GRAVITY)
    go = o.get("grav")[:n]
    rms = np.sqrt((go[:, :3] ** 2).sum(1).mean())
    err = np.linalg.norm(g[:, :3] - go[:, :3], axis=1) / np.maximum(np.linalg.norm(go[:, :3], axis=1), 1e-3 * rms)
    assert err.max() < 1e-5, err.max()
    np.testing.assert_allclose(g[:, 3], go[:, 3], rtol=2e-6)
    m = v64[:n, 3:4]
    ftot = np.abs((m * g[:, :3]).sum(0)).max()
    assert ftot <= 1e-5 * (m[:, 0] * np.linalg.norm(go[:, :3], axis=1)).sum(), ftot


@pytest.mark.parametrize("cfg,n,equal", [(3, 16384, True), (3, 16384, False), (3, 16000, True),
                                         (5, 262144, True)])
def test_mixed_gravity_far_field_against_fp64(cs, orc, cfg, n, equal):
    """The optional multipole far field (grav_far = 1: order-4 moments of cells,
    64-record sub-tiles and, by default, their 32-record halves, for
    wavefronts 8 group radii away), with and without the halves, against FP64
    gravity on the same positions, to the pairwise kernel's tolerance; equal
    masses (unit-mass moments) and masses spread over a factor of 3.  The
    uniform sphere needs N = 262144 for its sub-tiles to be small enough
    (R ~ 0.08 in a sphere of radius 1); there the reference is the FP64 form
    on the GPU (itself pinned to the FP64 oracle), the oracle's N^2 sum
    being too slow on the CPU at that size.  N = 16000 has padding records
    (and padding lanes), which must take no far-field term."""
    print(f"far field cfg={cfg} n={n} equal={equal}", flush=True)
    e = cs.make_engine("mixed", cfg, n, 29, ph_sort=True)
    v64 = e.get("vel64")
    if not equal:
        rng = np.random.default_rng(6)
        v64[:n, 3] *= rng.uniform(0.5, 1.5, n)
        e.set("vel64", v64)
        vel = e.get("vel")
        vel[:n, 3] = v64[:n, 3]
        e.set("vel", vel)
    e.run_pass(cs.GRAVITY)
    g_direct = e.get("grav")[:n].astype(np.float64)
    e.set_scalar("grav_far", 1)
    e.set_scalar("grav_far_half", 0)
    e.run_pass(cs.GRAVITY)
    g_whole = e.get("grav")[:n].astype(np.float64)
    e.set_scalar("grav_far_half", 1)   # the default: also 32-record halves
    e.run_pass(cs.GRAVITY)
    g = e.get("grav")[:n].astype(np.float64)
    assert not np.array_equal(g_whole, g_direct), "the far field was not taken"
    if n >= 262144:
        assert not np.array_equal(g, g_whole), "no 32-record half was taken"
    assert np.isfinite(e.get("grav")).all(), "padding records must stay finite"
    if n <= 16384:
        o = orc.Oracle(64, e.params.copy())
        P = e.get("pos").astype(np.float64)
        P[:n, :3] = e.get("pos64")[:n, :3]
        V = e.get("vel").astype(np.float64)
        V[:n] = v64[:n]
        o.set("pos", P)
        o.set("vel", V)
        o.run_pass(orc.GRAVITY)
        go = o.get("grav")[:n]
    else:
        f = cs.Engine("fp64", e.params.copy())
        P = f.get("pos")
        P[:n, :3] = e.get("pos64")[:n, :3]
        f.set("pos", P)
        V = f.get("vel")
        V[:n] = v64[:n]
        f.set("vel", V)
        f.run_pass(cs.GRAVITY)
        go = f.get("grav")[:n]
        f.close()
    rms = np.sqrt((go[:, :3] ** 2).sum(1).mean())
    for res in (g_whole, g):
        err = np.linalg.norm(res[:, :3] - go[:, :3], axis=1) / np.maximum(np.linalg.norm(go[:, :3], axis=1),
                                                                         1e-3 * rms)
        assert err.max() < 1e-5, err.max()
    np.testing.assert_allclose(g[:, 3], go[:, 3], rtol=2e-6)


def test_mixed_far_field_trajectory_against_fp64(cs, orc):
    """Three steps of the polytrope pair with the far field on: energy within
    1e-6 of the FP64 oracle's per step, positions to 1e-6 of their extent."""
    n = 16384
    e = cs.make_engine("mixed", 3, n, 31, ph_sort=True)
    e.set_scalar("grav_far", 1)
    o = orc.Oracle(64, e.params.copy())
    P = e.get("pos").astype(np.float64)
    P[:n, :3] = e.get("pos64")[:n, :3]
    H = e.get("hydro").astype(np.float64)
    H[:n, 2] = e.get("u64")[:n]
    o.set("pos", P)
    o.set("vel", e.get("vel64"))
    o.set("hydro", H)
    e.init()
    o.init()
    e0 = abs(o.stats()["E"])
    for _ in range(3):
        e.step()
        o.step()
    assert abs(e.stats()["E"] - o.stats()["E"]) / e0 / 3 <= 1e-6
    x, xo = e.get("pos64")[:n, :3], o.get("pos")[:n, :3]
    assert np.abs(x - xo).max() <= 1e-6 * np.abs(xo).max()


@pytest.mark.parametrize("form,far", [("plain", 0), ("mixed", 0), ("mixed", 1), ("fp64", 0)])
def test_gravity_stream_overlap_changes_nothing(cs, form, far):
    """Gravity on its own stream beside the SPH passes gives the same bits as
    the serial order (the default, overlap 1)."""
    n = 16384
    out = []
    for ov in (0, 1):
        e = cs.make_engine(form, 3, n, 9, ph_sort=True)
        e.set_scalar("overlap", ov)
        if far:
            e.set_scalar("grav_far", 1)
        e.init()
        e.step(3)
        out.append({k: e.get(k)[:n].copy() for k in ("pos", "vel", "hydro", "bdots", "grav")})
        e.close()
    for k in out[0]:
        assert_bits(out[1][k], out[0][k], k)


def walk_total(e):
    """nodes visited by the last list pass, summed over particles (the walk
    lengths it leaves in the hash keys)"""
    return int(e.get("hash")[:, 0].astype(np.int64).sum())


@pytest.mark.parametrize("form,cfg,n", [("mixed", 3, 16384), ("mixed", 2, 16000), ("mixed", 1, 16001),
                                        ("fp64", 2, 16384), ("fp64", 3, 16001)])
def test_box_pruned_walks_change_nothing(cs, form, cfg, n):
    """The mixed and FP64 forms' tree walks skip subtrees whose bounding box
    lies beyond the range.  Three steps with and without the boxes give the
    same bits everywhere (ranges, counts, every list slot in order, the
    state), while the pruned walks visit fewer nodes.  n = 16384 = n_dev puts
    the zero record at node n+1; 16000 (even) makes node n+1 reachable; 16001
    (odd) does not."""
    print(f"box-pruned walks {form} cfg={cfg} n={n}", flush=True)
    out, walks = [], []
    for boxes in (0, 1):
        e = cs.make_engine(form, cfg, n, 23, ph_sort=True)
        assert e.scalar("walk_boxes") == 1
        e.set_scalar("walk_boxes", boxes)
        e.init()
        e.step(3)
        r = {k: e.get(k)[:n].copy() for k in ("pos", "vel", "hydro", "dots", "bdots", "grav", "n_ngb", "ngb_count")}
        off = e.get("ngb_offset")
        r["ngb_offset"] = off.copy()
        r["ngb_list"] = e.get("ngb_list")[: int(off[-1])].copy()
        out.append(r)
        walks.append(walk_total(e))
        e.close()
    for k in out[0]:
        assert_bits(out[1][k], out[0][k], k)
    print(f"  nodes visited by the list pass: {walks[0]} -> {walks[1]}", flush=True)
    assert walks[1] < walks[0]


def test_state_file_round_trip(cs, tmp_path):
    n = 5000
    e = cs.make_engine("plain", 2, n, 3)
    e.init()
    e.step()
    f = tmp_path / "restart.bin"
    e.write_state(f)
    assert f.stat().st_size == 124 + 23 * 8 * n
    e2 = cs.Engine("plain", e.params)
    e2.read_state(f)
    for name in ("pos", "vel", "bdots", "grav"):
        assert np.array_equal(e2.get(name)[:n], e.get(name)[:n]), name
    h1, h2 = e.get("hydro")[:n], e2.get("hydro")[:n]
    assert np.array_equal(h1[:, [0, 2]], h2[:, [0, 2]])
    np.testing.assert_allclose(h2[:, 3], h1[:, 3], rtol=1e-6)   # mu goes through CGS amu and back
    assert e2.scalar("dt") == e.scalar("dt") and e2.scalar("iteration") == e.scalar("iteration")


def test_mixed_state_file_keeps_fp64_state(cs, tmp_path):
    """The mixed form integrates x, v and u in FP64: its state file carries
    those doubles (not their FP32 views), so a read gives back pos64, vel64
    and u64 bit for bit after two steps have moved them off FP32 values."""
    n = 5000
    e = cs.make_engine("mixed", 3, n, 3, ph_sort=True)
    e.init()
    e.step(2)
    p64, v64, u64 = e.get("pos64")[:n].copy(), e.get("vel64")[:n].copy(), e.get("u64")[:n].copy()
    assert (p64[:, :3] != p64[:, :3].astype(np.float32)).any(), "positions still FP32-representable"
    assert (u64 != u64.astype(np.float32)).any(), "u still FP32-representable"
    f = tmp_path / "restart.bin"
    e.write_state(f)
    rec = np.fromfile(f, dtype=np.float64, offset=124).reshape(n, 23)
    assert np.array_equal(rec[:, 0:3], p64[:, :3])
    assert np.array_equal(rec[:, 6:9], v64[:, :3])
    assert np.array_equal(rec[:, 3], v64[:, 3])
    assert np.array_equal(rec[:, 12], u64)
    e2 = cs.Engine("mixed", e.params)
    e2.read_state(f)
    assert np.array_equal(e2.get("pos64")[:n, :3], p64[:, :3])
    assert np.array_equal(e2.get("vel64")[:n], v64)
    assert np.array_equal(e2.get("u64")[:n], u64)
    assert np.array_equal(e2.get("pos")[:n, :3], e.get("pos")[:n, :3])
    e.close()
    e2.close()



@pytest.mark.parametrize("n", [1, 63, 64, 65, 127, 129, 4099, 16384, 16385, 65536, 65537])
@pytest.mark.parametrize("exact", [0, 1])
def test_plain_gravity_sums_apart_change_nothing(cs, n, exact):
    """The plain form's Plummer gravity with each particle's four sums in
    four wavefronts and in two gives the bits of the one-lane kernel
    (hardware and correctly rounded rsqrt), at sizes around the 64- and
    128-particle blocks and on either side of the automatic switches at
    16 384 and 65 536."""
    print(f"plain gravity sums n={n} exact={exact}", flush=True)
    out = []
    for sums in (0, 1, 2, -1):
        e = cs.make_engine("plain", 1, n, 4, ph_sort=True)
        assert e.params.grav_mode == 1
        e.set_scalar("exact_rsqrt", exact)
        e.set_scalar("grav_sums", sums)
        e.init()
        e.run_pass(cs.GRAVITY)
        out.append(e.get("grav")[:n].copy())
        e.close()
    for k, name in ((1, "four wavefronts"), (2, "two wavefronts"), (3, "automatic")):
        assert_bits(out[k], out[0], f"grav ({name})")


@pytest.mark.parametrize("n", [65536, 40001])
def test_mixed_gravity_parts_agree(cs, n):
    """The pairwise mixed kernel with its j-range in 8, 3 and 1 parts and the
    automatic count (4 at 65 536 particles, 6 at 40 001 on a 256-CU GPU): only
    the grouping of the compensated sums changes, so the accelerations agree
    within the form's tolerance (1e-5 of max(|a|, 1e-3 rms); particles whose
    forces nearly cancel differ most) and the potentials to 1e-6, and the
    default is the automatic count's bits."""
    e = cs.make_engine("mixed", 4, n, 5, ph_sort=True)
    auto = int(e.scalar("grav_parts_auto"))
    assert 1 <= auto <= 8
    out = {}
    for parts in (8, 3, 1, auto, 0):
        e.set_scalar("grav_parts", parts)
        e.run_pass(cs.GRAVITY)
        out[parts] = e.get("grav")[:n].astype(np.float64)
    e.close()
    assert_bits(out[0].astype(np.float32), out[auto].astype(np.float32), f"default = {auto} parts")
    b = out[8][:, :3]
    rms = np.sqrt((b ** 2).sum(1).mean())
    for parts in (3, 1, auto):
        a = out[parts][:, :3]
        err = np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-3 * rms)
        assert err.max() < 1e-5, (parts, err.max())
        pe = np.abs(out[parts][:, 3] - out[8][:, 3]) / np.abs(out[8][:, 3])
        assert pe.max() < 1e-6, (parts, pe.max())


@pytest.mark.parametrize("n", [65536, 40001])
def test_fp64_gravity_parts_agree(cs, n):
    """The FP64 forward gravity with its j-range in 8, 2 and 1 parts and the
    automatic count (2 at 65 536 particles, 4 at 40 001 on a 256-CU GPU)
    agrees to FP64 rounding; the default is the automatic count's bits."""
    e = cs.make_engine("fp64", 4, n, 5, ph_sort=True)
    auto = int(e.scalar("grav_parts_auto"))
    assert 1 <= auto <= 8
    out = {}
    for parts in (8, 2, 1, auto, 0):
        e.set_scalar("grav_parts", parts)
        e.run_pass(cs.GRAVITY)
        out[parts] = e.get("grav")[:n].copy()
    e.close()
    assert_bits(out[0], out[auto], f"default = {auto} parts")
    b = out[8]
    scale = np.abs(b).max(0)
    for parts in (2, 1, auto):
        assert (np.abs(out[parts] - b) <= 1e-12 * scale).all(), parts


@pytest.mark.parametrize("form,cfg,n", [("plain", 1, 16384), ("plain", 3, 5000), ("mixed", 1, 16384),
                                        ("mixed", 2, 30001), ("fp64", 4, 12000), ("plain", 2, 60000), ("fp64", 2, 40000)])
def test_solve_cache_changes_nothing(cs, form, cfg, n):
    """The h-solve with its candidate cache (one walk at 9/8 of the range,
    later evaluations filtered from it, or walked where a record's reach is
    not decided) against the walks alone: the same ranges and counts, and
    three whole steps the same in every bit."""
    print(f"solve cache {form} cfg={cfg} n={n}", flush=True)
    out = []
    for cache in (0, 1):
        e = cs.make_engine(form, cfg, n, 17, ph_sort=True)
        e.set_scalar("solve_cache", cache)
        e.init()
        r = [(e.get("pos")[:n, 3].copy(), e.get("n_ngb")[:n].copy())]
        e.step(3)
        r.append({k: e.get(k)[:n].copy() for k in ("pos", "vel", "hydro", "bdots", "n_ngb")})
        out.append(r)
        e.close()
    assert_bits(out[1][0][0], out[0][0][0], "init ranges")
    assert np.array_equal(out[1][0][1], out[0][0][1]), "init counts"
    for k in out[0][1]:
        assert_bits(out[1][1][k], out[0][1][k], k)
