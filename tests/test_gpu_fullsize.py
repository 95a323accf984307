"""GPU parity at the benchmark sizes: one SPH force evaluation (kd-tree,
h-solve, offsets, lists, density, EOS, hydro, reaction) of the plain form
against the CPU oracle, bit for bit, on the Plummer sphere at N = 262,144 and
the polytrope pair at N = 1,048,576 (lists compared slot by slot, in walk
order), plus a property that holds at any size: the hydro and reaction forces
conserve total momentum to rounding."""
import numpy as np
import pytest

from helpers import assert_bits

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,n", [(2, 262144), (3, 1048576)])
def test_sph_force_evaluation_bit_exact_at_full_size(cs, orc, cfg, n):
    """One force evaluation from the initial state: ranges, counts, offsets,
    walk lengths, every list slot in walk order and the SPH outputs
    bit-identical to the oracle's."""
    p = cs.preset_params(cfg, n)
    pos, vel, hyd = cs.generate_ic(cfg, n, 12345)
    p.n_threads = 16
    o = orc.Oracle(32, p)
    o.load_ic(pos, vel, hyd)
    o.run_pass(orc.SPH_ACC)
    print(f"full-size SPH evaluation cfg={cfg} n={n}: oracle done", flush=True)
    e = cs.Engine("plain", p)
    e.load_ic(pos, vel, hyd)
    e.run_pass(cs.SPH_ACC)
    check_force_evaluation(e, o, p, n, vel[:, 3])


def walk_lengths(sim, n):
    """nodes each particle's list walk visited (left in the hash keys)"""
    h = sim.get("hash")
    out = np.zeros(len(h), np.int64)
    out[h[:, 1]] = h[:, 0]
    return out[:n]


def check_force_evaluation(e, o, p, n, m):
    assert_bits(e.get("pos")[:n], o.get("pos")[:n], "pos (ranges)")
    assert np.array_equal(walk_lengths(e, n), walk_lengths(o, n)), "walk lengths differ"
    assert np.array_equal(e.get("n_ngb")[:n], o.get("n_ngb")[:n])
    cnt = e.get("ngb_count")
    assert np.array_equal(cnt[:n], o.get("ngb_count")[:n])
    off = e.get("ngb_offset")
    assert np.array_equal(off, o.get("ngb_offset"))
    tot = int(off[p.n_dev])
    le, lo = e.get("ngb_list")[:tot], o.get("ngb_list")[:tot]
    # every written slot of every particle (slots past a particle's count are padding)
    starts = off[:n]
    lens = cnt[:n]
    idx = np.repeat(starts - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens) + np.arange(lens.sum())
    assert np.array_equal(le[idx], lo[idx]), "neighbour lists differ"
    for name in ("hydro", "dots", "bdots"):
        assert_bits(e.get(name)[:n], o.get(name)[:n], name)
    assert_bits(e.get("hydro2")[:n, 0], o.get("hydro2")[:n, 0], "v_sig")
    # hydro + reaction conserve momentum to rounding
    bd = e.get("bdots")[:n].astype(np.float64)
    m = m[:n]
    total = (m[:, None] * bd[:, :3]).sum(0)
    scale = (m[:, None] * np.abs(bd[:, :3])).sum(0)
    assert np.all(np.abs(total) <= 1e-5 * scale), (total, scale)
    print(f"  mean list length {cnt[:n].mean():.2f}, ngb_tot {tot}", flush=True)


@pytest.mark.parametrize("cfg", [2, 3])
def test_box_pruned_walks_change_nothing_at_full_size(cs, cfg):
    """The mixed form with and without subtree-box pruning of its walks, one
    step from the same state at the config's full size: the same bits in the
    ranges, counts, every list slot and the state."""
    n = cs.preset_params(cfg, 0).n_bodies
    print(f"box-pruned walks at full size cfg={cfg} n={n}", flush=True)
    out, walks = [], []
    for boxes in (0, 1):
        e = cs.make_engine("mixed", cfg, n, 12345, ph_sort=True)
        e.set_scalar("walk_boxes", boxes)
        e.init()
        e.step()
        r = {k: e.get(k)[:n].copy() for k in ("pos", "vel", "hydro", "bdots", "n_ngb", "ngb_count")}
        off = e.get("ngb_offset")
        r["ngb_list"] = e.get("ngb_list")[: int(off[-1])].copy()
        out.append(r)
        walks.append(float(e.get("hash")[:, 0].astype(np.int64).sum()) / n)
        e.close()
    for k in out[0]:
        assert_bits(out[1][k], out[0][k], k)
    print(f"  nodes per list walk: {walks[0]:.1f} -> {walks[1]:.1f}", flush=True)
    assert walks[1] < walks[0]


def test_device_tree_same_neighbour_sets_at_full_size(cs):
    """At N = 2^20 the 24-bit coordinate draws tie many times per axis, so the
    device tree build (the mixed form's default) may pick other tied
    particles as nodes than the reference's quicksort build does.  The
    neighbour sets of every particle must still be the host tree's: lists
    built from both trees on the same positions and ranges, compared as
    sorted (particle, neighbour) keys."""
    n = 1048576
    print(f"full-size device tree n={n}", flush=True)
    p = cs.preset_params(3, n)
    pos, vel, hyd = cs.generate_ic(3, n, 12345)
    e = cs.Engine("plain", p)
    e.load_ic(pos, vel, hyd)

    def keys():
        off, cnt, lst = e.get("ngb_offset")[:n], e.get("ngb_count")[:n], e.get("ngb_list")
        owner = np.repeat(np.arange(n, dtype=np.int64), cnt)
        idx = np.repeat(off.astype(np.int64) - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt) + np.arange(cnt.sum())
        return np.sort(owner * (1 << 24) + lst[idx].astype(np.int64))

    e.set_scalar("tree_gpu", 0)
    e.run_pass(cs.TREE)
    e.run_pass(cs.SOLVE)
    e.run_pass(cs.OFFSETS)
    e.run_pass(cs.FIND)
    host = keys()
    e.set_scalar("tree_gpu", 1)
    e.run_pass(cs.TREE)
    e.run_pass(cs.OFFSETS)
    e.run_pass(cs.FIND)
    dev = keys()
    assert host.size == dev.size, (host.size, dev.size)
    assert np.array_equal(host, dev), "neighbour sets differ between the two tree builds"
    print(f"  {host.size} (particle, neighbour) pairs identical", flush=True)
    e.close()


@pytest.mark.parametrize("cfg", [3, 5])
def test_far_field_against_fp64_at_full_size(cs, cfg):
    """The mixed form's optional far field at the benchmark size, against FP64
    gravity on the same positions (the FP64 form on the GPU): the maximum
    per-particle acceleration error stays within the pairwise FP32 kernel's
    tolerance and the median within 1e-7."""
    n = 1048576
    print(f"full-size far field cfg={cfg} n={n}", flush=True)
    e = cs.make_engine("mixed", cfg, n, 12345, ph_sort=True, i_order=0)
    e.set_scalar("grav_far", 1)
    e.run_pass(cs.GRAVITY)
    g = e.get("grav")[:n].astype(np.float64)
    pos64, vel64 = e.get("pos64")[:n, :3].copy(), e.get("vel64")[:n].copy()
    p = e.params.copy()
    e.close()
    f = cs.Engine("fp64", p)
    P = f.get("pos")
    P[:n, :3] = pos64
    f.set("pos", P)
    V = f.get("vel")
    V[:n] = vel64
    f.set("vel", V)
    f.run_pass(cs.GRAVITY)
    go = f.get("grav")[:n]
    f.close()
    rms = np.sqrt((go[:, :3] ** 2).sum(1).mean())
    err = np.linalg.norm(g[:, :3] - go[:, :3], axis=1) / np.maximum(np.linalg.norm(go[:, :3], axis=1), 1e-3 * rms)
    print(f"  median {np.median(err):.2e} max {err.max():.2e}", flush=True)
    assert err.max() < 1e-5 and np.median(err) < 1e-7, (err.max(), np.median(err))


def test_energy_drift_against_fp64_at_full_size(cs):
    """The benchmark's accuracy condition at its own size: over init + 2 steps
    of the polytrope pair at N = 2^20, the mixed and plain forms keep
    |E_form(k) - E_fp64(k)| / |E_fp64(0)| / k below 1e-6."""
    n = 1048576
    print(f"full-size energy drift n={n}", flush=True)
    energies = {}
    for form in ("fp64", "mixed", "plain"):
        e = cs.make_engine(form, 3, n, 12345, ph_sort=True, i_order=0)
        e.init()
        es = [e.stats()["E"]]
        for _ in range(2):
            e.step()
            es.append(e.stats()["E"])
        energies[form] = np.array(es)
        e.close()
        print(f"  {form} done", flush=True)
    ref = energies["fp64"]
    for form in ("mixed", "plain"):
        drift = np.abs(energies[form][1:] - ref[1:]) / abs(ref[0]) / np.arange(1, len(ref))
        print(f"  {form} drift per step {drift.max():.2e}", flush=True)
        assert drift.max() < 1e-6, (form, drift)


def sampled_fp64_gravity(orc, params, pos4, vel4, windows, width):
    """FP64 oracle gravity (sph_grav_kernel.cu:14-111 restated in double)
    for i in the given windows of `width` particles, against all j."""
    n = params.n_bodies
    o = orc.Oracle(64, params)
    P = np.zeros((params.n_dev, 4))
    P[:n] = pos4[:n]
    V = np.zeros((params.n_dev, 4))
    V[:n] = vel4[:n]
    o.set("pos", P)
    o.set("vel", V)
    for i0 in windows:
        o.gravity_range(int(i0), int(i0) + width)
    g = o.get("grav")
    idx = np.concatenate([np.arange(i0, i0 + width) for i0 in windows])
    return idx, g[idx]


def gravity_errors(g, ref):
    """Per-particle |a - a_ref| / max(|a_ref|, 1e-3 rms |a_ref|) and the
    potential's relative error."""
    rms = np.sqrt((ref[:, :3] ** 2).sum(1).mean())
    err = np.linalg.norm(g[:, :3] - ref[:, :3], axis=1) / np.maximum(np.linalg.norm(ref[:, :3], axis=1), 1e-3 * rms)
    perr = np.abs(g[:, 3] - ref[:, 3]) / np.abs(ref[:, 3])
    return err, perr


@pytest.mark.parametrize("cfg", [3, 5])
def test_gravity_per_particle_against_fp64_oracle_at_full_size(cs, orc, cfg):
    """The headline kernel at the headline size: 2048 i-particles (16 windows
    of 128 spread over the Peano-Hilbert order) of N = 2^20 against the FP64
    oracle's direct sum over all j, for the mixed form's pairwise kernel, its
    far field, and the plain form's kernel with the hardware rsq the
    reference uses.  Bounds: max 1e-5 and median 1e-7 (mixed, far field),
    max 1e-5 and median 1e-6 (plain); potentials to 2e-6."""
    n = 1 << 20
    print(f"full-size gravity per particle cfg={cfg} n={n}", flush=True)
    width = 128
    windows = np.linspace(0, n - width, 16).astype(np.int64)
    e = cs.make_engine("mixed", cfg, n, 12345, ph_sort=True, i_order=0)
    e.run_pass(cs.GRAVITY)
    g_pair = e.get("grav")[:n].astype(np.float64)
    e.set_scalar("grav_far", 1)
    e.run_pass(cs.GRAVITY)
    g_far = e.get("grav")[:n].astype(np.float64)
    pos64, vel64 = e.get("pos64")[:n].copy(), e.get("vel64")[:n].copy()
    p = e.params.copy()
    e.close()
    p.n_threads = 16
    idx, ref = sampled_fp64_gravity(orc, p, pos64, vel64, windows, width)
    print("  oracle (mixed state) done", flush=True)
    for name, g, med_bound in (("pairwise mixed", g_pair, 1e-7), ("far field", g_far, 1e-7)):
        err, perr = gravity_errors(g[idx], ref)
        print(f"  {name}: median {np.median(err):.2e} max {err.max():.2e} potential max {perr.max():.2e}", flush=True)
        assert err.max() < 1e-5 and np.median(err) < med_bound, (name, err.max(), np.median(err))
        assert perr.max() < 2e-6, (name, perr.max())

    pl = cs.make_engine("plain", cfg, n, 12345, ph_sort=True, i_order=0)
    pl.set_scalar("exact_rsqrt", 0)
    pl.run_pass(cs.GRAVITY)
    g_plain = pl.get("grav")[:n].astype(np.float64)
    ppos, pvel = pl.get("pos")[:n].astype(np.float64), pl.get("vel")[:n].astype(np.float64)
    pp = pl.params.copy()
    pl.close()
    pp.n_threads = 16
    idx, ref = sampled_fp64_gravity(orc, pp, ppos, pvel, windows, width)
    err, perr = gravity_errors(g_plain[idx], ref)
    print(f"  plain (hardware rsq): median {np.median(err):.2e} max {err.max():.2e} "
          f"potential max {perr.max():.2e}", flush=True)
    assert err.max() < 1e-5 and np.median(err) < 1e-6, (err.max(), np.median(err))
    assert perr.max() < 2e-6, perr.max()


@pytest.mark.parametrize("cfg", [3, 5])
def test_mixed_sph_passes_against_fp64_oracle_at_full_size(cs, orc, cfg):
    """One SPH force evaluation of the mixed form at N = 2^20 against the FP64
    oracle run on the mixed form's own FP64 state and ranges (tree, lists,
    density, EOS, hydro, reaction in double), over the particles whose lists
    agree: rho and P to 2e-6; hydro accelerations, relative to
    max(|a|, 1e-3 rms), median 1e-6, 99.9th percentile 1e-5, max 1e-4 (the
    pressure forces of a star in hydrostatic balance cancel to a net force
    well below their pair terms, which the FP32 pair arithmetic rounds);
    du/dt the same where the state has one (the turbulent sphere: each
    polytrope moves rigidly at t = 0, so there du/dt vanishes to rounding),
    relative to the larger of |du/dt| and its gross rate (P/rho)(|div v| +
    |curl v|): where compression and shear cancel, du/dt is a small
    difference of terms of that size."""
    n = 1 << 20
    print(f"full-size mixed SPH passes cfg={cfg} n={n}", flush=True)
    e = cs.make_engine("mixed", cfg, n, 12345, ph_sort=True, i_order=0)
    e.run_pass(cs.SPH_ACC)
    p = e.params.copy()
    p.n_threads = 16
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
    for pid in (orc.TREE, orc.OFFSETS, orc.FIND, orc.DENSITY, orc.EOS, orc.HYDRO, orc.REACTION):
        o.run_pass(pid)
    print("  oracle done", flush=True)
    same = e.get("ngb_count")[:n] == o.get("ngb_count")[:n]
    assert same.mean() > 0.999, same.mean()
    he, ho = e.get("hydro")[:n].astype(np.float64), o.get("hydro")[:n]
    rho_err = (np.abs(he[:, 0] - ho[:, 0]) / ho[:, 0])[same]
    p_err = (np.abs(he[:, 1] - ho[:, 1]) / ho[:, 1])[same]
    off, cnt, lst = o.get("ngb_offset"), o.get("ngb_count"), o.get("ngb_list")
    ok = same.copy()
    for i in np.nonzero(~same)[0]:
        ok[lst[off[i]:off[i] + cnt[i]]] = False
    a, b = e.get("bdots")[:n].astype(np.float64)[ok], o.get("bdots")[:n][ok]
    floor = 1e-3 * np.sqrt((b[:, :3] ** 2).sum(1).mean())
    a_err = np.linalg.norm(a[:, :3] - b[:, :3], axis=1) / np.maximum(np.linalg.norm(b[:, :3], axis=1), floor)
    print(f"  lists equal {same.mean():.6f}; rho max {rho_err.max():.2e}, P max {p_err.max():.2e}, "
          f"a median {np.median(a_err):.2e} 99.9% {np.quantile(a_err, 0.999):.2e} max {a_err.max():.2e}", flush=True)
    assert rho_err.max() < 2e-6 and p_err.max() < 2e-6
    assert np.median(a_err) < 1e-6 and np.quantile(a_err, 0.999) < 1e-5 and a_err.max() < 1e-4
    if cfg == 5:
        dots, hyd = o.get("dots")[:n][ok], ho[ok]
        gross = hyd[:, 1] / hyd[:, 0] * (np.abs(dots[:, 0]) + dots[:, 1])
        u_err = np.abs(a[:, 3] - b[:, 3]) / np.maximum(np.maximum(np.abs(b[:, 3]), gross),
                                                        1e-3 * np.sqrt((b[:, 3] ** 2).mean()))
        print(f"  du/dt median {np.median(u_err):.2e} 99.9% {np.quantile(u_err, 0.999):.2e} max {u_err.max():.2e}",
              flush=True)
        assert np.median(u_err) < 1e-6 and np.quantile(u_err, 0.999) < 1e-5 and u_err.max() < 1e-4


def test_scale_workload_at_4m(cs, orc):
    """The scale workload at its own size, N = 4,194,304 (turbulent sphere,
    rms Mach 2, 64 +- 8 neighbours, eps = 0.001): the mixed form's pairwise
    gravity for 1024 sampled particles against the FP64 oracle's direct sum
    over all 4M (max 1e-5, median 1e-7); the multipole far field against the
    pairwise result (max 1e-5); one SPH force evaluation conserving momentum
    to rounding, with list lengths in the configuration's 64 +- 8 band."""
    n = 1 << 22
    print(f"scale workload n={n}", flush=True)
    e = cs.make_engine("mixed", 5, n, 12345, ph_sort=True, i_order=0)
    e.run_pass(cs.GRAVITY)
    g_pair = e.get("grav")[:n].astype(np.float64)
    e.set_scalar("grav_far", 1)
    e.run_pass(cs.GRAVITY)
    g_far = e.get("grav")[:n].astype(np.float64)
    e.set_scalar("grav_far", 0)
    e.run_pass(cs.SPH_ACC)
    bd = e.get("bdots")[:n].astype(np.float64)
    cnt = e.get("ngb_count")[:n]
    pos64, vel64 = e.get("pos64")[:n].copy(), e.get("vel64")[:n].copy()
    p = e.params.copy()
    e.close()
    print(f"  mean list length {cnt.mean():.2f}", flush=True)
    assert 56 <= cnt.mean() <= 72
    m = vel64[:, 3]
    total = (m[:, None] * bd[:, :3]).sum(0)
    scale = (m[:, None] * np.abs(bd[:, :3])).sum(0)
    assert np.all(np.abs(total) <= 1e-5 * scale), (total, scale)
    rms = np.sqrt((g_pair[:, :3] ** 2).sum(1).mean())
    far_err = np.linalg.norm(g_far[:, :3] - g_pair[:, :3], axis=1) / np.maximum(
        np.linalg.norm(g_pair[:, :3], axis=1), 1e-3 * rms)
    print(f"  far field vs pairwise: median {np.median(far_err):.2e} max {far_err.max():.2e}", flush=True)
    assert far_err.max() < 1e-5
    p.n_threads = 16
    windows = np.linspace(0, n - 128, 8).astype(np.int64)
    idx, ref = sampled_fp64_gravity(orc, p, pos64, vel64, windows, 128)
    err, perr = gravity_errors(g_pair[idx], ref)
    print(f"  pairwise vs FP64 oracle: median {np.median(err):.2e} max {err.max():.2e} "
          f"potential max {perr.max():.2e}", flush=True)
    assert err.max() < 1e-5 and np.median(err) < 1e-7
    assert perr.max() < 2e-6
