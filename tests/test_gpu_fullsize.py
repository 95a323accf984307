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
    print(f"full-size SPH evaluation cfg={cfg} n={n}", flush=True)
    p = cs.preset_params(cfg, n)
    pos, vel, hyd = cs.generate_ic(cfg, n, 12345)
    e = cs.Engine("plain", p)
    e.load_ic(pos, vel, hyd)
    e.run_pass(cs.SPH_ACC)
    p.n_threads = 16
    o = orc.Oracle(32, p)
    o.load_ic(pos, vel, hyd)
    o.run_pass(orc.SPH_ACC)
    print("  oracle done", flush=True)
    assert_bits(e.get("pos")[:n], o.get("pos")[:n], "pos (ranges)")
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
    m = vel[:n, 3]
    total = (m[:, None] * bd[:, :3]).sum(0)
    scale = (m[:, None] * np.abs(bd[:, :3])).sum(0)
    assert np.all(np.abs(total) <= 1e-5 * scale), (total, scale)
    print(f"  mean list length {cnt[:n].mean():.2f}, ngb_tot {tot}", flush=True)
