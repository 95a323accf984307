"""CPU-side checks of the product library without a GPU: every symbol declared
in include/cudasph.h is exported, parameter presets, the seeded initial
conditions, and the reference state-file layout."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "cudasph.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"^[\w\s\*]*?\b([A-Za-z_]\w*)\s*\(([^;{]*)\)\s*;", text, flags=re.M):
        if m.group(1) not in ("if", "while", "return", "sizeof"):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol(cs):
    lib = C.CDLL(str(cs.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 39
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    for ref in ("host_solve_range", "host_sph_accelerations", "compute_CUDA_gravity", "allocateCUDAarray",
                "zcopyArrayToDevice", "RadixSort", "get_device_count", "cuda_set_device0"):
        assert ref in names


def test_reference_defaults(cs):
    p = cs.default_params(1000)
    assert p.n_dev == 1024 and p.rel_err == pytest.approx(1e-3) and p.desired_ngb == 24
    assert p.max_iter == 20 and p.av_alpha == 1 and p.av_beta == 2 and p.grav_mode == 0
    assert p.eos_density_unit == 5.89994031541978 and p.eos_temp_base == 1e6
    assert p.ph_period == 100 and p.ph_bits == 18 and p.eos_newton == 0
    assert cs.default_params(256).n_dev == 256 and cs.default_params(257).n_dev == 512


@pytest.mark.parametrize("cfg,n,ngb,eps", [(1, 16384, 32, 0.01), (2, 262144, 50, 0.005),
                                          (3, 1048576, 64, 0.002), (4, 65536, 50, 0.005),
                                          (5, 4194304, 64, 0.001)])
def test_presets(cs, cfg, n, ngb, eps):
    p = cs.preset_params(cfg)
    assert p.n_bodies == n and p.n_dev == n
    assert p.grav_mode == 1 and p.grav_eps == pytest.approx(eps)
    assert p.desired_ngb == pytest.approx(0.7061074 * ngb, rel=1e-6)
    assert p.eos_rad_const == 1e-3 and p.eos_boltzmann == 1 and p.eos_newton == 1


def test_ic_is_seeded_and_deterministic(cs):
    a = cs.generate_ic(2, 4096, 77)
    b = cs.generate_ic(2, 4096, 77)
    c = cs.generate_ic(2, 4096, 78)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])


def test_ic_shapes(cs):
    n = 8192
    pos, vel, hyd = cs.generate_ic(1, n, 1)
    assert pos[:, :3].min() >= 0 and pos[:, :3].max() < 1 and np.all(vel[:, :3] == 0)
    assert np.all(pos[:, :3] * 2**24 == np.round(pos[:, :3] * 2**24)), "24-bit grid"
    assert vel[:, 3].sum() == pytest.approx(1.0) and np.all(hyd[:, 2] == 1.0) and np.all(hyd[:, 3] == 0.6)
    pos, vel, hyd = cs.generate_ic(2, n, 1)
    r = np.linalg.norm(pos[:, :3], axis=1)
    assert r.max() <= 10.0
    # Plummer cumulative mass M(<r) = r^3 / (1 + r^2)^1.5
    for rr in (0.5, 1.0, 2.0):
        assert (r < rr).mean() == pytest.approx(rr**3 / (1 + rr**2) ** 1.5 / (1000 / 101**1.5), abs=0.02)
    vesc = np.sqrt(2.0) * (1 + r * r) ** -0.25
    assert np.all(np.linalg.norm(vel[:, :3], axis=1) <= vesc + 1e-12)
    pos, vel, hyd = cs.generate_ic(3, n, 1)
    c1, c2 = pos[: n // 2, :3].mean(0), pos[n // 2:, :3].mean(0)
    np.testing.assert_allclose(c2 - c1, [-4, -3, 0], atol=0.05)
    v1, v2 = vel[: n // 2, :3].mean(0), vel[n // 2:, :3].mean(0)
    np.testing.assert_allclose(v2 - v1, [0.6, 0.2, 0], atol=1e-12)
    assert np.linalg.norm(pos[: n // 2, :3] - c1, axis=1).max() < 1.0 + 0.05
    pos, vel, hyd = cs.generate_ic(5, 32768, 1)
    assert np.linalg.norm(pos[:, :3], axis=1).max() <= 1.0
    vrms = np.sqrt((vel[:, :3] ** 2).sum(1).mean())
    assert vrms == pytest.approx(2 * np.sqrt(10 / 9), rel=1e-9)


def test_kd_tree_builder_thread_count_does_not_change_the_tree(cs):
    pos, _, _ = cs.generate_ic(3, 100000, 5)
    t1, m1 = cs.build_kd_tree(pos.astype(np.float32), None, 1)
    t8, m8 = cs.build_kd_tree(pos.astype(np.float32), None, 8)
    assert np.array_equal(t1.view(np.uint32), t8.view(np.uint32)) and np.array_equal(m1, m8)
    d1, _ = cs.build_kd_tree(pos, None, 1)
    d8, _ = cs.build_kd_tree(pos, None, 8)
    assert np.array_equal(d1, d8)
