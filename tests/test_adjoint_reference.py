"""Pins the FP64 PyTorch restatement of the adjoint's map (adjoint_reference.py)
on the CPU: its trajectory equals the FP64 oracle's when given the oracle's
own ranges and timesteps, and its autograd gradient agrees with central
finite differences of itself."""
import numpy as np
import pytest
import torch

import adjoint_reference as ar


def oracle_run(cs, orc, n, K, seed=4):
    p = cs.preset_params(4, n)
    pos, vel, hyd = cs.generate_ic(4, n, seed)
    o = orc.Oracle(64, p)
    o.load_ic(pos, vel, hyd)
    o.init()
    ranges, dts = [o.get("pos")[:n, 3].copy()], [o.scalar("dt")]
    for _ in range(K):
        o.step()
        ranges.append(o.get("pos")[:n, 3].copy())
        dts.append(o.scalar("dt"))
    return p, (pos, vel, hyd), o, ranges, dts


def test_forces_equal_the_fp64_oracle_passes(cs, orc):
    n = 768
    p, (pos, vel, hyd), o, ranges, dts = oracle_run(cs, orc, n, 0)
    c = ar.MapParams.of(p)
    t = torch.tensor
    a, udot, rho = ar.forces(t(pos[:n, :3]), t(vel[:n, :3]), t(hyd[:n, 2]), t(vel[:n, 3]), t(hyd[:n, 3]),
                             t(ranges[0]), c)
    bd = o.get("bdots")[:n]
    scale = np.abs(bd[:, :3]).max()
    assert np.abs(a.numpy() - bd[:, :3]).max() <= 1e-11 * scale
    assert np.abs(udot.numpy() - bd[:, 3]).max() <= 1e-11 * np.abs(bd[:, 3]).max()
    np.testing.assert_allclose(rho.numpy(), o.get("hydro")[:n, 0], rtol=1e-12)


def test_trajectory_equals_the_fp64_oracle(cs, orc):
    n, K = 768, 3
    p, (pos, vel, hyd), o, ranges, dts = oracle_run(cs, orc, n, K)
    c = ar.MapParams.of(p)
    t = torch.tensor
    loss, (x, v, u, rho) = ar.run_map(t(pos[:n, :3]), t(vel[:n, :3]), t(hyd[:n, 2]), t(vel[:n, 3]),
                                      t(hyd[:n, 3]), [t(r) for r in ranges], dts, c, return_state=True)
    P, V, H = o.get("pos")[:n], o.get("vel")[:n], o.get("hydro")[:n]
    assert np.abs(x.numpy() - P[:, :3]).max() <= 1e-12 * np.abs(P[:, :3]).max()
    assert np.abs(v.numpy() - V[:, :3]).max() <= 1e-10 * np.abs(V[:, :3]).max()
    np.testing.assert_allclose(u.numpy(), H[:, 2], rtol=1e-10)
    np.testing.assert_allclose(rho.numpy(), H[:, 0], rtol=1e-10)
    want = 0.5 * (V[:, 3] * (V[:, :3] ** 2).sum(1)).sum() + (H[:, 0] ** 2).sum()
    assert loss.item() == pytest.approx(want, rel=1e-12)


def test_autograd_matches_central_differences(cs, orc):
    """The restated map's gradient against its own central differences along
    seeded unit directions (step 1e-4)."""
    n, K = 256, 2
    p, (pos, vel, hyd), o, ranges, dts = oracle_run(cs, orc, n, K)
    loss, gx, gv, gu = ar.gradient(pos[:n], vel[:n], hyd[:n], ranges, dts, p)
    g = np.concatenate([gx.ravel(), gv.ravel(), gu])
    z = np.concatenate([pos[:n, :3].ravel(), vel[:n, :3].ravel(), hyd[:n, 2]])
    c = ar.MapParams.of(p)
    t = torch.tensor
    m, mu, rng = t(vel[:n, 3]), t(hyd[:n, 3]), [t(r) for r in ranges]

    def L(zz):
        x = t(zz[: 3 * n].reshape(n, 3))
        v = t(zz[3 * n: 6 * n].reshape(n, 3))
        u = t(zz[6 * n:])
        with torch.no_grad():
            return ar.run_map(x, v, u, m, mu, rng, dts, c).item()

    assert L(z) == pytest.approx(loss, rel=1e-14)
    rs = np.random.default_rng(11)
    h = 1e-4
    for _ in range(6):
        d = rs.standard_normal(z.size)
        d /= np.linalg.norm(d)
        fd = (L(z + h * d) - L(z - h * d)) / (2 * h)
        assert abs(g @ d - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.linalg.norm(g)), (g @ d, fd)
