"""GPU tests of the reverse mode (sph_adjoint_*): the hand-written adjoint of
the FP64 form against the FP64 CPU reference gradient (torch autograd through
adjoint_reference.py, given the GPU run's own ranges and timesteps), against
central finite differences of the recorded map (step 1e-4, 64 seeded unit
directions), checkpoint-interval independence, and the torch.autograd
wrapper."""
import numpy as np
import pytest
import torch

import adjoint_reference as ar
from torch_adjoint import SphRunLoss

pytestmark = pytest.mark.gpu


def fp64_engine(cs, n, seed=4):
    p = cs.preset_params(4, n)
    e = cs.Engine("fp64", p)
    pos, vel, hyd = cs.generate_ic(4, n, seed)
    e.load_ic(pos, vel, hyd)
    return e, p, (pos, vel, hyd)


def flat(gx, gv, gu):
    return np.concatenate([gx.ravel(), gv.ravel(), gu.ravel()])


@pytest.mark.parametrize("K,ck,n", [(4, 3, 1024), (0, 1, 1024), (3, 2, 1000)])
def test_adjoint_matches_fp64_cpu_reference_gradient(cs, K, ck, n):
    """n = 1000 is padded to 1024 records: the padding carries no mass and is
    in no neighbour list."""
    print(f"adjoint vs torch reference n={n} K={K} ck={ck}", flush=True)
    e, p, (pos, vel, hyd) = fp64_engine(cs, n)
    loss = e.adjoint_forward(K, ck)
    g = flat(*e.adjoint_backward())
    ranges, dts = e.adjoint_records()
    assert ranges.shape == (K + 1, n) and len(dts) == K + 1
    lref, gx, gv, gu = ar.gradient(pos[:n], vel[:n], hyd[:n], ranges, dts, p)
    assert loss == pytest.approx(lref, rel=1e-11)
    gr = flat(gx, gv, gu)
    err = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    assert err <= 1e-8, err
    # each block (x, v, u) on its own scale
    for a, b in ((g[:3 * n], gr[:3 * n]), (g[3 * n:6 * n], gr[3 * n:6 * n]), (g[6 * n:], gr[6 * n:])):
        assert np.abs(a - b).max() <= 1e-7 * np.abs(b).max()


def test_checkpoint_interval_does_not_change_the_gradient(cs):
    """Recomputation from checkpoints replays the recorded run bit for bit."""
    n, K = 4096, 6
    grads = []
    for ck in (1, 4, 6):
        print(f"checkpoint interval {ck}", flush=True)
        e, _, _ = fp64_engine(cs, n, seed=9)
        loss = e.adjoint_forward(K, ck)
        grads.append((loss, flat(*e.adjoint_backward())))
    for loss, g in grads[1:]:
        assert loss == grads[0][0]
        assert np.array_equal(g, grads[0][1])


def test_adjoint_vs_central_differences_64_directions(cs):
    """20 steps with checkpoints every 5 (the adjoint workload's schedule) at
    N = 8192: the directional derivative g.d against the central difference
    (L(z + h d) - L(z - h d)) / 2h of the recorded map, h = 1e-4, along 64
    seeded unit directions in (x0, v0, u0)."""
    n, K, ck = 8192, 20, 5
    print(f"adjoint FD n={n} K={K} ck={ck}", flush=True)
    e, p, (pos, vel, hyd) = fp64_engine(cs, n, seed=5)
    loss = e.adjoint_forward(K, ck)
    g = flat(*e.adjoint_backward())
    z = np.concatenate([pos[:n, :3].ravel(), vel[:n, :3].ravel(), hyd[:n, 2]])
    assert e.adjoint_replay() == loss, "replaying the recorded map reproduces the run bit for bit"

    def L(zz):
        return e.adjoint_replay(zz[: 3 * n].reshape(n, 3), zz[3 * n: 6 * n].reshape(n, 3), zz[6 * n:])

    rs = np.random.default_rng(2024)
    h = 1e-4
    errs = []
    for k in range(64):
        d = rs.standard_normal(z.size)
        d /= np.linalg.norm(d)
        fd = (L(z + h * d) - L(z - h * d)) / (2 * h)
        errs.append(abs(g @ d - fd) / abs(fd))
        if k % 16 == 15:
            print(f"  {k + 1} directions, max relative error {max(errs):.2e}", flush=True)
    errs = np.array(errs)
    print(f"relative error: max {errs.max():.2e} median {np.median(errs):.2e}", flush=True)
    assert errs.max() <= 1e-4, errs.max()


def test_adjoint_turbulent_sphere(cs):
    """The scale workload's physics at n = 2048 (uniform sphere, k^-4 velocity
    field at rms Mach 2, gas + radiation EOS): converging flows keep the
    artificial viscosity and its Balsara switch active on many pairs.  Two
    steps, checkpoint every step: the reverse mode against the FP64 CPU
    reference gradient and against central differences along 16 seeded unit
    directions (h = 1e-4)."""
    n, K, ck = 2048, 2, 1
    print(f"adjoint turbulent sphere n={n} K={K}", flush=True)
    p = cs.preset_params(5, n)
    e = cs.Engine("fp64", p)
    pos, vel, hyd = cs.generate_ic(5, n, 3)
    e.load_ic(pos, vel, hyd)
    loss = e.adjoint_forward(K, ck)
    g = flat(*e.adjoint_backward())
    ranges, dts = e.adjoint_records()
    lref, gx, gv, gu = ar.gradient(pos[:n], vel[:n], hyd[:n], ranges, dts, p)
    assert loss == pytest.approx(lref, rel=1e-11)
    gr = flat(gx, gv, gu)
    err = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    assert err <= 1e-8, err
    z = np.concatenate([pos[:n, :3].ravel(), vel[:n, :3].ravel(), hyd[:n, 2]])

    def L(zz):
        return e.adjoint_replay(zz[: 3 * n].reshape(n, 3), zz[3 * n: 6 * n].reshape(n, 3), zz[6 * n:])

    rs = np.random.default_rng(55)
    h = 1e-4
    errs = []
    for _ in range(16):
        d = rs.standard_normal(z.size)
        d /= np.linalg.norm(d)
        fd = (L(z + h * d) - L(z - h * d)) / (2 * h)
        errs.append(abs(g @ d - fd) / abs(fd))
    print(f"relative error: max {max(errs):.2e} median {np.median(errs):.2e}", flush=True)
    assert max(errs) <= 1e-4, max(errs)


def test_torch_autograd_function(cs):
    n = 2048
    e, p, (pos, vel, hyd) = fp64_engine(cs, n, seed=6)
    x0 = torch.tensor(pos[:n, :3], requires_grad=True)
    v0 = torch.tensor(vel[:n, :3], requires_grad=True)
    u0 = torch.tensor(hyd[:n, 2], requires_grad=True)
    loss = SphRunLoss.apply(x0, v0, u0, e, 3, 2)
    (2.0 * loss).backward()
    e2, _, _ = fp64_engine(cs, n, seed=6)
    l2 = e2.adjoint_forward(3, 2)
    gx, gv, gu = e2.adjoint_backward()
    assert loss.item() == l2
    assert np.array_equal(x0.grad.numpy(), 2.0 * gx)
    assert np.array_equal(v0.grad.numpy(), 2.0 * gv)
    assert np.array_equal(u0.grad.numpy(), 2.0 * gu)


def test_adjoint_errors(cs):
    e = cs.make_engine("plain", 4, 4096, 1)
    with pytest.raises(RuntimeError, match="FP64"):
        e.adjoint_forward(2, 1)
    f, _, _ = fp64_engine(cs, 4096)
    with pytest.raises(RuntimeError, match="adjoint_forward first"):
        f.adjoint_backward()
    with pytest.raises(RuntimeError, match="ckpt_every"):
        f.adjoint_forward(2, 0)


def test_step_after_backward_continues_the_recorded_run(cs):
    """adjoint_backward puts back the state adjoint_forward left: a step()
    afterwards gives the same bits, time and step count as init() plus K + 1
    plain steps."""
    n, K = 2048, 4
    print(f"step after backward n={n} K={K}", flush=True)
    a, _, _ = fp64_engine(cs, n, seed=13)
    a.adjoint_forward(K, 2)
    a.adjoint_backward()
    a.step()
    b, _, _ = fp64_engine(cs, n, seed=13)
    b.init()
    b.step(K + 1)
    for name in ("pos", "vel", "hydro", "bdots", "grav"):
        assert np.array_equal(a.get(name)[:n], b.get(name)[:n]), name
    for key in ("dt", "global_time", "iteration"):
        assert a.scalar(key) == b.scalar(key), key


def test_implicit_range_term_matches_the_resolved_map(cs):
    """With adj_implicit_h = 1 the backward adds each evaluation's h-solve
    derivative (dr_i/dx = -(dF_i/dx) / (dF_i/dr_i) for the smoothed-count
    equation of dev_solve_range.cu:17-89), so the gradient is that of the map
    whose ranges are solved from the positions (adj_free_ranges = 1 replays
    it; timesteps recorded).  The solve is converged tightly (rel_err 1e-10) so
    that this map is smooth; checked against its central differences along 16
    seeded directions, and the frozen-range gradient is shown to miss it."""
    n, K = 2048, 4
    print(f"implicit range term n={n} K={K}", flush=True)
    p = cs.preset_params(4, n)
    p.rel_err = 1e-10
    p.max_iter = 60
    pos, vel, hyd = cs.generate_ic(4, n, 11)
    e = cs.Engine("fp64", p)
    e.load_ic(pos, vel, hyd)
    e.adjoint_forward(K, 2)
    g_frozen = flat(*e.adjoint_backward())
    e.set_scalar("adj_implicit_h", 1)
    g = flat(*e.adjoint_backward())
    z = np.concatenate([pos[:n, :3].ravel(), vel[:n, :3].ravel(), hyd[:n, 2]])
    e.set_scalar("adj_free_ranges", 1)

    def L(zz):
        return e.adjoint_replay(zz[: 3 * n].reshape(n, 3), zz[3 * n: 6 * n].reshape(n, 3), zz[6 * n:])

    rs = np.random.default_rng(21)
    errs, frozen = [], []
    for _ in range(16):
        d = rs.standard_normal(z.size)
        d /= np.linalg.norm(d)
        fd = (L(z + 1e-4 * d) - L(z - 1e-4 * d)) / 2e-4
        errs.append(abs(g @ d - fd) / abs(fd))
        frozen.append(abs(g_frozen @ d - fd) / abs(fd))
    print(f"  implicit: max {max(errs):.2e} median {np.median(errs):.2e}; frozen: median {np.median(frozen):.2e}",
          flush=True)
    assert max(errs) < 1e-5, errs
    assert np.median(frozen) > 10 * max(errs), "the range term is not resolved by the check"
