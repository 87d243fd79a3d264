"""The HIP adjoint against the FP64 host adjoint (which central finite
differences pin, tests/test_adjoint_cpu.py).  The GPU differentiates its
FP32 forward, the host its FP64 forward, so gradients agree to the
trajectories' FP32 error over the horizon."""
import numpy as np
import pytest

from demso_amd import api
from demso_amd import scene as sc
from test_adjoint_cpu import scene_two_materials

pytestmark = pytest.mark.gpu
STEPS = 40


def gpu_grads(ctx, s, steps, W, V, K):
    import torch

    sim = api.Simulation(ctx, s)
    adj = api.Adjoint(sim, checkpoint_every=K)
    adj.forward(steps)
    gp = torch.from_numpy(W.astype(np.float32)).cuda()
    gv = torch.from_numpy(V.astype(np.float32)).cuda()
    gp0, gv0, gpar = adj.backward(gp, gv, num_types=len(s.materials))
    out = gp0.cpu().numpy().astype(np.float64), gv0.cpu().numpy().astype(np.float64), gpar.cpu().numpy()
    final = sim.get_by_id()[0].cpu().numpy()
    adj.close()
    sim.close()
    return out, final


def host_grads(oracle, s, steps, W, V):
    a = oracle.AdjointOracle(s)
    p, _ = a.forward(steps)
    out = a.backward(W, V)
    a.close()
    return out, p


def normwise(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def weights():
    s = scene_two_materials()
    rng = np.random.default_rng(0)
    return s, rng.normal(size=(s.n, 2)), rng.normal(size=(s.n, 2)) * 1e-3


@pytest.mark.parametrize("steps,K,rtol,ntol", [(5, 2, 1e-4, 1e-5), (STEPS, 16, 1e-2, 2e-3)])
def test_matches_fp64_host_adjoint(gpu_ctx, oracle, weights, steps, K, rtol, ntol):
    """Over a short horizon the FP32 and FP64 trajectories coincide to ~1e-7,
    so the two adjoints must agree tightly; over 40 steps of a jostling pack
    the FP32 drift moves a few damping/contact switches and the parameter
    gradients agree to ~1e-2 (the trajectory error of the forward)."""
    s, W, V = weights
    (gp0, gv0, gpar), final = gpu_grads(gpu_ctx, s, steps, W, V, K=K)
    (hp0, hv0, hpar), hfinal = host_grads(oracle, s, steps, W, V)
    assert np.abs(final - hfinal).max() < 1e-6           # same forward, FP32 vs FP64
    assert normwise(gp0, hp0) < ntol
    assert normwise(gv0, hv0) < ntol
    nz = hpar != 0.0
    assert np.all(np.abs(gpar[~nz]) <= 1e-12)
    assert np.allclose(gpar[nz], hpar[nz], rtol=rtol, atol=0)
    print(f"steps={steps}: max param rel diff {np.max(np.abs(gpar[nz] / hpar[nz] - 1)):.2e}, "
          f"x0 normwise {normwise(gp0, hp0):.2e}")


def test_checkpoint_interval_is_transparent(gpu_ctx, weights):
    s, W, V = weights
    a, _ = gpu_grads(gpu_ctx, s, STEPS, W, V, K=7)
    b, _ = gpu_grads(gpu_ctx, s, STEPS, W, V, K=STEPS)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.allclose(a[2], b[2], rtol=1e-12, atol=0)


def test_second_backward_recomputes_last_segment(gpu_ctx, weights):
    """The forward pass records its last segment for the first backward pass;
    a second backward pass after the same forward recomputes that segment from
    its checkpoint (the slots were reused), with the same gradients."""
    import torch

    s, W, V = weights
    sim = api.Simulation(gpu_ctx, s)
    adj = api.Adjoint(sim, checkpoint_every=7)
    adj.forward(STEPS)
    gp = torch.from_numpy(W.astype(np.float32)).cuda()
    gv = torch.from_numpy(V.astype(np.float32)).cuda()
    first = [t.cpu().numpy() for t in adj.backward(gp, gv, num_types=len(s.materials))]
    second = [t.cpu().numpy() for t in adj.backward(gp, gv, num_types=len(s.materials))]
    adj.close()
    sim.close()
    assert np.array_equal(first[0], second[0]) and np.array_equal(first[1], second[1])
    # parameter gradients are summed with FP64 atomics, in no fixed order
    assert np.allclose(first[2], second[2], rtol=1e-12, atol=0)


def test_ensemble_adjoint(gpu_ctx, oracle):
    s = sc.ensemble(instances=3, counts=[300, 500, 200], seed=6)
    for k in range(3):
        a, b = s.segments[k], s.segments[k + 1]
        s.pos[a:b] *= np.float32(0.86)
    rng = np.random.default_rng(4)
    s.vel[:] = (rng.normal(size=s.vel.shape) * 0.1).astype(np.float32)
    W = rng.normal(size=(s.n, 2))
    V = np.zeros((s.n, 2))
    (gp0, _, gpar), _ = gpu_grads(gpu_ctx, s, 30, W, V, K=10)
    (hp0, _, hpar), _ = host_grads(oracle, s, 30, W, V)
    assert normwise(gp0, hp0) < 1e-3
    nz = hpar != 0
    assert np.allclose(gpar[nz], hpar[nz], rtol=1e-3)
    assert np.all(np.abs(gpar[~nz]) <= 1e-12)


def test_torch_autograd_function(gpu_ctx, weights):
    """torch.autograd through the libtorch Function equals the C ABI adjoint."""
    import torch

    from demso_amd import torch_ext

    s, W, _ = weights
    traj = torch_ext.Trajectory(gpu_ctx, s, steps=STEPS, checkpoint_every=16)
    m = s.materials
    params = torch.tensor([[mm.kn, mm.kt, mm.cn, mm.mu] for mm in m], dtype=torch.float64, requires_grad=True)
    pos0 = torch.from_numpy(s.pos.copy()).cuda().requires_grad_(True)
    out = traj(pos0, params)
    loss = (out.double() * torch.from_numpy(W).cuda()).sum()
    loss.backward()
    (gp0, _, gpar), final = gpu_grads(gpu_ctx, s, STEPS, W, np.zeros_like(W), K=16)
    assert np.array_equal(out.detach().cpu().numpy(), final)
    assert np.allclose(pos0.grad.cpu().numpy(), gp0, rtol=1e-6, atol=1e-9)
    assert np.allclose(params.grad.numpy(), gpar, rtol=1e-9, atol=0)
    traj.close()


def test_torch_function_rejects_bad_inputs_and_replaced_trajectories(gpu_ctx, weights):
    """The autograd Function refuses initial positions whose row count is not
    the simulation's before anything is launched, and refuses to
    differentiate a trajectory that a later call on the same Trajectory has
    replaced (its checkpoints are gone); the latest call still works."""
    import torch

    from demso_amd import torch_ext

    s, W, _ = weights
    traj = torch_ext.Trajectory(gpu_ctx, s, steps=8, checkpoint_every=4)
    params = torch.tensor([[mm.kn, mm.kt, mm.cn, mm.mu] for mm in s.materials], dtype=torch.float64,
                          requires_grad=True)
    with pytest.raises(RuntimeError, match="rows"):
        traj(torch.from_numpy(s.pos[:-1].copy()).cuda(), params)
    with pytest.raises(RuntimeError, match="GPU tensor"):
        traj(torch.from_numpy(s.pos.copy()), params)
    a = torch.from_numpy(s.pos.copy()).cuda().requires_grad_(True)
    b = torch.from_numpy(s.pos.copy()).cuda().requires_grad_(True)
    out_a = traj(a, params)
    out_b = traj(b, params)
    with pytest.raises(RuntimeError, match="replaced"):
        out_a.sum().backward()
    out_b.sum().backward()
    assert b.grad is not None and torch.isfinite(b.grad).all()
    traj.close()


def fd_subcase(gpu_ctx):
    """Config-3 finite-difference sub-case: 512 discs (32 x 16 lattice with
    config-0 distributions) settled for 2000 steps, then 200 steps with
    loss = mean squared final position error against a target run
    (kn = 2e5, mu = 0.35, e = 0.6)."""
    from demso_amd import calibrate

    s = calibrate.settle(gpu_ctx, sc.calibration_column(seed=3, cols=32, rows=16), 2000)
    tgt = calibrate.set_params(s, [2e5], [0.35], [0.6])
    return s, tgt


def test_config3_fd_subcase(gpu_ctx, oracle):
    """dL/d(kn, kt, cn, mu) from the GPU adjoint (FP32 forward, FP64
    accumulation) and from the FP64 host adjoint against central finite
    differences of the FP64 oracle's loss (relative step 1e-4; the loss is
    smooth over that step, as the frozen contact set and stick/slip states
    assume).  Checked on every parameter whose log-sensitivity |p dL/dp| is
    at least 1e-3 of the largest (friction hardly acts in a settled column:
    nearly every contact sticks)."""
    from demso_amd import calibrate

    s, tgt = fd_subcase(gpu_ctx)
    steps, rel = 200, 1e-4
    a = oracle.AdjointOracle(tgt)
    target, _ = a.forward(steps)
    a.close()

    def loss(scene):
        o = oracle.AdjointOracle(scene)
        p, _ = o.forward(steps)
        o.close()
        return float(((p - target) ** 2).sum()) / scene.n

    run = calibrate.Runner(gpu_ctx, s, steps, checkpoint_every=50)
    res = run.run(s, target)
    run.close()
    host = oracle.AdjointOracle(s)
    pf, _ = host.forward(steps)
    _, _, ghost = host.backward(2.0 * (pf - target) / s.n)
    host.close()
    assert res.loss == pytest.approx(loss(s), rel=1e-3)
    names = ("kn", "kt", "cn", "mu")
    base = [getattr(s.materials[0], k) for k in names]
    sens = [abs(ghost[0, i] * base[i]) for i in range(4)]
    checked = 0
    for idx, name in enumerate(names):
        vals = []
        for sign in (1, -1):
            t = s.copy()
            setattr(t.materials[0], name, base[idx] * (1 + sign * rel))
            vals.append(loss(t))
        hp = float(np.float32(base[idx] * (1 + rel))) - float(np.float32(base[idx]))
        hm = float(np.float32(base[idx])) - float(np.float32(base[idx] * (1 - rel)))
        fd = (vals[0] - vals[1]) / (hp + hm)
        gpu, cpu = res.grad_params[0, idx], ghost[0, idx]
        print(f"{name}: gpu {gpu:.6e} host {cpu:.6e} fd {fd:.6e} "
              f"gpu/fd-1 {gpu / fd - 1 if fd else float('nan'):.2e} host/fd-1 {cpu / fd - 1 if fd else float('nan'):.2e}")
        if sens[idx] < 1e-3 * max(sens):
            continue
        assert cpu == pytest.approx(fd, rel=1e-3)
        assert gpu == pytest.approx(fd, rel=1e-3)
        checked += 1
    assert checked >= 2


def test_calibration_reduces_loss(gpu_ctx):
    from demso_amd import calibrate

    s = calibrate.settle(gpu_ctx, sc.calibration_column(seed=3, cols=24, rows=24), 2000)
    tgt = calibrate.set_params(s, [2e5], [0.35], [0.6])
    out = calibrate.calibrate(gpu_ctx, s, tgt, steps=150, iters=6, lr=0.1, checkpoint_every=50)
    losses = [h["loss"] for h in out["history"]]
    print("losses", losses)
    assert losses[-1] < losses[0]
