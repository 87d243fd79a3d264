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
