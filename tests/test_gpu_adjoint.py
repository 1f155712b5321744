"""GPU adjoint of the implicit step (fem_adjoint_step) against the oracle's reference gradient and
against central differences of the GPU rollout; the torch.autograd.Function harness."""
import sys

import numpy as np
import pytest

import oracle
from opensurgsim_amd import femgpu
from opensurgsim_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def problem(kind, n, steps, seed=3, dt=2e-3):
    """Block with heterogeneous E, top load, observed surface nodes and synthetic observations."""
    x0, tets, _ = mg.block_tet_mesh(*n)
    rng = np.random.default_rng(seed)
    ne = tets.shape[0]
    E = 3e4 * np.exp(0.3 * rng.normal(size=ne))
    fext = np.zeros(x0.size)
    fext[3 * mg.face_nodes(*n, 1, 1) + 1] = -0.05
    obs = mg.observation_nodes(*n, 12, seed)
    obs_dofs = (3 * obs[:, None] + np.arange(3)).reshape(-1)
    u_obs = 2e-4 * rng.normal(size=obs_dofs.size)
    return dict(kind=kind, n=n, x0=x0, tets=tets, E=E, fext=fext, obs_dofs=obs_dofs, u_obs=u_obs, dt=dt,
                steps=steps, bc=mg.face_nodes(*n, 1, 0))


def make_gpu(p, tol=1e-13):
    g = femgpu.FemGpu(0)
    g.upload_mesh(p["kind"], p["x0"], p["tets"], p["E"], 0.45, 1000.0)
    g.add_bc_nodes(p["bc"])
    g.set_options(dt=p["dt"], cg_tolerance=tol)
    g.set_external_force(p["fext"])
    return g


def oracle_gradient(p):
    s = oracle.Sim(p["kind"], p["x0"], p["tets"], p["E"], 0.45, 1000.0)
    for nd in p["bc"]:
        s.add_bc_node(nd)
    s.set_solver(tol=1e-15)
    s.set_external_force(p["fext"])
    xs, vs = [], []
    x, v = s.state()
    xs.append(x)
    vs.append(v)
    for _ in range(p["steps"]):
        s.step(p["dt"], 1)
        x, v = s.state()
        xs.append(x)
        vs.append(v)
    lx = np.zeros(x.size)
    lx[p["obs_dofs"]] = 2 * ((xs[-1] - p["x0"].reshape(-1))[p["obs_dofs"]] - p["u_obs"])
    lv = np.zeros(x.size)
    gE = np.zeros(p["E"].size)
    for k in range(p["steps"] - 1, -1, -1):
        lx, lv = s.adjoint_step(p["dt"], xs[k], vs[k], vs[k + 1], lx, lv, gE)
    return gE, lx, lv, xs[-1]


def gpu_gradient(p, k=10, tol=1e-13):
    import torch
    from opensurgsim_amd import adjoint
    g = make_gpu(p, tol)
    r = adjoint.Rollout(g, p["x0"], np.zeros(p["x0"].size), p["steps"], checkpoint_every=k)
    x, v = r.forward(p["E"])
    xr = torch.as_tensor(p["x0"].reshape(-1), device="cuda")
    od = torch.as_tensor(p["obs_dofs"], device="cuda")
    lx = torch.zeros_like(x)
    lx[od] = 2 * ((x - xr)[od] - torch.as_tensor(p["u_obs"], device="cuda"))
    gE, gx0, gv0 = r.backward(lx, None)
    out = (gE.cpu().numpy(), gx0.cpu().numpy(), gv0.cpu().numpy(), x.cpu().numpy())
    g.close()
    return out


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("k", [2, None])
@pytest.mark.parametrize("kind", [femgpu.TET_COROTATIONAL, femgpu.TET_LINEAR])
def test_adjoint_matches_oracle(kind, k):
    """Checkpoints every 2 steps, and the whole history kept (k None)."""
    p = problem(kind, (2, 3, 2), 5)
    say("adjoint vs oracle, kind", kind, "checkpoints", k)
    gE_o, lx_o, lv_o, x_o = oracle_gradient(p)
    gE, lx, lv, x = gpu_gradient(p, k=k)
    e = dict(x=rel(x, x_o), gE=rel(gE, gE_o), lx0=rel(lx, lx_o), lv0=rel(lv, lv_o))
    say("  rel err", e)
    assert e["x"] < 1e-10
    assert e["gE"] < 1e-6 and e["lx0"] < 1e-6 and e["lv0"] < 1e-6


def test_adjoint_vs_central_differences():
    import torch
    from opensurgsim_amd import adjoint
    p = problem(femgpu.TET_COROTATIONAL, (4, 4, 4), 12, seed=9)
    say("adjoint vs central differences, 4x4x4, 12 steps")
    gE, _, _, _ = gpu_gradient(p, k=5)
    g = make_gpu(p)
    r = adjoint.Rollout(g, p["x0"], np.zeros(p["x0"].size), p["steps"], checkpoint_every=100)
    xr = p["x0"].reshape(-1)

    def loss(E):
        x, _ = r.forward(E)
        u = x.cpu().numpy() - xr
        return float(np.sum((u[p["obs_dofs"]] - p["u_obs"]) ** 2))

    rng = np.random.default_rng(1)
    worst = 0.0
    for t in range(8):
        d = rng.normal(size=p["E"].size)
        d /= np.linalg.norm(d)
        h = 1e-3 * p["E"].mean()
        fd = (loss(p["E"] + h * d) - loss(p["E"] - h * d)) / (2 * h)
        ad = float(gE @ d)
        err = abs(ad - fd) / abs(fd)
        worst = max(worst, err)
        say("  dir %d: adjoint %.10e fd %.10e rel %.2e" % (t, ad, fd, err))
    g.close()
    assert worst < 1e-5


def test_checkpoint_interval_does_not_change_gradient():
    p = problem(femgpu.TET_COROTATIONAL, (3, 3, 3), 9)
    say("checkpoint intervals 1, 4, 20 and the whole history")
    ref = gpu_gradient(p, k=1)[0]
    for k in (4, 20, None):
        e = rel(gpu_gradient(p, k=k)[0], ref)
        say("  k", k, "rel", e)
        assert e < 1e-12


def test_autograd_function_and_identification():
    import torch
    from opensurgsim_amd import adjoint
    p = problem(femgpu.TET_COROTATIONAL, (3, 3, 3), 8)
    say("torch.autograd.Function and a short identification run")
    g = make_gpu(p, tol=1e-12)
    r = adjoint.Rollout(g, p["x0"], np.zeros(p["x0"].size), p["steps"], checkpoint_every=3)
    E = torch.tensor(p["E"], dtype=torch.float64, device="cuda", requires_grad=True)
    x, _ = adjoint.ImplicitEulerRollout.apply(E, r.x_init, r.v_init, r)
    xr = torch.as_tensor(p["x0"].reshape(-1), device="cuda")
    od = torch.as_tensor(p["obs_dofs"], device="cuda")
    L = adjoint.surface_loss(x, xr, od, torch.as_tensor(p["u_obs"], device="cuda"))
    L.backward()
    gE_ref = gpu_gradient(p, k=3, tol=1e-12)[0]
    assert rel(E.grad.cpu().numpy(), gE_ref) < 1e-12
    # identification: observations from the true moduli, start from a uniform guess
    with torch.no_grad():
        xt, _ = r.forward(p["E"])
    u_true = (xt - xr)[od]
    Ehat, losses = adjoint.recover_young_modulus(r, xr, od, u_true, np.full(p["E"].size, 3e4), iterations=6)
    say("  losses", losses)
    assert min(losses) < 0.5 * losses[0]
    g.close()
