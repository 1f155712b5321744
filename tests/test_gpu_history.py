"""Multi-step calls against step-at-a-time calls: fem_step(n) against n calls of fem_step(1) and
fem_step_record's state history against the states read after each single step (the same kernels in the same
order: bitwise equal, same iteration counts); fem_adjoint_steps against sequential fem_adjoint_step calls. Inside
one fem_adjoint_steps call every transpose solve after the first starts from the previous step's solution, so
its adjoints agree with the cold-started single steps to the CG tolerance (1e-12 relative residual here; the
bound is 1e-8 relative), and two calls with the same arguments agree bitwise.
"""
import sys

import numpy as np
import pytest

from opensurgsim_amd import femgpu
from opensurgsim_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(cfg, tol=1e-12):
    g = femgpu.FemGpu(0)
    g.upload_mesh(cfg.kind, cfg.x0, cfg.conn, cfg.E, cfg.nu, cfg.rho)
    g.add_bc_nodes(cfg.bc_nodes)
    g.set_options(dt=cfg.dt, cg_tolerance=tol)
    g.set_state(v=cfg.v0)
    return g


@pytest.mark.parametrize("kind", ["corotational", "linear"])
def test_multi_step_call_matches_single_steps(kind):
    if kind == "corotational":
        cfg = mg.config_corotational((6, 5, 7), seed=4)
    else:
        cfg = mg.config_linear_cpu((6, 5, 7))
        cfg.v0 = mg.config_corotational((6, 5, 7), seed=4).v0
    say("multi-step call against single steps,", kind)
    a, b = make(cfg), make(cfg)
    for call in range(2):
        for _ in range(6):
            assert a.step(1) == femgpu.FEM_OK
        assert b.step(6) == femgpu.FEM_OK
        xa, va = a.state()
        xb, vb = b.state()
        sa, sb = a.stats(), b.stats()
        say("  call", call, "iterations", sa["last_cg_iterations"], sb["last_cg_iterations"], "total",
            sa["total_cg_iterations"], sb["total_cg_iterations"])
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
        for key in ("steps", "last_cg_iterations", "total_cg_iterations", "converged", "last_cg_error"):
            assert sa[key] == sb[key], key
    a.close()
    b.close()


def test_step_record_history_and_adjoint_steps():
    import torch
    cfg = mg.config_adjoint((4, 4, 4))
    n = cfg.x0.size
    steps = 6
    say("step_record and adjoint_steps, %d DOF, %d steps" % (n, steps))

    def ctx():
        g = femgpu.FemGpu(0)
        g.upload_mesh(cfg.kind, cfg.x0, cfg.conn, cfg.E, cfg.nu, cfg.rho)
        g.add_bc_nodes(cfg.bc_nodes)
        g.set_options(dt=cfg.dt, cg_tolerance=1e-12)
        g.set_external_force(cfg.fext)
        g.set_stream(torch.cuda.current_stream().cuda_stream)
        return g

    a, b = ctx(), ctx()
    # history by single steps on a, recorded in one call on b
    hist_a = [a.state()]
    for _ in range(steps):
        assert a.step(1) == femgpu.FEM_OK
        hist_a.append(a.state())
    xh = torch.empty((steps + 1, n), dtype=torch.float64, device="cuda")
    vh = torch.empty_like(xh)
    assert b.step_record(steps, xh.data_ptr(), vh.data_ptr()) == femgpu.FEM_OK
    b.synchronize()
    for i in range(steps + 1):
        assert np.array_equal(xh[i].cpu().numpy(), hist_a[i][0]), i
        assert np.array_equal(vh[i].cpu().numpy(), hist_a[i][1]), i
    # adjoints: sequential fem_adjoint_step on a, fem_adjoint_steps on b, twice
    rng = np.random.default_rng(9)
    lx0 = torch.as_tensor(rng.normal(size=n), device="cuda")
    lv0 = torch.as_tensor(rng.normal(size=n), device="cuda")

    def rel(u, w):
        return float(torch.linalg.norm(u - w) / torch.linalg.norm(w))

    first = None
    for call in range(2):
        lxa, lva = lx0.clone(), lv0.clone()
        lxb, lvb = lx0.clone(), lv0.clone()
        ga = torch.zeros(cfg.E.size, dtype=torch.float64, device="cuda")
        gb = torch.zeros_like(ga)
        for j in range(steps - 1, -1, -1):
            a.adjoint_step(xh[j].data_ptr(), vh[j].data_ptr(), vh[j + 1].data_ptr(), lxa.data_ptr(), lva.data_ptr(),
                           lxa.data_ptr(), lva.data_ptr(), ga.data_ptr())
        b.adjoint_steps(steps, xh.data_ptr(), vh.data_ptr(), lxb.data_ptr(), lvb.data_ptr(), gb.data_ptr())
        a.synchronize()
        b.synchronize()
        e = dict(lx=rel(lxb, lxa), lv=rel(lvb, lva), gE=rel(gb, ga))
        say("  call", call, "|dL/dE|", float(torch.linalg.norm(ga)), "rel err", e)
        assert float(torch.linalg.norm(ga)) > 0
        assert max(e.values()) < 1e-8
        if first is None:
            first = (lxb.clone(), lvb.clone(), gb.clone())
        else:
            assert torch.equal(lxb, first[0]) and torch.equal(lvb, first[1]) and torch.equal(gb, first[2])
    a.close()
    b.close()
