"""The direct solver (FEM_LINEAR_SOLVER_LU: banded LU of the assembled operator) and the dense compliance matrix
against the FP64 oracle, whose banded LU and dense inverse are pinned in tests/test_oracle.py
(LinearSparseSolveAndInverseLU, Math/LinearSparseSolveAndInverse.cpp:31-97; OdeSolver::
computeComplianceMatrixFromSystemMatrix, Math/OdeSolver.cpp:69-76).

Both sides factor without pivoting, so the trajectories agree to rounding (bound 1e-11 relative); the GPU's
LU works on the operator CG sees (the lower-triangle view, equal to A for linear elements), so corotational
runs are compared with the oracle's CG on that same view at 1e-14.
"""
import sys

import numpy as np
import pytest

import oracle
from opensurgsim_amd import femgpu
from opensurgsim_amd import meshgen as mg

from helpers import make_gpu, make_oracle, rel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_linear_scheme_lu_vs_oracle_lu():
    """The reference's CPU-runnable case: linear tets on 10^3 cubes, OdeSolverLinearEulerImplicit with LU, 100
    steps: the operator is factored once, its dense compliance formed on the first step and kept (127.6 MB), and
    the later steps apply it; the trajectory matches the oracle's LU solves, and a second run matches the first
    bit for bit."""
    cfg = mg.config_linear_cpu()
    say("C1 linear scheme with LU, %d DOF" % cfg.x0.size)
    o = make_oracle(cfg, scheme=oracle.LINEAR_EULER_IMPLICIT, linsolver=oracle.LU)
    o.step(cfg.dt, cfg.steps)
    xr, vr = o.state()
    runs = []
    for _ in range(2):
        g = make_gpu(cfg, scheme=femgpu.LINEAR_EULER_IMPLICIT, linear_solver=femgpu.LINEAR_SOLVER_LU)
        assert g.step(cfg.steps) == femgpu.FEM_OK
        runs.append(g.state())
        st = g.stats()
        g.close()
    x, v = runs[0]
    x0 = cfg.x0.reshape(-1)
    e = dict(u=rel(x - x0, xr - x0), v=rel(v, vr))
    say("  rel err", e, "stats", st)
    assert e["u"] < 1e-11 and e["v"] < 1e-11
    assert st["last_cg_iterations"] == 0 and st["converged"] == 1
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])


@pytest.mark.parametrize("precision", [femgpu.FP64, femgpu.FP32_REFINED])
def test_implicit_corotational_lu_vs_oracle(precision):
    """Implicit Euler refactors every step; the precision option does not change the direct solve."""
    cfg = mg.config_corotational((7, 6, 8), seed=3)
    steps = 8
    say("corotational implicit Euler with LU, precision", precision)
    o = make_oracle(cfg, tol=1e-14)
    g = make_gpu(cfg, precision=precision, linear_solver=femgpu.LINEAR_SOLVER_LU)
    o.step(cfg.dt, steps)
    assert g.step(steps) == femgpu.FEM_OK
    xr, vr = o.state()
    x, v = g.state()
    x0 = cfg.x0.reshape(-1)
    e = dict(u=rel(x - x0, xr - x0), v=rel(v, vr))
    say("  rel err", e)
    assert e["u"] < 1e-11 and e["v"] < 1e-11


def test_compliance_through_lu_and_dense_compliance():
    cfg = mg.config_linear_cpu((5, 4, 6))
    o = make_oracle(cfg, scheme=oracle.LINEAR_EULER_IMPLICIT, linsolver=oracle.LU)
    o.set_compliance_at_first_linear_step(True)
    g = make_gpu(cfg, scheme=femgpu.LINEAR_EULER_IMPLICIT, linear_solver=femgpu.LINEAR_SOLVER_LU)
    o.step(cfg.dt, 3)
    assert g.step(3) == femgpu.FEM_OK
    Cr = o.compliance_matrix()
    C = g.compliance_matrix()
    scale = np.abs(Cr).max()
    say("dense compliance %d x %d: max abs err / max %.2e" % (C.shape[0], C.shape[1], np.abs(C - Cr).max() / scale))
    assert np.abs(C - Cr).max() <= 1e-12 * scale
    bc = (3 * np.asarray(cfg.bc_nodes)[:, None] + np.arange(3)).ravel()
    assert np.all(C[bc, :] == 0.0) and np.all(C[:, bc] == 0.0)
    # applyCompliance through the LU equals the dense compliance applied to the columns
    B = np.random.default_rng(5).normal(size=(cfg.x0.size, 11))
    X = g.apply_compliance(B)
    Xr = Cr @ B
    assert np.linalg.norm(X - Xr) <= 1e-12 * np.linalg.norm(Xr)
    # the device form writes the same matrix
    import torch
    Cd = torch.zeros((cfg.x0.size, cfg.x0.size), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    g.compliance_matrix_device(Cd.data_ptr())
    g.synchronize()
    assert np.array_equal(Cd.cpu().numpy().T, C)  # torch is row-major: its transpose is the column-major matrix
    # the allocation-free entry point in LU mode solves through the same factors
    bd = torch.from_numpy(np.ascontiguousarray(B.T)).cuda()
    x1 = torch.zeros_like(bd)
    x2 = torch.zeros_like(bd)
    torch.cuda.synchronize()
    g.apply_compliance_device(bd.data_ptr(), B.shape[1], x1.data_ptr())
    g.apply_compliance_device_noalloc(bd.data_ptr(), B.shape[1], x2.data_ptr())
    g.synchronize()
    assert torch.equal(x1, x2)
    g.close()


@pytest.mark.parametrize("n", [(1, 1, 1), (2, 1, 1), (3, 2, 2)])
def test_lu_small_meshes_vs_oracle(n):
    """Meshes below one 64-row block of the solves and one or a few 32-column panels of the factorization
    (24 to 108 DOF): the linear scheme's trajectory and its kept dense compliance against the oracle's LU."""
    cfg = mg.config_linear_cpu(n)
    say("small linear mesh with LU, %d DOF" % cfg.x0.size)
    o = make_oracle(cfg, scheme=oracle.LINEAR_EULER_IMPLICIT, linsolver=oracle.LU)
    o.set_compliance_at_first_linear_step(True)
    g = make_gpu(cfg, scheme=femgpu.LINEAR_EULER_IMPLICIT, linear_solver=femgpu.LINEAR_SOLVER_LU)
    o.step(cfg.dt, 5)
    assert g.step(5) == femgpu.FEM_OK
    xr, vr = o.state()
    x, v = g.state()
    x0 = cfg.x0.reshape(-1)
    e = dict(u=rel(x - x0, xr - x0), v=rel(v, vr))
    Cr = o.compliance_matrix()
    C = g.compliance_matrix()
    ec = np.abs(C - Cr).max() / np.abs(Cr).max()
    say("  rel err", e, "compliance %.2e" % ec)
    assert e["u"] < 1e-11 and e["v"] < 1e-11 and ec <= 1e-12
    g.close()


def test_lu_wide_band_vs_cg():
    """A 16^3 linear block (14,739 DOF, bandwidth 923: 15 update tiles per panel, 231 solve blocks), implicit Euler
    refactoring every step; its dense compliance (1.7 GB) is over the kept limit, so every step also solves through
    the band factors. The direct solve against FP64 CG at 1e-14 (the CG is pinned against the oracle elsewhere)."""
    cfg = mg.config_linear_cpu((16, 16, 16))
    say("wide-band LU, %d DOF" % cfg.x0.size)
    g1 = make_gpu(cfg, linear_solver=femgpu.LINEAR_SOLVER_LU)
    g2 = make_gpu(cfg, tol=1e-14)
    assert g1.step(3) == femgpu.FEM_OK and g2.step(3) == femgpu.FEM_OK
    x1, v1 = g1.state()
    x2, v2 = g2.state()
    x0 = cfg.x0.reshape(-1)
    e = dict(u=rel(x1 - x0, x2 - x0), v=rel(v1, v2))
    say("  rel err LU vs CG", e)
    assert e["u"] < 1e-10 and e["v"] < 1e-10
    g1.close()
    g2.close()


def test_dense_compliance_of_the_linear_case_with_cg_steps():
    """The dense compliance of the 10^3 linear case (3,993 DOF, 127.6 MB), which the reference's linear scheme
    forms on its first step, after CG steps: the LU is factored on demand."""
    cfg = mg.config_linear_cpu()
    o = make_oracle(cfg, scheme=oracle.LINEAR_EULER_IMPLICIT, linsolver=oracle.LU)
    o.set_compliance_at_first_linear_step(True)
    g = make_gpu(cfg, scheme=femgpu.LINEAR_EULER_IMPLICIT, tol=1e-13)
    o.step(cfg.dt, 1)
    assert g.step(1) == femgpu.FEM_OK
    Cr = o.compliance_matrix()
    C = g.compliance_matrix()
    err = np.abs(C - Cr).max() / np.abs(Cr).max()
    say("C1 dense compliance: max abs err / max %.2e" % err)
    assert err <= 1e-12
    g.close()


def test_direct_solver_rejections():
    hexcfg = mg.config_hex((3, 2, 2))
    g = femgpu.FemGpu(0)
    g.upload_mesh(femgpu.HEX, hexcfg.x0, hexcfg.conn, hexcfg.E, hexcfg.nu, hexcfg.rho)
    with pytest.raises(femgpu.FemError) as e:
        g.set_options(dt=1e-3, linear_solver=femgpu.LINEAR_SOLVER_LU)
    assert e.value.code == femgpu.FEM_ERR_INVALID
    g.close()
    cfgs = mg.config_batched(2, seed=3, lo=2, hi=3)
    b = mg.concat_instances(cfgs)
    g = femgpu.FemGpu(0)
    g.upload_mesh(b.kind, b.x0, b.conn, b.E, b.nu, b.rho)
    g.set_options(dt=1e-3, linear_solver=femgpu.LINEAR_SOLVER_LU)
    with pytest.raises(femgpu.FemError) as e:
        g.set_instances(b.node_offsets, b.dt)
    assert e.value.code == femgpu.FEM_ERR_INVALID
    g.close()
