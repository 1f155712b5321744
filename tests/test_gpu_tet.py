"""GPU parity of the tetrahedron path against the FP64 oracle (through the C ABI)."""
import sys

import numpy as np
import pytest

import oracle
from opensurgsim_amd import femgpu
from opensurgsim_amd import meshgen as mg

from helpers import lower_view_csr, make_gpu, make_oracle, rel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


@pytest.mark.parametrize("n", [(10, 10, 10), (4, 7, 5)])
def test_pattern_bit_exact(n):
    say("pattern", n)
    cfg = mg.config_corotational(n)
    o = make_oracle(cfg)
    g = make_gpu(cfg)
    outer, inner = o.pattern()
    rowptr, colidx = g.pattern()
    assert rowptr.dtype == np.int32 and colidx.dtype == np.int32
    assert np.array_equal(rowptr, outer)
    assert np.array_equal(colidx, inner)


@pytest.mark.parametrize("kind", [oracle.TET_COROTATIONAL, oracle.TET_LINEAR])
def test_element_kernels_vs_oracle(kind):
    say("elements kind", kind)
    rng = np.random.default_rng(3)
    cfg = mg.config_corotational((3, 3, 3))
    g = femgpu.FemGpu(0)
    g.upload_mesh(kind, cfg.x0, cfg.conn, cfg.E, cfg.nu, cfg.rho)
    x0 = cfg.x0
    # a rigid rotation + translation + small random deformation
    Rg = random_rotation(rng)
    x = (x0 @ Rg.T + np.array([0.01, -0.02, 0.005]) + 3e-4 * rng.normal(size=x0.shape)).reshape(-1)
    R, K, f = g.evaluate_elements(x)
    worst = 0.0
    for e in range(cfg.conn.shape[0]):
        ids = cfg.conn[e]
        xe0 = x0[ids].reshape(-1)
        xe = x.reshape(-1, 3)[ids].reshape(-1)
        if kind == oracle.TET_COROTATIONAL:
            Ro, Mo, Ko, fo, flags = oracle.coro_element(xe0, xe, cfg.E[e], cfg.nu[e], cfg.rho[e])
            assert flags == 0
            assert np.abs(R[e] - Ro).max() < 1e-13
            # The oracle's force has the reference's cancellation error eps*|K||x|; compare on that scale.
            assert np.abs(f[e] - fo).max() <= 1e-10 * np.abs(Ko).max() * np.abs(xe).max() + 1e-12 * np.abs(fo).max()
        else:
            Ko, Mo, vol = oracle.tet_linear_km(xe0, cfg.E[e], cfg.nu[e], cfg.rho[e])
            fo = -Ko @ (xe - xe0)
            assert np.abs(f[e] - fo).max() <= 1e-11 * np.abs(fo).max() + 1e-14 * np.abs(Ko).max()
        err = np.abs(K[e] - Ko).max() / np.abs(Ko).max()
        worst = max(worst, err)
    say("worst K rel err", worst)
    assert worst < 1e-12


def test_element_kernels_large_stretch_vs_oracle():
    """Stretches far from a rotation (|F^T F - I| > 1/2): the polar decomposition takes the scaled Newton
    iteration instead of the Newton-Schulz one used near rotations."""
    rng = np.random.default_rng(5)
    cfg = mg.config_corotational((3, 3, 3))
    g = femgpu.FemGpu(0)
    g.upload_mesh(oracle.TET_COROTATIONAL, cfg.x0, cfg.conn, cfg.E, cfg.nu, cfg.rho)
    x0 = cfg.x0
    Rg = random_rotation(rng)
    stretch = np.diag([1.7, 0.6, 1.25])
    x = (x0 @ stretch.T @ Rg.T + 2e-4 * rng.normal(size=x0.shape)).reshape(-1)
    R, K, f = g.evaluate_elements(x)
    worst = 0.0
    for e in range(cfg.conn.shape[0]):
        ids = cfg.conn[e]
        xe0 = x0[ids].reshape(-1)
        xe = x.reshape(-1, 3)[ids].reshape(-1)
        Ro, Mo, Ko, fo, flags = oracle.coro_element(xe0, xe, cfg.E[e], cfg.nu[e], cfg.rho[e])
        assert flags == 0
        assert np.abs(R[e] - Ro).max() < 1e-13
        worst = max(worst, np.abs(K[e] - Ko).max() / np.abs(Ko).max())
    say("large stretch: worst K rel err", worst)
    assert worst < 1e-12


def test_system_matrix_and_rhs_one_step():
    say("system matrix 6x5x4")
    cfg = mg.config_corotational((6, 5, 4))
    o = make_oracle(cfg)
    g = make_gpu(cfg)
    o.step(cfg.dt, 1)
    g.step(1)
    outer, inner = o.pattern()
    _, _, Sref = lower_view_csr(outer, inner, o.matrix_values(2))
    S = g.system_matrix()
    scale = np.abs(Sref).max()
    assert np.abs(S - Sref).max() <= 1e-12 * scale
    b = g.rhs()
    bref = o.rhs()
    assert rel(b, bref) < 1e-12


def trajectory_errors(cfg, steps, precision=femgpu.FP64, scheme=femgpu.EULER_IMPLICIT,
                      oscheme=oracle.EULER_IMPLICIT, olin=oracle.CG, otol=None, gtol=None, **gopts):
    o = make_oracle(cfg, scheme=oscheme, linsolver=olin, tol=otol)
    g = make_gpu(cfg, precision=precision, scheme=scheme, tol=gtol, **gopts)
    o.step(cfg.dt, steps)
    rc = g.step(steps)
    assert rc == femgpu.FEM_OK
    xr, vr = o.state()
    x, v = g.state()
    x0 = cfg.x0.reshape(-1)
    return dict(x=rel(x, xr), u=rel(x - x0, xr - x0), v=rel(v, vr), stats=g.stats(), oracle_cg=o.total_cg_iters())


def test_trajectory_corotational_fp64():
    cfg = mg.config_corotational((8, 8, 8))
    e = trajectory_errors(cfg, 20)
    say("coro fp64 errors", e)
    assert e["x"] < 1e-9 and e["u"] < 1e-9 and e["v"] < 1e-9


def test_trajectory_linear_scheme_vs_lu():
    # config of the CPU-runnable case (10^3 linear tets, LinearEulerImplicit + LU in the reference);
    # the GPU solves the same constant operator with CG to a 1e-13 relative residual.
    cfg = mg.config_linear_cpu()
    e = trajectory_errors(cfg, 100, scheme=femgpu.LINEAR_EULER_IMPLICIT, oscheme=oracle.LINEAR_EULER_IMPLICIT,
                          olin=oracle.LU, gtol=1e-13)
    say("linear scheme errors", e)
    assert e["x"] < 1e-9 and e["u"] < 1e-9 and e["v"] < 1e-9


def test_trajectory_fp32_refined():
    cfg = mg.config_corotational((8, 8, 8))
    e = trajectory_errors(cfg, 20, precision=femgpu.FP32_REFINED)
    say("fp32 refined errors", e)
    assert e["u"] < 1e-6 and e["v"] < 1e-6


@pytest.mark.parametrize("precision", [femgpu.FP64, femgpu.FP32_REFINED])
def test_trajectory_medium_mesh(precision):
    """16^3 cubes (20,480 tets): every CG workgroup has rows, so cross-workgroup ordering is exercised."""
    cfg = mg.config_corotational((16, 16, 16))
    steps = 8
    say("medium mesh, precision", precision)
    o = make_oracle(cfg, tol=1e-10)
    g = make_gpu(cfg, precision=precision, tol=1e-10)
    o.step(cfg.dt, steps)
    assert g.step(steps) == femgpu.FEM_OK
    xr, vr = o.state()
    x, v = g.state()
    x0 = cfg.x0.reshape(-1)
    e = dict(u=rel(x - x0, xr - x0), v=rel(v, vr))
    say("  rel err", e)
    bound = 1e-9 if precision == femgpu.FP64 else 1e-6
    assert e["u"] < bound and e["v"] < bound


@pytest.mark.parametrize("precision", [femgpu.FP64, femgpu.FP32_REFINED])
def test_boundary_conditions_added_between_steps(precision):
    """Nodes and single DOFs fixed after some steps (OdeState::addBoundaryCondition at any time): the
    assembly's per-block boundary bits are rebuilt before the next step."""
    cfg = mg.config_corotational((6, 6, 6))
    o = make_oracle(cfg, tol=1e-10)
    g = make_gpu(cfg, precision=precision, tol=1e-10)
    o.step(cfg.dt, 4)
    assert g.step(4) == femgpu.FEM_OK
    x0 = cfg.x0.reshape(-1, 3)
    top = np.flatnonzero(np.isclose(x0[:, 2], x0[:, 2].max()))
    for n in top[::3]:
        o.add_bc_node(int(n))
        g.add_bc_node(int(n))
    for n in top[1::3]:
        o.add_bc_dof(3 * int(n) + 1)
        g.add_bc_dof(3 * int(n) + 1)
    say("added BCs on", len(top[::3]), "nodes and", len(top[1::3]), "single DOFs")
    o.step(cfg.dt, 6)
    assert g.step(6) == femgpu.FEM_OK
    xr, vr = o.state()
    x, v = g.state()
    xs = cfg.x0.reshape(-1)
    e = dict(u=rel(x - xs, xr - xs), v=rel(v, vr))
    say("  rel err", e)
    bound = 1e-9 if precision == femgpu.FP64 else 1e-6
    assert e["u"] < bound and e["v"] < bound


@pytest.mark.parametrize("kind", [femgpu.TET_LINEAR, femgpu.TET_COROTATIONAL])
def test_tet_repeating_a_node_rejected(kind):
    cfg = mg.config_corotational((2, 2, 2))
    conn = cfg.conn.copy()
    conn[3, 2] = conn[3, 0]
    g = femgpu.FemGpu(0)
    with pytest.raises(femgpu.FemError) as err:
        g.upload_mesh(kind, cfg.x0, conn, cfg.E, cfg.nu, cfg.rho)
    say("rejected:", err.value)
    assert err.value.code == femgpu.FEM_ERR_INVALID


@pytest.mark.parametrize("precision", [femgpu.FP64, femgpu.FP32_REFINED])
def test_node_in_no_element(precision):
    """A node that no element uses: no mass, no stiffness and no pattern entries, so its rows are empty (the
    Jacobi rule of Eigen's preconditioner takes 1 for the missing diagonal, the block preconditioner falls
    back to it). It stays where it is while the rest of the block follows the oracle."""
    cfg = mg.config_corotational((4, 3, 5), seed=21)
    n0 = cfg.x0.shape[0]
    cfg["x0"] = np.vstack([cfg.x0, [[0.5, 0.5, 0.5]]])
    cfg["v0"] = np.concatenate([cfg.v0, np.zeros(3)])
    if cfg.fext is not None:
        cfg["fext"] = np.concatenate([cfg.fext, np.zeros(3)])
    say("node in no element, precision", precision)
    e = trajectory_errors(cfg, 6, precision=precision)
    say("  errors", e)
    bound = 1e-9 if precision == femgpu.FP64 else 1e-6
    assert e["u"] < bound and e["v"] < bound
    g = make_gpu(cfg, precision=precision)
    assert g.step(3) == femgpu.FEM_OK
    x, v = g.state()
    assert np.array_equal(x[3 * n0:], cfg.x0[n0]) and not v[3 * n0:].any()
    assert g.stats()["active"] == 1
