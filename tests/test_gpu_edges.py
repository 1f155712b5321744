"""GPU parity on the corotational element's degenerate branches and on the Newton-Raphson loop, against the
FP64 oracle through the C ABI.

- Inverted (det F < 0) and nearly flat (0 < det F <= 1e-6 |F|^3) elements take the polar decomposition's
  eigen-decomposition path on the GPU; Eigen's computeRotationScaling flips the largest singular value of an
  inverted F (Physics/Fem3DElementCorotationalTetrahedron.cpp:146-154). With distinct singular values R is
  unique, so R, K and f are compared with oracle.coro_element directly.
- A singular G = (tr S I - S) R^T (:205-216: K = R K0 R^T, no correction terms) needs two principal stretches
  that sum to zero, i.e. tied singular values, where R itself is not unique (any rotation in the tied plane is
  a valid polar factor, and Eigen's choice depends on its SVD's internals). Those cases check the GPU's R as a
  polar factor (proper, R^T F symmetric, the stretches with the largest flipped) and everything downstream of
  R against the oracle given that same R (oracle.coro_element_given_rotation), plus the step's g_singular count.
- The implicit step and its adjoint through a state with inverted elements (the adjoint's dual-number
  rotation tangent on the eigen path) against the oracle's step and its reference gradient.
- Newton-Raphson with several iterations (Math/OdeSolverEulerImplicit.cpp:84-115), the oracle's loop pinned
  on the CPU by the reference's OdeComplexNonLinear test (tests/test_oracle.py).
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


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def affine_state(cfg, F, rng, noise=0.0):
    """Every node mapped by x = F x0 + t (+ noise relative to the element size): every element gets F."""
    x0 = cfg.x0
    c = x0.mean(axis=0)
    x = (x0 - c) @ F.T + c + np.array([0.003, -0.001, 0.002])
    if noise:
        x = x + noise * cfg.h * rng.normal(size=x.shape)
    return x.reshape(-1)


def element_blocks(cfg, x, e):
    ids = cfg.conn[e]
    return cfg.x0[ids].reshape(-1), x.reshape(-1, 3)[ids].reshape(-1)


@pytest.mark.parametrize("case", ["inverted", "inverted-noisy", "flat"])
def test_degenerate_elements_vs_oracle(case):
    rng = np.random.default_rng({"inverted": 1, "inverted-noisy": 2, "flat": 3}[case])
    cfg = mg.config_corotational((3, 3, 3))
    Q1, Q2 = random_rotation(rng), random_rotation(rng)
    if case == "flat":
        F = Q1 @ np.diag([1.2, 0.9, 1e-7]) @ Q2.T      # det F = 1.1e-7 > 0 but <= 1e-6 |F|^3
    else:
        F = Q1 @ np.diag([-1.3, 0.8, 0.6]) @ Q2.T      # det F < 0, singular values 1.3 > 0.8 > 0.6
    x = affine_state(cfg, F, rng, noise=1e-3 if case == "inverted-noisy" else 0.0)
    g = femgpu.FemGpu(0)
    g.upload_mesh(femgpu.TET_COROTATIONAL, cfg.x0, cfg.conn, cfg.E, cfg.nu, cfg.rho)
    R, K, f = g.evaluate_elements(x)
    worst_k = worst_r = worst_f = 0.0
    for e in range(cfg.conn.shape[0]):
        xe0, xe = element_blocks(cfg, x, e)
        Ro, Mo, Ko, fo, flags = oracle.coro_element(xe0, xe, cfg.E[e], cfg.nu[e], cfg.rho[e])
        assert flags == 0
        assert abs(np.linalg.det(Ro) - 1.0) < 1e-12
        worst_r = max(worst_r, np.abs(R[e] - Ro).max())
        worst_k = max(worst_k, np.abs(K[e] - Ko).max() / np.abs(Ko).max())
        worst_f = max(worst_f, np.abs(f[e] - fo).max() / (np.abs(Ko).max() * np.abs(xe).max()))
    say(case, "worst R abs err %.2e, K rel err %.2e, f err %.2e" % (worst_r, worst_k, worst_f))
    assert worst_r < 1e-12
    assert worst_k < 1e-11
    assert worst_f < 1e-11
    g.close()


@pytest.mark.parametrize("delta", [0.0, 1e-13])
def test_singular_g_vs_oracle_given_gpu_rotation(delta):
    rng = np.random.default_rng(11)
    cfg = mg.config_corotational((3, 3, 3))
    Q1, Q2 = random_rotation(rng), random_rotation(rng)
    F = Q1 @ np.diag([-1.0, 1.0 - delta, 0.5]) @ Q2.T  # flipped largest + middle stretch sum to -delta
    x = affine_state(cfg, F, rng)
    g = femgpu.FemGpu(0)
    g.upload_mesh(femgpu.TET_COROTATIONAL, cfg.x0, cfg.conn, cfg.E, cfg.nu, cfg.rho)
    R, K, f = g.evaluate_elements(x)
    stretches = np.sort([-1.0, 1.0 - delta, 0.5])
    worst_k = worst_f = 0.0
    for e in range(cfg.conn.shape[0]):
        Re = R[e]
        # the GPU's R is a polar factor of F with Eigen's convention
        assert np.abs(Re.T @ Re - np.eye(3)).max() < 1e-13
        assert abs(np.linalg.det(Re) - 1.0) < 1e-13
        S = Re.T @ F
        assert np.abs(S - S.T).max() < 1e-12
        assert np.abs(np.sort(np.linalg.eigvalsh(0.5 * (S + S.T))) - stretches).max() < 1e-12
        xe0, xe = element_blocks(cfg, x, e)
        Mo, Ko, fo, flags = oracle.coro_element_given_rotation(xe0, xe, cfg.E[e], cfg.nu[e], cfg.rho[e], Re)
        assert flags == 2  # G singular: K = R K0 R^T without the correction terms
        worst_k = max(worst_k, np.abs(K[e] - Ko).max() / np.abs(Ko).max())
        worst_f = max(worst_f, np.abs(f[e] - fo).max() / (np.abs(Ko).max() * np.abs(xe).max()))
    say("singular G delta %g: K rel err %.2e, f err %.2e" % (delta, worst_k, worst_f))
    assert worst_k < 1e-12 and worst_f < 1e-12
    # the step's element kernel flags every element
    g.add_bc_nodes(cfg.bc_nodes)
    g.set_options(dt=cfg.dt, cg_tolerance=1e-10)
    g.set_state(x=x)
    assert g.step(1) == femgpu.FEM_OK
    st = g.stats()
    assert st["g_singular"] == cfg.conn.shape[0] and st["det_r_invalid"] == 0
    g.close()


def inverted_interior_state(cfg, rng):
    """A 3x3x3 block with one interior node pushed through the opposite face of its cube: the tets around it
    invert, the rest deform mildly."""
    x = cfg.x0.copy()
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    node = mg.node_id(1, 1, 1, ny, nz)
    x[node] += np.array([1.35, 1.2, 1.25]) * cfg.h
    x += 2e-4 * cfg.h * rng.normal(size=x.shape)
    return x.reshape(-1), node


def test_step_and_adjoint_through_inverted_elements():
    import torch
    rng = np.random.default_rng(4)
    cfg = mg.config_corotational((3, 3, 3), seed=4)
    cfg["E"] = np.full(cfg.conn.shape[0], 2e4)
    x, node = inverted_interior_state(cfg, rng)
    # the state really holds inverted elements
    dets = []
    for e in range(cfg.conn.shape[0]):
        p = x.reshape(-1, 3)[cfg.conn[e]]
        dets.append(np.linalg.det(np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]])) /
                    np.linalg.det(np.stack([cfg.x0[cfg.conn[e]][k] - cfg.x0[cfg.conn[e]][0] for k in (1, 2, 3)])))
    ninv = int(np.sum(np.array(dets) < 0))
    say("inverted elements:", ninv, "of", cfg.conn.shape[0])
    assert ninv >= 2
    v = 0.01 * rng.normal(size=x.size)
    for n in cfg.bc_nodes:
        v[3 * n:3 * n + 3] = 0.0
    dt = cfg.dt
    o = make_oracle(cfg, tol=1e-14)
    o.set_state(x=x, v=v)
    o.step(dt, 1)
    xo, vo = o.state()
    g = make_gpu(cfg, tol=1e-14)
    g.set_state(x=x, v=v)
    assert g.step(1) == femgpu.FEM_OK
    xg, vg = g.state()
    e = dict(x=rel(xg, xo), dv=rel(vg - v, vo - v))
    say("step through inverted elements:", e, "stats", g.stats())
    assert e["x"] < 1e-12 and e["dv"] < 1e-9
    # adjoint of that step for a loss on every free DOF
    lx = rng.normal(size=x.size)
    lv = rng.normal(size=x.size)
    gE_o = np.zeros(cfg.conn.shape[0])
    lxo, lvo = o.adjoint_step(dt, x, v, vo, lx, lv, gE_o)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    X, V, VN, LX, LV = dev(x), dev(v), dev(vg), dev(lx), dev(lv)
    LXO, LVO, GE = torch.zeros_like(X), torch.zeros_like(X), torch.zeros(cfg.conn.shape[0], dtype=torch.float64,
                                                                         device="cuda")
    torch.cuda.synchronize()
    g.adjoint_step(X.data_ptr(), V.data_ptr(), VN.data_ptr(), LX.data_ptr(), LV.data_ptr(), LXO.data_ptr(),
                   LVO.data_ptr(), GE.data_ptr())
    g.synchronize()
    ea = dict(lx=rel(LXO.cpu().numpy(), lxo), lv=rel(LVO.cpu().numpy(), lvo), gE=rel(GE.cpu().numpy(), gE_o))
    say("adjoint through inverted elements:", ea)
    # the oracle's reference gradient takes its element derivatives by Richardson-extrapolated central
    # differences (accurate to about 1e-9 here)
    assert ea["lx"] < 1e-6 and ea["lv"] < 1e-6 and ea["gE"] < 1e-6
    g.close()


@pytest.mark.parametrize("precision", [femgpu.FP64, femgpu.FP32_REFINED])
@pytest.mark.parametrize("max_iter,eps", [(3, 1e-5), (10, 1e-8)])
def test_newton_raphson_vs_oracle(precision, max_iter, eps):
    """A soft block with large top velocities, so that one Newton iteration is far from the implicit
    solution: the trajectory and every step's iteration count match the oracle's loop."""
    cfg = mg.config_corotational((6, 5, 7), seed=13)
    cfg["E"] = np.full(cfg.conn.shape[0], 3e3)
    cfg["v0"] = cfg.v0 * 40.0
    cfg["dt"] = 4e-3
    say("newton max_iter", max_iter, "eps", eps, "precision", precision)
    o = make_oracle(cfg, tol=1e-13)
    o.set_solver(tol=1e-13, newton_max_iter=max_iter, newton_eps=eps)
    g = make_gpu(cfg, precision=precision, tol=1e-13, newton_max_iterations=max_iter, newton_epsilon=eps)
    counts = []
    for k in range(6):
        o.step(cfg.dt, 1)
        assert g.step(1) == femgpu.FEM_OK
        counts.append((o.last_newton_iters(), g.stats()["last_newton_iterations"]))
    xr, vr = o.state()
    x, v = g.state()
    x0 = cfg.x0.reshape(-1)
    e = dict(u=rel(x - x0, xr - x0), v=rel(v, vr))
    say("  newton iterations (oracle, gpu):", counts, "rel err", e)
    assert all(a == b for a, b in counts)
    assert max(a for a, _ in counts) > 1
    bound = 1e-9 if precision == femgpu.FP64 else 1e-6
    assert e["u"] < bound and e["v"] < bound
    g.close()
