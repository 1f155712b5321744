"""GPU parity, through the C ABI, of the integration schemes other than implicit Euler, Rayleigh damping,
mass lumping and localized external forces against the FP64 oracle (Math/OdeSolver*.cpp,
Physics/FemRepresentation.cpp:479-629, Physics/Fem3DRepresentation.cpp:133-198).

Both sides solve with Eigen-semantics Jacobi CG at a 1e-12 relative residual; the bounds are relative
errors of displacement and velocity after the steps. The oracle's schemes, damping, lumping and external
terms are pinned by tests/test_oracle_schemes.py (modal closed forms and the reference's
FemRepresentationTests checks).
"""
import sys

import numpy as np
import pytest

import oracle
from opensurgsim_amd import femgpu
from opensurgsim_amd import meshgen as mg

from helpers import make_gpu, make_oracle, rel

pytestmark = pytest.mark.gpu

AM, AK = 2.0, 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def soft_config(kind, n=(4, 4, 4), dt=2e-4):
    """Soft block (E 5 kPa) with the corotational config's tangential top velocities and a small dt, so the
    explicit schemes stay bounded over the steps."""
    cfg = mg.config_corotational(n, seed=31)
    cfg["kind"] = kind
    cfg["E"] = np.full(cfg.conn.shape[0], 5e3)
    cfg["dt"] = dt
    cfg["cg_tol"] = 1e-12
    return cfg


def run_pair(cfg, steps, scheme, precision=femgpu.FP64, am=0.0, ak=0.0, lumping=False, before_step=None):
    o = make_oracle(cfg, scheme=scheme, tol=1e-12)
    o.set_damping(am, ak, lumping)
    g = make_gpu(cfg, precision=precision, scheme=scheme, tol=1e-12, rayleigh_mass=am, rayleigh_stiffness=ak,
                 mass_lumping=int(lumping))
    for k in range(steps):
        if before_step:
            before_step(k, o, g)
        assert o.step(cfg.dt, 1) == 0
        assert g.step(1) == femgpu.FEM_OK
    xr, vr = o.state()
    x, v = g.state()
    x0 = cfg.x0.reshape(-1)
    return dict(u=rel(x - x0, xr - x0), v=rel(v, vr) if np.linalg.norm(vr) > 0 else np.linalg.norm(v))


NONLINEAR = [femgpu.STATIC, femgpu.EULER_EXPLICIT, femgpu.EULER_EXPLICIT_MODIFIED, femgpu.RUNGE_KUTTA_4]
LINEAR = [femgpu.LINEAR_STATIC, femgpu.LINEAR_EULER_EXPLICIT, femgpu.LINEAR_EULER_EXPLICIT_MODIFIED,
          femgpu.LINEAR_RUNGE_KUTTA_4]


@pytest.mark.parametrize("scheme", NONLINEAR)
@pytest.mark.parametrize("damped", [False, True])
def test_corotational_schemes_vs_oracle(scheme, damped):
    cfg = soft_config(oracle.TET_COROTATIONAL)
    steps = 2 if scheme == femgpu.STATIC else 10
    say("corotational scheme", scheme, "damped", damped)
    e = run_pair(cfg, steps, scheme, am=AM if damped else 0.0, ak=AK if damped else 0.0)
    say("  rel err", e)
    assert e["u"] < 1e-9 and e["v"] < 1e-9


@pytest.mark.parametrize("scheme", LINEAR)
def test_linear_variants_vs_oracle(scheme):
    cfg = soft_config(oracle.TET_LINEAR)
    steps = 3 if scheme == femgpu.LINEAR_STATIC else 10
    say("linear-tet scheme", scheme)
    e = run_pair(cfg, steps, scheme, am=AM, ak=AK)
    say("  rel err", e)
    assert e["u"] < 1e-9 and e["v"] < 1e-9


@pytest.mark.parametrize("precision", [femgpu.FP64, femgpu.FP32_REFINED])
@pytest.mark.parametrize("scheme", NONLINEAR + LINEAR)
def test_hex_schemes_vs_oracle(scheme, precision):
    """The static, explicit, modified explicit and Runge-Kutta 4 schemes and their linear variants on hexahedra
    (the element-by-element operator is K, or M / dt never lumped), damped, with the generator's loads."""
    cfg = mg.config_hex((4, 3, 4))
    cfg["E"] = np.full(cfg.conn.shape[0], 5e3)
    cfg["dt"] = 1e-4
    steps = 2 if scheme in (femgpu.STATIC, femgpu.LINEAR_STATIC) else 8
    say("hex scheme", scheme, "precision", precision)
    e = run_pair(cfg, steps, scheme, precision=precision, am=AM, ak=AK)
    say("  rel err", e)
    bound = 1e-9 if precision == femgpu.FP64 else 1e-8
    assert e["u"] < bound and e["v"] < bound


@pytest.mark.parametrize("scheme", [femgpu.EULER_EXPLICIT, femgpu.RUNGE_KUTTA_4])
def test_explicit_schemes_fp32_refined(scheme):
    cfg = soft_config(oracle.TET_COROTATIONAL)
    say("fp32 refined scheme", scheme)
    e = run_pair(cfg, 8, scheme, precision=femgpu.FP32_REFINED, am=AM, ak=AK)
    say("  rel err", e)
    assert e["u"] < 1e-8 and e["v"] < 1e-8


@pytest.mark.parametrize("scheme", [femgpu.EULER_EXPLICIT, femgpu.EULER_EXPLICIT_MODIFIED, femgpu.RUNGE_KUTTA_4,
                                    femgpu.LINEAR_EULER_EXPLICIT, femgpu.LINEAR_RUNGE_KUTTA_4, femgpu.STATIC])
def test_lumped_rayleigh_mass_other_schemes_vs_oracle(scheme):
    """Mass lumping with Rayleigh mass damping on the non-implicit schemes: the first F takes -alpha_M M v element
    by element, every later one -alpha_M diag(M) v with the consistent diagonal (computeF's global form,
    Physics/FemRepresentation.cpp:396-410, 583-587; pinned in tests/test_oracle_schemes.py)."""
    kind = oracle.TET_LINEAR if scheme in (femgpu.LINEAR_EULER_EXPLICIT, femgpu.LINEAR_RUNGE_KUTTA_4) else \
        oracle.TET_COROTATIONAL
    cfg = soft_config(kind)
    steps = 3 if scheme == femgpu.STATIC else 10
    say("lumped Rayleigh mass, scheme", scheme)
    e = run_pair(cfg, steps, scheme, am=AM, ak=AK, lumping=True)
    say("  rel err", e)
    assert e["u"] < 1e-9 and e["v"] < 1e-9


@pytest.mark.parametrize("scheme", [femgpu.EULER_EXPLICIT, femgpu.RUNGE_KUTTA_4, femgpu.LINEAR_EULER_EXPLICIT_MODIFIED])
def test_hex_lumped_rayleigh_mass_other_schemes_vs_oracle(scheme):
    """The same on hexahedra: diag(M) per DOF from the element quadrature, consistent or the row sums."""
    cfg = mg.config_hex((4, 3, 4))
    cfg["E"] = np.full(cfg.conn.shape[0], 5e3)
    cfg["dt"] = 1e-4
    say("hex lumped Rayleigh mass, scheme", scheme)
    e = run_pair(cfg, 8, scheme, am=AM, ak=AK, lumping=True)
    say("  rel err", e)
    assert e["u"] < 1e-9 and e["v"] < 1e-9


def test_lumped_rayleigh_mass_after_implicit_steps_vs_oracle():
    """Implicit steps with lumping, then explicit ones: the first explicit F takes the row sums of the implicit
    step's lumped M, the later ones the consistent diagonal of computeM. Rayleigh stiffness damping is left out:
    after an implicit step the reference damps explicit steps with that step's global K, stale for corotational
    elements, which the GPU rejects (checked at the end)."""
    cfg = soft_config(oracle.TET_COROTATIONAL)
    o = make_oracle(cfg, scheme=oracle.EULER_IMPLICIT, tol=1e-12)
    o.set_damping(AM, 0.0, True)
    g = make_gpu(cfg, scheme=femgpu.EULER_IMPLICIT, tol=1e-12, rayleigh_mass=AM, mass_lumping=1)
    for k in range(8):
        if k == 3:
            o.set_solver(scheme=oracle.EULER_EXPLICIT, tol=1e-12)
            g.set_options(dt=cfg.dt, cg_tolerance=1e-12, scheme=femgpu.EULER_EXPLICIT, gravity=cfg.gravity,
                          rayleigh_mass=AM, mass_lumping=1)
        assert o.step(cfg.dt, 1) == 0
        assert g.step(1) == femgpu.FEM_OK
    xr, vr = o.state()
    x, v = g.state()
    x0 = cfg.x0.reshape(-1)
    e = dict(u=rel(x - x0, xr - x0), v=rel(v, vr))
    say("implicit then explicit, lumped Rayleigh mass", e)
    assert e["u"] < 1e-9 and e["v"] < 1e-9
    g.set_options(dt=cfg.dt, cg_tolerance=1e-12, scheme=femgpu.EULER_EXPLICIT, gravity=cfg.gravity,
                  rayleigh_mass=AM, rayleigh_stiffness=AK, mass_lumping=1)
    with pytest.raises(femgpu.FemError) as err:
        g.step(1)
    assert err.value.code == femgpu.FEM_ERR_INVALID
    g.close()


@pytest.mark.parametrize("precision", [femgpu.FP64, femgpu.FP32_REFINED])
@pytest.mark.parametrize("lumping", [False, True])
def test_implicit_rayleigh_and_lumping_vs_oracle(precision, lumping):
    cfg = mg.config_corotational((6, 6, 6), seed=5)
    cfg["dt"] = 1e-3
    say("implicit Rayleigh, lumping", lumping, "precision", precision)
    e = run_pair(cfg, 10, femgpu.EULER_IMPLICIT, precision=precision, am=AM, ak=AK, lumping=lumping)
    say("  rel err", e)
    bound = 1e-9 if precision == femgpu.FP64 else 1e-6
    assert e["u"] < bound and e["v"] < bound


def test_linear_implicit_rayleigh_lumping_vs_oracle():
    cfg = soft_config(oracle.TET_LINEAR, dt=1e-3)
    e = run_pair(cfg, 10, femgpu.LINEAR_EULER_IMPLICIT, am=AM, ak=AK, lumping=True)
    say("linear implicit Rayleigh + lumping", e)
    assert e["u"] < 1e-9 and e["v"] < 1e-9


@pytest.mark.parametrize("precision", [femgpu.FP64, femgpu.FP32_REFINED])
@pytest.mark.parametrize("lumping", [False, True])
def test_hex_rayleigh_and_lumping_vs_oracle(precision, lumping):
    cfg = mg.config_hex((5, 3, 4))
    cfg["cg_tol"] = 1e-12 if precision == femgpu.FP64 else 1e-10
    say("hex Rayleigh, lumping", lumping, "precision", precision)
    o = make_oracle(cfg, tol=1e-12)
    o.set_damping(AM, AK, lumping)
    g = make_gpu(cfg, precision=precision, rayleigh_mass=AM, rayleigh_stiffness=AK, mass_lumping=int(lumping))
    o.step(cfg.dt, 6)
    assert g.step(6) == femgpu.FEM_OK
    xr, vr = o.state()
    x, v = g.state()
    x0 = cfg.x0.reshape(-1)
    e = dict(u=rel(x - x0, xr - x0), v=rel(v, vr))
    say("  rel err", e)
    bound = 1e-9 if precision == femgpu.FP64 else 1e-6
    assert e["u"] < bound and e["v"] < bound


def _localized(k, o, g, with_kd=True):
    """Localized terms before the first and third steps (each lasts one step)."""
    if k not in (0, 2):
        return
    rng = np.random.default_rng(100 + k)
    for e in (3, 41, 77):
        coord = rng.dirichlet(np.ones(4))
        F = rng.normal(size=3) * 0.05
        K = rng.normal(size=(3, 3)) * 20.0 + 60.0 * np.eye(3) if with_kd else None
        D = rng.normal(size=(3, 3)) * 0.05 + 0.1 * np.eye(3) if with_kd else None
        o.add_localized_force(e, coord, F, K, D)
        g.add_external_generalized_force(e, coord, F, K, D)


@pytest.mark.parametrize("precision", [femgpu.FP64, femgpu.FP32_REFINED])
def test_localized_force_stiffness_damping_implicit(precision):
    cfg = mg.config_corotational((5, 5, 5), seed=8)
    say("localized external terms, implicit, precision", precision)
    e = run_pair(cfg, 4, femgpu.EULER_IMPLICIT, precision=precision, am=AM, ak=AK, before_step=_localized)
    say("  rel err", e)
    bound = 1e-9 if precision == femgpu.FP64 else 1e-6
    assert e["u"] < bound and e["v"] < bound


@pytest.mark.parametrize("scheme", [femgpu.STATIC, femgpu.EULER_EXPLICIT, femgpu.RUNGE_KUTTA_4])
def test_localized_terms_other_schemes(scheme):
    cfg = soft_config(oracle.TET_COROTATIONAL, n=(5, 5, 5))
    say("localized external terms, scheme", scheme)
    e = run_pair(cfg, 4, scheme, before_step=_localized)
    say("  rel err", e)
    assert e["u"] < 1e-9 and e["v"] < 1e-9


def test_hex_localized_force():
    cfg = mg.config_hex((4, 3, 4))
    o = make_oracle(cfg, tol=1e-12)
    g = make_gpu(cfg, tol=1e-12)
    for k in range(3):
        if k == 0:
            for e, coord in ((5, np.full(8, 0.125)), (30, np.eye(8)[6])):
                o.add_localized_force(e, coord, [0.0, -2.0, 0.5])
                g.add_external_generalized_force(e, coord, [0.0, -2.0, 0.5])
        o.step(cfg.dt, 1)
        assert g.step(1) == femgpu.FEM_OK
    xr, vr = o.state()
    x, v = g.state()
    x0 = cfg.x0.reshape(-1)
    e = dict(u=rel(x - x0, xr - x0), v=rel(v, vr))
    say("hex localized force", e)
    assert e["u"] < 1e-9 and e["v"] < 1e-9


def _hex_localized(k, o, g):
    """Localized force, stiffness and damping on three hexahedra before the first and third steps."""
    if k not in (0, 2):
        return
    rng = np.random.default_rng(200 + k)
    for e in (3, 20, 41):
        coord = rng.dirichlet(np.ones(8))
        F = rng.normal(size=3) * 0.05
        K = rng.normal(size=(3, 3)) * 20.0 + 60.0 * np.eye(3)
        D = rng.normal(size=(3, 3)) * 0.05 + 0.1 * np.eye(3)
        o.add_localized_force(e, coord, F, K, D)
        g.add_external_generalized_force(e, coord, F, K, D)


@pytest.mark.parametrize("scheme,precision", [(femgpu.EULER_IMPLICIT, femgpu.FP64),
                                              (femgpu.EULER_IMPLICIT, femgpu.FP32_REFINED),
                                              (femgpu.STATIC, femgpu.FP64), (femgpu.RUNGE_KUTTA_4, femgpu.FP64)])
def test_hex_localized_stiffness_damping(scheme, precision):
    """Localized stiffness and damping on hexahedra: the lower-triangle view of D + dt K (implicit Euler) or K
    (static) applied beside the element-by-element operator and in its Jacobi diagonal, -dt K v in the implicit
    right-hand side; the explicit schemes take the force alone. Rayleigh damping on the implicit cases."""
    cfg = mg.config_hex((4, 3, 4))
    cfg["E"] = np.full(cfg.conn.shape[0], 5e3)
    cfg["dt"] = 1e-4 if scheme == femgpu.RUNGE_KUTTA_4 else 1e-3
    damped = scheme == femgpu.EULER_IMPLICIT
    say("hex localized stiffness and damping, scheme", scheme, "precision", precision)
    e = run_pair(cfg, 4, scheme, precision=precision, am=AM if damped else 0.0, ak=AK if damped else 0.0,
                 before_step=_hex_localized)
    say("  rel err", e)
    bound = 1e-9 if precision == femgpu.FP64 else 1e-8
    assert e["u"] < bound and e["v"] < bound


@pytest.mark.parametrize("precision", [femgpu.FP64, femgpu.FP32_REFINED])
def test_hex_compliance_after_localized_step(precision):
    """The compliance after a hexahedral step with localized stiffness and damping uses that step's operator
    with the external entries (applyCompliance of the last system matrix)."""
    cfg = mg.config_hex((4, 3, 4))
    o = make_oracle(cfg, tol=1e-12)
    g = make_gpu(cfg, precision=precision, tol=1e-12)
    _hex_localized(0, o, g)
    assert o.step(cfg.dt, 1) == 0
    assert g.step(1) == femgpu.FEM_OK
    rng = np.random.default_rng(6)
    B = rng.normal(size=(cfg.x0.size, 3))
    X = g.apply_compliance(B)
    Xr = o.apply_compliance(B)
    err = np.linalg.norm(X - Xr) / np.linalg.norm(Xr)
    say("hex compliance after a localized step, precision", precision, "rel err", err)
    assert err < (1e-9 if precision == femgpu.FP64 else 1e-6)


def test_unsupported_combinations_rejected():
    cfg = mg.config_hex((2, 2, 2))
    g = make_gpu(cfg, rayleigh_stiffness=AK)
    g.add_external_generalized_force(0, np.full(8, 0.125), [0, 0, 1.0], np.eye(3), None)
    assert g.step(1) == femgpu.FEM_OK  # localized stiffness on hexahedra
    g.set_options(dt=cfg.dt, scheme=femgpu.RUNGE_KUTTA_4, rayleigh_stiffness=AK)
    with pytest.raises(femgpu.FemError):
        g.step(1)  # Rayleigh stiffness force from the reference's global K holding that localized stiffness
    g.set_options(dt=cfg.dt, scheme=femgpu.RUNGE_KUTTA_4)
    assert g.step(1) == femgpu.FEM_OK  # the other schemes step hexahedra too
    t = make_gpu(soft_config(oracle.TET_LINEAR))
    t.add_external_generalized_force(0, np.full(4, 0.25), [0, 0, 1.0], np.eye(3), None)
    t.set_options(dt=1e-3, scheme=femgpu.LINEAR_EULER_IMPLICIT)
    with pytest.raises(femgpu.FemError):
        t.step(1)  # stiffness queued before the switch to a linear scheme: rejected and dropped
    assert t.step(1) == femgpu.FEM_OK
    with pytest.raises(femgpu.FemError):
        t.add_external_generalized_force(0, np.full(4, 0.25), [0, 0, 1.0], np.eye(3), None)
    with pytest.raises(femgpu.FemError):
        t.add_external_generalized_force(10 ** 6, np.full(4, 0.25), [0, 0, 1.0])


def _fixed_point_h(cfg, element, coord):
    """H of a fixed-point constraint at natural coordinates of an element (three rows, one per axis), as
    Physics/FemConstraintFixedPoint.cpp builds it; returned as the ndof x 3 matrix H^T."""
    nn = cfg.conn.shape[1]
    Ht = np.zeros((cfg.x0.size, 3))
    for a in range(nn):
        for k in range(3):
            Ht[3 * cfg.conn[element, a] + k, k] = coord[a]
    return Ht


@pytest.mark.parametrize("precision", [femgpu.FP64, femgpu.FP32_REFINED])
@pytest.mark.parametrize("kind", ["corotational", "hex"])
def test_apply_compliance_vs_oracle(precision, kind):
    if kind == "hex":
        cfg = mg.config_hex((4, 3, 5))
        coord = np.full(8, 0.125)
    else:
        cfg = mg.config_corotational((6, 6, 6), seed=3)
        coord = np.array([0.1, 0.2, 0.3, 0.4])
    say("apply compliance", kind, "precision", precision)
    o = make_oracle(cfg, tol=1e-12)
    g = make_gpu(cfg, precision=precision, tol=1e-12)
    with pytest.raises(femgpu.FemError):
        g.apply_compliance(np.zeros(cfg.x0.size))  # no system matrix before the first step
    o.step(cfg.dt, 3)
    assert g.step(3) == femgpu.FEM_OK
    stats = g.stats()
    rng = np.random.default_rng(4)
    B = np.concatenate([rng.normal(size=(cfg.x0.size, 3)), _fixed_point_h(cfg, 7, coord)], axis=1)
    X = g.apply_compliance(B)
    Xr = o.apply_compliance(B)
    err = np.linalg.norm(X - Xr) / np.linalg.norm(Xr)
    say("  rel err", err)
    assert err < (1e-9 if precision == femgpu.FP64 else 1e-8)
    bc = (3 * np.asarray(cfg.bc_nodes)[:, None] + np.arange(3)).reshape(-1)
    assert np.all(X[bc] == 0.0)
    assert g.stats() == stats  # the step's statistics are untouched
    # the state moves on unchanged by the compliance solves
    o.step(cfg.dt, 1)
    assert g.step(1) == femgpu.FEM_OK
    xr, _ = o.state()
    x, _ = g.state()
    x0 = cfg.x0.reshape(-1)
    assert rel(x - x0, xr - x0) < (1e-9 if precision == femgpu.FP64 else 1e-6)


@pytest.mark.parametrize("precision", [femgpu.FP64, femgpu.FP32_REFINED])
def test_apply_compliance_many_columns(precision):
    """Columns go through the solver up to 8 at a time (here 8 + 3: padded to 4 in FP64, one by one through the
    FP32 step solver in the FP32 + refinement mode): a zero column returns zero, equal columns of one batch
    return equal solutions, every column matches the oracle's solve."""
    cfg = mg.config_corotational((5, 6, 4), seed=6)
    o = make_oracle(cfg, tol=1e-12)
    g = make_gpu(cfg, precision=precision, tol=1e-12)
    bound = 1e-9 if precision == femgpu.FP64 else 1e-8
    o.step(cfg.dt, 2)
    assert g.step(2) == femgpu.FEM_OK
    rng = np.random.default_rng(8)
    B = rng.normal(size=(cfg.x0.size, 11))
    B[:, 4] = 0.0
    B[:, 6] = B[:, 1]
    X = g.apply_compliance(B)
    Xr = o.apply_compliance(B)
    err = np.linalg.norm(X - Xr) / np.linalg.norm(Xr)
    say("many columns: precision", precision, "rel err", err)
    assert err < bound
    assert np.all(X[:, 4] == 0.0)
    assert np.array_equal(X[:, 6], X[:, 1])
    x1 = g.apply_compliance(B[:, 9])  # a single column on its own
    assert np.linalg.norm(x1 - Xr[:, 9]) <= bound * np.linalg.norm(Xr[:, 9])


def test_apply_compliance_device_pointers():
    import torch
    cfg = mg.config_corotational((4, 5, 4), seed=9)
    g = make_gpu(cfg, tol=1e-12)
    assert g.step(2) == femgpu.FEM_OK
    B = np.random.default_rng(2).normal(size=(cfg.x0.size, 5))
    X = g.apply_compliance(B)
    bd = torch.from_numpy(np.ascontiguousarray(B.T)).cuda()
    xd = torch.zeros_like(bd)
    torch.cuda.synchronize()
    g.apply_compliance_device(bd.data_ptr(), B.shape[1], xd.data_ptr())
    g.synchronize()
    assert np.array_equal(xd.cpu().numpy().T, X)


def _localized_h(cfg, element, coord, dirs, scale):
    """H^T (ndof x m) of a constraint localized in an element: row j of H is scale * coord[a] * dirs[j] at node
    a's DOFs (Physics/FemConstraintFixedPoint.cpp:84-106, FemConstraintFrictionlessContact.cpp:84-112)."""
    Ht = np.zeros((cfg.x0.size, len(dirs)))
    for j, d in enumerate(dirs):
        for a in range(cfg.conn.shape[1]):
            if coord[a] != 0.0:
                for k in range(3):
                    Ht[3 * cfg.conn[element, a] + k, j] = coord[a] * d[k] * scale
    return Ht


@pytest.mark.parametrize("kind,precision", [("corotational", femgpu.FP64), ("corotational", femgpu.FP32_REFINED),
                                            ("hex", femgpu.FP64)])
def test_constraint_compliance(kind, precision):
    """fem_constraint_compliance builds H^T on the device and returns C H^T and H C H^T: C H^T equals
    fem_apply_compliance of the same dense columns bit for bit and the oracle's solve to the CG tolerance;
    H C H^T equals H (C H^T); boundary rows are zero; the step's state and statistics are untouched. In the
    FP32 + refinement mode the 1 or 3 columns go one by one through the step's FP32 solver."""
    if kind == "hex":
        cfg = mg.config_hex((4, 3, 5))
        coord = np.array([0.2, 0.0, 0.1, 0.15, 0.05, 0.3, 0.1, 0.1])
    else:
        cfg = mg.config_corotational((6, 5, 6), seed=12)
        coord = np.array([0.5, 0.0, 0.3, 0.2])
    say("constraint compliance", kind, "precision", precision)
    o = make_oracle(cfg, tol=1e-12)
    g = make_gpu(cfg, precision=precision, tol=1e-12)
    bound = 1e-9 if precision == femgpu.FP64 else 1e-8
    with pytest.raises(femgpu.FemError):
        g.constraint_compliance(0, coord, np.eye(3), 1.0)  # no system matrix before the first step
    o.step(cfg.dt, 2)
    assert g.step(2) == femgpu.FEM_OK
    stats = g.stats()
    bcset = set(int(n) for n in cfg.bc_nodes)
    # an element with fixed nodes and at least one free node of non-zero coordinate
    touching = [e for e in range(cfg.conn.shape[0]) if bcset & set(int(n) for n in cfg.conn[e])
                and any(coord[a] != 0.0 and int(cfg.conn[e, a]) not in bcset for a in range(cfg.conn.shape[1]))]
    free = [e for e in range(cfg.conn.shape[0]) if not bcset & set(int(n) for n in cfg.conn[e])]
    bc = (3 * np.asarray(cfg.bc_nodes)[:, None] + np.arange(3)).reshape(-1)
    normal = np.array([0.3, -0.8, 0.5]) / np.linalg.norm([0.3, -0.8, 0.5])
    scale = cfg.dt * 1.3
    for element, dirs in ((free[len(free) // 2], np.eye(3)), (touching[0], np.eye(3)),
                          (free[len(free) // 3], normal[None, :])):
        cht, hcht = g.constraint_compliance(element, coord, dirs, scale)
        Ht = _localized_h(cfg, element, coord, dirs, scale)
        assert np.array_equal(cht, g.apply_compliance(Ht))
        Xr = o.apply_compliance(Ht)
        err = np.linalg.norm(cht - Xr) / np.linalg.norm(Xr)
        href = Ht.T @ Xr
        herr = np.abs(hcht - href).max() / np.abs(href).max()
        say("  element", element, "directions", len(dirs), "rel err", err, "H C H^T rel err", herr)
        assert err < bound and herr < bound
        assert np.abs(hcht - Ht.T @ cht).max() <= 1e-13 * np.abs(hcht).max()
        assert np.all(cht[bc] == 0.0)
        assert np.all(np.linalg.eigvalsh(0.5 * (hcht + hcht.T)) > 0.0)
        _, h_only = g.constraint_compliance(element, coord, dirs, scale, want_cht=False)
        assert np.array_equal(h_only, hcht)
    assert g.stats() == stats
    with pytest.raises(femgpu.FemError):
        g.constraint_compliance(cfg.conn.shape[0], coord, np.eye(3), scale)  # element out of range
    with pytest.raises(femgpu.FemError):
        g.constraint_compliance(0, coord, np.zeros((9, 3)), scale)  # more than 8 directions
    with pytest.raises(ValueError):
        g.constraint_compliance(0, coord[:-1], np.eye(3), scale)
    # a zero localization gives zero terms
    cht, hcht = g.constraint_compliance(free[0], np.zeros_like(coord), np.eye(3), scale)
    assert not cht.any() and not hcht.any()
