"""Pins of the oracle's Rayleigh damping, mass lumping, localized external force and integration schemes
other than implicit Euler (Math/OdeSolver*.cpp, Physics/FemRepresentation.cpp:479-629,
Physics/Fem3DRepresentation.cpp:133-198).

The schemes are checked against modal closed forms. For linear tetrahedra f = -K u - D v with
D = alpha_M M + alpha_K K, so every M-orthonormal mode of the free DOFs (K phi = lam M phi) evolves by the
scalar recurrence of the scheme with acceleration a(q, qd) = -lam q - (alpha_M + alpha_K lam) qd. The
modes come from a dense generalized eigen-solve (scipy) of K and M assembled here from the element
matrices (oracle.tet_linear_km, pinned by the reference's element known-answer tests); the oracle
assembles sparse, applies boundary conditions and solves with its direct solver. A dropped damping term,
a wrong sign or a swapped position/velocity update moves the result by O(1) of the modal motion.

The assembled damping, lumped mass and localized external terms are checked the way the reference's
FemRepresentationTests (Physics/UnitTests/FemRepresentationTests.cpp:240-410, 529-560) check them:
D = alpha_M M + alpha_K K, the damping force -D v, the lumped diagonal equal to the row sums (and, for
tetrahedra, to the node masses), and F, K, D plus the external terms scattered with the localization's
coordinates.
"""
import numpy as np
import pytest
import scipy.linalg as sla

import oracle
from opensurgsim_amd import meshgen as mg

DT = 1e-4
AM, AK = 3.0, 2e-4


def small_block():
    x0, tets, _ = mg.block_tet_mesh(2, 2, 2)
    bc = mg.face_nodes(2, 2, 2, 1, 0)
    return x0, tets, bc


def dense_km(x0, tets, E, nu, rho):
    n = x0.size
    K = np.zeros((n, n))
    M = np.zeros((n, n))
    for t in tets:
        Ke, Me, _ = oracle.tet_linear_km(x0[t].reshape(-1), E, nu, rho)
        dofs = (3 * t[:, None] + np.arange(3)[None]).reshape(-1)
        K[np.ix_(dofs, dofs)] += Ke
        M[np.ix_(dofs, dofs)] += Me
    return K, M


MAT = (1e3, 0.3, 1000.0)


def setup(seed=0):
    x0, tets, bc = small_block()
    K, M = dense_km(x0, tets, *MAT)
    bcd = (3 * bc[:, None] + np.arange(3)).reshape(-1)
    free = np.setdiff1d(np.arange(x0.size), bcd)
    rng = np.random.default_rng(seed)
    u = np.zeros(x0.size)
    v = np.zeros(x0.size)
    u[free] = 1e-4 * rng.normal(size=free.size)
    v[free] = 1e-3 * rng.normal(size=free.size)
    return x0, tets, bc, K, M, free, u, v


def make_sim(x0, tets, bc, scheme, am=0.0, ak=0.0, lumping=False):
    s = oracle.Sim(oracle.TET_LINEAR, x0, tets, MAT[0], MAT[1], MAT[2])
    for n in bc:
        s.add_bc_node(int(n))
    s.set_gravity((0.0, -9.81, 0.0), False)
    s.set_damping(am, ak, lumping)
    s.set_solver(scheme=scheme, linsolver=oracle.LU)
    return s


def modal(K, M, free, u, v):
    lam, Phi = sla.eigh(K[np.ix_(free, free)], M[np.ix_(free, free)])
    Mf = M[np.ix_(free, free)]
    return lam, Phi, Phi.T @ Mf @ u[free], Phi.T @ Mf @ v[free]


def modal_steps(scheme, lam, q, qd, am, ak, steps):
    c = am + ak * lam

    def acc(q_, qd_):
        return -lam * q_ - c * qd_

    for _ in range(steps):
        if scheme in (oracle.EULER_IMPLICIT, oracle.LINEAR_EULER_IMPLICIT):
            dqd = (-lam * q - c * qd - DT * lam * qd) / (1.0 / DT + c + DT * lam)
            qd = qd + dqd
            q = q + DT * qd
        elif scheme in (oracle.EULER_EXPLICIT, oracle.LINEAR_EULER_EXPLICIT):
            q, qd = q + DT * qd, qd + DT * acc(q, qd)
        elif scheme in (oracle.EULER_EXPLICIT_MODIFIED, oracle.LINEAR_EULER_EXPLICIT_MODIFIED):
            qd = qd + DT * acc(q, qd)
            q = q + DT * qd
        elif scheme in (oracle.RUNGE_KUTTA_4, oracle.LINEAR_RUNGE_KUTTA_4):
            k1 = (qd, acc(q, qd))
            k2 = (qd + k1[1] * DT / 2, acc(q + k1[0] * DT / 2, qd + k1[1] * DT / 2))
            k3 = (qd + k2[1] * DT / 2, acc(q + k2[0] * DT / 2, qd + k2[1] * DT / 2))
            k4 = (qd + k3[1] * DT, acc(q + k3[0] * DT, qd + k3[1] * DT))
            q = q + DT / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
            qd = qd + DT / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1])
        elif scheme in (oracle.STATIC, oracle.LINEAR_STATIC):
            # K dx = f = -K u - D v: q' = -c qd / lam, qd' = 0
            q = q + (-lam * q - c * qd) / lam
            qd = np.zeros_like(qd)
    return q, qd


SCHEMES = [oracle.EULER_IMPLICIT, oracle.LINEAR_EULER_IMPLICIT, oracle.EULER_EXPLICIT,
           oracle.LINEAR_EULER_EXPLICIT, oracle.EULER_EXPLICIT_MODIFIED, oracle.LINEAR_EULER_EXPLICIT_MODIFIED,
           oracle.RUNGE_KUTTA_4, oracle.LINEAR_RUNGE_KUTTA_4, oracle.STATIC, oracle.LINEAR_STATIC]


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("damping", [(0.0, 0.0), (AM, AK)])
def test_schemes_match_modal_closed_form(scheme, damping):
    am, ak = damping
    x0, tets, bc, K, M, free, u, v = setup()
    steps = 2 if scheme in (oracle.STATIC, oracle.LINEAR_STATIC) else 12
    s = make_sim(x0, tets, bc, scheme, am, ak)
    s.set_state(x=x0.reshape(-1) + u, v=v)
    assert s.step(DT, steps) == 0
    x, vv = s.state()
    lam, Phi, q, qd = modal(K, M, free, u, v)
    q, qd = modal_steps(scheme, lam, q, qd, am, ak, steps)
    ue = np.zeros_like(u)
    ve = np.zeros_like(v)
    ue[free] = Phi @ q
    ve[free] = Phi @ qd
    du = x - x0.reshape(-1)
    scale_u = max(np.linalg.norm(u), np.linalg.norm(ue))
    scale_v = max(np.linalg.norm(v), 1e-300)
    assert np.linalg.norm(du - ue) <= 1e-9 * scale_u
    assert np.linalg.norm(vv - ve) <= 1e-9 * scale_v
    # boundary DOFs never move
    bcd = np.setdiff1d(np.arange(x0.size), free)
    assert np.all(du[bcd] == 0.0) and np.all(vv[bcd] == 0.0)


def test_static_with_rayleigh_first_step_closed_form():
    """Static from a moving state: K dx = -K u - D v, so the new displacement is -K^-1 D v."""
    x0, tets, bc, K, M, free, u, v = setup(1)
    s = make_sim(x0, tets, bc, oracle.STATIC, AM, AK)
    s.set_state(x=x0.reshape(-1) + u, v=v)
    s.step(DT, 1)
    x, vv = s.state()
    D = AM * M + AK * K
    expect = -np.linalg.solve(K[np.ix_(free, free)], (D @ v)[free])
    assert np.allclose((x - x0.reshape(-1))[free], expect, rtol=0, atol=1e-10 * np.abs(expect).max())
    assert np.all(vv == 0.0)


@pytest.mark.parametrize("scheme", [oracle.EULER_IMPLICIT, oracle.LINEAR_EULER_IMPLICIT])
def test_implicit_with_mass_lumping_matches_lumped_modal_form(scheme):
    x0, tets, bc, K, M, free, u, v = setup(2)
    Ml = np.diag(M.sum(axis=1))
    s = make_sim(x0, tets, bc, scheme, AM, AK, lumping=True)
    s.set_state(x=x0.reshape(-1) + u, v=v)
    s.step(DT, 10)
    x, vv = s.state()
    lam, Phi, q, qd = modal(K, Ml, free, u, v)
    q, qd = modal_steps(scheme, lam, q, qd, AM, AK, 10)
    assert np.linalg.norm((x - x0.reshape(-1))[free] - Phi @ q) <= 1e-9 * np.linalg.norm(u)
    assert np.linalg.norm(vv[free] - Phi @ qd) <= 1e-9 * np.linalg.norm(v)


def test_explicit_ignores_mass_lumping():
    """computeM never lumps (Physics/FemRepresentation.cpp:412-421): explicit schemes use the consistent M."""
    x0, tets, bc, K, M, free, u, v = setup(3)
    out = []
    for lump in (False, True):
        s = make_sim(x0, tets, bc, oracle.EULER_EXPLICIT, 0.0, AK, lumping=lump)
        s.set_state(x=x0.reshape(-1) + u, v=v)
        s.step(DT, 5)
        out.append(s.state())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_explicit_lumped_rayleigh_mass_uses_the_global_diagonal():
    """computeF calls addRayleighDampingForce with useGlobalMassMatrix = hasF() (Physics/FemRepresentation.cpp:
    396-410, 565-593): a representation's first F takes -alpha_M M v element by element; once F exists the global
    M is used, and with lumping only its diagonal, which is the consistent one after computeM (explicit schemes,
    :412-421) and the row sums after an implicit step's lumping computeFMDK (:499-508). The force differences of
    two oracles that differ only in alpha_M, at the same state, against those closed forms (M assembled here from
    the pinned element matrices)."""
    x0, tets, bc, K, M, free, u, v = setup(5)
    x = x0.reshape(-1) + u

    def force_difference(first_scheme, steps_before):
        fs = []
        for am in (AM, 0.0):
            s = make_sim(x0, tets, [], first_scheme, am, 0.0, lumping=True)
            for _ in range(steps_before):
                s.set_state(x=x, v=v)
                s.step(DT, 1)
            s.set_solver(scheme=oracle.EULER_EXPLICIT, linsolver=oracle.LU)
            s.set_state(x=x, v=v)
            s.step(DT, 1)
            fs.append(s.force())
        return fs[0] - fs[1]

    scale = AM * np.abs(M @ v).max()
    # the first F of the representation: -alpha_M M v (consistent, element by element)
    d = force_difference(oracle.EULER_EXPLICIT, 0)
    assert np.abs(d - (-AM * (M @ v))).max() <= 1e-12 * scale
    # after an explicit step (computeM, never lumped): -alpha_M diag(M) v
    d = force_difference(oracle.EULER_EXPLICIT, 1)
    assert np.abs(d - (-AM * np.diag(M) * v)).max() <= 1e-12 * scale
    assert np.abs(d - (-AM * (M @ v))).max() > 1e-3 * scale  # the diagonal form differs from M v
    # after an implicit step (lumping computeFMDK): -alpha_M rowsum(M) v
    d = force_difference(oracle.EULER_IMPLICIT, 1)
    assert np.abs(d - (-AM * M.sum(axis=1) * v)).max() <= 1e-12 * scale


def test_rayleigh_terms_and_lumped_mass_assembled():
    x0, tets, bc, K, M, free, u, v = setup(4)
    base = make_sim(x0, tets, [], oracle.EULER_IMPLICIT)
    damp = make_sim(x0, tets, [], oracle.EULER_IMPLICIT, AM, AK)
    lump = make_sim(x0, tets, [], oracle.EULER_IMPLICIT, 0.0, 0.0, lumping=True)
    for s in (base, damp, lump):
        s.set_state(x=x0.reshape(-1) + u, v=v)
        s.update_fmdk()
    outer, inner = base.pattern()
    n = x0.size

    def dense(vals):
        A = np.zeros((n, n))
        cols = np.repeat(np.arange(n), np.diff(outer))
        A[inner, cols] = vals
        return A

    Mb, Kb = dense(base.matrix_values(0)), dense(base.matrix_values(1))
    assert np.allclose(Mb, M, rtol=0, atol=1e-15 * np.abs(M).max())
    assert np.allclose(Kb, K, rtol=0, atol=1e-12 * np.abs(K).max())
    Dd = dense(damp.matrix_values(3))
    assert np.allclose(Dd, AM * M + AK * K, rtol=0, atol=1e-12 * np.abs(AK * K).max())
    assert np.allclose(damp.force() - base.force(), -(AM * M + AK * K) @ v, rtol=0,
                       atol=1e-12 * np.abs((AK * K) @ v).max())
    Ml = dense(lump.matrix_values(0))
    assert np.allclose(Ml, np.diag(M.sum(axis=1)), rtol=0, atol=1e-16)
    # tetrahedra: the lumped diagonal is each node's share m_e / 4 of its elements' masses
    assert np.allclose(np.diag(Ml).reshape(-1, 3), lump.mass_per_node()[:, None], rtol=1e-14, atol=0)


def test_localized_external_force_stiffness_damping():
    x0, tets, bc, K, M, free, u, v = setup(5)
    n = x0.size
    a = make_sim(x0, tets, bc, oracle.EULER_IMPLICIT)
    b = make_sim(x0, tets, bc, oracle.EULER_IMPLICIT)
    for s in (a, b):
        s.set_state(x=x0.reshape(-1) + u, v=v)
    rng = np.random.default_rng(9)
    e = 17
    coord = np.array([0.1, 0.2, 0.3, 0.4])
    F = np.array([0.3, -1.2, 0.5])
    Kl = rng.normal(size=(3, 3)) * 50.0
    Dl = rng.normal(size=(3, 3)) * 0.2
    a.add_localized_force(e, coord, F, Kl, Dl)
    a.update_fmdk()
    b.update_fmdk()
    outer, inner = a.pattern()
    cols = np.repeat(np.arange(n), np.diff(outer))

    def dense(vals):
        A = np.zeros((n, n))
        A[inner, cols] = vals
        return A

    ids = tets[e]
    fexp = np.zeros(n)
    Kexp = np.zeros((n, n))
    Dexp = np.zeros((n, n))
    for i, ni in enumerate(ids):
        fexp[3 * ni:3 * ni + 3] += coord[i] * F
        for j, nj in enumerate(ids):
            Kexp[3 * ni:3 * ni + 3, 3 * nj:3 * nj + 3] += coord[i] * coord[j] * Kl
            Dexp[3 * ni:3 * ni + 3, 3 * nj:3 * nj + 3] += coord[i] * coord[j] * Dl
    df = a.force() - b.force()
    assert np.allclose(df, fexp, rtol=0, atol=1e-15)
    # the force distribution carries F and no moment about the localized point
    p = coord @ x0[ids]
    assert np.allclose(df.reshape(-1, 3).sum(axis=0), F, atol=1e-15)
    assert np.allclose(np.cross(x0 - p, df.reshape(-1, 3)).sum(axis=0), 0.0, atol=1e-15)
    assert np.allclose(dense(a.matrix_values(1)) - dense(b.matrix_values(1)), Kexp, rtol=0, atol=1e-10)
    assert np.allclose(dense(a.matrix_values(3)) - dense(b.matrix_values(3)), Dexp, rtol=0, atol=1e-15)
    # one step: the system matrix and right-hand side hold the external terms (OdeSolverEulerImplicit.cpp:148-158)
    a.step(DT, 1)
    Ma, Da, Ka = (dense(a.matrix_values(w)) for w in (0, 3, 1))
    A = dense(a.matrix_values(2))
    bcd = np.setdiff1d(np.arange(n), free)
    Aexp = Ma / DT + Da + Ka * DT
    Aexp[bcd, :] = 0.0
    Aexp[:, bcd] = 0.0
    Aexp[bcd, bcd] = 1.0
    assert np.allclose(A, Aexp, rtol=0, atol=1e-12 * np.abs(Aexp).max())
    # afterUpdate resets the external terms: the next update at the same state equals one that never had them
    b.set_state(*a.state())
    a.update_fmdk()
    b.update_fmdk()
    assert np.array_equal(a.force(), b.force())
    assert np.array_equal(a.matrix_values(1), b.matrix_values(1))
    assert np.array_equal(a.matrix_values(3), b.matrix_values(3))


@pytest.mark.parametrize("scheme", [oracle.EULER_IMPLICIT, oracle.EULER_EXPLICIT, oracle.STATIC])
def test_apply_compliance_equals_dense_compliance(scheme):
    """applyCompliance (zero boundary rows, solve, zero boundary rows; Physics/DeformableRepresentation.cpp:175-190)
    equals the dense compliance matrix of OdeSolver::computeComplianceMatrixFromSystemMatrix (the inverse of the
    system matrix with boundary rows and columns zeroed, Math/OdeSolver.cpp:69-76) applied to the columns."""
    x0, tets, bc, K, M, free, u, v = setup(6)
    s = make_sim(x0, tets, bc, scheme, AM, AK)
    s.set_state(x=x0.reshape(-1) + u, v=v)
    s.step(DT, 1)
    n = x0.size
    outer, inner = s.pattern()
    A = np.zeros((n, n))
    A[inner, np.repeat(np.arange(n), np.diff(outer))] = s.matrix_values(2)
    C = np.linalg.inv(A)
    bcd = np.setdiff1d(np.arange(n), free)
    C[bcd, :] = 0.0
    C[:, bcd] = 0.0
    B = np.random.default_rng(1).normal(size=(n, 5))
    X = s.apply_compliance(B)
    assert np.allclose(X, C @ B, rtol=0, atol=1e-10 * np.abs(C @ B).max())
    assert np.array_equal(s.apply_compliance(B[:, 2]), X[:, 2])
