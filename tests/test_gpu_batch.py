"""Batched independent instances (fem_set_instances): every instance must step exactly as its own
representation would, against the oracle run per instance."""
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


def small_batch(count=5, seed=21):
    cfgs = mg.config_batched(count, seed=seed, lo=2, hi=5)
    for i, c in enumerate(cfgs):  # give every instance some motion besides gravity
        top = mg.face_nodes(c.nx, c.ny, c.nz, 1, 1)
        v0 = np.zeros((c.x0.shape[0], 3))
        v0[top, 0] = 0.01 * (i + 1)
        c["v0"] = v0.reshape(-1)
    return cfgs, mg.concat_instances(cfgs)


def run_batch(b, steps, precision=femgpu.FP64, tol=1e-12, tuning=None, **opts):
    g = femgpu.FemGpu(0)
    g.upload_mesh(b.kind, b.x0, b.conn, b.E, b.nu, b.rho)
    g.add_bc_nodes(b.bc_nodes)
    g.set_options(dt=1e-3, cg_tolerance=tol, precision=precision, **opts)
    g.set_state(v=b.v0)
    g.set_instances(b.node_offsets, b.dt)
    if tuning:
        g.set_batch_tuning(*tuning)
    assert g.step(steps) == femgpu.FEM_OK
    x, v = g.state()
    it, err, conv = g.instance_stats()
    tun = g.batch_tuning()
    g.close()
    return x, v, it, conv, tun


def run_oracle(c, steps, tol=1e-13, damping=None):
    s = oracle.Sim(c.kind, c.x0, c.conn, c.E, c.nu, c.rho)
    for n in c.bc_nodes:
        s.add_bc_node(n)
    s.set_solver(tol=tol)
    if damping:
        s.set_damping(*damping)
    s.set_state(v=c.v0)
    s.step(c.dt, steps)
    return s.state(), s.last_cg()


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_batch_matches_oracle_per_instance():
    cfgs, b = small_batch()
    steps = 8
    say("batch of %d instances, %d steps, dts %s" % (len(cfgs), steps, list(b.dt)))
    x, v, it, conv, tun = run_batch(b, steps, tol=1e-12)
    say("  tuning", tun, "iterations", it)
    assert conv.all()
    off = b.node_offsets
    for i, c in enumerate(cfgs):
        (xo, vo), _ = run_oracle(c, steps, tol=1e-12)
        sl = slice(3 * off[i], 3 * off[i + 1])
        x0 = c.x0.reshape(-1)
        e = dict(u=rel(x[sl] - x0, xo - x0), v=rel(v[sl], vo))
        say("  instance", i, c.conn.shape[0], "tets, dt", c.dt, e)
        assert e["u"] < 1e-9 and e["v"] < 1e-9


def test_batch_rayleigh_damping_and_lumping_per_instance():
    """Rayleigh damping and mass lumping with a time step per instance: the element kernel forms each
    instance's M (1/dt_i + alpha_M) + K (dt_i + alpha_K)."""
    cfgs, b = small_batch(4, seed=17)
    steps = 6
    x, v, it, conv, _ = run_batch(b, steps, tol=1e-12, rayleigh_mass=2.0, rayleigh_stiffness=1e-4, mass_lumping=1)
    assert conv.all()
    off = b.node_offsets
    for i, c in enumerate(cfgs):
        (xo, vo), _ = run_oracle(c, steps, tol=1e-12, damping=(2.0, 1e-4, True))
        sl = slice(3 * off[i], 3 * off[i + 1])
        x0 = c.x0.reshape(-1)
        e = dict(u=rel(x[sl] - x0, xo - x0), v=rel(v[sl], vo))
        say("  damped instance", i, "dt", c.dt, e)
        assert e["u"] < 1e-9 and e["v"] < 1e-9


def test_batch_matches_single_instance_runs():
    """Each instance of the batch against the same mesh stepped alone through the single-mesh path."""
    cfgs, b = small_batch(4, seed=5)
    steps = 6
    x, v, it, conv, _ = run_batch(b, steps, tol=1e-10)
    off = b.node_offsets
    for i, c in enumerate(cfgs):
        g = femgpu.FemGpu(0)
        g.upload_mesh(c.kind, c.x0, c.conn, c.E, c.nu, c.rho)
        g.add_bc_nodes(c.bc_nodes)
        g.set_options(dt=c.dt, cg_tolerance=1e-10)
        g.set_state(v=c.v0)
        g.step(steps)
        xs, vs = g.state()
        its = g.stats()["last_cg_iterations"]
        g.close()
        sl = slice(3 * off[i], 3 * off[i + 1])
        say("  instance", i, "iterations batch", it[i], "single", its, "v rel", rel(v[sl], vs))
        assert rel(v[sl], vs) < 1e-9
        assert abs(int(it[i]) - int(its)) <= 1


def test_batch_fp32_refined_close_to_fp64():
    cfgs, b = small_batch(6, seed=8)
    steps = 6
    x64, v64, _, _, _ = run_batch(b, steps, femgpu.FP64, tol=1e-10)
    x32, v32, it, conv, _ = run_batch(b, steps, femgpu.FP32_REFINED, tol=1e-10)
    x0 = b.x0.reshape(-1)
    e = dict(u=rel(x32 - x0, x64 - x0), v=rel(v32, v64))
    say("  fp32+refine vs fp64", e, "inner iterations", it)
    assert conv.all()
    assert e["u"] < 1e-6 and e["v"] < 1e-6


def test_batch_graphs_and_shapes_agree():
    cfgs, b = small_batch(5, seed=13)
    steps = 4
    ref = run_batch(b, steps, tol=1e-11, tuning=(128, 8, True))
    plain = run_batch(b, steps, tol=1e-11, tuning=(128, 8, False))
    assert np.array_equal(ref[0], plain[0]) and np.array_equal(ref[1], plain[1])  # graph replay = launches
    for shape in [(64, 4, True), (512, 16, True)]:
        other = run_batch(b, steps, tol=1e-11, tuning=shape)
        say("  shape", shape, "v rel", rel(other[1], ref[1]))
        assert rel(other[1], ref[1]) < 1e-10


def test_batch_rejects_elements_across_instances():
    cfgs, b = small_batch(3)
    g = femgpu.FemGpu(0)
    g.upload_mesh(b.kind, b.x0, b.conn, b.E, b.nu, b.rho)
    bad = b.node_offsets.copy()
    bad[1] += 1  # one node of instance 1 moved into instance 0: its elements straddle
    with pytest.raises(femgpu.FemError):
        g.set_instances(bad, b.dt)
    g.close()


def test_non_finite_instance_is_deactivated_alone():
    """An instance whose state becomes non-finite is deactivated by itself (as one representation would
    be); the others keep stepping and still match the oracle."""
    cfgs, b = small_batch(3, seed=4)
    bad = 1
    v0 = b.v0.copy()
    v0[3 * (b.node_offsets[bad] + 5) + 1] = np.nan
    g = femgpu.FemGpu(0)
    g.upload_mesh(b.kind, b.x0, b.conn, b.E, b.nu, b.rho)
    g.add_bc_nodes(b.bc_nodes)
    g.set_options(dt=1e-3, cg_tolerance=1e-12)
    g.set_state(v=v0)
    g.set_instances(b.node_offsets, b.dt)
    steps = 4
    assert g.step(steps) == femgpu.FEM_OK
    g.instance_stats()
    active = g.last_instance_active.tolist()
    say("  active", active)
    assert active == [1, 0, 1]
    x, v = g.state()
    g.close()
    off = b.node_offsets
    for i in (0, 2):
        (xo, vo), _ = run_oracle(cfgs[i], steps, tol=1e-12)
        sl = slice(3 * off[i], 3 * off[i + 1])
        assert rel(v[sl], vo) < 1e-9


AM, AK = 2.0, 1e-4
SCHEMES = [(femgpu.STATIC, 2), (femgpu.EULER_EXPLICIT, 8), (femgpu.EULER_EXPLICIT_MODIFIED, 8),
           (femgpu.RUNGE_KUTTA_4, 6), (femgpu.LINEAR_STATIC, 3), (femgpu.LINEAR_EULER_EXPLICIT, 8),
           (femgpu.LINEAR_EULER_EXPLICIT_MODIFIED, 8), (femgpu.LINEAR_RUNGE_KUTTA_4, 6),
           (femgpu.LINEAR_EULER_IMPLICIT, 6)]
LINEAR_SCHEMES = {femgpu.LINEAR_STATIC, femgpu.LINEAR_EULER_EXPLICIT, femgpu.LINEAR_EULER_EXPLICIT_MODIFIED,
                  femgpu.LINEAR_RUNGE_KUTTA_4, femgpu.LINEAR_EULER_IMPLICIT}


def soft_batch(count, seed, kind):
    """Soft instances (E 2-8 kPa) with time steps of 0.1, 0.2 and 0.4 ms, so that the explicit schemes stay
    bounded over the steps, and a tangential velocity on one face of each."""
    cfgs = mg.config_batched(count, seed=seed, lo=2, hi=4)
    for i, c in enumerate(cfgs):
        c["kind"] = kind
        c["E"] = np.full(c.conn.shape[0], 2e3 * (1 + i % 4))
        c["dt"] = [1e-4, 2e-4, 4e-4][i % 3]
        top = mg.face_nodes(c.nx, c.ny, c.nz, 1, 1)
        v0 = np.zeros((c.x0.shape[0], 3))
        v0[top, 0] = 0.02 * (i + 1)
        c["v0"] = v0.reshape(-1)
        rng = np.random.default_rng(seed + i)
        c["fext"] = 0.01 * rng.normal(size=c.x0.size)
    return cfgs, mg.concat_instances(cfgs)


@pytest.mark.parametrize("lumping", [False, True])
@pytest.mark.parametrize("scheme,steps", SCHEMES)
def test_batch_other_schemes_per_instance(scheme, steps, lumping):
    """The static, explicit, modified explicit and Runge-Kutta 4 schemes, their linear variants and linear
    implicit Euler on batched instances, each instance with its own time step, Rayleigh damping and a
    constant external force: every instance against the oracle stepped alone with that scheme
    (Math/OdeSolver*.cpp). The Runge-Kutta right-hand side F / dt and the explicit updates take each
    instance's dt."""
    kind = oracle.TET_LINEAR if scheme in LINEAR_SCHEMES else oracle.TET_COROTATIONAL
    cfgs, b = soft_batch(4, seed=29, kind=kind)
    say("batched scheme", scheme, "lumping", lumping, "dts", list(b.dt))
    g = femgpu.FemGpu(0)
    g.upload_mesh(b.kind, b.x0, b.conn, b.E, b.nu, b.rho)
    g.add_bc_nodes(b.bc_nodes)
    g.set_options(dt=1e-3, cg_tolerance=1e-12, scheme=scheme, rayleigh_mass=AM, rayleigh_stiffness=AK,
                  mass_lumping=int(lumping))
    g.set_external_force(b.fext)
    g.set_state(v=b.v0)
    g.set_instances(b.node_offsets, b.dt)
    for _ in range(steps):
        assert g.step(1) == femgpu.FEM_OK
    x, v = g.state()
    g.instance_stats()
    assert g.last_instance_active.tolist() == [1] * len(cfgs)
    g.close()
    off = b.node_offsets
    for i, c in enumerate(cfgs):
        s = oracle.Sim(c.kind, c.x0, c.conn, c.E, c.nu, c.rho)
        for n in c.bc_nodes:
            s.add_bc_node(n)
        s.set_solver(scheme=scheme, tol=1e-12)
        s.set_damping(AM, AK, lumping)
        s.set_external_force(c.fext)
        s.set_state(v=c.v0)
        for _ in range(steps):
            assert s.step(c.dt, 1) == 0
        xo, vo = s.state()
        sl = slice(3 * off[i], 3 * off[i + 1])
        x0 = c.x0.reshape(-1)
        e = dict(u=rel(x[sl] - x0, xo - x0), v=rel(v[sl], vo) if np.linalg.norm(vo) > 0 else np.linalg.norm(v[sl]))
        say("  instance", i, "dt", c.dt, e)
        assert e["u"] < 1e-9 and e["v"] < 1e-9


def test_batch_linear_scheme_keeps_operator_until_instances_change():
    """Linear implicit Euler on batched instances keeps the first step's operator; declaring the instances
    again (new time steps) forms it anew, so the next steps match the oracle at the new time steps."""
    cfgs, b = soft_batch(3, seed=41, kind=oracle.TET_LINEAR)
    g = femgpu.FemGpu(0)
    g.upload_mesh(b.kind, b.x0, b.conn, b.E, b.nu, b.rho)
    g.add_bc_nodes(b.bc_nodes)
    g.set_options(dt=1e-3, cg_tolerance=1e-12, scheme=femgpu.LINEAR_EULER_IMPLICIT)
    g.set_state(v=b.v0)
    g.set_instances(b.node_offsets, b.dt)
    assert g.step(3) == femgpu.FEM_OK
    dt2 = b.dt * 2.0
    g.set_instances(b.node_offsets, dt2)
    assert g.step(3) == femgpu.FEM_OK
    x, v = g.state()
    g.close()
    off = b.node_offsets
    for i, c in enumerate(cfgs):
        s = oracle.Sim(c.kind, c.x0, c.conn, c.E, c.nu, c.rho)
        for n in c.bc_nodes:
            s.add_bc_node(n)
        s.set_solver(scheme=femgpu.LINEAR_EULER_IMPLICIT, tol=1e-12)
        s.set_state(v=c.v0)
        s.step(c.dt, 3)
        # the oracle's linear solver keeps its matrix; a fresh one at the new dt continues from the state
        xo, vo = s.state()
        s2 = oracle.Sim(c.kind, c.x0, c.conn, c.E, c.nu, c.rho)
        for n in c.bc_nodes:
            s2.add_bc_node(n)
        s2.set_solver(scheme=femgpu.LINEAR_EULER_IMPLICIT, tol=1e-12)
        s2.set_state(x=xo, v=vo)
        s2.step(dt2[i], 3)
        xo, vo = s2.state()
        sl = slice(3 * off[i], 3 * off[i + 1])
        x0 = c.x0.reshape(-1)
        e = dict(u=rel(x[sl] - x0, xo - x0), v=rel(v[sl], vo))
        say("  instance", i, "dt", c.dt, "then", dt2[i], e)
        assert e["u"] < 1e-9 and e["v"] < 1e-9


def test_fp32_batch_switches_inadequate_instances_to_fp64():
    """In the FP32 + refinement mode an instance whose FP32 inner solve takes more than a quarter of its
    iteration cap (or whose refinement stops short of the tolerance) is solved in FP64 from then on; the
    others stay FP32. Here the cap is 1.5 times the FP64 iteration count of a stiff instance, so its FP32
    inner solve (about half of that per sweep) passes a quarter of the cap, while a soft instance needs far
    fewer iterations. Both instances stay within the FP32 bound of the FP64 batch, every step converges, and
    after the switch the stiff instance takes the FP64 run's iteration counts."""
    cfgs = mg.config_batched(1, seed=7, lo=24, hi=24) + mg.config_batched(1, seed=8, lo=6, hi=6)
    cfgs[0]["E"] = np.full(cfgs[0].conn.shape[0], 1e6)
    cfgs[0]["dt"] = 2e-3
    cfgs[1]["E"] = np.full(cfgs[1].conn.shape[0], 3e3)
    cfgs[1]["dt"] = 5e-4
    b = mg.concat_instances(cfgs)

    def run(precision, cap, steps):
        g = femgpu.FemGpu(0)
        g.upload_mesh(b.kind, b.x0, b.conn, b.E, b.nu, b.rho)
        g.add_bc_nodes(b.bc_nodes)
        g.set_options(dt=1e-3, cg_tolerance=1e-8, precision=precision, cg_max_iterations=cap)
        g.set_instances(b.node_offsets, b.dt)
        its, flags = [], []
        for _ in range(steps):
            assert g.step(1) == femgpu.FEM_OK
            it, _, conv = g.instance_stats()
            assert conv.all()
            its.append(it.copy())
            flags.append(g.last_instance_fp64.copy())
        x, v = g.state()
        g.close()
        return x, v, np.array(its), np.array(flags)

    _, _, its64, _ = run(femgpu.FP64, 100000, 1)
    n64 = int(its64[0, 0])
    say("FP64 iterations of the stiff instance", n64, "of the soft one", int(its64[0, 1]))
    assert n64 >= 160 and its64[0, 1] < 64
    cap = int(1.5 * n64)
    steps = 4
    x64, v64, its64, flags64 = run(femgpu.FP64, cap, steps)
    x32, v32, its32, flags32 = run(femgpu.FP32_REFINED, cap, steps)
    say("  cap", cap, "FP64 its", its64.tolist(), "mixed its", its32.tolist(), "flags", flags32.tolist())
    assert not flags64.any()
    assert flags32[0].tolist() == [1, 0] and flags32[-1].tolist() == [1, 0]
    for k in range(1, steps):  # after the switch the stiff instance solves in FP64 from x = 0
        assert abs(int(its32[k, 0]) - int(its64[k, 0])) <= 1
    off = b.node_offsets
    x0 = b.x0.reshape(-1)
    for i in range(2):
        sl = slice(3 * off[i], 3 * off[i + 1])
        e = dict(u=rel(x32[sl] - x0[sl], x64[sl] - x0[sl]), v=rel(v32[sl], v64[sl]))
        say("  instance", i, e)
        assert e["u"] < 1e-6 and e["v"] < 1e-6
