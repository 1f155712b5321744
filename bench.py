"""Benchmark: implicit-Euler DOF-steps/s of the corotational-tet block on one MI355X, with every other
single-GPU configuration as a sub-record of the same JSON line.

Headline (one GPU): 40x40x40 cubes split into 5 corotational tetrahedra each (320,000 tets, 68,921 nodes,
206,763 DOF), E=5e4 Pa, nu=0.45, rho=1000, bottom face fixed, seeded tangential top-face velocities
N(0, 0.01 m/s), gravity, implicit Euler dt=1 ms, Jacobi-CG to 1e-8. A step is one implicit-Euler step (element
kernels, assembly, right-hand side, CG, update) over the whole mesh; inputs are resident in HBM when the timed
region starts. Both precisions run: FP64, and FP32 operator/CG inside FP64 iterative refinement. The headline is
the faster of the runs whose displacement and velocity agree with the FP64 run to <= 1e-6 relative (the FP64
run always qualifies). The FP64 path is checked against the CPU oracle over the baseline's first sample.

CPU baseline: the FP64 oracle (portable -O3 -march=x86-64-v3 build), one core, one warm-up step and 3 timed
samples (median), with the host's CPU model; plus the faithful variant with the reference's O(#BC x D) boundary
zeroing.

Sub-records ("configs"), each at its own configuration's step count: the linear 10^3 case with the CPU's LU
and first-step dense compliance, the 64-instance batch, the adjoint gradient (gradients/s and the 32-direction
finite-difference check) and the 128x64x128 hexahedral block; each with its dominant kernel's roofline.
--headline-only skips them; --workload runs one of them alone.

Prints one JSON line on stdout (everything else goes to stderr).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from opensurgsim_amd import femgpu  # noqa: E402
from opensurgsim_amd import meshgen as mg  # noqa: E402

METRIC = "implicit-Euler DOF-steps/s at <=1e-6 rel err vs FP64 oracle; adjoint grads/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def device_sync(g):
    """All work of a context is ordered on its stream (torch is not involved outside the adjoint workload)."""
    g.synchronize()


def make_gpu(cfg, precision):
    g = femgpu.FemGpu(0)
    g.upload_mesh(cfg.kind, cfg.x0, cfg.conn, cfg.E, cfg.nu, cfg.rho)
    g.add_bc_nodes(cfg.bc_nodes)
    g.set_options(dt=cfg.dt, cg_tolerance=cfg.cg_tol, precision=precision, gravity=cfg.gravity)
    g.set_state(v=cfg.v0)
    return g


def phase_pass(g, steps, timer_every):
    """Per-phase event times from `steps` further steps, one in `timer_every` event-timed, after the timed region
    (the events' own cost stays out of the measured steps)."""
    if not timer_every:
        return {}
    g.enable_phase_timers(timer_every)
    g.step(steps)
    device_sync(g)
    phases = g.phase_times()
    g.enable_phase_timers(0)
    return phases


def run_timed(cfg, precision, warmup, steps, timer_every):
    g = make_gpu(cfg, precision)
    if warmup:
        g.step(warmup)
    device_sync(g)
    it0 = g.stats()["total_cg_iterations"]
    device_sync(g)
    t0 = time.perf_counter()
    rc = g.step(steps)
    device_sync(g)
    t1 = time.perf_counter()
    if rc != femgpu.FEM_OK:
        raise RuntimeError("GPU state became non-finite")
    st = g.stats()
    x, v = g.state()
    phases = phase_pass(g, steps, timer_every)
    op_ms = g.time_operator(100)
    g.close()
    return dict(seconds=t1 - t0, cg_iters=st["total_cg_iterations"] - it0, stats=st, phases=phases, x=x, v=v,
                op_ms=op_ms)


FUSED_SUMS, CG_BLOCKS = 7, 512


def cg_kernel_name(w):
    """The dominant kernel of the CG iteration: the library runs one fused kernel per iteration unless
    FEM_CG_KERNELS=split selects the operator + vector kernel pair (its operator kernel is then timed)."""
    t = "float" if w == 4 else "double"
    split = os.environ.get("FEM_CG_KERNELS", "") == "split"
    return ("cg_apply_kernel<%s>" if split else "cg_fused_kernel<%s>") % t


def spmv_bytes(cfg_sizes, w):
    """Algorithmic bytes of one launch of the dominant CG kernel.
    cg_fused_kernel: BSR values and indices once; for its own rows it reads r, q, z, w = B q, p and x
    (6 vectors) and the preconditioner B (FP64: the diagonal inverse, 1 value per DOF; FP32: the 3x3
    block-diagonal inverse, 9 values per node = 3 per DOF), and writes r, z, p, q, w and x (6 vectors);
    the gathers of z, w, p at the columns read the same vectors; 7 block sums read and written.
    cg_apply_kernel: BSR values and indices once, z and p_old read (p = z + beta p_old is formed on the
    fly), p_new and q written."""
    N, D, Bk = cfg_sizes
    Z = 9 * Bk
    if cg_kernel_name(w).startswith("cg_apply"):
        return Z * w + Bk * 4 + (N + 1) * 4 + 4 * D * w
    precond = 3 * D if w == 4 else D
    return Z * w + Bk * 4 + (N + 1) * 4 + (12 * D + precond) * w + 2 * FUSED_SUMS * CG_BLOCKS * 8


def phase_summary(run):
    """Average microseconds per launch and launch count of every phase of the event-timed steps."""
    return {k: [round(1e3 * v[0] / v[1], 2), v[1]] for k, v in run["phases"].items() if v[1]}


def copy_probe(mib, reps=3):
    """A plain device copy of known size (probe_copy_kernel), for scaling the FETCH_SIZE / WRITE_SIZE
    counters and for the achievable copy bandwidth."""
    nbytes = mib * (1 << 20)
    ms = femgpu.probe_copy(0, nbytes, reps)
    gbs = 2 * nbytes / (ms * 1e-3) / 1e9
    log("copy probe: %d MiB, %.3f ms per copy, %.0f GB/s (read + write)" % (mib, ms, gbs))
    return gbs


def cpu_model():
    """The host CPU's model string and logical core count (the CPU baseline's hardware)."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": name, "logical_cores": os.cpu_count()}


ORACLE_BUILD = ("-O3 -march=x86-64-v3 (AVX2, FMA), the portable prebuilt oracle: the image's gcc 11.4 predates the "
                "GPU hosts' CPU, and its -march=native build crashed there")


def oracle_sim(cfg, scheme=None, linsolver=None, tol=None, faithful_bc=False):
    import oracle
    s = oracle.Sim(cfg.kind, cfg.x0, cfg.conn, cfg.E, cfg.nu, cfg.rho)
    for n in cfg.bc_nodes:
        s.add_bc_node(n)
    s.set_gravity(cfg.gravity, True)
    if cfg.fext is not None:
        s.set_external_force(cfg.fext)
    s.set_solver(scheme=oracle.EULER_IMPLICIT if scheme is None else scheme,
                 linsolver=oracle.CG if linsolver is None else linsolver,
                 tol=cfg.cg_tol if tol is None else tol, faithful_bc=faithful_bc)
    s.set_state(v=cfg.v0)
    return s


def cpu_baseline(cfg, sample_steps, runs=3):
    """Oracle (FP64 C++ restatement, one core) on the same workload: one warm-up step, then
    `runs` timed samples of `sample_steps` steps each; the median is the baseline. Also checks the GPU FP64
    path against the oracle state after the first 1 + sample_steps steps."""
    s = oracle_sim(cfg)
    s.step(cfg.dt, 1)  # warm-up step
    times = []
    ref = None
    for r in range(runs):
        t0 = time.perf_counter()
        s.step(cfg.dt, sample_steps)
        times.append(time.perf_counter() - t0)
        if r == 0:
            ref = s.state()
    xr, vr = ref
    g = make_gpu(cfg, femgpu.FP64)
    g.step(1 + sample_steps)
    x, v = g.state()
    g.close()
    x0 = cfg.x0.reshape(-1)
    err = dict(x=float(np.linalg.norm(x - xr) / np.linalg.norm(xr)),
               u=float(np.linalg.norm((x - x0) - (xr - x0)) / np.linalg.norm(xr - x0)),
               v=float(np.linalg.norm(v - vr) / np.linalg.norm(vr)))
    t = float(np.median(times))
    return x0.size * sample_steps / t, times, err


def cpu_faithful(cfg, steps):
    """The oracle with the reference's own boundary-condition zeroing (zeroRow / zeroColumn through coeff()
    look-ups of every row and column, O(#BC x D), Math/SparseMatrix.h:352-377), one core, after one warm-up step:
    the faithful CPU variant, to set beside the restated-efficient one."""
    s = oracle_sim(cfg, faithful_bc=True)
    s.step(cfg.dt, 1)
    t0 = time.perf_counter()
    s.step(cfg.dt, steps)
    t = time.perf_counter() - t0
    return cfg.x0.size * steps / t, t


def bench_hex(args):
    """Hexahedral block (Fem3DElementCube), 128 x 64 x 128 cubes by default: FP32 element-by-element CG inside
    FP64 refinement to 1e-6 with one cached 24x24 element matrix per element (the headline form), the same with
    the table of distinct cached matrices a structured mesh allows (identical results), and the FP64 matrix-free
    path at 1e-10 as the reference for the errors."""
    n = (args.hex_n, args.hex_n // 2, args.hex_n)
    cfg = mg.config_hex(n)
    N = cfg.x0.shape[0]
    D = 3 * N
    E = cfg.conn.shape[0]
    steps = args.steps if args.steps_given else cfg.steps
    log("workload: hexahedra %s, %d elements, %d nodes, %d DOF, %d steps" % (n, E, N, D, steps))

    def run(precision, tol, table=True):
        saved = os.environ.get("FEM_HEX_TABLE")
        if not table:
            os.environ["FEM_HEX_TABLE"] = "0"
        try:
            g = femgpu.FemGpu(0)
            t0 = time.perf_counter()
            g.upload_mesh(cfg.kind, cfg.x0, cfg.conn, cfg.E, cfg.nu, cfg.rho)
            g.add_bc_nodes(cfg.bc_nodes)
            g.set_options(dt=cfg.dt, cg_tolerance=tol, precision=precision, gravity=cfg.gravity)
            g.set_external_force(cfg.fext)
            log("  set-up %.1f s" % (time.perf_counter() - t0))
            if args.warmup:
                g.step(args.warmup)
            device_sync(g)
            it0 = g.stats()["total_cg_iterations"]
            t0 = time.perf_counter()
            rc = g.step(steps)
            device_sync(g)
            t = time.perf_counter() - t0
            assert rc == femgpu.FEM_OK
            st = g.stats()
            x, v = g.state()
            ph = phase_pass(g, steps, args.timer_every)
            distinct = g.hex_distinct_matrices()
            op_ms = g.time_operator(20)
            g.close()
        finally:
            if saved is None:
                os.environ.pop("FEM_HEX_TABLE", None)
            else:
                os.environ["FEM_HEX_TABLE"] = saved
        return dict(seconds=t, x=x, v=v, cg=st["total_cg_iterations"] - it0, stats=st, phases=ph, distinct=distinct,
                    op_ms=op_ms)

    r32 = run(femgpu.FP32_REFINED, 1e-6, table=False)
    log("fp32+refine, per-element cache: %.3f s, %d CG iterations, last sweeps %d, operator %.3f ms" % (
        r32["seconds"], r32["cg"], r32["stats"]["last_refinements"], r32["op_ms"]))
    rt = run(femgpu.FP32_REFINED, 1e-6, table=True)
    log("fp32+refine, distinct-matrix table (%d): %.3f s, operator %.3f ms" % (rt["distinct"], rt["seconds"],
                                                                               rt["op_ms"]))
    r64 = run(femgpu.FP64, 1e-10)
    log("fp64 (1e-10): %.3f s, %d CG iterations, operator %.3f ms" % (r64["seconds"], r64["cg"], r64["op_ms"]))
    x0 = cfg.x0.reshape(-1)

    def err(r):
        return dict(u=float(np.linalg.norm(r["x"] - r64["x"]) / np.linalg.norm(r64["x"] - x0)),
                    v=float(np.linalg.norm(r["v"] - r64["v"]) / np.linalg.norm(r64["v"])))
    e32, et = err(r32), err(rt)
    log("fp32+refine vs fp64 rel err", e32, "table", et, "identical", bool(np.array_equal(r32["x"], rt["x"])))
    # one application of the FP32 operator (8 colours on a block mesh): per element the cached packed upper
    # triangle (300 floats, 1,200 B; table: a 4 B index), 32 B of node ids and its 4 B slot; per colour p read
    # and q read and written over the DOFs
    colours = 8
    cache_bytes = E * (1200 + 32 + 4) + colours * 12 * D
    table_bytes = E * (4 + 32 + 4) + colours * 12 * D
    return {"metric": METRIC, "value": D * steps / r32["seconds"], "unit": "DOF-steps/s", "n_gpus": 1,
            "steps": steps, "warmup": args.warmup, "ms_per_step": 1e3 * r32["seconds"] / steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64refine",
            "data": "synthetic",
            "config": {"workload": "hex-block %dx%dx%d Fem3DElementCube, EbE FP32 with one cached 24x24 Ke per "
                                   "element + FP64 refinement 1e-6, dt=1ms" % n, "elements": int(E), "nodes": int(N),
                       "dof": int(D), "parallelism": "single-gpu"},
            "rel_err_vs_fp64": e32,
            "roofline": roofline("hex_apply32_kernel<false> (all 8 colours)", cache_bytes, r32["op_ms"] / 1e3),
            "table_variant": {"value": D * steps / rt["seconds"], "ms_per_step": 1e3 * rt["seconds"] / steps,
                              "distinct_element_matrices": rt["distinct"], "rel_err_vs_fp64": et,
                              "identical_to_per_element_cache": bool(np.array_equal(r32["x"], rt["x"]) and
                                                                     np.array_equal(r32["v"], rt["v"])),
                              "roofline": roofline("hex_apply32_kernel<true> (all 8 colours)", table_bytes,
                                                   rt["op_ms"] / 1e3)},
            "fp64": {"value": D * steps / r64["seconds"], "ms_per_step": 1e3 * r64["seconds"] / steps,
                     "cg_iters_per_step": r64["cg"] / steps, "operator_ms": r64["op_ms"]},
            "fp32_refined": {"cg_iters_per_step": r32["cg"] / steps},
            "phases_fp32": {k: {"ms": round(v[0], 3), "launches": v[1]} for k, v in r32["phases"].items()}}


SCHEME_NAMES = {femgpu.EULER_EXPLICIT: "euler_explicit", femgpu.EULER_EXPLICIT_MODIFIED: "euler_explicit_modified",
                femgpu.RUNGE_KUTTA_4: "runge_kutta_4", femgpu.STATIC: "static",
                femgpu.LINEAR_EULER_EXPLICIT: "linear_euler_explicit",
                femgpu.LINEAR_RUNGE_KUTTA_4: "linear_runge_kutta_4", femgpu.LINEAR_STATIC: "linear_static"}


def bench_schemes(args):
    """The other integration schemes on the 40^3 corotational block: explicit Euler, modified explicit Euler,
    Runge-Kutta 4 (dt 0.1 ms, within their stability limit) and static, with their linear variants; Jacobi-CG
    to 1e-8 on M / dt (K for static), FP64. Rayleigh damping alpha_M = 1, alpha_K = 1e-4."""
    n = (args.n, args.n, args.n)
    cfg = mg.config_corotational(n)
    N = cfg.x0.shape[0]
    D = 3 * N
    log("workload: schemes on corotational tets %s, %d DOF" % (n, D))
    res = {}
    for sc, name in SCHEME_NAMES.items():
        static = sc in (femgpu.STATIC, femgpu.LINEAR_STATIC)
        dt = 1e-3 if static else 1e-4
        steps = 4 if static else 20
        g = femgpu.FemGpu(0)
        g.upload_mesh(cfg.kind, cfg.x0, cfg.conn, cfg.E, cfg.nu, cfg.rho)
        g.add_bc_nodes(cfg.bc_nodes)
        g.set_options(dt=dt, cg_tolerance=cfg.cg_tol, precision=femgpu.FP64, gravity=cfg.gravity, scheme=sc,
                      rayleigh_mass=1.0, rayleigh_stiffness=1e-4)
        g.set_state(v=cfg.v0)
        g.step(2)
        device_sync(g)
        it0 = g.stats()["total_cg_iterations"]
        t0 = time.perf_counter()
        rc = g.step(steps)
        device_sync(g)
        t = time.perf_counter() - t0
        assert rc == femgpu.FEM_OK, (name, rc)
        it = g.stats()["total_cg_iterations"] - it0
        g.close()
        res[name] = {"dof_steps_per_s": D * steps / t, "ms_per_step": 1e3 * t / steps, "dt": dt,
                     "cg_iters_per_step": it / steps}
        log("  %-26s %8.3f ms/step, %7.1f CG iterations/step" % (name, 1e3 * t / steps, it / steps))
    main_name = "runge_kutta_4"
    out = {"metric": METRIC, "value": res[main_name]["dof_steps_per_s"], "unit": "DOF-steps/s", "n_gpus": 1,
           "steps": 20, "warmup": 2, "ms_per_step": res[main_name]["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "corotational-tet-block %dx%dx%d, Runge-Kutta 4 dt=0.1ms (other schemes below), "
                                  "Rayleigh 1 / 1e-4, Jacobi-CG 1e-8" % n, "dof": int(D),
                      "parallelism": "single-gpu"},
           "schemes": res}
    return out


def bench_compliance(args):
    """Compliance solves on the 40^3 corotational block after warm-up steps: C H^T for 8 fixed-point
    constraints (24 columns, the getComplianceMatrix() * H^T of Physics/FemConstraintFixedPoint.cpp:107) through
    fem_apply_compliance, FP64 and FP32 + refinement; checked by the residual of the system matrix."""
    import scipy.sparse as sps
    import torch  # device buffers for the device-pointer variant (started in main before the library's runtime)
    n = (args.n, args.n, args.n)
    cfg = mg.config_corotational(n)
    N = cfg.x0.shape[0]
    D = 3 * N
    ncon = 8
    rng = np.random.default_rng(3)
    elems = rng.choice(cfg.conn.shape[0], ncon, replace=False)
    Ht = np.zeros((D, 3 * ncon))
    coords = []
    for i, e in enumerate(elems):
        coord = rng.dirichlet(np.ones(4))
        coords.append(coord)
        for a in range(4):
            for k in range(3):
                Ht[3 * cfg.conn[e, a] + k, 3 * i + k] = coord[a]
    log("workload: compliance C H^T, %d columns, corotational tets %s, %d DOF" % (Ht.shape[1], n, D))
    res = {}
    for prec, name in ((femgpu.FP64, "fp64"), (femgpu.FP32_REFINED, "fp32_refined")):
        g = make_gpu(cfg, prec)
        g.step(args.warmup)
        X = g.apply_compliance(Ht)  # warm
        device_sync(g)
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            X = g.apply_compliance(Ht)
        t_host = (time.perf_counter() - t0) / reps
        bd = torch.from_numpy(np.ascontiguousarray(Ht.T)).cuda()
        xd = torch.empty_like(bd)
        torch.cuda.synchronize()
        g.apply_compliance_device(bd.data_ptr(), Ht.shape[1], xd.data_ptr())
        device_sync(g)
        t0 = time.perf_counter()
        for _ in range(reps):
            g.apply_compliance_device(bd.data_ptr(), Ht.shape[1], xd.data_ptr())
        device_sync(g)
        t = (time.perf_counter() - t0) / reps
        dev_diff = float(np.abs(xd.cpu().numpy().T - X).max() / np.abs(X).max())
        # the same constraints one by one through fem_constraint_compliance (H^T built on the device; C H^T and
        # H C H^T returned, or H C H^T only)
        for want in (True, False):
            g.constraint_compliance(elems[0], coords[0], np.eye(3), 1.0, want_cht=want)
        t_con = {}
        hdiff = 0.0
        for want in (True, False):
            t0 = time.perf_counter()
            for _ in range(reps):
                for i, e in enumerate(elems):
                    g.constraint_compliance(e, coords[i], np.eye(3), 1.0, want_cht=want)
            t_con[want] = (time.perf_counter() - t0) / (reps * ncon)
        for i, e in enumerate(elems):
            cht, _ = g.constraint_compliance(e, coords[i], np.eye(3), 1.0)
            hdiff = max(hdiff, float(np.abs(cht - X[:, 3 * i:3 * i + 3]).max() / np.abs(X).max()))
        rowptr, colidx = g.pattern()
        A = sps.csr_matrix((g.system_matrix(), colidx, rowptr), shape=(D, D))
        bc = (3 * np.asarray(cfg.bc_nodes)[:, None] + np.arange(3)).reshape(-1)
        B = Ht.copy()
        B[bc] = 0.0
        resid = float(np.linalg.norm(A @ X - B) / np.linalg.norm(B))
        g.close()
        res[name] = {"ms_per_column": 1e3 * t / Ht.shape[1], "columns_per_s": Ht.shape[1] / t,
                     "ms_per_column_host_pointers": 1e3 * t_host / Ht.shape[1], "relative_residual": resid,
                     "device_vs_host_max_rel_diff": dev_diff,
                     "ms_per_fixed_point_constraint": 1e3 * t_con[True],
                     "ms_per_fixed_point_constraint_hcht_only": 1e3 * t_con[False],
                     "constraint_vs_dense_max_rel_diff": hdiff}
        log("  %s: %.3f ms per column (device pointers), %.3f (host pointers), residual %.2e; "
            "fem_constraint_compliance %.3f ms per fixed point (%.3f with H C H^T only)" %
            (name, 1e3 * t / Ht.shape[1], 1e3 * t_host / Ht.shape[1], resid, 1e3 * t_con[True], 1e3 * t_con[False]))
    best = min(res, key=lambda k: res[k]["ms_per_column"])
    out = {"metric": "compliance columns/s (C H^T, applyCompliance)", "value": res[best]["columns_per_s"],
           "unit": "columns/s", "n_gpus": 1, "steps": 3, "warmup": args.warmup,
           "ms_per_step": res[best]["ms_per_column"] * Ht.shape[1], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64refine" if best == "fp32_refined" else "f64", "data": "synthetic",
           "config": {"workload": "C H^T of %d fixed-point constraints (%d columns) on corotational-tet-block "
                                  "%dx%dx%d after %d implicit steps, CG 1e-8" % ((ncon, 3 * ncon) + n + (args.warmup,)),
                      "dof": int(D), "parallelism": "single-gpu"},
           "precisions": res}
    return out


def bench_adjoint(args):
    """Adjoint material identification workload: 20x20x20 corotational tet block (40,000 tets), per-cube
    log-normal ground-truth E, top pressure 200 Pa, 50 implicit steps at dt = 2 ms; loss = squared
    displacement mismatch at 500 seeded surface nodes; full dL/dE (all element moduli) by the adjoint over
    the whole state history kept in HBM (22 MB; --adj-checkpoint k recomputes k-step segments from
    checkpoints instead), at a uniform guess E = 3e4 Pa; checked against central differences along
    32 seeded unit directions."""
    log("importing torch (the adjoint harness is a torch.autograd.Function)")
    import torch
    from opensurgsim_amd import adjoint
    torch.cuda.synchronize()
    log("torch ready")
    n = (args.adj_n, args.adj_n, args.adj_n)
    cfg = mg.config_adjoint(n)
    steps = cfg.steps
    log("workload: adjoint, corotational tets %s, %d elements, %d DOF, %d steps" % (n, cfg.conn.shape[0],
                                                                                  cfg.x0.size, steps))
    g = femgpu.FemGpu(0)
    g.upload_mesh(cfg.kind, cfg.x0, cfg.conn, cfg.E, cfg.nu, cfg.rho)
    g.add_bc_nodes(cfg.bc_nodes)
    g.set_options(dt=cfg.dt, cg_tolerance=args.adj_tol)
    g.set_external_force(cfg.fext)
    r = adjoint.Rollout(g, cfg.x0, np.zeros(cfg.x0.size), steps, checkpoint_every=args.adj_checkpoint or None)
    keep = "whole state history in HBM" if r.full_history else "checkpoints every %d" % r.k
    xr = torch.as_tensor(cfg.x0.reshape(-1), device="cuda")
    obs = torch.as_tensor((3 * cfg.obs_nodes[:, None] + np.arange(3)).reshape(-1), device="cuda")
    with torch.no_grad():
        xt, _ = r.forward(cfg.E)
    u_obs = (xt - xr)[obs].clone()
    E0 = np.full(cfg.E.size, 3e4)

    def loss_and_grad(E):
        x, _ = r.forward(E)
        d = (x - xr)[obs] - u_obs
        lx = torch.zeros_like(x)
        lx[obs] = 2 * d
        gE, _, _ = r.backward(lx, None)
        return float(torch.sum(d * d)), gE

    def loss(E):
        x, _ = r.forward(E)
        d = (x - xr)[obs] - u_obs
        return float(torch.sum(d * d))

    L0, gE = loss_and_grad(E0)  # warm-up
    device_sync(g)
    reps = args.steps if args.steps_given else 5
    t0 = time.perf_counter()
    for _ in range(reps):
        L0, gE = loss_and_grad(E0)
    device_sync(g)
    t = time.perf_counter() - t0
    gE = gE.cpu().numpy()
    log("adjoint: %d gradients in %.3f s (%.1f ms each), loss %.6e, |dL/dE| %.3e" % (reps, t, 1e3 * t / reps, L0,
                                                                                   np.linalg.norm(gE)))
    # central differences along 32 seeded unit directions (FP64)
    dirs = mg.normal(32 * E0.size, 4242, stream=11).reshape(32, E0.size)
    h = args.adj_fd_h * E0.mean()
    errs = []
    for k in range(args.adj_fd_dirs):
        d = dirs[k] / np.linalg.norm(dirs[k])
        fd = (loss(E0 + h * d) - loss(E0 - h * d)) / (2 * h)
        ad = float(gE @ d)
        errs.append(abs(ad - fd) / abs(fd))
    log("fd check: %d directions, max rel err %.3e, median %.3e" % (len(errs), max(errs), float(np.median(errs))))
    ident = None
    if args.adj_ident:
        Ehat, losses = adjoint.recover_young_modulus(r, xr, obs, u_obs, E0, iterations=args.adj_ident)
        ident = {"lbfgs_iterations": args.adj_ident, "evaluations": len(losses), "loss_first": losses[0],
                 "loss_best": min(losses),
                 "rel_err_logE_first": float(np.linalg.norm(np.log(E0) - np.log(cfg.E)) / np.linalg.norm(np.log(cfg.E))),
                 "rel_err_logE_last": float(np.linalg.norm(np.log(Ehat.cpu().numpy()) - np.log(cfg.E)) /
                                            np.linalg.norm(np.log(cfg.E)))}
        log("identification:", ident)
    st = g.stats()
    op_ms = g.time_operator(100)
    N = cfg.x0.shape[0]
    Bk = g.num_nonzeros() // 9
    roof = roofline("cg_fused_kernel<double>", spmv_bytes((N, 3 * N, Bk), 8), op_ms / 1e3,
                    "the forward and transpose solves' CG iteration; operator and vectors (11.5 MB) stay in the MALL")
    out = {"metric": METRIC, "value": reps / t, "unit": "gradients/s", "n_gpus": 1, "steps": reps,
           "roofline": roof,
           "warmup": 1, "ms_per_step": 1e3 * t / reps, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "adjoint dL/dE, corotational tet block %dx%dx%d, %d implicit steps dt=2ms, "
                                  "%s, CG %.0e" % (n + (steps, keep, args.adj_tol)),
                      "elements": int(cfg.conn.shape[0]), "dof": int(cfg.x0.size), "observed_nodes":
                      int(cfg.obs_nodes.size), "parallelism": "single-gpu"},
           "fd_check": {"directions": len(errs), "step_rel": args.adj_fd_h, "max_rel_err": max(errs),
                        "median_rel_err": float(np.median(errs))},
           "identification": ident, "cg_iterations_total": st["total_cg_iterations"]}
    g.close()
    return out


def bench_batch(args):
    """64 independent corotational tet blocks with cube counts ~ U{6..48} per axis (6.3M tets, 4.2M DOF in
    total), E ~ logU[1e3, 1e6], nu ~ U[0.30, 0.49], dt per instance in {0.5, 1, 2} ms, one random fixed face
    each, gravity; all advanced together (per-instance CG to 1e-8), FP32 + FP64 refinement and FP64."""
    cfgs = mg.config_batched(args.batch_count)
    b = mg.concat_instances(cfgs)
    N = b.x0.shape[0]
    D = 3 * N
    steps = args.steps if args.steps_given else 100
    log("workload: batch of %d instances, %d tets, %d nodes, %d DOF, %d steps" % (len(cfgs), b.conn.shape[0], N,
                                                                                  D, steps))

    def run(precision, tuning=None):
        g = femgpu.FemGpu(0)
        t0 = time.perf_counter()
        g.upload_mesh(b.kind, b.x0, b.conn, b.E, b.nu, b.rho)
        g.add_bc_nodes(b.bc_nodes)
        g.set_options(dt=1e-3, cg_tolerance=1e-8, precision=precision, cg_max_iterations=args.batch_maxit)
        g.set_state(v=b.v0)
        g.set_instances(b.node_offsets, b.dt)
        if tuning:
            g.set_batch_tuning(*tuning)
        nnz = g.num_nonzeros()
        log("  set-up %.1f s" % (time.perf_counter() - t0))
        if args.warmup:
            g.step(args.warmup)
        device_sync(g)
        log("  warm-up done (%d steps), tuning %s" % (args.warmup, g.batch_tuning()))
        it0 = g.stats()["total_cg_iterations"]
        inst_dof = 3 * np.diff(b.node_offsets)
        dof_steps = 0
        t0 = time.perf_counter()
        done = 0
        while done < steps:  # blocks of 10 steps; DOF-steps count the instances still active after a block
            k = min(10, steps - done)
            rc = g.step(k)
            g.instance_stats()
            dof_steps += int(inst_dof[g.last_instance_active > 0].sum()) * k
            done += k
        device_sync(g)
        t = time.perf_counter() - t0
        assert rc == femgpu.FEM_OK
        st = g.stats()
        x, v = g.state()
        tun = g.batch_tuning()
        its, errs, conv = g.instance_stats()
        active = g.last_instance_active.copy()
        unconv = g.last_instance_unconverged_steps.copy()
        g.close()
        log("  %d of %d instances active at the end" % (int(active.sum()), active.size))
        return dict(seconds=t, x=x, v=v, cg=st["total_cg_iterations"] - it0, tuning=tun, last_iters=its.tolist(),
                    converged=bool(conv.all()), active=active, dof_steps=dof_steps, unconv=unconv, nnz=nnz)

    r32 = run(femgpu.FP32_REFINED)
    log("fp32+refine: %.3f s (%.2f ms/step), %d instance-iterations, tuning %s" % (
        r32["seconds"], 1e3 * r32["seconds"] / steps, r32["cg"], r32["tuning"]))
    r64 = run(femgpu.FP64)
    log("fp64: %.3f s (%.2f ms/step), %d instance-iterations, tuning %s" % (
        r64["seconds"], 1e3 * r64["seconds"] / steps, r64["cg"], r64["tuning"]))
    # FP64 again with another chunk shape (another summation order): instances whose FP64 answer moves
    # under a change of rounding order are physically unstable over these 100 steps (large deformation,
    # inverted elements); no precision can be compared on them
    alt = (512, 16, True) if r64["tuning"]["chunk_rows"] != 512 else (128, 8, True)
    r64b = run(femgpu.FP64, alt)
    log("fp64, chunk shape %s: %.3f s" % (alt, r64b["seconds"]))
    x0 = b.x0.reshape(-1)
    off = b.node_offsets
    worst = {"u": 0.0, "v": 0.0}
    # the gate covers instances whose solves met their tolerance at every step in both runs; an instance
    # whose CG stops at the iteration cap (large deformation, inverted elements) has no converged answer to
    # compare, in the reference either
    sens = np.zeros(len(cfgs))
    umax = np.zeros(len(cfgs))
    for i in range(len(cfgs)):
        sl = slice(3 * off[i], 3 * off[i + 1])
        sens[i] = max(np.linalg.norm(r64b["x"][sl] - r64["x"][sl]) / max(np.linalg.norm(r64["x"][sl] - x0[sl]), 1e-300),
                      np.linalg.norm(r64b["v"][sl] - r64["v"][sl]) / max(np.linalg.norm(r64["v"][sl]), 1e-300))
        umax[i] = np.abs(r64["x"][sl] - x0[sl]).max()
    unconverged = np.flatnonzero((r32["unconv"] > 0) | (r64["unconv"] > 0) | (r32["active"] == 0) |
                                 (r64["active"] == 0) | (sens > 1e-9))
    both = np.setdiff1d(np.arange(len(cfgs)), unconverged)
    log("excluded instances (id, fp64 order sensitivity, max |u| m, capped steps fp64/fp32):",
        [(int(i), float("%.2e" % sens[i]), float("%.3g" % umax[i]), int(r64["unconv"][i]), int(r32["unconv"][i]))
         for i in unconverged])
    for i in both:
        sl = slice(3 * off[i], 3 * off[i + 1])
        du = np.linalg.norm(r32["x"][sl] - r64["x"][sl]) / max(np.linalg.norm(r64["x"][sl] - x0[sl]), 1e-300)
        dv = np.linalg.norm(r32["v"][sl] - r64["v"][sl]) / max(np.linalg.norm(r64["v"][sl]), 1e-300)
        worst["u"] = max(worst["u"], float(du))
        worst["v"] = max(worst["v"], float(dv))
    log("fp32+refine vs fp64, worst instance:", worst)
    use32 = worst["u"] <= 1e-6 and worst["v"] <= 1e-6 and r32["seconds"] < r64["seconds"]
    head = r32 if use32 else r64
    # one batched CG iteration over every instance (operator kernel: BSR values and indices, z and p_old
    # gathered, p and q written; vector kernel: x and r read and written, p and q read, z written, the
    # preconditioner: FP32 3x3 blocks, 3 values per DOF, FP64 diagonal, 1)
    w = 4 if use32 else 8
    Bk = head["nnz"] // 9
    it_bytes = 9 * Bk * w + 4 * Bk + 4 * (N + len(cfgs)) + (11 * D + (3 * D if use32 else D)) * w
    out = {"metric": METRIC, "value": head["dof_steps"] / head["seconds"], "unit": "DOF-steps/s", "n_gpus": 1,
           "iteration_bytes": it_bytes,
           "steps": steps, "warmup": args.warmup, "ms_per_step": 1e3 * head["seconds"] / steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64refine" if use32 else "f64", "data": "synthetic",
           "config": {"workload": "batch of %d corotational tet blocks (U{6..48}^3 cubes), per-instance dt and CG "
                                  "1e-8" % len(cfgs), "instances": len(cfgs), "elements": int(b.conn.shape[0]),
                      "nodes": int(N), "dof": int(D), "parallelism": "single-gpu"},
           "rel_err_fp32_vs_fp64_worst_instance": worst,
           "instances_active_at_end": {"fp32_refined": int(r32["active"].sum()), "fp64": int(r64["active"].sum())},
           "instances_excluded_from_gate": {"ids": unconverged.tolist(),
                                            "reason": "FP64 answer moves by > 1e-9 under another summation order "
                                                      "(unstable dynamics), CG stopped at its cap, or deactivated",
                                            "fp64_order_sensitivity": [float(sens[i]) for i in unconverged],
                                            "max_displacement_m": [float(umax[i]) for i in unconverged],
                                            "capped_steps_fp64": r64["unconv"][unconverged].tolist(),
                                            "capped_steps_fp32": r32["unconv"][unconverged].tolist()},
           "instances_in_gate": int(both.size),
           "cg_max_iterations": args.batch_maxit,
           "fp32_refined": {"value": r32["dof_steps"] / r32["seconds"], "ms_per_step": 1e3 * r32["seconds"] / steps,
                            "tuning": r32["tuning"], "instance_iterations": r32["cg"]},
           "fp64": {"value": r64["dof_steps"] / r64["seconds"], "ms_per_step": 1e3 * r64["seconds"] / steps,
                    "tuning": r64["tuning"], "instance_iterations": r64["cg"]}}
    return out


def bench_corotational(args):
    """The headline: the 40^3 corotational block in FP64 and FP32 + refinement, the dominant kernel's roofline,
    and the CPU baseline (median of 3 samples, plus the faithful boundary-condition variant)."""
    n = (args.n, args.n, args.n)
    cfg = mg.config_corotational(n)
    N = cfg.x0.shape[0]
    D = 3 * N
    log("workload: corotational tets %s, %d elements, %d nodes, %d DOF" % (n, cfg.conn.shape[0], N, D))
    probe = make_gpu(cfg, femgpu.FP64)
    Bk = probe.num_nonzeros() // 9
    probe.close()

    r64 = run_timed(cfg, femgpu.FP64, args.warmup, args.steps, args.timer_every)
    log("fp64: %.3f s for %d steps, cg iters %d" % (r64["seconds"], args.steps, r64["cg_iters"]))
    r32 = run_timed(cfg, femgpu.FP32_REFINED, args.warmup, args.steps, args.timer_every)
    log("fp32+refine: %.3f s for %d steps, cg iters %d" % (r32["seconds"], args.steps, r32["cg_iters"]))
    x0 = cfg.x0.reshape(-1)
    err32 = dict(x=float(np.linalg.norm(r32["x"] - r64["x"]) / np.linalg.norm(r64["x"])),
                 u=float(np.linalg.norm(r32["x"] - r64["x"]) / np.linalg.norm(r64["x"] - x0)),
                 v=float(np.linalg.norm(r32["v"] - r64["v"]) / np.linalg.norm(r64["v"])))
    log("fp32 vs fp64 rel err", err32)
    # headline: the faster of the two runs that meet the 1e-6 gate (the FP64 run is the oracle-parity path)
    use32 = err32["u"] <= 1e-6 and err32["v"] <= 1e-6 and r32["seconds"] < r64["seconds"]
    head = r32 if use32 else r64
    value = D * args.steps / head["seconds"]

    w = 4 if use32 else 8
    # average launch of the dominant kernel: 100 back-to-back launches on the final state, each timed by
    # the start/stop events of its own dispatch (execution only, as a kernel trace reports it; the event
    # pairs inside the steps bracket a launch from the host side and are reported under "phases")
    sp_ms, sp_n = head["phases"]["spmv"]
    sp_avg_s = head["op_ms"] / 1e3
    bytes_per = spmv_bytes((N, D, Bk), w)
    achieved = bytes_per / sp_avg_s / 1e9
    # HBM bytes per dispatch from the FETCH_SIZE / WRITE_SIZE counter runs, scaled by a copy kernel of
    # known size (profiles/hbm_traffic_corotational_40x40x40.md); only valid for the 40^3 block
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "cg_traffic.json")
    if os.path.exists(pmc) and n == (40, 40, 40):
        try:
            with open(pmc) as fh:
                t = json.load(fh)
            traffic = t["bytes_per_dispatch"].get(cg_kernel_name(w))
        except (OSError, ValueError, KeyError):
            traffic = None

    cpu = None
    if not args.skip_cpu:
        flags = ORACLE_BUILD
        cps, times, err_or = cpu_baseline(cfg, args.cpu_steps)
        log("cpu oracle: %.1f DOF-steps/s (median of %s s over %d steps); gpu fp64 vs oracle %s" % (
            cps, ["%.2f" % t for t in times], args.cpu_steps, err_or))
        cpu = {"value": cps, "unit": "DOF-steps/s", "cores": 1, "kind": "port",
               "sample": "same workload: 1 warm-up step, then 3 timed samples of %d implicit steps; median; FP64 "
                         "C++ oracle on one host core (literal 12x12 corotational products, O(nnz) boundary "
                         "conditions)" % args.cpu_steps,
               "sample_seconds": times, "build": flags, "cpu": cpu_model(),
               "gpu_fp64_vs_oracle_rel_err": err_or}
        if args.cpu_faithful_steps:
            fps, ft = cpu_faithful(cfg, args.cpu_faithful_steps)
            log("cpu oracle with the reference's O(#BC x D) boundary zeroing: %.1f DOF-steps/s (%.1f s)" % (fps, ft))
            cpu["faithful_bc_variant"] = {"value": fps, "unit": "DOF-steps/s", "cores": 1,
                                          "sample": "%d step(s) after 1 warm-up step; the reference's zeroRow / "
                                                    "zeroColumn through coeff() look-ups (O(#BC x D))"
                                                    % args.cpu_faithful_steps}

    phases = {k: {"ms": round(v[0], 3), "launches": v[1]} for k, v in head["phases"].items()}
    copy_gbs = copy_probe(args.copy_probe or 512)
    return {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * head["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64refine" if use32 else "f64",
        "data": "synthetic",
        "config": {"workload": "corotational-tet-block %dx%dx%d (5 tets/cube), implicit Euler dt=1ms, "
                               "Jacobi-CG 1e-8" % n,
                   "elements": int(cfg.conn.shape[0]), "nodes": int(N), "dof": int(D), "nnz": int(9 * Bk),
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "kernel": cg_kernel_name(w), "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": bytes_per, "avg_launch_us": sp_avg_s * 1e6,
                     "avg_launch_us_in_step_events": 1e3 * sp_ms / max(sp_n, 1), "in_step_launches": sp_n,
                     "measured_copy_GBps": copy_gbs},
        "cpu_baseline": cpu,
        "fp64": {"value": D * args.steps / r64["seconds"], "ms_per_step": 1e3 * r64["seconds"] / args.steps,
                 "cg_iters_per_step": r64["cg_iters"] / args.steps, "phases_us_per_launch": phase_summary(r64)},
        "fp32_refined": {"value": D * args.steps / r32["seconds"], "ms_per_step": 1e3 * r32["seconds"] / args.steps,
                         "cg_iters_per_step": r32["cg_iters"] / args.steps, "rel_err_vs_fp64": err32,
                         "refinements_last_step": r32["stats"]["last_refinements"],
                         "phases_us_per_launch": phase_summary(r32)},
        "phases": phases,
    }


def roofline(kernel, nbytes, seconds, note=None):
    a = nbytes / seconds / 1e9
    out = {"bound": "hbm", "kernel": kernel, "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": a / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": nbytes, "avg_launch_us": seconds * 1e6}
    if note:
        out["note"] = note
    return out


def bench_linear_case(args):
    """The reference's CPU-runnable case: linear tets on 10 x 10 x 10 cubes (3,993 DOF), bottom fixed, gravity,
    OdeSolverLinearEulerImplicit with the direct solver, dt 1 ms, 100 steps, the dense compliance formed on the
    first step (Math/OdeSolverLinearEulerImplicit.cpp:43-64, Math/OdeSolver.cpp:69-76). GPU: the banded LU of the
    operator, factored on step 1, and the dense compliance from it, kept and applied on the later steps; CPU: the
    oracle's faithful linear
    implicit Euler with its banded LU and the same dense compliance. Both time all 100 steps."""
    import oracle
    import torch
    cfg = mg.config_linear_cpu()
    steps = cfg.steps
    N = cfg.x0.shape[0]
    D = 3 * N
    log("workload C1: linear tets 10^3, %d DOF, linear implicit Euler with LU, %d steps" % (D, steps))

    def gpu_run():
        g = femgpu.FemGpu(0)
        g.upload_mesh(cfg.kind, cfg.x0, cfg.conn, cfg.E, cfg.nu, cfg.rho)
        g.add_bc_nodes(cfg.bc_nodes)
        g.set_options(dt=cfg.dt, scheme=femgpu.LINEAR_EULER_IMPLICIT, linear_solver=femgpu.LINEAR_SOLVER_LU,
                      gravity=cfg.gravity)
        g.set_state(v=cfg.v0)
        C = torch.empty((D, D), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.step(1)  # assembles and factors the operator
        g.compliance_matrix_device(C.data_ptr())
        device_sync(g)
        t1 = time.perf_counter()
        assert g.step(steps - 1) == femgpu.FEM_OK
        device_sync(g)
        t2 = time.perf_counter()
        x, v = g.state()
        return g, C, t2 - t0, t1 - t0, t2 - t1, x, v

    g, C, _, _, _, _, _ = gpu_run()  # warm-up (code objects, allocations)
    g.close()
    del C
    g, C, t, t_first, t_rest, x, v = gpu_run()
    del C
    # the solve kernel's own time: event pairs around its launches on further steps
    g.enable_phase_timers(1)
    g.step(10)
    device_sync(g)
    ms, cnt = g.phase_times()["cg_init"]
    g.close()
    # every step applies the dense compliance kept from step 1. The phase events ride on the first launch of the
    # phase, dense_gemv_part_kernel: C (D x D) read once, x read, the 128-column chunks' partials written (its
    # second kernel, dense_gemv_sum_kernel, reads the partials and writes y: about 4.5 us in a kernel trace)
    chunks = (D + 127) // 128
    gemv_bytes = D * D * 8 + D * 8 + chunks * D * 8
    times = []
    for r in range(3):
        s = oracle_sim(cfg, scheme=oracle.LINEAR_EULER_IMPLICIT, linsolver=oracle.LU)
        s.set_compliance_at_first_linear_step(True)
        t1 = time.perf_counter()
        s.step(cfg.dt, steps)
        times.append(time.perf_counter() - t1)
        if r == 0:
            xr, vr = s.state()
        del s
    x0 = cfg.x0.reshape(-1)
    err = dict(u=float(np.linalg.norm((x - x0) - (xr - x0)) / np.linalg.norm(xr - x0)),
               v=float(np.linalg.norm(v - vr) / np.linalg.norm(vr)))
    tc = float(np.median(times))
    log("C1: gpu %.2f ms for %d steps (step 1 with factorization and dense compliance %.2f ms, then %.3f ms/step), "
        "cpu %.2f s (median of %s), rel err %s" % (1e3 * t, steps, 1e3 * t_first, 1e3 * t_rest / (steps - 1), tc,
                                                  ["%.2f" % a for a in times], err))
    return {"value": D * steps / t, "unit": "DOF-steps/s", "ms_per_step": 1e3 * t / steps, "steps": steps,
            "dtype": "f64",
            "config": {"workload": "linear-tet-block 10x10x10, linear implicit Euler dt=1ms, %d steps, banded LU "
                                   "factored on step 1, dense compliance (%d x %d) formed on step 1 and applied on "
                                   "the later steps" % (steps, D, D),
                       "dof": int(D)},
            "step1_ms": 1e3 * t_first, "later_steps_ms_per_step": 1e3 * t_rest / (steps - 1),
            "rel_err_vs_oracle_lu": err,
            "roofline": roofline("dense_gemv_part_kernel", gemv_bytes, ms / max(cnt, 1) / 1e3,
                                 "every step applies the dense compliance the linear scheme forms on its first step "
                                 "(127.6 MB, read once per step); the chunk sum that follows is a second, 1 MB "
                                 "kernel"),
            "cpu_baseline": {"value": D * steps / tc, "unit": "DOF-steps/s", "cores": 1, "kind": "port",
                             "sample": "the whole 100-step run, 3 runs, median: OdeSolverLinearEulerImplicit with "
                                       "the banded LU of the first system matrix and its dense compliance "
                                       "(3993 x 3993) at step 1", "sample_seconds": times, "build": ORACLE_BUILD,
                             "cpu": cpu_model()}}


def bench_sub_batch(args):
    out = bench_batch(args)
    tun = out["fp32_refined" if out["dtype"].startswith("f32") else "fp64"]["tuning"]
    sub = {k: out[k] for k in ("value", "unit", "ms_per_step", "steps", "dtype", "config",
                               "rel_err_fp32_vs_fp64_worst_instance", "instances_in_gate",
                               "instances_excluded_from_gate", "fp32_refined", "fp64")}
    if out.get("iteration_bytes") and tun.get("ms_per_iteration"):
        sub["roofline"] = roofline("bcg_apply_kernel + bcg_vector_kernel (one CG iteration, every instance)",
                                   out["iteration_bytes"], tun["ms_per_iteration"] / 1e3,
                                   "time per iteration from the autotuner's replay on the real operator")
    return sub


def bench_sub_adjoint(args):
    out = bench_adjoint(args)
    sub = {k: out[k] for k in ("value", "unit", "ms_per_step", "dtype", "config", "fd_check",
                               "cg_iterations_total") if k in out}
    if "roofline" in out:
        sub["roofline"] = out["roofline"]
    return sub


def bench_sub_hex(args):
    out = bench_hex(args)
    keep = ("value", "unit", "ms_per_step", "steps", "dtype", "config", "rel_err_vs_fp64", "fp64", "fp32_refined",
            "table_variant", "roofline")
    return {k: out[k] for k in keep if k in out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=40, help="cubes per axis")
    ap.add_argument("--cpu-steps", type=int, default=3, help="oracle steps per timed sample (3 samples)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-faithful-steps", type=int, default=1,
                    help="steps of the oracle with the reference's O(#BC x D) boundary zeroing (0: skip)")
    ap.add_argument("--timer-every", type=int, default=10,
                    help="phase breakdown: after the timed steps, as many more steps with the kernels of one in this "
                         "many event-timed (0: off)")
    ap.add_argument("--workload", default="corotational",
                    help="corotational (headline and every config's sub-record), linear, batch, adjoint, hex, "
                         "schemes, compliance")
    ap.add_argument("--headline-only", action="store_true", help="corotational: skip the other configs")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--batch-maxit", type=int, default=2000,
                    help="per-instance CG iteration cap of the batch workload")
    ap.add_argument("--adj-n", type=int, default=20, help="adjoint block: n^3 cubes")
    ap.add_argument("--adj-tol", type=float, default=1e-12, help="adjoint workload CG tolerance")
    ap.add_argument("--adj-fd-dirs", type=int, default=32)
    ap.add_argument("--adj-fd-h", type=float, default=1e-4, help="FD step relative to mean E")
    ap.add_argument("--adj-ident", type=int, default=0, help="identification iterations to run (0: none)")
    ap.add_argument("--adj-checkpoint", type=int, default=0,
                    help="adjoint: checkpoint every k steps and recompute segments (0: keep the whole history)")
    ap.add_argument("--hex-n", type=int, default=128, help="hex block: n x n/2 x n cubes")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also time a copy of this many MiB (probe_copy_kernel)")
    args = ap.parse_args()
    args.steps_given = "--steps" in sys.argv
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; the multi-GPU sharding is a design (no --gpus > 1).",
              file=sys.stderr)
        sys.exit(2)
    if args.workload in ("adjoint", "compliance", "linear") or (args.workload == "corotational" and
                                                                 not args.headline_only):
        # PyTorch's own HIP runtime (device buffers of the adjoint harness and of the dense compliance) has to
        # start before the library's system ROCm runtime in one process: started the other way round it finds
        # no GPU and the library's later calls fail
        import torch
        torch.cuda.init()
        if not torch.cuda.is_available():
            raise RuntimeError("torch sees no GPU")
    single = {"adjoint": bench_adjoint, "schemes": bench_schemes, "compliance": bench_compliance,
              "batch": bench_batch, "hex": bench_hex, "linear": bench_linear_case}
    if args.workload in single:
        out = single[args.workload](args)
        if args.workload == "hex" and args.copy_probe:
            copy_probe(args.copy_probe)
        print(json.dumps(out), flush=True)
        return
    out = bench_corotational(args)
    if not args.headline_only:
        # every other configuration that fits one GPU, at its own step count (the headline's --steps and
        # --warmup do not apply to them)
        sub_args = argparse.Namespace(**vars(args))
        sub_args.steps_given = False
        sub_args.warmup = 2
        configs = {}
        for name, fn in (("C1_linear_tets_10x10x10", bench_linear_case), ("C3_batch_64_instances", bench_sub_batch),
                         ("C4_adjoint_20x20x20", bench_sub_adjoint), ("C5_hex_128x64x128", bench_sub_hex)):
            t0 = time.perf_counter()
            log("--- sub-record %s" % name)
            try:
                configs[name] = fn(sub_args)
            except Exception as e:  # noqa: BLE001 - a failed sub-record is reported, the headline still prints
                log("sub-record %s failed: %r" % (name, e))
                configs[name] = {"error": repr(e)}
            configs[name]["wall_s"] = round(time.perf_counter() - t0, 1)
        out["configs"] = configs
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
