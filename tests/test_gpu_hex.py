"""GPU parity of the hexahedral (Fem3DElementCube) path against the FP64 oracle, through the C ABI."""
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


def test_hex_pattern_bit_exact():
    say("hex pattern 3x4x2")
    cfg = mg.config_hex((3, 4, 2))
    o = make_oracle(cfg)
    g = make_gpu(cfg)
    outer, inner = o.pattern()
    rowptr, colidx = g.pattern()
    assert np.array_equal(rowptr, outer) and np.array_equal(colidx, inner)


def test_hex_element_matrices_vs_oracle():
    say("hex elements")
    rng = np.random.default_rng(7)
    cfg = mg.config_hex((3, 2, 2))
    x0 = cfg.x0 + 1e-4 * rng.normal(size=cfg.x0.shape)  # distorted rest shapes
    g = femgpu.FemGpu(0)
    g.upload_mesh(femgpu.HEX, x0, cfg.conn, cfg.E, cfg.nu, cfg.rho)
    x = x0 + 2e-4 * rng.normal(size=x0.shape)
    R, K, f = g.evaluate_elements(x.reshape(-1))
    for e in range(cfg.conn.shape[0]):
        Ko, Mo, vol = oracle.cube_km(x0[cfg.conn[e]], cfg.E[e], cfg.nu[e], cfg.rho[e])
        assert np.abs(K[e] - Ko).max() <= 1e-12 * np.abs(Ko).max()
        fo = -Ko @ (x[cfg.conn[e]] - x0[cfg.conn[e]]).reshape(-1)
        assert np.abs(f[e] - fo).max() <= 1e-12 * np.abs(Ko).max() * 1e-3


def trajectory(cfg, steps, precision, otol, gtol):
    o = make_oracle(cfg, tol=otol)
    g = make_gpu(cfg, precision=precision, tol=gtol)
    o.step(cfg.dt, steps)
    assert g.step(steps) == femgpu.FEM_OK
    xr, vr = o.state()
    x, v = g.state()
    x0 = cfg.x0.reshape(-1)
    return dict(x=rel(x, xr), u=rel(x - x0, xr - x0), v=rel(v, vr), stats=g.stats())


def test_hex_trajectory_fp64():
    cfg = mg.config_hex((6, 4, 5))
    e = trajectory(cfg, 15, femgpu.FP64, 1e-12, 1e-12)
    say("hex fp64", e)
    assert e["x"] < 1e-9 and e["u"] < 1e-9 and e["v"] < 1e-9


def test_hex_trajectory_fp32_refined():
    cfg = mg.config_hex((6, 4, 5))
    e = trajectory(cfg, 15, femgpu.FP32_REFINED, 1e-12, 1e-9)
    say("hex fp32+refine", e)
    assert e["u"] < 1e-6 and e["v"] < 1e-6
    assert e["stats"]["last_refinements"] >= 1


def test_hex_invalid_element_rejected():
    cfg = mg.config_hex((1, 1, 1))
    conn = cfg.conn[:, [4, 5, 6, 7, 0, 1, 2, 3]]  # mirrored faces: negative volume
    g = femgpu.FemGpu(0)
    with pytest.raises(femgpu.FemError):
        g.upload_mesh(femgpu.HEX, cfg.x0, conn, cfg.E, cfg.nu, cfg.rho)


def test_hex_distinct_matrix_table(monkeypatch):
    """The FP32 operator of a structured hex block reads a table of its few distinct element matrices (the
    table holds exactly the floats of the per-element cache), so the trajectory equals the per-element
    cache's bit for bit. A mesh with perturbed nodes has more than 32 distinct matrices and keeps the
    per-element cache, and still follows the oracle."""
    cfg = mg.config_hex((8, 5, 7))
    runs = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("FEM_HEX_TABLE", flag)
        say("hex table", flag)
        g = make_gpu(cfg, precision=femgpu.FP32_REFINED, tol=1e-9)
        assert g.hex_distinct_matrices() == 0  # built with the first step's operator
        assert g.step(6) == femgpu.FEM_OK
        runs[flag] = (g.state(), g.hex_distinct_matrices(), g.stats()["total_cg_iterations"])
        g.close()
    (x1, v1), nd1, it1 = runs["1"]
    (x0, v0), nd0, it0 = runs["0"]
    say("  distinct matrices", nd1, "iterations", it1, it0)
    assert 1 <= nd1 <= 32 and nd0 == 0
    assert np.array_equal(x1, x0) and np.array_equal(v1, v0) and it1 == it0
    monkeypatch.delenv("FEM_HEX_TABLE")
    irr = mg.config_hex((4, 3, 4))
    rng = np.random.default_rng(11)
    irr["x0"] = irr.x0 + rng.uniform(-0.1, 0.1, irr.x0.shape) * irr.h
    e = trajectory(irr, 4, femgpu.FP32_REFINED, 1e-12, 1e-9)
    g = make_gpu(irr, precision=femgpu.FP32_REFINED, tol=1e-9)
    assert g.step(1) == femgpu.FEM_OK
    say("  perturbed mesh: distinct", g.hex_distinct_matrices(), "errors", e)
    assert g.hex_distinct_matrices() == 0
    assert e["u"] < 1e-6 and e["v"] < 1e-6
