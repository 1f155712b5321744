"""The C-ABI library loads and exports every entry point declared in include/fem_gpu.h (no GPU calls)."""
import ctypes
import os
import re

from opensurgsim_amd import femgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fem_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(fem_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_header_symbols():
    L = femgpu.load()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert sorted(femgpu.EXPORTS) == names


def test_default_options_match_reference_defaults():
    o = femgpu.default_options()
    assert o.dt > 0
    assert list(o.gravity) == [0.0, -9.81, 0.0]          # Physics/Representation.cpp:36
    assert o.gravity_enabled == 1
    assert o.newton_max_iterations == 1                    # Math/OdeSolverEulerImplicit.cpp:26
    assert o.cg_tolerance == 2.220446049250313e-16         # Eigen: NumTraits<double>::epsilon()
    assert o.cg_max_iterations < 0                         # Eigen: 2 * cols
    assert o.rayleigh_mass == 0.0 and o.rayleigh_stiffness == 0.0  # Physics/FemRepresentation.cpp:38-43
    assert o.mass_lumping == 0
    assert o.newton_epsilon == 1e-5                        # Math/OdeSolverEulerImplicit.cpp:26
    assert o.linear_solver == femgpu.LINEAR_SOLVER_CG       # the C ABI's default (the adapters keep LU)
    assert o.scheme == femgpu.EULER_IMPLICIT


def test_options_struct_layout():
    assert ctypes.sizeof(femgpu.FemOptions) == 120
    assert ctypes.sizeof(femgpu.FemStats) == 80
    assert ctypes.sizeof(femgpu.FemInstanceStats) == 40
