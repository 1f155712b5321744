"""Runs the C++ adapter test program (tests/cpp/fem3d_representation_test.cpp): the C++ classes with the
reference's names (Fem3DRepresentation, OdeState, LinearSparseSolveAndInverseCG, ...) over the C ABI."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "fem3d_representation_test")


def _binary():
    # make rebuilds the program when the adapter header or the libraries changed
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")])
    return BIN


def _run(mode):
    print("C++ adapter tests", mode, flush=True)
    p = subprocess.run([_binary(), mode, os.path.join(ROOT, "tests", "golden", "ply")], capture_output=True,
                       text=True, timeout=300)
    sys.stdout.write(p.stdout)
    sys.stdout.write(p.stderr)
    assert p.returncode == 0, "C++ adapter tests failed (%s)" % mode


def test_cpp_adapters_cpu():
    _run("--cpu")


@pytest.mark.gpu
def test_cpp_adapters_gpu():
    _run("--gpu")
