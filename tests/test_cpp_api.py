"""Runs the C++ host-layer tests (tests/cpp, built by make into
hybrid-sparse-pde-solver_amd/build/tests) — the reference API surface
(SparseMatrix, Vector, preconditioners, CG/FGMRES, SolverFactory,
AdaptiveSelector, MatrixMarketIO) driving the device path."""
import os
import subprocess

import pytest

from conftest import ROOT

BIN = os.path.join(ROOT, "hybrid-sparse-pde-solver_amd", "build", "tests")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["test_core", "test_solvers", "test_adaptive"])
def test_cpp_suite(name):
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), f"{exe} not built (make -C hybrid-sparse-pde-solver_amd)"
    print(f"running {name}", flush=True)
    r = subprocess.run([exe], cwd=ROOT, capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, HSPS_DATA=os.path.join(ROOT, "tests", "golden", "data")))
    print(r.stdout[-4000:])
    print(r.stderr[-2000:])
    assert r.returncode == 0, r.stdout[-2000:]
