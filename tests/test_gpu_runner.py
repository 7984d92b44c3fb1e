"""LSQR runner end to end on the GPU: MatrixMarket problem files in, the
reference runner's report lines out (lsqr_precision_runner.cu:692-715).
The problem is the C1 generator's (A = G diag(d), d log-spaced to 1e-3,
b = A x_true); the thresholds are the solver's own stopping rule plus the
conditioning of the problem."""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.test_runner_io import RUNNER, write_array

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(RUNNER), reason="runner not built (__graft_entry__.build())")]


@pytest.fixture(scope="module")
def problem(tmp_path_factory):
    d = tmp_path_factory.mktemp("runner")
    rng = np.random.default_rng(11)
    m, n = 4000, 60
    A = rng.standard_normal((m, n)) * np.logspace(0, -3, n)
    A = A.astype(np.float32).astype(np.float64)  # exactly representable in FP32
    xt = rng.standard_normal(n)
    b = A @ xt
    write_array(d / "A.mtx", A)
    write_array(d / "x.mtx", xt)
    write_array(d / "b.mtx", b)
    return d, A, xt, b


def run(problem, *extra):
    d = problem[0]
    r = subprocess.run([RUNNER, "--matrix", str(d / "A.mtx"), "--x-true", str(d / "x.mtx"), "--rhs",
                        str(d / "b.mtx"), *extra], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    get = lambda key: float(re.search(key + r":\s+([-0-9.eE+]+)", out).group(1))
    return int(get("iterations")), get("final residual"), get("solution error"), out


@pytest.mark.parametrize("extra,res_max,err_max", [
    (("--mv-type", "fp64", "--value-type", "fp64", "--refine-type", "fp64", "--tol", "1e-8",
      "--max-relres", "1e-8", "--max-iters", "3000"), 1.5e-8, 1e-4),
    (("--mv-type", "fp32", "--tol", "1e-5", "--max-relres", "1e-5", "--max-iters", "3000"), 2e-5, 1e-1),
    (("--mv-type", "fp64", "--value-type", "fp64", "--refine-type", "fp64", "--precond", "--tol", "1e-10",
      "--max-relres", "1e-10"), 1e-9, 1e-6),
    (("--mv-type", "fp32_bf16", "--implicit-duplicates", "--refine-type", "fp64", "--precond-type", "fp64",
      "--precond", "--tol", "1e-6", "--max-relres", "1e-6", "--max-iters", "8", "--restarts", "20"), 1.5e-6, 1e-2),
    (("--implicit-duplicates", "--refine-type", "fp64", "--precond-type", "fp64", "--precond", "--tol", "1e-6",
      "--max-relres", "1e-6", "--max-iters", "10", "--restarts", "20", "--two-pass"), 1.5e-6, 1e-2),
    (("--mv-type", "fp64", "--value-type", "fp64", "--refine-type", "fp64", "--normalize-columns",
      "--tol", "1e-8", "--max-relres", "1e-8", "--max-iters", "3000"), 1.5e-8, 1e-4),
])
def test_runner_solves(problem, extra, res_max, err_max):
    it, res, err, out = run(problem, *extra)
    assert res <= res_max and err <= err_max, out
    assert it >= 1


@pytest.mark.parametrize("mv", ["fp16", "bf16"])
def test_native_16bit_storage_and_explicit_duplicate(problem, mv):
    """A stored as IEEE half or bf16: the solve alone stops near the 16-bit
    rounding of A; with an FP32 duplicate of A (--explicit-duplicates) the
    restarts recompute the residual from an exact copy of A at the refine
    precision (FP64 here) and the run reaches the requested tolerance against
    the file's A (lsqr_precision_runner.cu:559-580)."""
    common = ("--mv-type", mv, "--refine-type", "fp64", "--precond-type", "fp64", "--precond", "--tol", "1e-6",
              "--max-relres", "1e-6", "--max-iters", "8", "--restarts", "30")
    it, res, err, out = run(problem, *common, "--explicit-duplicates")
    assert "explicit FP64-precision duplicate" in out, out
    assert res <= 1.5e-6 and err <= 1e-2, out
    it16, res16, err16, out16 = run(problem, *common)
    unit = 2.0 ** -11 if mv == "fp16" else 2.0 ** -8
    assert res16 <= 4 * unit and res16 > res, out16


def test_residual_history_file(problem, tmp_path):
    h = tmp_path / "hist.txt"
    it, res, err, out = run(problem, "--mv-type", "fp64", "--value-type", "fp64", "--refine-type", "fp64",
                            "--precond", "--tol", "1e-10", "--max-relres", "1e-10", "--residual-history", str(h))
    rows = [line.split() for line in h.read_text().splitlines()]
    assert len(rows) >= it and [int(r[0]) for r in rows] == list(range(1, len(rows) + 1))
    vals = np.array([float(r[1]) for r in rows])
    assert vals[-1] < 1e-9 and np.all(vals > 0)
