"""MatrixMarket input of the LSQR runner (host only: --check-inputs stops
before any GPU work).  Values are pinned through the runner's checksum line
against numpy on the same matrices: array (column-major), coordinate
(1-based, repeated entries summed), symmetric (mirrored), comments, and the
error paths."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# TOMOS_RUNNER selects another build (the sanitizer build of tests/run_sanitizers.sh)
RUNNER = os.environ.get("TOMOS_RUNNER") or os.path.join(ROOT, "tomos_amd", "tomos_lsqr_runner")

pytestmark = pytest.mark.skipif(not os.path.exists(RUNNER), reason="runner not built (__graft_entry__.build())")


def write_array(path, M, banner="%%MatrixMarket matrix array real general"):
    M = np.asarray(M, dtype=np.float64)
    if M.ndim == 1:
        M = M[:, None]
    with open(path, "w") as f:
        f.write(banner + "\n% written by the test\n\n")
        f.write("%d %d\n" % M.shape)
        for v in M.T.ravel():
            f.write("%.17g\n" % v)


def write_coord(path, shape, entries, symmetric=False):
    sym = "symmetric" if symmetric else "general"
    with open(path, "w") as f:
        f.write("%%%%MatrixMarket matrix coordinate real %s\n%% c\n" % sym)
        f.write("%d %d %d\n" % (shape[0], shape[1], len(entries)))
        for i, j, v in entries:
            f.write("%d %d %.17g\n" % (i + 1, j + 1, v))


def check(tmp_path, A_file, A, x, b, extra=()):
    write_array(tmp_path / "x.mtx", x)
    write_array(tmp_path / "b.mtx", b)
    r = subprocess.run([RUNNER, "--matrix", str(A_file), "--x-true", str(tmp_path / "x.mtx"), "--rhs",
                        str(tmp_path / "b.mtx"), "--check-inputs", *extra], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    got = dict(re.findall(r"(\w+)=([-0-9.eE+]+)", r.stdout.split("checksums:")[1].splitlines()[0]))
    m, n = A.shape
    w = (np.arange(1, m + 1)[:, None] + 7 * np.arange(1, n + 1)[None, :]).astype(np.float64)
    assert float(got["sum_A"]) == pytest.approx(A.sum(), rel=1e-14, abs=1e-12)
    assert float(got["weighted_A"]) == pytest.approx((A * w).sum(), rel=1e-14, abs=1e-9)
    assert float(got["sum_x"]) == pytest.approx(np.sum(x), rel=1e-14, abs=1e-12)
    assert float(got["sum_b"]) == pytest.approx(np.sum(b), rel=1e-14, abs=1e-12)
    return r.stdout


def test_array_file(tmp_path):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((37, 5))
    write_array(tmp_path / "A.mtx", A)
    out = check(tmp_path, tmp_path / "A.mtx", A, rng.standard_normal(5), rng.standard_normal(37),
                ["--mv-type", "fp32_bf16", "--precond"])
    assert "A: 37 x 5 (array, 185 entries)" in out and "mv_storage=FP32_BF16" in out


def test_coordinate_sums_repeats(tmp_path):
    A = np.zeros((6, 4))
    entries = [(0, 0, 1.5), (5, 3, -2.0), (2, 1, 0.25), (2, 1, 0.5), (4, 2, 3.0)]
    for i, j, v in entries:
        A[i, j] += v
    write_coord(tmp_path / "A.mtx", A.shape, entries)
    out = check(tmp_path, tmp_path / "A.mtx", A, np.ones(4), np.arange(6.0))
    assert "(coordinate, 5 entries)" in out


def test_symmetric_array_and_coordinate(tmp_path):
    rng = np.random.default_rng(2)
    L = np.tril(rng.standard_normal((4, 4)))
    S = L + np.tril(L, -1).T
    with open(tmp_path / "S.mtx", "w") as f:
        f.write("%%MatrixMarket matrix array real symmetric\n4 4\n")
        for j in range(4):
            for i in range(j, 4):
                f.write("%.17g\n" % S[i, j])
    check(tmp_path, tmp_path / "S.mtx", S, np.ones(4), np.ones(4))
    write_coord(tmp_path / "C.mtx", (4, 4), [(i, j, S[i, j]) for j in range(4) for i in range(j, 4)], True)
    check(tmp_path, tmp_path / "C.mtx", S, np.ones(4), np.ones(4))


@pytest.mark.parametrize("content,msg", [
    ("not a banner\n2 2\n1\n2\n3\n4\n", "no %%MatrixMarket banner"),
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n", "truncated"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n", "out of range"),
    ("%%MatrixMarket matrix array complex general\n1 1\n1 0\n", "not supported"),
    ("%%MatrixMarket matrix array real general\n1 1\nabc\n", "bad number"),
])
def test_malformed_files(tmp_path, content, msg):
    (tmp_path / "A.mtx").write_text(content)
    write_array(tmp_path / "x.mtx", np.ones(2))
    write_array(tmp_path / "b.mtx", np.ones(2))
    r = subprocess.run([RUNNER, "--matrix", str(tmp_path / "A.mtx"), "--x-true", str(tmp_path / "x.mtx"),
                        "--rhs", str(tmp_path / "b.mtx"), "--check-inputs"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and msg in r.stderr, r.stderr


def test_option_errors(tmp_path):
    A = np.ones((4, 2))
    write_array(tmp_path / "A.mtx", A)
    write_array(tmp_path / "x.mtx", np.ones(2))
    write_array(tmp_path / "b.mtx", np.ones(4))
    base = [RUNNER, "--matrix", str(tmp_path / "A.mtx"), "--x-true", str(tmp_path / "x.mtx"),
            "--rhs", str(tmp_path / "b.mtx"), "--check-inputs"]
    for extra, msg in ((["--mv-type", "fp8"], "unknown --mv-type"),
                       (["--explicit-duplicates", "--implicit-duplicates"], "mutually exclusive"),
                       (["--value-type", "fp16"], "supports fp64 and fp32"),
                       (["--bogus"], "unrecognized argument")):
        r = subprocess.run(base + extra, capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and msg in r.stderr, (extra, r.stderr)
    r = subprocess.run(base[:-1] + ["--device", "cpu"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "runs LSQR on the GPU" in r.stderr
    for mv, name in (("fp16", "FP16"), ("half", "FP16"), ("bf16", "BF16")):
        r = subprocess.run(base + ["--mv-type", mv, "--explicit-duplicates"], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 0 and "mv_storage=%s mv_compute=%s" % (name, name) in r.stdout, r.stderr
