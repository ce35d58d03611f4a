"""Initial guesses of the demag solvers from their kept solved pairs: the
least-squares helpers (backward differences, scaled Cholesky) against a
dense least-squares solve in numpy, on a smooth sequence of right-hand sides
(CPU; the device kernels call the same helpers)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("winls") / "win_ls_check")
    src = os.path.join(ROOT, "tests", "native", "win_ls_check.cpp")
    subprocess.run([HIPCC, "-O2", "-std=c++17", f"-I{ROOT}/include", f"-I{ROOT}/tetmag_amd/csrc", src, "-o", exe],
                   check=True, capture_output=True)
    return exe


def run(exe, tmp_path, dt):
    out = str(tmp_path / f"seq_{dt}.bin")
    subprocess.run([exe, out, str(dt)], check=True)
    raw = open(out, "rb").read()
    n, steps, K = np.frombuffer(raw[:12], np.int32)
    rec = []
    off = 12
    for _ in range(steps):
        cnt = int(np.frombuffer(raw[off:off + 4], np.int32)[0]); off += 4
        a = np.frombuffer(raw[off:off + 8 * K], np.float64).copy(); off += 8 * K
        b = np.frombuffer(raw[off:off + 8 * n], np.float64).copy(); off += 8 * n
        rec.append((cnt, a, b))
    return K, rec


@pytest.mark.parametrize("dt", [0.02, 0.3])
def test_projection_is_the_least_squares_fit(driver, tmp_path, dt):
    K, rec = run(driver, tmp_path, dt)
    assert K >= 4
    for st, (cnt, a, b) in enumerate(rec):
        assert cnt == min(st, K)
        if cnt == 0:
            continue
        kept = np.stack([rec[st - 1 - i][2] for i in range(cnt)], 1)  # newest first
        res = np.linalg.norm(b - kept @ a[:cnt])
        # the best fit over the span of the kept vectors, by SVD
        c, *_ = np.linalg.lstsq(kept, b, rcond=None)
        best = np.linalg.norm(b - kept @ c)
        warm = np.linalg.norm(b - kept[:, 0])
        nb = np.linalg.norm(b)
        assert res <= warm * (1 + 1e-12), (st, res, warm)
        assert res <= best + 1e-10 * nb, (st, res, best)
        assert np.all(a[cnt:] == 0.0)
    # with six kept vectors of a 7-mode sequence at a small step the fit is tight
    if dt == 0.02:
        cnt, a, b = rec[-1]
        kept = np.stack([rec[-2 - i][2] for i in range(cnt)], 1)
        assert np.linalg.norm(b - kept @ a[:cnt]) < 1e-8 * np.linalg.norm(b)


def test_repeated_right_hand_side_gives_the_newest_solution(driver, tmp_path):
    # b equal to every kept vector: zero differences (dropped directions) and
    # b - b_0 = 0, so the start is the newest solution alone, a = e_0
    K, rec = run(driver, tmp_path, 0.0)
    for st, (cnt, a, b) in enumerate(rec):
        if cnt:
            np.testing.assert_array_equal(a, np.eye(K)[0])
