"""bench.py's host-side helpers (no GPU): the committed counter profiles
it reads for the roofline traffic, the step quantiles and the byte model."""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def test_traffic_profiles_are_read():
    t = bench.measured_traffic("cuboid-128x128x16-h3", 40962)
    assert t is not None and 1.0 < t / (8.0 * 40962 ** 2) < 1.1
    own = bench.measured_gemm_traffic("cuboid-128x128x16-h3", 40962, True)
    lib = bench.measured_gemm_traffic("cuboid-128x128x16-h3", 40962, False)
    assert own is not None and lib is not None
    assert 1.0 < lib / (8.0 * 40962 ** 2) < 1.1 and own > lib
    assert bench.measured_traffic("cuboid-128x128x16-h3", 123) is None
    assert bench.measured_gemm_traffic("no-such-workload-h1", 40962) is None


def test_step_quantiles_and_byte_model():
    class Ev:
        def __init__(self, t):
            self.t = t

        def elapsed_time(self, other):
            return other.t - self.t

    q = bench.step_quantiles([Ev(x) for x in (0.0, 1.0, 3.0, 4.0, 8.0)])
    assert q["median"] == pytest.approx(1.5) and q["p10"] <= q["median"] <= q["p90"]  # steps 1, 2, 1, 4 ms
    b0 = bench.algorithmic_bytes(1000, 10000, 100, 0, 0)
    b1 = bench.algorithmic_bytes(1000, 10000, 100, 5, 5)
    assert b1 > b0 > 8.0 * 100 ** 2
