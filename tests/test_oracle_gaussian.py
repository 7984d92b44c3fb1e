"""Pins the oracle's dense Gaussian sketch (orc_gaussian_fill_f64) without a
GPU.  The reference fills S with i.i.d. N(0,1) values from a clock-seeded
generator (src/kernels/cpu/sketch/gaussian_sketch_kernels.cpp:10-18,
src/kernels/cuda/sketch/gaussian_sketch_cuda_kernels.cu:21-45), so what the
library shares with it is the distribution; the seeded counter layout is
this library's own and is pinned here by:
  * SplitMix64's published first outputs from state 0 (0xe220a8397b1dcdaf,
    0x6e789e6aa1b965f4) through an independent pure-Python mixer;
  * the pair layout (rows 2q, 2q+1 of column j share pair j*ceil(rows/2)+q)
    read back from the oracle for odd and even row counts;
  * the distribution: Kolmogorov-Smirnov against the normal CDF, the first
    four moments, and the correlation between pair members, between
    neighbouring pairs and between their squares (a Box-Muller pair built
    from shared bits, or a cos/sin mix-up, shows up there)."""
import math

import numpy as np
import pytest
from scipy import stats

M64 = (1 << 64) - 1
GAMMA = 0x9E3779B97F4A7C15


def splitmix_out(state):
    z = state & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def stream(seed, p):
    """p-th output (0-based) of SplitMix64 started at `seed`."""
    return splitmix_out(seed + (p + 1) * GAMMA)


def pair(seed, p):
    z = stream(seed, p)
    u1 = ((z >> 41) + 0.5) * 2.0 ** -23
    u2 = ((z >> 8) & 0xFFFFFF) * 2.0 ** -24
    r = math.sqrt(-2.0 * math.log(u1))
    return r * math.cos(2 * math.pi * u2), r * math.sin(2 * math.pi * u2)


def test_splitmix64_known_outputs():
    assert stream(0, 0) == 0xE220A8397B1DCDAF
    assert stream(0, 1) == 0x6E789E6AA1B965F4


@pytest.mark.parametrize("rows,cols,seed", [(7, 3, 0), (8, 2, 12345), (1, 5, 2**63 + 11), (2, 1, M64)])
def test_counter_layout(orc, rows, cols, seed):
    S = orc.gaussian_fill(rows, cols, seed)
    half = (rows + 1) // 2
    for j in range(cols):
        for q in range(half):
            a, b = pair(seed, j * half + q)
            assert S[2 * q, j] == pytest.approx(a, rel=1e-13, abs=1e-13)
            if 2 * q + 1 < rows:
                assert S[2 * q + 1, j] == pytest.approx(b, rel=1e-13, abs=1e-13)


def test_distribution(orc):
    S = orc.gaussian_fill(1000, 400, 7)  # 400k samples
    v = S.ravel(order="F")
    N = v.size
    assert abs(v.mean()) < 5 / math.sqrt(N)
    assert abs(v.var() - 1) < 5 * math.sqrt(2 / N)
    assert abs(stats.skew(v)) < 5 * math.sqrt(6 / N)
    assert abs(stats.kurtosis(v)) < 5 * math.sqrt(24 / N)
    assert stats.kstest(v, "norm").pvalue > 1e-4
    ev, od = S[0::2].ravel(), S[1::2].ravel()
    lim = 5 / math.sqrt(ev.size)
    assert abs(np.corrcoef(ev, od)[0, 1]) < lim              # pair members
    assert abs(np.corrcoef(ev ** 2, od ** 2)[0, 1]) < lim    # independent, not just uncorrelated
    assert abs(np.corrcoef(S[0:-2:2].ravel(), S[2::2].ravel())[0, 1]) < lim  # neighbouring pairs
    assert abs(np.corrcoef(S[:, :-1].ravel(), S[:, 1:].ravel())[0, 1]) < 5 / math.sqrt(S[:, 1:].size)


def test_seeds_differ(orc):
    a, b = orc.gaussian_fill(64, 4, 1), orc.gaussian_fill(64, 4, 2)
    assert not np.any(a == b)
    assert np.array_equal(a, orc.gaussian_fill(64, 4, 1))
