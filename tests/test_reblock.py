"""Reblocking against the reference analysis tool's own fixtures
(tools/pyblock/pyblock/tests/base.py: the correlated series
convolve(RandomState(5).randn(1024), ones(64)/10, 'same') and its reblock
table), plus the ratio estimator on a closed form."""
import math

import numpy as np

from hande_amd import reblock

# (level, ndata, mean, variance, std_err, std_err_err) -- tools/pyblock/pyblock/tests/base.py reblock_1D
REBLOCK_1D = [
    (0, 1024, 0.20848489810011175, 0.5059733357706058, 0.022228688348921586, 0.0004914292726987686),
    (1, 512, 0.20848489810011178, 0.5015874592634239, 0.0312995687889446, 0.0009790681131713157),
    (2, 256, 0.20848489810011184, 0.4956240227095556, 0.04400035612025432, 0.001948368517458123),
    (3, 128, 0.20848489810011173, 0.48304884454343255, 0.06143141784132584, 0.003854549974240774),
    (4, 64, 0.2084848981001117, 0.46303627895395055, 0.08505846141716576, 0.007577610011170585),
    (5, 32, 0.20848489810011178, 0.40641824428237183, 0.11269680622725792, 0.014312508703377615),
    (6, 16, 0.20848489810011178, 0.3388777084790293, 0.14553300924511708, 0.026570570675052838),
    (7, 8, 0.20848489810011178, 0.228782900550399, 0.16910902568698064, 0.04519628822370252),
    (8, 4, 0.20848489810011178, 0.21072809492758346, 0.229525649398702, 0.09370345398462808),
    (9, 2, 0.20848489810011175, 0.011498881312406542, 0.07582506614704185, 0.05361641845649181),
]
REBLOCK_1D_OPT = 9
REBLOCK_1D_DDOF0_VAR = [0.5054792211848923, 0.50060779625705, 0.4936879913708464]


def data_1d():
    rand = np.random.RandomState(seed=5)
    return np.convolve(rand.randn(2 ** 10), np.ones(2 ** 6) / 10, "same")


def test_reblock_matches_fixture():
    stats = reblock.reblock(data_1d())
    assert len(stats) == len(REBLOCK_1D)
    for got, ref in zip(stats, REBLOCK_1D):
        level, n, mean, cov, se, see = got
        assert (level, n) == ref[:2]
        assert math.isclose(mean[0], ref[2], rel_tol=1e-12)
        assert math.isclose(cov[0, 0], ref[3], rel_tol=1e-10)
        assert math.isclose(se[0], ref[4], rel_tol=1e-10)
        assert math.isclose(see[0], ref[5], rel_tol=1e-10)
    assert reblock.find_optimal_block(1024, stats) == [REBLOCK_1D_OPT]


def test_reblock_ddof0_matches_fixture():
    stats = reblock.reblock(data_1d(), ddof=0)
    for got, ref in zip(stats, REBLOCK_1D_DDOF0_VAR):
        assert math.isclose(got[3][0, 0], ref, rel_tol=1e-10)


def test_ratio_closed_form():
    # uncorrelated A, B: relative errors add in quadrature
    m, se = reblock.ratio(2.0, 4.0, 0.02, 0.04, 0.0, 100)
    assert m == 0.5
    assert math.isclose(se, 0.5 * math.sqrt(0.01 ** 2 + 0.01 ** 2))
    # perfectly correlated A = 2B: ratio has no error
    rng = np.random.default_rng(0)
    b = 5 + rng.normal(size=4096)
    out = reblock.projected_energy(2 * b, b)
    assert math.isclose(out["energy"], 2.0) and out["error"] < 1e-6


def test_unconverged_series_reports_nan_error():
    """A series much shorter than its correlation time has no optimal block;
    its error is NaN rather than the last level's (two or three blocks)."""
    rng = np.random.default_rng(4)
    x = np.cumsum(rng.normal(size=64)) * 0.01 + 1.0  # random walk: correlated throughout
    out = reblock.mean_error(x)
    assert not out["converged"] and math.isnan(out["error"])
    assert math.isclose(out["mean"], x.mean())
    pe = reblock.projected_energy(x * 2.0 + rng.normal(size=64) * 1e-3, np.full(64, 1.0) + x * 0.1)
    assert not pe["converged"] and math.isnan(pe["error"])
