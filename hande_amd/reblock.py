"""Reblocking analysis of QMC report-table columns.

Restates the analysis the reference's tools apply to FCIQMC output:
  * Flyvbjerg--Petersen reblocking (tools/pyblock/pyblock/blocking.py:9-169):
    repeatedly average neighbouring pairs, recording mean, covariance,
    standard error and its error at each level (the last point is dropped
    when the length is odd; covariance normalised by N - ddof, ddof = 1);
  * the optimal block of Wolff and Lee et al. (blocking.py:171-259): the
    smallest level with B^3 > 2 n (SE_B/SE_0)^4, B = 2^level; when no level
    satisfies it the series is too short for its correlation time and the
    error is reported as NaN (pyblock likewise finds no optimal block), not
    read off the last level's two or three blocks;
  * the ratio estimator with covariance (tools/pyblock/pyblock/error.py:ratio)
    used for E_proj = <sum H_0j N_j> / <N_0>
    (tools/pyhande/pyhande/analysis.py:14-62);
  * the statistical efficiency 1/(sigma^2 t) and the reference's
    inefficiency sigma * sqrt(N_w * iterations * tau)
    (tools/pyhande/pyhande/analysis.py:256-320).
"""
import math

import numpy as np


def reblock(data, ddof=1):
    """data: (nvar, n) or (n,).  Returns a list of
    (level, ndata, mean, cov, std_err, std_err_err)."""
    x = np.atleast_2d(np.asarray(data, dtype=float))
    stats = []
    level = 0
    while x.shape[1] >= 2:
        n = x.shape[1]
        mean = x.mean(axis=1)
        cov = np.atleast_2d(np.cov(x, ddof=ddof))
        se = np.sqrt(np.diag(cov) / n)
        se_err = se / np.sqrt(2.0 * (n - ddof))
        stats.append((level, n, mean, cov, se, se_err))
        half = n // 2
        x = 0.5 * (x[:, 0:2 * half:2] + x[:, 1:2 * half:2])
        level += 1
    return stats


def find_optimal_block(ndata, stats):
    """Per variable, the first level at which correlation has been removed
    (NaN if the data are too short to tell)."""
    se0 = np.atleast_1d(stats[0][4])
    opt = [float("nan")] * se0.size
    for (level, _n, _m, _c, se, _e) in reversed(stats):
        b3 = 2 ** (3 * level)
        for i, v in enumerate(np.atleast_1d(se)):
            if b3 > 2 * ndata * (v / se0[i]) ** 4:
                opt[i] = level
    return opt


def ratio(mean_a, mean_b, se_a, se_b, cov_ab, ndata):
    """Mean and standard error of A/B with correlated A and B."""
    m = mean_a / mean_b
    se = abs(m * math.sqrt((se_a / mean_a) ** 2 + (se_b / mean_b) ** 2 - 2 * cov_ab / (ndata * mean_a * mean_b)))
    return m, se


def projected_energy(sum_h0j_nj, n0):
    """E_proj = <sum H_0j N_j>/<N_0> and its error at the optimal block of the
    pair (the larger of the two variables' optimal levels); the error is NaN
    when either variable has no optimal level."""
    data = np.vstack([np.asarray(sum_h0j_nj, float), np.asarray(n0, float)])
    stats = reblock(data)
    opt = find_optimal_block(data.shape[1], stats)
    converged = not any(math.isnan(o) for o in opt)
    level = int(max(opt)) if converged else len(stats) - 1
    _lvl, n, mean, cov, se, _e = stats[level]
    e, err = ratio(mean[0], mean[1], se[0], se[1], cov[0, 1], n)
    if not converged:
        e, err = stats[0][2][0] / stats[0][2][1], float("nan")
    return {"energy": e, "error": err, "level": level, "converged": converged, "ndata": n}


def mean_error(series):
    """Reblocked mean and standard error of one column."""
    stats = reblock(np.asarray(series, float))
    opt = find_optimal_block(len(series), stats)[0]
    converged = not math.isnan(opt)
    level = int(opt) if converged else len(stats) - 1
    _lvl, n, mean, _cov, se, _e = stats[level]
    if not converged:
        return {"mean": float(stats[0][2][0]), "error": float("nan"), "level": level, "converged": False}
    return {"mean": float(mean[0]), "error": float(se[0]), "level": level, "converged": True}


def efficiency(error, seconds):
    """Statistical efficiency 1/(sigma^2 t)."""
    return 1.0 / (error * error * seconds)


def inefficiency(error, walkers, iterations, tau):
    """The reference's inefficiency: sigma * sqrt(N_w * iterations * tau)."""
    return error * math.sqrt(walkers * iterations * tau)
