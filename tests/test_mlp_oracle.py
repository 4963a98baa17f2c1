"""The learned-proposal oracle (oracle/mlp_oracle.py) pinned by mathematics:
  * bf16 rounding equals torch's float32 -> bfloat16 conversion (RNE);
  * the analytic gradient of the training loss equals central differences
    of the loss itself;
  * the closed-form mixture probability q of every double equals the sum over
    the sampling paths of both mixture components, and the paths' total
    probability (doubles + null attempts) is 1;
  * with a flat hole-particle affinity the network component factorises into
    a hole-pair and a particle-pair probability (the product proposal)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 10, 5000), rng.normal(0, 1e-3, 1000), [0.0, -0.0, 1.0, 65504.0, 1e30]])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    assert np.array_equal(M.bf16(x), ref)


def _random_case(nb, nel, nsamp, seed, nsym=1):
    rng = np.random.default_rng(seed)
    orb_sym = [int(rng.integers(nsym)) for _ in range((nb + 1) // 2) for _ in range(2)][:nb]
    eps = np.sort(rng.uniform(-2, 3, nb))
    occ = np.zeros((nsamp, nb))
    samples = []
    while len(samples) < nsamp:
        s = len(samples)
        al = rng.choice(np.arange(1, nb + 1, 2), nel // 2, replace=False)
        be = rng.choice(np.arange(2, nb + 1, 2), nel // 2, replace=False)
        ol = sorted(int(o) for o in np.concatenate([al, be]))
        occset = set(ol)
        i, j = sorted(rng.choice(ol, 2, replace=False))
        unocc = [o for o in range(1, nb + 1) if o not in occset]
        a = int(rng.choice(unocc))
        cb = M.second_class(occset, orb_sym, nb, int(i), int(j), a)
        if not cb:
            continue
        b = int(rng.choice(cb))
        occ[s, [o - 1 for o in ol]] = 1.0
        samples.append((int(i), int(j), min(a, b), max(a, b)))
    return occ, eps, orb_sym, samples, rng.uniform(0.1, 2.0, nsamp)


@pytest.mark.parametrize("nsym,with_flat,variance", [(1, False, False), (2, False, False), (2, True, False),
                                                     (1, False, True), (2, True, True)])
def test_loss_gradient_matches_central_differences(nsym, with_flat, variance):
    nb, nel = 12, 4
    occ, eps, orb_sym, samples, w = _random_case(nb, nel, 24, 3 + nsym, nsym)
    flat = (np.arange(len(w)) % 3 == 0) if with_flat else None
    habs = np.random.default_rng(3).uniform(0.01, 0.5, len(w)) if variance else None
    params = M.init_params(nb, 11).astype(np.float64)
    loss, grad = M.loss_and_grad(params, occ, eps, orb_sym, samples, w, flat, habs)
    assert abs(loss - M.loss_only(params, occ, eps, orb_sym, samples, w, flat, habs)) < 1e-12 * max(1.0, abs(loss))
    rng = np.random.default_rng(1)
    _d_in, _d_out, off, n = M.sizes(nb)
    # probe entries of every block
    idx = np.concatenate([off[k][0] + rng.choice(off[k][1], 12, replace=False) for k in off])
    h = 1e-5
    for k in idx:
        p1 = params.copy()
        p1[k] += h
        p2 = params.copy()
        p2[k] -= h
        fd = (M.loss_only(p1, occ, eps, orb_sym, samples, w, flat, habs)
              - M.loss_only(p2, occ, eps, orb_sym, samples, w, flat, habs)) / (2 * h)
        assert abs(fd - grad[k]) < 1e-7 * max(1.0, abs(loss)) + 1e-5 * abs(fd), (k, fd, grad[k])


@pytest.mark.parametrize("nsym,mix,flat", [(1, 0.1, False), (2, 0.1, False), (2, 0.5, False), (2, 0.1, True)])
def test_mixture_probability_matches_path_sum(nsym, mix, flat):
    nb, nel = 12, 4
    rng = np.random.default_rng(7 + nsym)
    orb_sym = [int(rng.integers(nsym)) for _ in range(nb // 2) for _ in range(2)]
    occ_list = sorted([1, 2] + [int(rng.choice(range(3, nb + 1, 2))), int(rng.choice(range(4, nb + 1, 2)))])
    wrow = np.exp(rng.normal(0, 1.5, 2 * nb)).astype(np.float32)
    aff = None if flat else np.exp(rng.normal(0, 1.5, (nb, nb))).astype(np.float32)
    out, null = M.enumerate_doubles(wrow, occ_list, orb_sym, nb, mix, aff)
    assert abs(sum(out.values()) + null - 1.0) < 1e-12
    assert len(out) > 10
    for (i, j, a, b), p in out.items():
        q, q_mlp, q_unif = M.q_double(wrow, occ_list, orb_sym, nb, i, j, a, b, mix, aff)
        assert abs(q - p) <= 1e-12 * p, ((i, j, a, b), q, p)
        assert q_unif > 0.0


def test_flat_affinity_is_the_product_proposal():
    nb = 12
    rng = np.random.default_rng(5)
    orb_sym = [int(rng.integers(2)) for _ in range(nb // 2) for _ in range(2)]
    occ_list = [1, 2, 5, 8]
    occset = set(occ_list)
    unocc = [o for o in range(1, nb + 1) if o not in occset]
    wrow = np.exp(rng.normal(0, 1.5, 2 * nb))
    e, f = wrow[:nb], wrow[nb:]
    Zh = e[[o - 1 for o in occ_list]].sum()
    Zp = f[[o - 1 for o in unocc]].sum()
    out, _null = M.enumerate_doubles(wrow, occ_list, orb_sym, nb, 0.0)
    for (i, j, a, b), p in out.items():
        hp = e[i - 1] / Zh * e[j - 1] / (Zh - e[i - 1]) + e[j - 1] / Zh * e[i - 1] / (Zh - e[j - 1])
        zb = sum(f[o - 1] for o in M.second_class(occset, orb_sym, nb, i, j, a))
        za = sum(f[o - 1] for o in M.second_class(occset, orb_sym, nb, i, j, b))
        pp = f[a - 1] / Zp * f[b - 1] / zb + f[b - 1] / Zp * f[a - 1] / za
        assert abs(p - hp * pp) <= 1e-12 * p
