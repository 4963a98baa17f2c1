"""Learned double-excitation proposal and the excitation probe on the GPU.

* Inference (bf16 MFMA, fp32 accumulation) against the oracle's bf16-emulating
  float64 forward pass, at a padded (nbasis 52) and an unpadded (16) width.
* Training loss and gradient (fp32 kernels) against the oracle's float64
  loss and gradient, which are themselves pinned by central differences.
* Exactness of the samplers: every pgen the GPU generator reports equals the
  exact probability (the oracle's closed form for the learned mixture, path
  enumeration for heat-bath), and the observed frequencies of 4e5 draws match
  those probabilities (chi-square).
* Training reduces the loss, and the trained mixture spends less variance per
  attempt than the uniform fallback alone: sum_ex |H_ex|^2 / q(ex), computed
  exactly over every double from the weights, drops.
* Production runs with the learned generator reproduce the exact ground
  state energy of a small synthetic system: the mean of eight independent
  runs within 2 standard errors."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import exact_small
from hande_amd import api, synthetic
from helpers import replicate_estimates
from hande_amd import _native as N

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu

P = C.POINTER


def _system(tmp_path, norb, nel, seed):
    path = synthetic.write_fcidump(os.path.join(str(tmp_path), f"FCIDUMP_{norb}_{nel}"), norb, nel, seed)
    return api.read_in(int_file=path, nel=nel, ms=0)


def _engine(system, excit_gen=N.FQ_EXCIT_LEARNED, seed=5, cap=1 << 16, mix=0.1, min_attempts=0):
    lib = N.lib()
    p = N.FqParams()
    p.nspaces = 1
    p.initiator = 0
    p.real_amplitudes = 1
    p.pop_bit_shift = 31
    p.spawn_cutoff = int(np.ceil(0.01 * (1 << 31)))
    p.excit_gen = excit_gen
    p.rng_mode = N.FQ_RNG_PHILOX
    p.philox_seed = seed
    p.main_capacity = cap
    p.spawn_capacity = cap
    p.device = 0
    p.proposal_mix = mix
    p.proposal_min_attempts = min_attempts
    eng = C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(eng)), "fq_create")
    return lib, eng


def _orb_info(system):
    v = system.view
    nb = v.nbasis
    eps = np.array([v.sp_eigv[k] for k in range(nb)])
    orb_sym = [v.orb_sym[k] for k in range(nb)]
    return nb, eps, orb_sym


def _random_dets(nb, nel, n, rng):
    dets = np.zeros((n, (nb + 63) // 64), dtype=np.uint64)
    occs = []
    for k in range(n):
        al = rng.choice(np.arange(1, nb + 1, 2), nel // 2, replace=False)
        be = rng.choice(np.arange(2, nb + 1, 2), nel // 2, replace=False)
        ol = sorted(int(o) for o in np.concatenate([al, be]))
        for o in ol:
            dets[k, (o - 1) // 64] |= np.uint64(1) << np.uint64((o - 1) % 64)
        occs.append(ol)
    return dets, occs


def _forward(lib, eng, dets, nb):
    out = np.zeros((len(dets), 2 * nb), dtype=np.float32)
    d = np.ascontiguousarray(dets)
    N.check(lib.fq_mlp_forward(eng, d.ctypes.data_as(P(C.c_uint64)), len(dets), out.ctypes.data_as(P(C.c_float))),
            "fq_mlp_forward")
    return out


def _affinity(lib, eng, nb):
    out = np.zeros((nb, nb), dtype=np.float32)
    N.check(lib.fq_mlp_affinity(eng, out.ctypes.data_as(P(C.c_float)), nb * nb), "fq_mlp_affinity")
    return out


def _set_params(lib, eng, params):
    params = np.ascontiguousarray(params, dtype=np.float32)
    N.check(lib.fq_mlp_set_params(eng, params.ctypes.data_as(P(C.c_float)), len(params)), "fq_mlp_set_params")


@pytest.mark.parametrize("norb,nel", [(8, 4), (26, 10)])
def test_mlp_inference_matches_oracle(tmp_path, norb, nel):
    print(f"[mlp] inference {norb}o{nel}e", flush=True)
    system = _system(tmp_path, norb, nel, 300 + norb)
    nb, eps, _ = _orb_info(system)
    lib, eng = _engine(system)
    try:
        n = C.c_int64()
        N.check(lib.fq_mlp_nparams(eng, C.byref(n)), "nparams")
        assert n.value == M.sizes(nb)[3]
        params = M.init_params(nb, 17)
        _set_params(lib, eng, params)
        back = np.zeros(n.value, dtype=np.float32)
        N.check(lib.fq_mlp_get_params(eng, back.ctypes.data_as(P(C.c_float)), n.value), "get")
        assert np.array_equal(back, params)
        rng = np.random.default_rng(3)
        dets, _occs = _random_dets(nb, nel, 200, rng)
        got = _forward(lib, eng, dets, nb)
        z = M.forward(params, M.occupations(dets, nb), eps, bf16_inference=True)
        ref = np.clip(z, -M.CLIP, M.CLIP)
        lg = np.log(got.astype(np.float64))
        err = np.abs(lg - ref)
        print(f"  max |dlogit| {err.max():.3e}, mean {err.mean():.3e}", flush=True)
        assert err.max() < 3e-2 and err.mean() < 2e-3
    finally:
        lib.fq_destroy(eng)


def _random_samples(nb, nel, orb_sym, n, rng):
    dets, occs = _random_dets(nb, nel, n, rng)
    exc = np.zeros((n, 4), dtype=np.int32)
    for k, ol in enumerate(occs):
        occset = set(ol)
        while True:
            i, j = sorted(int(x) for x in rng.choice(ol, 2, replace=False))
            a = int(rng.choice([o for o in range(1, nb + 1) if o not in occset]))
            cb = M.second_class(occset, orb_sym, nb, i, j, a)
            if cb:
                b = int(rng.choice(cb))
                exc[k] = (i, j, min(a, b), max(a, b))
                break
    w = rng.uniform(0.0, 2.0, n)
    w[::7] = 0.0  # rejected samples carry no weight
    return dets, occs, exc, w


def test_mlp_loss_and_gradient_match_oracle(tmp_path):
    print("[mlp] loss/grad", flush=True)
    system = _system(tmp_path, 8, 4, 808)
    nb, eps, orb_sym = _orb_info(system)
    lib, eng = _engine(system)
    try:
        params = M.init_params(nb, 23)
        _set_params(lib, eng, params)
        rng = np.random.default_rng(9)
        dets, occs, exc, w = _random_samples(nb, 4, orb_sym, 300, rng)
        flat = (np.arange(len(w)) % 5 == 2).astype(np.int32)  # some samples scored with flat weights
        loss = C.c_double()
        grad = np.zeros(len(params), dtype=np.float32)
        N.check(lib.fq_mlp_loss_grad(eng, np.ascontiguousarray(dets).ctypes.data_as(P(C.c_uint64)),
                                     exc.ctypes.data_as(P(C.c_int32)), w.ctypes.data_as(P(C.c_double)),
                                     flat.ctypes.data_as(P(C.c_int32)), len(w),
                                     C.byref(loss), grad.ctypes.data_as(P(C.c_float))), "fq_mlp_loss_grad")
        ref_loss, ref_grad = M.loss_and_grad(params.astype(np.float64), M.occupations(dets, nb), eps, orb_sym,
                                             [tuple(int(x) for x in e) for e in exc], w, flat.astype(bool))
        print(f"  loss {loss.value:.8f} oracle {ref_loss:.8f}", flush=True)
        assert abs(loss.value - ref_loss) < 1e-5 * abs(ref_loss)
        _d_in, _d_out, off, _n = M.sizes(nb)
        for name, (o, cnt) in off.items():
            g, r = grad[o:o + cnt].astype(np.float64), ref_grad[o:o + cnt]
            scale = np.abs(r).max()
            print(f"  {name}: max |g - ref| / max|ref| = {np.abs(g - r).max() / scale:.2e}", flush=True)
            assert np.abs(g - r).max() <= 2e-4 * scale + 1e-9
    finally:
        lib.fq_destroy(eng)


def test_mlp_variance_objective_matches_oracle(tmp_path):
    """fq_mlp_variance_grad (the spawning-variance objective sum w |H| / q)
    against the oracle's float64 loss and gradient, which central differences
    pin; the network's fp32 matrix-core training products at 2e-4 of each
    block's largest gradient."""
    print("[mlp] variance objective loss/grad", flush=True)
    system = _system(tmp_path, 8, 4, 808)
    nb, eps, orb_sym = _orb_info(system)
    lib, eng = _engine(system)
    try:
        params = M.init_params(nb, 29)
        _set_params(lib, eng, params)
        rng = np.random.default_rng(13)
        dets, occs, exc, w = _random_samples(nb, 4, orb_sym, 300, rng)
        habs = rng.uniform(0.001, 0.3, len(w))
        flat = (np.arange(len(w)) % 5 == 1).astype(np.int32)
        loss = C.c_double()
        grad = np.zeros(len(params), dtype=np.float32)
        N.check(lib.fq_mlp_variance_grad(eng, np.ascontiguousarray(dets).ctypes.data_as(P(C.c_uint64)),
                                         exc.ctypes.data_as(P(C.c_int32)), w.ctypes.data_as(P(C.c_double)),
                                         habs.ctypes.data_as(P(C.c_double)), flat.ctypes.data_as(P(C.c_int32)),
                                         len(w), C.byref(loss), grad.ctypes.data_as(P(C.c_float))),
                "fq_mlp_variance_grad")
        ref_loss, ref_grad = M.loss_and_grad(params.astype(np.float64), M.occupations(dets, nb), eps, orb_sym,
                                             [tuple(int(x) for x in e) for e in exc], w, flat.astype(bool), habs)
        print(f"  loss {loss.value:.8f} oracle {ref_loss:.8f}", flush=True)
        assert abs(loss.value - ref_loss) < 1e-5 * abs(ref_loss)
        _d_in, _d_out, off, _n = M.sizes(nb)
        for name, (o, cnt) in off.items():
            g, r = grad[o:o + cnt].astype(np.float64), ref_grad[o:o + cnt]
            scale = np.abs(r).max()
            print(f"  {name}: max |g - ref| / max|ref| = {np.abs(g - r).max() / scale:.2e}", flush=True)
            assert np.abs(g - r).max() <= 2e-4 * scale + 1e-9
    finally:
        lib.fq_destroy(eng)


def test_online_training_runs_inside_cycles_and_lowers_variance(tmp_path):
    """fq_mlp_online: production cycles train the proposal from their own
    walker lists (objective: spawning variance); the parameters change only on
    the training cycles, two engines with the same seeds follow the same
    chain, and the variance objective on fresh heat-bath samples falls."""
    print("[mlp] online training", flush=True)
    system = _system(tmp_path, 10, 6, 1010)
    nb, eps, orb_sym = _orb_info(system)
    engs = []
    try:
        outs = []
        for rep in range(2):
            lib, eng = _engine(system)
            engs.append(eng)
            N.check(lib.fq_mlp_online(eng, 5, 2, 512, 3e-3, 1), "fq_mlp_online")
            f0 = system.f0()
            pop = np.array([200 << 31], dtype=np.int64)  # real amplitudes, 31 fraction bits
            N.check(lib.fq_load(eng, f0.ctypes.data_as(P(C.c_uint64)), pop.ctypes.data_as(P(C.c_int64)),
                                np.zeros(1).ctypes.data_as(P(C.c_double)), 1), "fq_load")
            n = C.c_int64()
            N.check(lib.fq_mlp_nparams(eng, C.byref(n)), "nparams")
            p_prev = np.zeros(n.value, dtype=np.float32)
            N.check(lib.fq_mlp_get_params(eng, p_prev.ctypes.data_as(P(C.c_float)), n.value), "get")
            rows = []
            for cyc in range(20):
                o = N.FqCycleOut()
                N.check(lib.fq_cycle(eng, (C.c_double * 2)(0.0, 0.0), (C.c_double * 2)(0.0, 0.0), 0.01, C.byref(o)),
                        "fq_cycle")
                p = np.zeros(n.value, dtype=np.float32)
                N.check(lib.fq_mlp_get_params(eng, p.ctypes.data_as(P(C.c_float)), n.value), "get")
                changed = not np.array_equal(p, p_prev)
                assert changed == (cyc % 5 == 0), cyc
                p_prev = p
                rows.append((o.nstates, o.nattempts, o.nspawn_events, o.nparticles[0], o.proj_energy[0]))
            outs.append(rows)
        assert outs[0] == outs[1]
        # the objective on fresh samples: trained further it falls
        losses = np.zeros(60)
        N.check(N.lib().fq_mlp_train(engs[0], 512, 60, 3e-3, 77, losses.ctypes.data_as(P(C.c_double))), "train")
        print(f"  variance objective: first {losses[:10].mean():.4g}, last {losses[-10:].mean():.4g}", flush=True)
        assert losses[-10:].mean() < losses[:10].mean()
    finally:
        for eng in engs:
            N.lib().fq_destroy(eng)


def _probe(lib, eng, det, n, seed):
    exc = np.zeros((n, 6), dtype=np.int32)
    pg = np.zeros(n)
    hm = np.zeros(n)
    d = np.ascontiguousarray(det, dtype=np.uint64)
    N.check(lib.fq_probe_excitations(eng, d.ctypes.data_as(P(C.c_uint64)), n, seed, exc.ctypes.data_as(P(C.c_int32)),
                                     pg.ctypes.data_as(P(C.c_double)), hm.ctypes.data_as(P(C.c_double))), "probe")
    return exc, pg, hm


def _chi_square(counts, probs, n):
    """Pearson statistic over outcomes with expected count >= 5, as a z-score."""
    keys = [k for k in probs if n * probs[k] >= 5]
    stat = sum((counts.get(k, 0) - n * probs[k]) ** 2 / (n * probs[k]) for k in keys)
    dof = max(len(keys) - 1, 1)
    return (stat - dof) / np.sqrt(2 * dof)


@pytest.mark.parametrize("excit_gen", ["learned", "learned_flat", "heat_bath", "renorm"])
def test_sampler_pgen_exact_and_frequencies(tmp_path, excit_gen):
    """learned_flat: a determinant without a network row (flat weights, the
    learned affinity)."""
    print(f"[probe] {excit_gen}", flush=True)
    norb, nel = 6, 4
    system = _system(tmp_path, norb, nel, 611)
    nb, _eps, orb_sym = _orb_info(system)
    flat = excit_gen == "learned_flat"
    if flat:
        excit_gen = "learned"
    gen = {"learned": N.FQ_EXCIT_LEARNED, "heat_bath": N.FQ_EXCIT_HEAT_BATH, "renorm": N.FQ_EXCIT_RENORM}[excit_gen]
    lib, eng = _engine(system, gen, min_attempts=2 if flat else 0)
    try:
        if excit_gen == "learned":
            _set_params(lib, eng, M.init_params(nb, 41) * 3.0)  # strongly non-uniform weights
        rng = np.random.default_rng(12)
        dets, occs = _random_dets(nb, nel, 2, rng)
        f0 = int(system.f0()[0])
        dets = np.vstack([np.array([[f0]], dtype=np.uint64), dets])
        occs = [[o for o in range(1, nb + 1) if (f0 >> (o - 1)) & 1]] + occs
        nsamp = 400000
        pdouble = system.view.pattempt_double
        for k in range(len(dets)):
            exc, pg, hm = _probe(lib, eng, dets[k], nsamp, 100 + k)
            ok = exc[:, 0] == 1
            assert ok.sum() > nsamp // 4
            keys = [(int(e[1]), int(e[2]), int(e[3]), int(e[4]), int(e[5])) for e in exc[ok]]
            claimed = {}
            for key, p in zip(keys, pg[ok]):
                if key in claimed:
                    # pgen is a function of the excitation (heat-bath sums its four
                    # ordered paths starting from the drawn one: last-bit differences)
                    assert abs(claimed[key] - p) <= 1e-14 * p, key
                claimed[key] = p
            if excit_gen == "learned":
                wrow = np.ones(2 * nb, dtype=np.float32) if flat else _forward(lib, eng, dets[k:k + 1], nb)[0]
                aff = _affinity(lib, eng, nb)
                assert np.array_equal(aff, M.affinity(M.init_params(nb, 41) * 3.0, nb))
                for key, p in claimed.items():
                    if key[0] == 2:
                        q, _qm, _qu = M.q_double(wrow, occs[k], orb_sym, nb, key[1], key[2], key[3], key[4], 0.1,
                                                 aff)
                        assert abs(p - pdouble * q) <= 1e-12 * p, (key, p, pdouble * q)
            if excit_gen == "heat_bath":
                import helpers
                import test_heat_bath as T
                orc = helpers.oracle()
                p_ = N.FqParams()
                p_.nspaces, p_.excit_gen, p_.rng_mode = 1, N.FQ_EXCIT_HEAT_BATH, N.FQ_RNG_REPLAY
                p_.main_capacity = p_.spawn_capacity = 1 << 10
                h = orc.orc_create(system.view, C.byref(p_), 7)
                try:
                    tabs = T._tables(orc, h)
                finally:
                    orc.orc_destroy(h)
                h1, eri = synthetic.synthetic_integrals(norb, 611)[1:]
                outcomes, _null = T._enumerate(tabs, h1, eri, occs[k])
                for key, p in claimed.items():
                    nx = key[0]
                    ek = (nx, tuple(key[1:1 + nx]), tuple(key[3:3 + nx]))
                    assert abs(p - outcomes[ek]) <= 1e-10 * p, (key, p, outcomes[ek])
            counts = {}
            for key in keys:
                counts[key] = counts.get(key, 0) + 1
            z = _chi_square(counts, claimed, nsamp)
            print(f"  det {k}: {len(claimed)} excitations, accepted {ok.mean():.3f}, chi-square z = {z:.2f}",
                  flush=True)
            assert abs(z) < 5.0
    finally:
        lib.fq_destroy(eng)


def _all_doubles(occ, nb, orb_sym):
    occset = set(occ)
    for x in range(len(occ)):
        for y in range(x + 1, len(occ)):
            i, j = occ[x], occ[y]
            for a in range(1, nb + 1):
                if a in occset:
                    continue
                for b in M.second_class(occset, orb_sym, nb, i, j, a):
                    if b > a:
                        yield i, j, a, b


def _variance_cost(rows, aff, occs, nb, orb_sym, h1, eri, mix):
    """sum over determinants and doubles of |H|^2 / q: the second moment of the
    spawning estimator per attempt (lower is better; q ~ |H| is optimal)."""
    import exact_small as X
    total = 0.0
    for wrow, occ in zip(rows, occs):
        for (i, j, a, b) in _all_doubles(occ, nb, orb_sym):
            hval = abs(X._anti(eri, a, b, i, j))
            if hval == 0.0:
                continue
            q = M.q_double(wrow, occ, orb_sym, nb, i, j, a, b, mix, aff)[0]
            total += hval * hval / q
    return total


def test_training_reduces_loss_and_spawning_variance(tmp_path):
    print("[mlp] training", flush=True)
    norb, nel, seed = 10, 6, 1006
    system = _system(tmp_path, norb, nel, seed)
    nb, _eps, orb_sym = _orb_info(system)
    lib, eng = _engine(system, cap=1 << 18)
    try:
        # a walker distribution: 40 cycles from 200 walkers on the reference
        rf = 1 << 31
        f0 = system.f0()
        pop0 = np.array([200 * rf], dtype=np.int64)
        dat0 = np.zeros(1)
        N.check(lib.fq_load(eng, f0.ctypes.data_as(P(C.c_uint64)), pop0.ctypes.data_as(P(C.c_int64)),
                            dat0.ctypes.data_as(P(C.c_double)), 1), "fq_load")
        sh = (C.c_double * 2)(0.0, 0.0)
        eo = (C.c_double * 2)(0.0, 0.0)
        out = N.FqCycleOut()
        for _ in range(40):
            N.check(lib.fq_cycle(eng, sh, eo, 0.01, C.byref(out)), "fq_cycle")
        print(f"  walkers {out.nparticles[0]:.1f} on {out.nstates} determinants", flush=True)
        rng = np.random.default_rng(2)
        dets, occs = _random_dets(nb, nel, 12, rng)
        f0w = int(f0[0])
        dets = np.vstack([np.array([[f0w]], dtype=np.uint64), dets])
        occs = [[o for o in range(1, nb + 1) if (f0w >> (o - 1)) & 1]] + occs
        _h, h1, eri = synthetic.synthetic_integrals(norb, seed)
        before = _forward(lib, eng, dets, nb)
        aff_before = _affinity(lib, eng, nb)
        assert np.all(aff_before == 1.0)
        steps = 300
        losses = np.zeros(steps)
        N.check(lib.fq_mlp_train(eng, 4096, steps, 2e-3, 77, losses.ctypes.data_as(P(C.c_double))), "fq_mlp_train")
        print(f"  loss first 20 {losses[:20].mean():.4f} last 20 {losses[-20:].mean():.4f}", flush=True)
        assert losses[-20:].mean() < losses[:20].mean() - 0.05
        after = _forward(lib, eng, dets, nb)
        aff_after = _affinity(lib, eng, nb)
        cost_unif = _variance_cost(before, aff_before, occs, nb, orb_sym, h1, eri, 1.0)
        cost_init = _variance_cost(before, aff_before, occs, nb, orb_sym, h1, eri, 0.1)
        cost_trained = _variance_cost(after, aff_after, occs, nb, orb_sym, h1, eri, 0.1)
        print(f"  sum |H|^2/q: uniform {cost_unif:.4f}, untrained mix {cost_init:.4f}, trained mix "
              f"{cost_trained:.4f} (gain over uniform {cost_unif / cost_trained:.2f}x)", flush=True)
        assert cost_trained < 0.9 * cost_unif
        assert cost_trained < cost_init
    finally:
        lib.fq_destroy(eng)


@pytest.mark.parametrize("min_attempts", [0, 4])
def test_learned_projected_energy_matches_exact(tmp_path, min_attempts):
    """min_attempts 4: most determinants skip the network and draw doubles with
    flat weights and the affinity; the estimate stays unbiased."""
    norb, nel, seed = 8, 4, 600
    exact, _hf, weight = exact_small.ground_state(norb, nel, seed, with_weight=True)
    assert weight > 0.5
    system = _system(tmp_path, norb, nel, seed)
    print(f"[production] synthetic 8o4e learned, min_attempts {min_attempts}", flush=True)
    q = dict(tau=0.01, init_pop=10, mc_cycles=20, nreports=400, target_population=2000,
             shift_damping=0.05, state_size=100000, spawned_state_size=100000, excit_gen="learned",
             proposal_min_attempts=min_attempts)
    e, se, _sh, _sse = replicate_estimates(system, q, range(99, 107))
    print(f"  E_proj={e:.6f} +- {se:.6f} (8 runs), exact {exact:.6f}", flush=True)
    assert abs(e - exact) < 2 * se + 1e-4
