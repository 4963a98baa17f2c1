"""Each benchmark configuration's own workload under a parity check.

* mol26 i-FCIQMC (26 spatial orbitals, 10 electrons, real_64, initiator 3,
  tau 1e-3, renorm): production lockstep against the oracle from 1e5 walkers
  for 30 cycles (walker lists, counts and reference population identical
  after every cycle); then the full 1e7-walker list grown on the device checked
  for what the reference's list guarantees (sorted and unique, no empty
  entries, nparticles = sum |N_i|, K_ii = <D|H|D> - H00) and a sampled spawned
  list annihilated into it identically by the engine and the oracle.
* Batched Hubbard-k (64 rings of 10-18 sites, fillings near half, U/t
  uniform on [1, 8], integer populations): the batched engine follows all 64
  solo chains bit for bit for 30 cycles.
* mol40 learned proposal (40 orbitals, 14 electrons: two-word determinants,
  nbasis 80): every pgen of the sampler equals the exact mixture probability,
  the draw frequencies match it (chi-square over the excitations and over
  their hole pairs), and inference, loss and gradient match the oracle.
* mol60 replica pair (60 orbitals, 20 electrons: nbasis 120, two words,
  two walker spaces): production lockstep against the oracle.

The reference cycle these follow is src/fciqmc.f90:307-399; the initiator and
replica rules are src/spawn_data.F90:937-1101."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from hande_amd import _native as N
from hande_amd import api, synthetic
from helpers import download, oracle, odownload, production_lockstep

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu

P = C.POINTER
RF = 1 << 31


def _mol(tmp_path, name):
    norb, nel, seed = synthetic.MOLECULAR_CONFIGS[name]
    path = synthetic.write_fcidump(os.path.join(str(tmp_path), f"FCIDUMP_{name}"), norb, nel, seed)
    return api.read_in(int_file=path, nel=nel, ms=0)


def _params(seed, excit_gen=N.FQ_EXCIT_RENORM, spaces=1, cap=1 << 18, real=True, initiator=True):
    p = N.FqParams()
    p.nspaces = spaces
    p.initiator = int(initiator)
    p.initiator_pop = 3.0
    p.quadrature_initiator = 1
    p.real_amplitudes = int(real)
    p.pop_bit_shift = 31 if real else 0
    p.spawn_cutoff = int(np.ceil(0.01 * RF)) if real else 0
    p.excit_gen = excit_gen
    p.rng_mode = N.FQ_RNG_PHILOX
    p.philox_seed = seed
    p.main_capacity = cap
    p.spawn_capacity = cap
    p.device = 0
    return p


def _ptr(a, t):
    return a.ctypes.data_as(P(t))


# ---------------------------------------------------------------------------
# mol26 (i-FCIQMC, real_64, initiator 3, renorm, tau 1e-3)
# ---------------------------------------------------------------------------
def test_mol26_production_lockstep(tmp_path):
    system = _mol(tmp_path, "mol26")
    print("[config] mol26 lockstep from 1e5 walkers", flush=True)
    outs = production_lockstep(system, _params(26), 30, 100000, tau=1e-3, label="mol26", every=10)
    assert max(o.nspawn_events for o in outs) > 2 * 4096  # several spawning, sort and run tiles


def _grow(lib, system, p, target, tau=1e-3, max_cycles=4000):
    eng = C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(eng)), "fq_create")
    f0 = system.f0()
    pop0 = np.array([int(target * 1e-3) * RF], dtype=np.int64)
    N.check(lib.fq_load(eng, _ptr(f0, C.c_uint64), _ptr(pop0, C.c_int64), _ptr(np.zeros(1), C.c_double), 1), "load")
    # zero shift, then a shift that holds the population once it passes the
    # target (preparation only: any shift keeps the list's invariants)
    sh = (C.c_double * 2)(0.0, 0.0)
    eo = (C.c_double * 2)(0.0, 0.0)
    out = N.FqCycleOut()
    for i in range(max_cycles):
        N.check(lib.fq_cycle(eng, sh, eo, tau, C.byref(out)), "fq_cycle")
        assert not out.spawn_overflow and not out.main_overflow
        if out.nparticles[0] >= target:
            break
    return eng, out, i + 1


def test_mol26_full_size_list_properties(tmp_path):
    """The 1e7-walker list: invariants of the reference's main list, then one
    annihilation of a sampled spawned list against the oracle."""
    lib, orc = N.lib(), oracle()
    system = _mol(tmp_path, "mol26")
    p = _params(2626, cap=1 << 25)
    eng, out, cycles = _grow(lib, system, p, 1.0e7)
    try:
        print(f"[config] mol26 grown to N={out.nparticles[0]:.4g} on {out.nstates} determinants in {cycles} cycles",
              flush=True)
        assert out.nparticles[0] >= 1.0e7
        st, po, da = download(lib, eng, 1, 1)
    finally:
        lib.fq_destroy(eng)
    n = len(da)
    assert n == out.nstates and n > 1_000_000
    assert np.all(st[1:] > st[:-1]), "main list not sorted and unique"
    pc = np.array([bin(int(x)).count("1") for x in st[:: max(1, n // 20000)]])
    assert np.all(pc == 10)
    assert np.all(po != 0), "empty entry kept"
    assert np.abs(po).sum() / RF == out.nparticles[0]
    rng = np.random.default_rng(7)
    sample = rng.choice(n, 2000, replace=False)
    kii = np.array([system.diagonal(st[i:i + 1]) - system.H00 for i in sample])
    np.testing.assert_allclose(da[sample], kii, rtol=0, atol=1e-11)
    # spawned list: 3e5 entries, half on occupied determinants (repeats
    # included), half on single and double excitations of them; populations
    # of 0.5 to 3 walkers (some below one: rounding draws), random initiator flags
    ns = 300000
    hit = st[rng.integers(0, n, ns // 2)]
    base = st[rng.integers(0, n, ns - ns // 2)]
    fresh = np.empty_like(base)
    for k, d in enumerate(base):
        d = int(d)
        occ = [b for b in range(52) if (d >> b) & 1]
        for o in rng.choice(occ, 1 + (k % 2), replace=False):  # spin-conserving singles and doubles
            vir = [b for b in range(int(o) & 1, 52, 2) if not (d >> b) & 1]
            d ^= (1 << int(o)) | (1 << int(rng.choice(vir)))
        fresh[k] = d
    sk = np.concatenate([hit, fresh])
    perm = rng.permutation(ns)
    sk = np.ascontiguousarray(sk[perm])
    sp = (rng.uniform(0.5, 3.0, ns) * RF).astype(np.int64) * rng.choice([-1, 1], ns)
    sf = rng.integers(0, 2, ns).astype(np.int64)
    p2 = _params(2627, cap=1 << 25)
    eng = C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p2), C.byref(eng)), "fq_create")
    h = orc.orc_create(system.view, C.byref(p2), 1)
    orc.orc_set_production(h, p2.philox_seed)
    orc.orc_set_stable_order(h, 1)
    try:
        args = (_ptr(st, C.c_uint64), _ptr(po, C.c_int64), _ptr(da, C.c_double), n)
        N.check(lib.fq_load(eng, *args), "fq_load")
        orc.orc_load(h, *args)
        o1, o2 = N.FqCycleOut(), N.FqCycleOut()
        N.check(lib.fq_annihilate(eng, _ptr(sk, C.c_uint64), _ptr(sp, C.c_int64), _ptr(sf, C.c_int64), ns,
                                  C.byref(o1)), "fq_annihilate")
        orc.orc_annihilate(h, _ptr(sk, C.c_uint64), _ptr(sp, C.c_int64), _ptr(sf, C.c_int64), ns, C.byref(o2))
        print(f"  annihilation: {n} + {ns} spawned -> {o1.nstates} (oracle {o2.nstates})", flush=True)
        assert o1.nstates == o2.nstates
        # the oracle keeps the reference's running FP64 total, inexact above
        # 2^53 encoded units; the engine sums the encoded populations exactly
        np.testing.assert_allclose(o1.nparticles[0], o2.nparticles[0], rtol=1e-12)
        a, b = download(lib, eng, 1, 1), odownload(orc, h, 1, 1)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert o1.nstates > n
    finally:
        lib.fq_destroy(eng)
        orc.orc_destroy(h)


# ---------------------------------------------------------------------------
# batched Hubbard-k: 64 rings of 10-18 sites
# ---------------------------------------------------------------------------
def _c3_instances(n=64, seed=2024):
    """n Hubbard-k rings of 10-18 sites, fillings near half, U/t uniform on
    [1, 8], targets log-uniform on [1e4, 1e6] (the batched workload)."""
    rng = np.random.default_rng(seed)
    decks, targets = [], []
    for _ in range(n):
        L = int(rng.choice([10, 12, 14, 16, 18]))
        nel = L + int(rng.choice([-2, 0, 0, 2]))
        U = float(rng.uniform(1.0, 8.0))
        decks.append(dict(electrons=nel, lattice=[[L]], ms=0, sym=1, U=U, t=1.0))
        targets.append(float(10 ** rng.uniform(4, 6)))
    return decks, np.array(targets)


def test_batched_64_instances_follow_solo_chains():
    lib = N.lib()
    decks, targets = _c3_instances()
    systems = [api.hubbard_k(**d) for d in decks]
    K = len(systems)
    assert K == 64 and {s.view.nbasis for s in systems} == {20, 24, 28, 32, 36}
    arr = (N.FqSystem * K)(*[s.view for s in systems])
    seed = 6464
    beng = C.c_void_p()
    N.check(lib.fq_create_batch(arr, K, C.byref(_params(seed, cap=1 << 24, real=False, initiator=False)),
                                C.byref(beng)), "fq_create_batch")
    solo = []
    try:
        inst = np.arange(K, dtype=np.int32)
        f0 = np.array([int(s.f0()[0]) for s in systems], dtype=np.uint64)
        pops = np.maximum(10, (targets / 1000).astype(np.int64))
        dat = np.zeros(K)
        N.check(lib.fq_load_batch(beng, _ptr(inst, C.c_int32), _ptr(f0, C.c_uint64), _ptr(pops, C.c_int64),
                                  _ptr(dat, C.c_double), K), "load")
        for k, s in enumerate(systems):
            eng = C.c_void_p()
            N.check(lib.fq_create(s.view, C.byref(_params(seed + k, cap=1 << 21, real=False, initiator=False)),
                                  C.byref(eng)), "fq_create")
            solo.append(eng)
            one = f0[k:k + 1].copy()
            pk = pops[k:k + 1].copy()
            N.check(lib.fq_load(eng, _ptr(one, C.c_uint64), _ptr(pk, C.c_int64), _ptr(dat, C.c_double), 1), "load")
        rng = np.random.default_rng(5)
        taus = np.full(K, 0.005)
        shifts = np.zeros(K)
        eolds = rng.uniform(-3.0, 0.0, K)
        hold = 5000.0  # walkers per instance at which the shift starts to vary
        vary = np.zeros(K, bool)
        n_old = np.zeros(K)
        outs = (N.FqCycleOut * K)()
        for cyc in range(30):
            if cyc > 0:
                # a shift update every cycle (damping 0.5) once an instance passes
                # `hold` (strongly correlated rings double each cycle at zero
                # shift): the same function of the same outputs for the batched
                # and the solo chains
                n = np.array([outs[k].nparticles[0] for k in range(K)])
                upd = vary & (n_old > 0)
                shifts[upd] -= np.log(n[upd] / n_old[upd]) * 0.5 / 0.005
                vary |= n > hold
                n_old[:] = n
            N.check(lib.fq_cycle_batch(beng, _ptr(shifts, C.c_double), _ptr(eolds, C.c_double),
                                       _ptr(taus, C.c_double), outs), "fq_cycle_batch")
            bdet, bpop, bkii = download(lib, beng, 1, 1)
            binst = bdet >> np.uint64(48)
            for k, eng in enumerate(solo):
                o = N.FqCycleOut()
                N.check(lib.fq_cycle(eng, (C.c_double * 2)(shifts[k], 0.0), (C.c_double * 2)(eolds[k], 0.0),
                                     taus[k], C.byref(o)), "fq_cycle")
                b = outs[k]
                assert not (b.spawn_overflow or b.main_overflow or o.spawn_overflow or o.main_overflow), (cyc, k)
                assert (b.nstates, b.nspawn_events, b.ndeath, b.nattempts) == \
                       (o.nstates, o.nspawn_events, o.ndeath, o.nattempts), (cyc, k)
                assert b.nparticles[0] == o.nparticles[0], (cyc, k)
                assert b.proj_energy[0] == o.proj_energy[0] and b.D0_population[0] == o.D0_population[0], (cyc, k)
                sdet, spop, skii = download(lib, eng, 1, 1)
                sel = binst == np.uint64(k)
                assert np.array_equal(bdet[sel] & np.uint64((1 << 48) - 1), sdet), (cyc, k)
                assert np.array_equal(bpop[sel], spop) and np.array_equal(bkii[sel], skii), (cyc, k)
            if cyc % 10 == 0:
                print(f"  cycle {cyc}: {len(bdet)} determinants over 64 instances", flush=True)
        assert all(outs[k].nstates > 20 for k in range(K))
    finally:
        lib.fq_destroy(beng)
        for eng in solo:
            lib.fq_destroy(eng)


# ---------------------------------------------------------------------------
# mol40 learned proposal: two-word determinants, nbasis 80
# ---------------------------------------------------------------------------
def _learned_engine(system, seed=40, mix=0.1):
    lib = N.lib()
    p = _params(seed, N.FQ_EXCIT_LEARNED, cap=1 << 16, initiator=False)
    p.proposal_mix = mix
    p.proposal_min_attempts = 0
    eng = C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(eng)), "fq_create")
    return lib, eng


def _occ_of(det_words, nb):
    return [o for o in range(1, nb + 1) if (int(det_words[(o - 1) // 64]) >> ((o - 1) % 64)) & 1]


def _dets_from_occ(occs, nb):
    d = np.zeros((len(occs), (nb + 63) // 64), dtype=np.uint64)
    for k, ol in enumerate(occs):
        for o in ol:
            d[k, (o - 1) // 64] |= np.uint64(1) << np.uint64((o - 1) % 64)
    return d


def test_mol40_learned_sampler_exact_two_words(tmp_path):
    system = _mol(tmp_path, "mol40")
    v = system.view
    nb, nel = v.nbasis, v.nel
    assert nb == 80 and v.nwords == 2
    orb_sym = [v.orb_sym[k] for k in range(nb)]
    lib, eng = _learned_engine(system)
    try:
        params = M.init_params(nb, 41) * 3.0  # strongly non-uniform weights and affinity
        params = np.ascontiguousarray(params, dtype=np.float32)
        N.check(lib.fq_mlp_set_params(eng, _ptr(params, C.c_float), len(params)), "set_params")
        aff = np.zeros((nb, nb), dtype=np.float32)
        N.check(lib.fq_mlp_affinity(eng, _ptr(aff, C.c_float), nb * nb), "affinity")
        # the reference determinant and one whose occupied set spans both words
        hf = _occ_of(system.f0(), nb)
        rng = np.random.default_rng(40)
        al = sorted(int(x) for x in rng.choice(np.arange(1, nb + 1, 2), nel // 2, replace=False))
        be = sorted(int(x) for x in rng.choice(np.arange(2, nb + 1, 2), nel // 2, replace=False))
        mixed = sorted(al + be)
        assert any(o <= 64 for o in mixed) and any(o > 64 for o in mixed)
        pdouble = v.pattempt_double
        nsamp = 2_000_000
        for k, occ in enumerate([hf, mixed]):
            det = _dets_from_occ([occ], nb)
            wrow = np.zeros((1, 2 * nb), dtype=np.float32)
            N.check(lib.fq_mlp_forward(eng, _ptr(det, C.c_uint64), 1, _ptr(wrow, C.c_float)), "forward")
            exact, null = M.enumerate_doubles(wrow[0], occ, orb_sym, nb, 0.1, aff)
            exc = np.zeros((nsamp, 6), dtype=np.int32)
            pg = np.zeros(nsamp)
            hm = np.zeros(nsamp)
            N.check(lib.fq_probe_excitations(eng, _ptr(det, C.c_uint64), nsamp, 500 + k, _ptr(exc, C.c_int32),
                                             _ptr(pg, C.c_double), _ptr(hm, C.c_double)), "probe")
            dbl = (exc[:, 0] == 1) & (exc[:, 1] == 2)
            keys = [(int(e[2]), int(e[3]), int(e[4]), int(e[5])) for e in exc[dbl]]
            for key, p in zip(keys[:200000], pg[dbl][:200000]):
                q = exact[key]
                assert abs(p - pdouble * q) <= 1e-12 * p, (key, p, pdouble * q)
            # frequencies: doubles drawn per double attempt; null (rejected) draws included
            ndbl_att = int(round(nsamp * pdouble))
            counts = {}
            for key in keys:
                counts[key] = counts.get(key, 0) + 1
            big = [kk for kk, q in exact.items() if ndbl_att * q >= 5]
            stat = sum((counts.get(kk, 0) - ndbl_att * exact[kk]) ** 2 / (ndbl_att * exact[kk]) for kk in big)
            z1 = (stat - (len(big) - 1)) / np.sqrt(2 * (len(big) - 1))
            # hole-pair marginals over every double
            pair_p, pair_c = {}, {}
            for kk, q in exact.items():
                pair_p[kk[:2]] = pair_p.get(kk[:2], 0.0) + q
            for kk, c in counts.items():
                pair_c[kk[:2]] = pair_c.get(kk[:2], 0) + c
            stat2 = sum((pair_c.get(kk, 0) - ndbl_att * q) ** 2 / (ndbl_att * q) for kk, q in pair_p.items())
            z2 = (stat2 - (len(pair_p) - 1)) / np.sqrt(2 * (len(pair_p) - 1))
            print(f"  det {k}: {len(exact)} doubles, {len(counts)} drawn, null {null:.3f}; chi-square z "
                  f"{z1:.2f} over {len(big)} excitations, {z2:.2f} over {len(pair_p)} hole pairs", flush=True)
            assert abs(z1) < 4.0 and abs(z2) < 4.0
            assert abs(len(keys) / ndbl_att - (1.0 - null)) < 5 * np.sqrt(null * (1 - null) / ndbl_att)
    finally:
        lib.fq_destroy(eng)


def test_mol40_inference_and_gradient_match_oracle(tmp_path):
    system = _mol(tmp_path, "mol40")
    v = system.view
    nb, nel = v.nbasis, v.nel
    eps = np.array([v.sp_eigv[k] for k in range(nb)])
    orb_sym = [v.orb_sym[k] for k in range(nb)]
    lib, eng = _learned_engine(system)
    try:
        params = np.ascontiguousarray(M.init_params(nb, 43), dtype=np.float32)
        N.check(lib.fq_mlp_set_params(eng, _ptr(params, C.c_float), len(params)), "set_params")
        rng = np.random.default_rng(44)
        occs = []
        for _ in range(256):
            al = rng.choice(np.arange(1, nb + 1, 2), nel // 2, replace=False)
            be = rng.choice(np.arange(2, nb + 1, 2), nel // 2, replace=False)
            occs.append(sorted(int(x) for x in np.concatenate([al, be])))
        dets = _dets_from_occ(occs, nb)
        got = np.zeros((len(dets), 2 * nb), dtype=np.float32)
        N.check(lib.fq_mlp_forward(eng, _ptr(dets, C.c_uint64), len(dets), _ptr(got, C.c_float)), "forward")
        z = M.forward(params, M.occupations(dets, nb), eps, bf16_inference=True)
        err = np.abs(np.log(got.astype(np.float64)) - np.clip(z, -M.CLIP, M.CLIP))
        print(f"[config] mol40 inference: max |dlogit| {err.max():.3e}, mean {err.mean():.3e}", flush=True)
        assert err.max() < 3e-2 and err.mean() < 2e-3
        exc = np.zeros((len(occs), 4), dtype=np.int32)
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
        w = rng.uniform(0.0, 2.0, len(occs))
        flat = (np.arange(len(w)) % 4 == 1).astype(np.int32)
        loss = C.c_double()
        grad = np.zeros(len(params), dtype=np.float32)
        N.check(lib.fq_mlp_loss_grad(eng, _ptr(dets, C.c_uint64), _ptr(exc, C.c_int32), _ptr(w, C.c_double),
                                     _ptr(flat, C.c_int32), len(w), C.byref(loss), _ptr(grad, C.c_float)),
                "fq_mlp_loss_grad")
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


# ---------------------------------------------------------------------------
# mol60 replica pair: nbasis 120, two words, two walker spaces
# ---------------------------------------------------------------------------
def test_mol60_replica_pair_production_lockstep(tmp_path):
    system = _mol(tmp_path, "mol60")
    assert system.view.nbasis == 120 and system.view.nwords == 2
    print("[config] mol60 replica pair lockstep from 2e4 walkers per replica", flush=True)
    outs = production_lockstep(system, _params(60, spaces=2), 30, 20000, tau=1e-3, label="mol60", every=10)
    assert outs[-1].nstates > 4000
    assert max(o.nspawn_events for o in outs) > 4096
