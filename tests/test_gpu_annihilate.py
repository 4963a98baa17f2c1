"""Annihilation given an identical spawned list is bit-exact against the
oracle (sort, self-annihilation with the plain and initiator rules, merge into
the main list, rounding and compaction, insertion with new K_ii).

Replay mode follows the reference's quicksort order of equal keys and draws
the rounding numbers from the dSFMT stream; production mode uses the radix
sort with the earliest-spawned entry first in each run, compared against the
oracle's stable-order variant on lists whose outcome needs no random numbers."""
import ctypes as C

import numpy as np
import pytest

from hande_amd import api, _native as N
from helpers import oracle, engine_params

pytestmark = pytest.mark.gpu


def _mk(system, initiator, real, rng_mode):
    q = {"tau": 0.01, "rng_seed": 11, "initiator": initiator, "real_amplitudes": real}
    st = api.qmc_in(q, rng_mode=rng_mode)
    p = engine_params(system.view, st)
    p.main_capacity = 1 << 15
    p.spawn_capacity = 1 << 15
    return st, p


def _random_state(rng, system, nmain, nspawn, rf, real, big):
    """Sorted main list and a spawned list that hits it, misses it and repeats keys."""
    nb, nel = system.view.nbasis, system.view.nel
    keys = set()
    while len(keys) < nmain + nspawn:
        occ = rng.choice(nb, nel, replace=False)
        keys.add(int(sum(1 << int(o) for o in occ)))
    keys = sorted(keys)
    main = np.array(sorted(rng.choice(keys, nmain, replace=False)), dtype=np.uint64)
    pops_main = rng.integers(1, 20, nmain) * rng.choice([-1, 1], nmain)
    if real:
        # big: multiples of 3 walkers, so sums with the +-1.5k-walker spawns below
        # are 0 or >= 1.5 walkers and never need a rounding draw
        pops_main = pops_main * (3 * rf) if big else pops_main * rf + rng.integers(-rf // 2, rf // 2, nmain)
    # spawned: draw with repetition from main keys and fresh keys
    pool = np.array(keys, dtype=np.uint64)
    sk = rng.choice(pool, nspawn, replace=True)
    sp = rng.integers(1, 4, nspawn) * rng.choice([-1, 1], nspawn)
    if real:
        if big:
            sp = sp * (3 * rf // 2)  # >= rf: no rounding; low 32 bits may read negative
        else:
            sp = sp * rng.integers(rf // 100, 3 * rf, nspawn)
    flags = rng.integers(0, 2, nspawn)
    return main, pops_main.astype(np.int64), sk, sp.astype(np.int64), flags.astype(np.int64)


def _run(lib, orc, system, st, p, main, pops, sk, sp, sf, stable):
    eng = C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(eng)), "fq_create")
    h = orc.orc_create(system.view, C.byref(p), st.rng_seed)
    orc.orc_set_stable_order(h, 1 if stable else 0)
    kii = np.array([system.diagonal(np.array([m], dtype=np.uint64)) - system.H00 for m in main])
    ptr = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    N.check(lib.fq_load(eng, ptr(main, C.c_uint64), ptr(pops, C.c_int64), ptr(kii, C.c_double), len(main)), "load")
    orc.orc_load(h, ptr(main, C.c_uint64), ptr(pops, C.c_int64), ptr(kii, C.c_double), len(main))
    if p.rng_mode == N.FQ_RNG_REPLAY:
        stream = np.zeros(1 << 16)
        lib.fqh_dsfmt_fill(st.rng_seed, ptr(stream, C.c_double), len(stream))
        N.check(lib.fq_push_rng(eng, ptr(stream, C.c_double), len(stream)), "push")
    o1, o2 = N.FqCycleOut(), N.FqCycleOut()
    N.check(lib.fq_annihilate(eng, ptr(sk, C.c_uint64), ptr(sp, C.c_int64), ptr(sf, C.c_int64), len(sk),
                              C.byref(o1)), "fq_annihilate")
    orc.orc_annihilate(h, ptr(sk, C.c_uint64), ptr(sp, C.c_int64), ptr(sf, C.c_int64), len(sk), C.byref(o2))
    assert o1.nstates == o2.nstates
    assert o1.nparticles[0] == o2.nparticles[0]
    assert o1.rng_consumed == o2.rng_consumed
    n = o1.nstates
    a = [np.zeros(n, np.uint64), np.zeros(n, np.int64), np.zeros(n)]
    got = C.c_int64()
    N.check(lib.fq_download(eng, ptr(a[0], C.c_uint64), ptr(a[1], C.c_int64), ptr(a[2], C.c_double), n,
                            C.byref(got)), "download")
    b = [np.zeros(n, np.uint64), np.zeros(n, np.int64), np.zeros(n)]
    orc.orc_download(h, ptr(b[0], C.c_uint64), ptr(b[1], C.c_int64), ptr(b[2], C.c_double))
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])  # new K_ii from the device Slater-Condon diagonal
    lib.fq_destroy(eng)
    orc.orc_destroy(h)
    return n


@pytest.fixture(scope="module")
def hub():
    return api.hubbard_k(electrons=6, lattice=[[3, 0], [0, 3]], ms=0, sym="tot_sym", U=2)


@pytest.mark.parametrize("initiator,real", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_annihilate_replay_matches_oracle(hub, initiator, real, seed):
    lib, orc = N.lib(), oracle()
    rng = np.random.default_rng(seed)
    st, p = _mk(hub, initiator, real, "replay")
    rf = 1 << p.pop_bit_shift
    main, pops, sk, sp, sf = _random_state(rng, hub, 300, 900, rf, real, big=False)
    assert _run(lib, orc, hub, st, p, main, pops, sk, sp, sf, stable=False) > 0


@pytest.mark.parametrize("initiator", [False, True])
@pytest.mark.parametrize("seed", [4, 5])
def test_annihilate_production_matches_stable_oracle(hub, initiator, seed):
    """Radix sort + run reduction; real_64 populations >= 1 walker (no rounding
    draws) with bit 31 set, so the first-entry event sign matters."""
    lib, orc = N.lib(), oracle()
    rng = np.random.default_rng(seed)
    st, p = _mk(hub, initiator, True, "philox")
    rf = 1 << p.pop_bit_shift
    main, pops, sk, sp, sf = _random_state(rng, hub, 400, 3000, rf, True, big=True)
    assert _run(lib, orc, hub, st, p, main, pops, sk, sp, sf, stable=True) > 0


@pytest.mark.parametrize("mode", ["philox", "replay"])
def test_annihilate_long_runs_across_tiles(hub, mode):
    """Runs far longer than a reduction tile (1024 sorted entries): keys spawned
    1023, 1024, 1025, 2500 and 7000 times, beside single spawns, with random
    initiator flags and signs (first-entry event quirk included): each run's
    totals cross tile boundaries, and a tile may hold no run head at all."""
    print(f"[annihilate] long runs, {mode}", flush=True)
    lib, orc = N.lib(), oracle()
    rng = np.random.default_rng(77)
    st, p = _mk(hub, True, True, mode)
    p.spawn_capacity = 1 << 16
    rf = 1 << p.pop_bit_shift
    main, pops, sk, _, _ = _random_state(rng, hub, 300, 2000, rf, True, big=True)
    heavy = rng.choice(np.unique(sk), 5, replace=False)
    reps = np.repeat(heavy, [1023, 1024, 1025, 2500, 7000])
    sk = rng.permutation(np.concatenate([sk, reps]))
    ns = len(sk)
    sp = (rng.integers(1, 4, ns) * rng.choice([-1, 1], ns) * (3 * rf // 2)).astype(np.int64)
    sf = rng.integers(0, 2, ns).astype(np.int64)
    assert _run(lib, orc, hub, st, p, main, pops, sk, sp, sf, stable=(mode == "philox")) > 0


def test_annihilate_molecular_new_kii(tmp_path):
    """New determinants get K_ii = <D|H|D> - H00 from the device diagonal rule."""
    from helpers import deck
    lib, orc = N.lib(), oracle()
    system, _ = deck("H2O-RHF-cc-pVTZ", str(tmp_path))
    st, p = _mk(system, False, False, "replay")
    rng = np.random.default_rng(9)
    main, pops, sk, sp, sf = _random_state(rng, system, 200, 600, 1, False, big=False)
    assert _run(lib, orc, system, st, p, main, pops, sk, sp, sf, stable=False) > 0


def _random_keys(rng, nbits, nel, count):
    """count distinct random nbits-bit keys with nel bits set (vectorised rejection)."""
    table = np.array([bin(b).count("1") for b in range(256)], dtype=np.int64)
    out = np.zeros(0, dtype=np.uint64)
    while len(out) < count:
        x = rng.integers(0, 1 << nbits, 8 * count, dtype=np.uint64)
        c = table[x.view(np.uint8).reshape(-1, 8)].sum(axis=1)
        out = np.unique(np.concatenate([out, x[c == nel]]))
    return rng.permutation(out)[:count]


@pytest.mark.parametrize("nmain,nspawn", [(600_000, 1_200_000), (600_000, 3_000)])
def test_annihilate_production_large_lists(nmain, nspawn):
    """Lists of 1e6 entries: every prefix sum spans many tiles and several
    chunks of the tile-sum scan; radix sort over 32-bit keys in 4 passes.
    With few spawns the main-list window of a merge block is too wide to stage
    in LDS (searched in global memory); with many it is staged."""
    print(f"[annihilate] production, {nmain} main + {nspawn} spawned", flush=True)
    lib, orc = N.lib(), oracle()
    system = api.hubbard_k(electrons=10, lattice=[[4, 0], [0, 4]], ms=0, sym="tot_sym", U=4)
    rng = np.random.default_rng(21)
    st, p = _mk(system, True, True, "philox")
    p.main_capacity = 1 << 22
    p.spawn_capacity = 1 << 22
    rf = 1 << p.pop_bit_shift
    keys = _random_keys(rng, 32, 10, nmain + nspawn // 2)
    main = np.sort(keys[:nmain])
    pops = (rng.integers(1, 20, nmain) * rng.choice([-1, 1], nmain) * (3 * rf)).astype(np.int64)
    sk = rng.choice(keys, nspawn, replace=True)
    sp = (rng.integers(1, 4, nspawn) * rng.choice([-1, 1], nspawn) * (3 * rf // 2)).astype(np.int64)
    sf = rng.integers(0, 2, nspawn).astype(np.int64)
    eng = C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(eng)), "fq_create")
    h = orc.orc_create(system.view, C.byref(p), st.rng_seed)
    orc.orc_set_stable_order(h, 1)
    ptr = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    kii = np.zeros(nmain)
    try:
        N.check(lib.fq_load(eng, ptr(main, C.c_uint64), ptr(pops, C.c_int64), ptr(kii, C.c_double), nmain), "load")
        orc.orc_load(h, ptr(main, C.c_uint64), ptr(pops, C.c_int64), ptr(kii, C.c_double), nmain)
        o1, o2 = N.FqCycleOut(), N.FqCycleOut()
        N.check(lib.fq_annihilate(eng, ptr(sk, C.c_uint64), ptr(sp, C.c_int64), ptr(sf, C.c_int64), nspawn,
                                  C.byref(o1)), "fq_annihilate")
        orc.orc_annihilate(h, ptr(sk, C.c_uint64), ptr(sp, C.c_int64), ptr(sf, C.c_int64), nspawn, C.byref(o2))
        assert o1.nstates == o2.nstates and o1.nparticles[0] == o2.nparticles[0]
        n = o1.nstates
        a = [np.zeros(n, np.uint64), np.zeros(n, np.int64), np.zeros(n)]
        got = C.c_int64()
        N.check(lib.fq_download(eng, ptr(a[0], C.c_uint64), ptr(a[1], C.c_int64), ptr(a[2], C.c_double), n,
                                C.byref(got)), "download")
        b = [np.zeros(n, np.uint64), np.zeros(n, np.int64), np.zeros(n)]
        orc.orc_download(h, ptr(b[0], C.c_uint64), ptr(b[1], C.c_int64), ptr(b[2], C.c_double))
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        print(f"  {n} determinants after annihilation", flush=True)
    finally:
        lib.fq_destroy(eng)
        orc.orc_destroy(h)


def test_annihilate_production_two_word_rank_keys():
    """128 spin orbitals and 20 electrons: the production sort orders 78-bit
    combinatorial ranks held in two words."""
    print("[annihilate] production, 128 spin orbitals, two-word rank keys", flush=True)
    lib, orc = N.lib(), oracle()
    system = api.hubbard_k(electrons=20, lattice=[[64]], ms=0, sym="tot_sym", U=2)
    assert system.view.nwords == 2 and system.view.nbasis == 128
    rng = np.random.default_rng(33)
    st, p = _mk(system, True, True, "philox")
    p.main_capacity = 1 << 18
    p.spawn_capacity = 1 << 18
    rf = 1 << p.pop_bit_shift
    nmain, nspawn = 20_000, 60_000
    keys = set()
    while len(keys) < nmain + nspawn // 2:
        occ = rng.choice(128, 20, replace=False)
        keys.add(sum(1 << int(o) for o in occ))
    keys = sorted(keys)
    main_keys = sorted(rng.choice(len(keys), nmain, replace=False))
    main = np.array([[keys[i] & ((1 << 64) - 1), keys[i] >> 64] for i in main_keys], dtype=np.uint64)
    pops = (rng.integers(1, 20, nmain) * rng.choice([-1, 1], nmain) * (3 * rf)).astype(np.int64)
    pick = rng.integers(0, len(keys), nspawn)
    sk = np.array([[keys[i] & ((1 << 64) - 1), keys[i] >> 64] for i in pick], dtype=np.uint64)
    sp = (rng.integers(1, 4, nspawn) * rng.choice([-1, 1], nspawn) * (3 * rf // 2)).astype(np.int64)
    sf = rng.integers(0, 2, nspawn).astype(np.int64)
    eng = C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(eng)), "fq_create")
    h = orc.orc_create(system.view, C.byref(p), st.rng_seed)
    orc.orc_set_stable_order(h, 1)
    ptr = lambda a, t: np.ascontiguousarray(a).ctypes.data_as(C.POINTER(t))
    main = np.ascontiguousarray(main)
    sk = np.ascontiguousarray(sk)
    kii = np.zeros(nmain)
    try:
        N.check(lib.fq_load(eng, ptr(main, C.c_uint64), ptr(pops, C.c_int64), ptr(kii, C.c_double), nmain), "load")
        orc.orc_load(h, ptr(main, C.c_uint64), ptr(pops, C.c_int64), ptr(kii, C.c_double), nmain)
        o1, o2 = N.FqCycleOut(), N.FqCycleOut()
        N.check(lib.fq_annihilate(eng, ptr(sk, C.c_uint64), ptr(sp, C.c_int64), ptr(sf, C.c_int64), nspawn,
                                  C.byref(o1)), "fq_annihilate")
        orc.orc_annihilate(h, ptr(sk, C.c_uint64), ptr(sp, C.c_int64), ptr(sf, C.c_int64), nspawn, C.byref(o2))
        assert o1.nstates == o2.nstates and o1.nparticles[0] == o2.nparticles[0]
        n = o1.nstates
        a = [np.zeros(2 * n, np.uint64), np.zeros(n, np.int64), np.zeros(n)]
        got = C.c_int64()
        N.check(lib.fq_download(eng, a[0].ctypes.data_as(C.POINTER(C.c_uint64)),
                                a[1].ctypes.data_as(C.POINTER(C.c_int64)), a[2].ctypes.data_as(C.POINTER(C.c_double)),
                                n, C.byref(got)), "download")
        b = [np.zeros(2 * n, np.uint64), np.zeros(n, np.int64), np.zeros(n)]
        orc.orc_download(h, b[0].ctypes.data_as(C.POINTER(C.c_uint64)), b[1].ctypes.data_as(C.POINTER(C.c_int64)),
                         b[2].ctypes.data_as(C.POINTER(C.c_double)))
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    finally:
        lib.fq_destroy(eng)
        orc.orc_destroy(h)
