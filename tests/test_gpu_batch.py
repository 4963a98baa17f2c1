"""Batched independent instances (fq_create_batch / fq_cycle_batch).

A batch of Hubbard-k systems of different sizes, fillings and U/t propagated
in one engine must follow, instance by instance and bit for bit, the chains of
solo engines keyed with philox_seed + k: the same walker lists (determinants,
populations, K_ii) and the same per-cycle estimators, spawn/death counts and
attempts."""
import ctypes as C

import numpy as np
import pytest

from hande_amd import api
from hande_amd import _native as N
from helpers import download

pytestmark = pytest.mark.gpu

P = C.POINTER
MASK = (1 << 48) - 1

DECKS = [
    dict(electrons=4, lattice=[[6]], ms=0, sym=1, U=1.5, t=1.0),
    dict(electrons=6, lattice=[[8]], ms=0, sym=1, U=4.0, t=1.0),
    dict(electrons=4, lattice=[[2, 0], [0, 3]], ms=0, sym=1, U=2.5, t=1.0),
    dict(electrons=8, lattice=[[10]], ms=0, sym=1, U=6.0, t=1.0),
    dict(electrons=6, lattice=[[12]], ms=0, sym=1, U=3.0, t=1.0),
]


def _params(seed, cap=1 << 17):
    p = N.FqParams()
    p.nspaces = 1
    p.initiator = 0
    p.real_amplitudes = 0
    p.pop_bit_shift = 0
    p.spawn_cutoff = 0
    p.excit_gen = N.FQ_EXCIT_RENORM
    p.rng_mode = N.FQ_RNG_PHILOX
    p.philox_seed = seed
    p.main_capacity = cap
    p.spawn_capacity = cap
    p.device = 0
    return p


@pytest.mark.parametrize("initiator", [False, True])
def test_batch_follows_solo_chains(initiator):
    print(f"[batch] {len(DECKS)} instances, initiator={initiator}", flush=True)
    lib = N.lib()
    systems = [api.hubbard_k(**d) for d in DECKS]
    K = len(systems)
    arr = (N.FqSystem * K)(*[s.view for s in systems])
    seed = 4242
    pb = _params(seed)
    pb.initiator = int(initiator)
    pb.initiator_pop = 2.0
    beng = C.c_void_p()
    N.check(lib.fq_create_batch(arr, K, C.byref(pb), C.byref(beng)), "fq_create_batch")
    solo = []
    try:
        inst = np.arange(K, dtype=np.int32)
        f0 = np.array([int(s.f0()[0]) for s in systems], dtype=np.uint64)
        pops = np.array([12, 8, 20, 5, 10], dtype=np.int64)
        dat = np.zeros(K)
        N.check(lib.fq_load_batch(beng, inst.ctypes.data_as(P(C.c_int32)), f0.ctypes.data_as(P(C.c_uint64)),
                                  pops.ctypes.data_as(P(C.c_int64)), dat.ctypes.data_as(P(C.c_double)), K), "load")
        for k, s in enumerate(systems):
            p = _params(seed + k)
            p.initiator = int(initiator)
            p.initiator_pop = 2.0
            eng = C.c_void_p()
            N.check(lib.fq_create(s.view, C.byref(p), C.byref(eng)), "fq_create")
            solo.append(eng)
            one = np.array([f0[k]], dtype=np.uint64)
            pk = np.array([pops[k]], dtype=np.int64)
            N.check(lib.fq_load(eng, one.ctypes.data_as(P(C.c_uint64)), pk.ctypes.data_as(P(C.c_int64)),
                                dat.ctypes.data_as(P(C.c_double)), 1), "fq_load")
        taus = np.array([0.01, 0.02, 0.015, 0.01, 0.02])
        shifts = np.array([0.1, -0.2, 0.0, 0.3, -0.1])
        eolds = np.array([-0.5, -1.0, -0.3, -2.0, -0.7])
        outs = (N.FqCycleOut * K)()
        for cyc in range(60):
            N.check(lib.fq_cycle_batch(beng, shifts.ctypes.data_as(P(C.c_double)), eolds.ctypes.data_as(P(C.c_double)),
                                       taus.ctypes.data_as(P(C.c_double)), outs), "fq_cycle_batch")
            blist = download(lib, beng, 1, 1)
            bdet = blist[0]
            for k, eng in enumerate(solo):
                sh = (C.c_double * 2)(shifts[k], 0.0)
                eo = (C.c_double * 2)(eolds[k], 0.0)
                o = N.FqCycleOut()
                N.check(lib.fq_cycle(eng, sh, eo, taus[k], C.byref(o)), "fq_cycle")
                b = outs[k]
                assert (b.nstates, b.nspawn_events, b.ndeath, b.nattempts) == \
                       (o.nstates, o.nspawn_events, o.ndeath, o.nattempts), (cyc, k)
                assert b.nparticles[0] == o.nparticles[0], (cyc, k)
                assert b.proj_energy[0] == o.proj_energy[0] and b.D0_population[0] == o.D0_population[0], (cyc, k)
                sdet, spop, skii = download(lib, eng, 1, 1)
                sel = (bdet >> np.uint64(48)) == np.uint64(k)
                assert np.array_equal(bdet[sel] & np.uint64(MASK), sdet), (cyc, k)
                assert np.array_equal(blist[1][sel], spop), (cyc, k)
                assert np.array_equal(blist[2][sel], skii), (cyc, k)
            if cyc % 20 == 0:
                print(f"  cycle {cyc}: states {[outs[k].nstates for k in range(K)]}", flush=True)
        assert all(outs[k].nstates > 5 for k in range(K))
    finally:
        lib.fq_destroy(beng)
        for eng in solo:
            lib.fq_destroy(eng)


def _cycle_fields(o):
    return (o.nstates, o.nspawn_events, o.ndeath, o.nattempts, o.nparticles[0], o.proj_energy[0],
            o.D0_population[0], o.spawn_overflow, o.main_overflow)


def test_graph_cycles_match_ordinary_cycles():
    """fq_batch_graph: the same batch run as graph launches (through the
    tuner's candidate schedules, then its choice, across several launch-bound
    steps as the populations grow) and as ordinary launches gives the same
    outputs and walker lists every cycle.  The ordinary engine's cycles and
    both list downloads run between graph replays, so each replay follows
    eager work and host synchronisations."""
    print(f"[graph] {len(DECKS)} instances, graph vs ordinary cycles", flush=True)
    lib = N.lib()
    systems = [api.hubbard_k(**d) for d in DECKS]
    K = len(systems)
    arr = (N.FqSystem * K)(*[s.view for s in systems])
    engs = []
    try:
        for graph in (0, 1):
            eng = C.c_void_p()
            N.check(lib.fq_create_batch(arr, K, C.byref(_params(99, cap=1 << 20)), C.byref(eng)), "fq_create_batch")
            engs.append(eng)
            N.check(lib.fq_batch_graph(eng, graph), "fq_batch_graph")
            inst = np.arange(K, dtype=np.int32)
            f0 = np.array([int(s.f0()[0]) for s in systems], dtype=np.uint64)
            pops = np.array([1200, 800, 2000, 500, 1000], dtype=np.int64)
            N.check(lib.fq_load_batch(eng, inst.ctypes.data_as(P(C.c_int32)), f0.ctypes.data_as(P(C.c_uint64)),
                                      pops.ctypes.data_as(P(C.c_int64)), np.zeros(K).ctypes.data_as(P(C.c_double)),
                                      K), "load")
        taus = np.array([0.01, 0.02, 0.015, 0.01, 0.02])
        shifts = np.array([0.3, 0.2, 0.3, 0.4, 0.3])  # growing populations: the launch bounds change
        eolds = np.array([-0.5, -1.0, -0.3, -2.0, -0.7])
        outs = [(N.FqCycleOut * K)(), (N.FqCycleOut * K)()]
        for cyc in range(80):
            if cyc == 50:
                shifts[:] = -0.5  # and shrink
            for g in (0, 1):
                N.check(lib.fq_cycle_batch(engs[g], shifts.ctypes.data_as(P(C.c_double)),
                                           eolds.ctypes.data_as(P(C.c_double)), taus.ctypes.data_as(P(C.c_double)),
                                           outs[g]), "fq_cycle_batch")
            for k in range(K):
                assert _cycle_fields(outs[0][k]) == _cycle_fields(outs[1][k]), (cyc, k)
            a, b = download(lib, engs[0], 1, 1), download(lib, engs[1], 1, 1)
            for x, y in zip(a, b):
                assert np.array_equal(x, y), cyc
            if cyc % 20 == 0:
                print(f"  cycle {cyc}: {len(a[0])} determinants", flush=True)
        info = N.FqScheduleOut()
        N.check(lib.fq_batch_schedule(engs[1], C.byref(info)), "fq_batch_schedule")
        print(f"  graph cycles {info.graph_cycles}, ordinary {info.eager_cycles}, graphs {info.graphs}, "
              f"octaves tuned {info.tuned_octaves}, last octave {info.octave} chose {info.chosen} "
              f"({info.digit_bits}-bit digits, spawn blocks {info.spawn_blocks}), "
              f"candidates ms {[round(info.cand_ms[q], 3) for q in range(info.ncand)]}", flush=True)
        assert info.graph == 1 and info.tuned_octaves >= 1 and -1 <= info.chosen < info.ncand
        # the first cycle after the load runs ordinary launches (its attempt
        # bound is unknown), as do cycles whose attempts pass the staging capacity
        assert info.eager_cycles >= 1 and info.graph_cycles + info.eager_cycles == 80
        assert info.graph_cycles >= 60 and info.graphs >= info.ncand
    finally:
        for eng in engs:
            lib.fq_destroy(eng)
