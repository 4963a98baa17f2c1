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
