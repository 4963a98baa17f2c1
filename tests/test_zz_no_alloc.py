"""fq_cycle_noalloc (the cycle with its scratch in the library's fixed device
array, no device allocation) against fq_cycle, bit for bit.

Two engines with the same parameters and seed hold the same walker list; one
cycles through fq_cycle, the other through fq_cycle_noalloc.  Cycle outputs
and walker lists must be identical after each call: first on a list large
enough for the multi-tile prefix sums and several radix-sort tiles, then, on
the next call, on a smaller list loaded into both engines."""
import ctypes as C
import os

import numpy as np
import pytest

from hande_amd import _native as N
from hande_amd import api, synthetic
from helpers import download

pytestmark = pytest.mark.gpu


def _params(seed, real, initiator, excit_gen=N.FQ_EXCIT_RENORM, cap=1 << 20):
    p = N.FqParams()
    p.nspaces = 1
    p.initiator = int(initiator)
    p.initiator_pop = 3.0
    p.quadrature_initiator = 0
    p.real_amplitudes = int(real)
    p.pop_bit_shift = 31 if real else 0
    p.spawn_cutoff = int(np.ceil(0.01 * (1 << 31))) if real else 0
    p.excit_gen = excit_gen
    p.rng_mode = N.FQ_RNG_PHILOX
    p.philox_seed = seed
    p.main_capacity = cap
    p.spawn_capacity = cap
    p.device = 0
    return p


def _grow(lib, system, p, init_pop, target, tau):
    """A walker list of about `target` walkers, grown from the reference
    determinant at zero shift (the list both engines then load)."""
    W = system.view.nwords
    eng = C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(eng)), "fq_create")
    try:
        rf = 1 << p.pop_bit_shift
        f0 = np.array([system.view.f0[w] for w in range(W)], dtype=np.uint64)
        pop0 = np.array([int(init_pop) * rf], dtype=np.int64)
        N.check(lib.fq_load(eng, f0.ctypes.data_as(C.POINTER(C.c_uint64)), pop0.ctypes.data_as(C.POINTER(C.c_int64)),
                            np.zeros(1).ctypes.data_as(C.POINTER(C.c_double)), 1), "fq_load")
        sh = (C.c_double * 2)(0.0, 0.0)
        out = N.FqCycleOut()
        for _ in range(3000):
            N.check(lib.fq_cycle(eng, sh, sh, tau, C.byref(out)), "fq_cycle")
            if out.nparticles[0] > target:
                break
        return download(lib, eng, W, 1)
    finally:
        lib.fq_destroy(eng)


def _load(lib, eng, lst):
    st, po, da = lst
    N.check(lib.fq_load(eng, st.ctypes.data_as(C.POINTER(C.c_uint64)), po.ctypes.data_as(C.POINTER(C.c_int64)),
                        da.ctypes.data_as(C.POINTER(C.c_double)), len(da)), "fq_load")


def _same(lib, a, b, W, o1, o2, what):
    assert (o1.nspawn_events, o1.ndeath, o1.nattempts, o1.nstates) == \
           (o2.nspawn_events, o2.ndeath, o2.nattempts, o2.nstates), what
    assert o1.proj_energy[0] == o2.proj_energy[0] and o1.D0_population[0] == o2.D0_population[0], what
    assert o1.nparticles[0] == o2.nparticles[0], what
    for x, y in zip(download(lib, a, W, 1), download(lib, b, W, 1)):
        assert np.array_equal(x, y), what


@pytest.mark.parametrize("kind", ["hubbard_integer", "molecular_real_initiator"])
def test_noalloc_cycle_matches_cycle(tmp_path, kind):
    lib = N.lib()
    if kind == "hubbard_integer":
        system = api.hubbard_k(electrons=8, lattice=[[12]], ms=0, sym=1, U=4.0, t=1.0)
        p = _params(211, real=False, initiator=False)
        lst = _grow(lib, system, p, 50, 2.0e5, 0.01)
        tau = 0.01
    else:
        path = synthetic.write_fcidump(os.path.join(tmp_path, "FCIDUMP"), 12, 6, 1206, chol_scale=0.1)
        system = api.read_in(int_file=path, nel=6, ms=0)
        p = _params(212, real=True, initiator=True)
        lst = _grow(lib, system, p, 3000, 2.0e5, 0.02)
        tau = 0.02
    W = system.view.nwords
    small = tuple(x[: len(lst[2]) // 3 * (len(x) // len(lst[2]))] for x in lst)
    print(f"[noalloc] {kind}: large list {len(lst[2])} determinants, small list {len(small[2])}", flush=True)
    ra, rb = C.c_void_p(), C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(ra)), "fq_create")
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(rb)), "fq_create")
    try:
        sh = (C.c_double * 2)(-0.5, -0.5)
        eo = (C.c_double * 2)(0.0, 0.0)
        for label, data in (("large", lst), ("small", small)):
            _load(lib, ra, data)
            _load(lib, rb, data)
            o1, o2 = N.FqCycleOut(), N.FqCycleOut()
            N.check(lib.fq_cycle(ra, sh, eo, tau, C.byref(o1)), "fq_cycle")
            N.check(lib.fq_cycle_noalloc(rb, sh, eo, tau, C.byref(o2)), "fq_cycle_noalloc")
            print(f"  {label}: attempts={o1.nattempts} spawned={o1.nspawn_events} nstates={o1.nstates}", flush=True)
            _same(lib, ra, rb, W, o1, o2, label)
            if label == "large":
                assert o1.nspawn_events > 2 * 2048 and o1.nattempts > 8192  # several sort tiles, tiled scans
    finally:
        lib.fq_destroy(ra)
        lib.fq_destroy(rb)


def test_noalloc_rejects_replay():
    lib = N.lib()
    system = api.hubbard_k(electrons=4, lattice=[[6]], ms=0, sym=1, U=2.0, t=1.0)
    p = _params(213, real=False, initiator=False, cap=1 << 12)
    p.rng_mode = N.FQ_RNG_REPLAY
    eng = C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(eng)), "fq_create")
    try:
        sh = (C.c_double * 2)(0.0, 0.0)
        out = N.FqCycleOut()
        assert lib.fq_cycle_noalloc(eng, sh, sh, 0.01, C.byref(out)) == -1  # FQ_ERR_ARG, before any device work
    finally:
        lib.fq_destroy(eng)
