"""Production mode against the oracle's production mode, bit for bit.

Every random decision of a production cycle reads a Philox4x32-10 stream keyed
by the seed, the cycle and an index (determinant and space for the attempt
count and death, the global attempt index for a spawning attempt, the list
position or unique-key run for stochastic rounding), so a serial CPU loop over
the same streams reproduces the parallel kernels exactly: the walker lists
(determinants, populations, K_ii), the spawned/death/attempt counts and the
reference population must be identical after every cycle.  This pins the
parallel spawning path (attempt numbering, parent search in the attempt
prefix sum, per-tile compaction of the successful attempts, packing in
attempt order) and the production annihilation against a restatement of the
reference's cycle (src/fciqmc.f90:307-399)."""
import ctypes as C
import os

import numpy as np
import pytest

from hande_amd import _native as N
from hande_amd import api, synthetic
from helpers import production_lockstep

pytestmark = pytest.mark.gpu


def _params(seed, excit_gen=N.FQ_EXCIT_RENORM, real=False, initiator=False, spaces=1, cap=1 << 18):
    p = N.FqParams()
    p.nspaces = spaces
    p.initiator = int(initiator)
    p.initiator_pop = 3.0
    p.quadrature_initiator = 1 if initiator else 0
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


def _mol(tmp_path, norb, nel, seed, cs=0.1):
    path = synthetic.write_fcidump(os.path.join(tmp_path, "FCIDUMP"), norb, nel, seed, chol_scale=cs)
    return api.read_in(int_file=path, nel=nel, ms=0)


@pytest.mark.parametrize("real,initiator", [(False, False), (True, True)])
def test_hubbard_k(real, initiator):
    system = api.hubbard_k(electrons=8, lattice=[[12]], ms=0, sym=1, U=4.0, t=1.0)
    print(f"[prod-lockstep] hubbard_k 12 sites 8e real={real} initiator={initiator}", flush=True)
    outs = production_lockstep(system, _params(101, real=real, initiator=initiator), 120, 50, shift=0.0,
                               tau=0.01, label="hub")
    assert outs[-1].nstates > 2000


def test_hubbard_k_replica_pair():
    system = api.hubbard_k(electrons=6, lattice=[[10]], ms=0, sym=1, U=3.0, t=1.0)
    print("[prod-lockstep] hubbard_k replica pair", flush=True)
    outs = production_lockstep(system, _params(102, real=True, initiator=True, spaces=2), 100, 30, tau=0.01,
                               label="replica")
    assert outs[-1].nstates > 200


@pytest.mark.parametrize("excit_gen", ["renorm", "heat_bath"])
def test_molecular_real_initiator(tmp_path, excit_gen):
    """The benchmark's options (real_64, initiator 3, spawn cutoff 0.01) on a
    12-orbital, 6-electron synthetic Hamiltonian."""
    system = _mol(tmp_path, 12, 6, 1206)
    code = N.FQ_EXCIT_RENORM if excit_gen == "renorm" else N.FQ_EXCIT_HEAT_BATH
    print(f"[prod-lockstep] 12o6e {excit_gen}", flush=True)
    outs = production_lockstep(system, _params(103, code, real=True, initiator=True), 80, 3000, shift=0.0,
                               tau=0.02, label=excit_gen)
    assert outs[-1].nstates > 3000
    assert max(o.nspawn_events for o in outs) > 5 * 1024  # several spawning tiles per cycle


def test_two_word_determinants(tmp_path):
    system = _mol(tmp_path, 34, 6, 3406)
    assert system.view.nwords == 2
    print("[prod-lockstep] 34o6e two-word renorm", flush=True)
    outs = production_lockstep(system, _params(104, real=True, initiator=True), 40, 2000, tau=0.01, label="W2")
    assert outs[-1].nstates > 2000


def test_sparse_attempts_and_dense_parents(tmp_path):
    """A loaded list of 40000 determinants holding 0.03 walkers each (most get
    no attempt, so a tile of attempts spans tens of thousands of entries) next
    to the reference determinant holding 20000 walkers (one parent spanning
    many tiles)."""
    lib = N.lib()
    import helpers
    orc = helpers.oracle()
    system = _mol(tmp_path, 12, 6, 1207)
    p = _params(105, real=True, initiator=True)
    rf = 1 << 31
    nb = system.view.nbasis
    rng = np.random.default_rng(3)
    alpha = np.arange(0, nb, 2)
    beta = np.arange(1, nb, 2)
    dets = set()
    while len(dets) < 40000:
        bits = np.concatenate([rng.choice(alpha, 3, replace=False), rng.choice(beta, 3, replace=False)])
        dets.add(int(np.sum(np.left_shift(np.uint64(1), bits.astype(np.uint64)))))
    f0 = int(system.view.f0[0])
    dets.add(f0)
    st = np.array(sorted(dets), dtype=np.uint64)
    po = np.full(len(st), int(0.03 * rf), dtype=np.int64)
    po[st == np.uint64(f0)] = 20000 * rf
    da = np.zeros(len(st))
    eng = C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(eng)), "fq_create")
    h = orc.orc_create(system.view, C.byref(p), 1)
    orc.orc_set_production(h, p.philox_seed)
    print(f"[prod-lockstep] sparse list of {len(st)} determinants", flush=True)
    try:
        args = (st.ctypes.data_as(C.POINTER(C.c_uint64)), po.ctypes.data_as(C.POINTER(C.c_int64)),
                da.ctypes.data_as(C.POINTER(C.c_double)), len(st))
        N.check(lib.fq_load(eng, *args), "fq_load")
        orc.orc_load(h, *args)
        sh = (C.c_double * 2)(0.0, 0.0)
        eo = (C.c_double * 2)(0.0, 0.0)
        for cyc in range(6):
            o1, o2 = N.FqCycleOut(), N.FqCycleOut()
            N.check(lib.fq_cycle(eng, sh, eo, 0.02, C.byref(o1)), "fq_cycle")
            orc.orc_cycle(h, sh, eo, 0.02, C.byref(o2))
            assert (o1.nspawn_events, o1.ndeath, o1.nattempts, o1.nstates) == \
                   (o2.nspawn_events, o2.ndeath, o2.nattempts, o2.nstates), cyc
            a, b = helpers.download(lib, eng, 1, 1), helpers.odownload(orc, h, 1, 1)
            for x, y in zip(a, b):
                assert np.array_equal(x, y), cyc
            print(f"  cycle {cyc}: attempts={o1.nattempts} spawned={o1.nspawn_events} nstates={o1.nstates}",
                  flush=True)
        assert o1.nattempts > 20000
    finally:
        lib.fq_destroy(eng)
        orc.orc_destroy(h)
