"""Replay mode on the GPU: the engine consumes the reference's dSFMT stream and
must reproduce the reference's Markov chain bit for bit.

* whole report tables against the golden benchmark outputs (1e-10);
* walker lists (determinants, populations, K_ii) against the oracle after
  every cycle, bit-exact."""
import ctypes as C
import sys

import numpy as np
import pytest

from hande_amd import api, _native as N
from helpers import goldens, oracle, compare_tables, engine_params

pytestmark = pytest.mark.gpu

SMALL = ["hubbard_k_1d_n4", "hubbard_k_nontilted_2d_fciqmc", "hubbard_k_1d_ifciqmc",
         "hubbard_k_nontilted_2d_ifciqmc", "hubbard_k_1d_ifciqmc_cas", "hubbard_k_1d_ifciqmc_real_32"]


@pytest.mark.parametrize("name", SMALL)
def test_replay_reproduces_golden_table(name):
    e = goldens()[name]
    print(f"[replay golden] {name}", flush=True)
    system, opts = api.load_deck(e["input"])
    got = api.fciqmc(system, opts["qmc"], opts.get("fciqmc"), rng_mode="replay")
    compare_tables(got, e["report"])


def _download(lib, eng, W, S):
    n = C.c_int64()
    lib.fq_nstates(eng, C.byref(n))
    cap = max(n.value, 1)
    st = np.zeros(cap * W, dtype=np.uint64)
    po = np.zeros(cap * S, dtype=np.int64)
    da = np.zeros(cap)
    out = C.c_int64()
    N.check(lib.fq_download(eng, st.ctypes.data_as(C.POINTER(C.c_uint64)), po.ctypes.data_as(C.POINTER(C.c_int64)),
                            da.ctypes.data_as(C.POINTER(C.c_double)), cap, C.byref(out)), "fq_download")
    k = out.value
    return st[:k * W], po[:k * S], da[:k]


def _odownload(orc, h, W, S):
    k = orc.orc_nstates(h)
    st = np.zeros(max(k, 1) * W, dtype=np.uint64)
    po = np.zeros(max(k, 1) * S, dtype=np.int64)
    da = np.zeros(max(k, 1))
    orc.orc_download(h, st.ctypes.data_as(C.POINTER(C.c_uint64)), po.ctypes.data_as(C.POINTER(C.c_int64)),
                     da.ctypes.data_as(C.POINTER(C.c_double)))
    return st[:k * W], po[:k * S], da[:k]


@pytest.mark.parametrize("name,ncycles", [("hubbard_k_1d_ifciqmc_real_32", 400), ("hubbard_k_1d_renorm", 60)])
def test_replay_walker_lists_match_oracle_every_cycle(name, ncycles):
    """Cycle-level lock-step: same stream, same shift, bit-identical lists."""
    e = goldens()[name]
    print(f"[replay lockstep] {name}", flush=True)
    lib = N.lib()
    orc = oracle()
    system, opts = api.load_deck(e["input"])
    st = api.qmc_in(opts["qmc"], opts.get("fciqmc"), rng_mode="replay")
    p = engine_params(system.view, st)
    W, S = system.view.nwords, p.nspaces
    eng = C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(eng)), "fq_create")
    h = orc.orc_create(system.view, C.byref(p), st.rng_seed)
    rf = 1 << p.pop_bit_shift
    f0 = np.array([system.view.f0[w] for w in range(W)], dtype=np.uint64)
    pop0 = np.full(S, int(round(st.init_pop)) * rf, dtype=np.int64)
    dat0 = np.zeros(1)
    args = (f0.ctypes.data_as(C.POINTER(C.c_uint64)), pop0.ctypes.data_as(C.POINTER(C.c_int64)),
            dat0.ctypes.data_as(C.POINTER(C.c_double)), 1)
    N.check(lib.fq_load(eng, *args), "fq_load")
    orc.orc_load(h, *args)
    stream = np.zeros(1 << 22)
    lib.fqh_dsfmt_fill(st.rng_seed, stream.ctypes.data_as(C.POINTER(C.c_double)), len(stream))
    N.check(lib.fq_push_rng(eng, stream.ctypes.data_as(C.POINTER(C.c_double)), len(stream)), "fq_push_rng")
    shift = (C.c_double * 2)(0.0, 0.0)
    eold = (C.c_double * 2)(0.0, 0.0)
    consumed = 0
    for cyc in range(ncycles):
        o1, o2 = N.FqCycleOut(), N.FqCycleOut()
        N.check(lib.fq_cycle(eng, shift, eold, st.tau, C.byref(o1)), "fq_cycle")
        orc.orc_cycle(h, shift, eold, st.tau, C.byref(o2))
        consumed += o1.rng_consumed
        assert o1.rng_consumed == o2.rng_consumed, cyc
        assert o1.nspawn_events == o2.nspawn_events and o1.ndeath == o2.ndeath, cyc
        assert o1.nstates == o2.nstates, cyc
        for s in range(S):
            assert o1.nparticles[s] == o2.nparticles[s], cyc
            assert o1.D0_population[s] == o2.D0_population[s], cyc
            assert o1.proj_energy[s] == o2.proj_energy[s], cyc
        a, b = _download(lib, eng, W, S), _odownload(orc, h, W, S)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), cyc
        if cyc % 50 == 0:
            print(f"  cycle {cyc}: nstates={o1.nstates} N={o1.nparticles[0]}", flush=True)
    assert consumed == orc.orc_rng_consumed(h)
    lib.fq_destroy(eng)
    orc.orc_destroy(h)


MOLECULAR = {"H2O-RHF-cc-pVTZ": 30, "Ne-RHF-aug-cc-pVDZ": 40}


@pytest.mark.parametrize("name", sorted(MOLECULAR))
def test_replay_reproduces_molecular_golden(name, tmp_path):
    """Molecular renorm generator, Slater--Condon rules and (Ne) the initiator
    rules on the GPU, replaying the reference's stream: the first report rows of
    the golden table, bit-exact at 1e-10."""
    from helpers import deck
    e = goldens()[name]
    print(f"[replay molecular] {name}", flush=True)
    system, opts = deck(name, str(tmp_path))
    q = dict(opts["qmc"])
    q["nreports"] = MOLECULAR[name]
    got = api.fciqmc(system, q, opts.get("fciqmc"), rng_mode="replay")
    compare_tables(got, e["report"][:MOLECULAR[name] + 1])
