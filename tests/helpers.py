"""Shared test utilities: golden fixtures, oracle bindings, table comparison.

The oracle (oracle/liboracle.so) is the checker; it is built from oracle/ with
g++ when missing (make -C oracle liboracle.so)."""
import ctypes as C
import json
import os
import subprocess

import numpy as np

from hande_amd import _native as N
from hande_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

_goldens = None


def goldens():
    global _goldens
    if _goldens is None:
        with open(os.path.join(GOLDEN, "qmc_goldens.json")) as f:
            _goldens = json.load(f)
    return _goldens


def oracle():
    src = os.path.join(ROOT, "oracle", "fqmc_oracle.cpp")
    if not os.path.exists(ORACLE_SO) or (os.path.exists(src) and os.path.getmtime(src) > os.path.getmtime(ORACLE_SO)):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), os.path.join(ROOT, "oracle", "liboracle.so")],
                       check=True)
    lib = C.CDLL(ORACLE_SO)
    P = C.POINTER
    vp = C.c_void_p
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [P(N.FqSystem), P(N.FqParams), C.c_int64]
    lib.orc_destroy.argtypes = [vp]
    lib.orc_load.argtypes = [vp, P(C.c_uint64), P(C.c_int64), P(C.c_double), C.c_int64]
    lib.orc_cycle.argtypes = [vp, P(C.c_double), P(C.c_double), C.c_double, P(N.FqCycleOut)]
    lib.orc_annihilate.argtypes = [vp, P(C.c_uint64), P(C.c_int64), P(C.c_int64), C.c_int64, P(N.FqCycleOut)]
    lib.orc_nstates.restype = C.c_int64
    lib.orc_nstates.argtypes = [vp]
    lib.orc_download.argtypes = [vp, P(C.c_uint64), P(C.c_int64), P(C.c_double)]
    lib.orc_rng_consumed.restype = C.c_uint64
    lib.orc_rng_consumed.argtypes = [vp]
    lib.orc_set_stable_order.argtypes = [vp, C.c_int]
    lib.orc_set_production.argtypes = [vp, C.c_uint64]
    lib.orc_philox4x32_10.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
    lib.orc_fciqmc.argtypes = [P(N.FqSystem), P(N.FqhQmcIn), P(N.FqhReportRow), C.c_int, P(C.c_int)]
    lib.orc_fciqmc_ranks.argtypes = [P(N.FqSystem), P(N.FqhQmcIn), C.c_int, P(N.FqhReportRow), C.c_int, P(C.c_int)]
    lib.orc_fciqmc_restart.argtypes = [P(N.FqSystem), P(N.FqhQmcIn), P(N.FqhRestart), P(N.FqhRestart),
                                       P(N.FqhReportRow), C.c_int, P(C.c_int)]
    lib.orc_murmur2.restype = C.c_uint32
    lib.orc_murmur2.argtypes = [C.c_char_p, C.c_int, C.c_uint32]
    lib.orc_owner_rank.argtypes = [P(C.c_uint64), C.c_int, C.c_int]
    lib.orc_hb_biased.argtypes = [vp]
    lib.orc_hb_table.restype = C.c_int64
    lib.orc_hb_table.argtypes = [vp, C.c_char_p, P(C.c_double), C.c_int64]
    lib.orc_alias_tables.argtypes = [C.c_int, P(C.c_double), C.c_double, P(C.c_double), P(C.c_int)]
    lib.orc_hb_generate.restype = C.c_int64
    lib.orc_hb_generate.argtypes = [vp, P(C.c_uint64), P(C.c_double), C.c_int64, P(C.c_int), P(C.c_double)]
    lib.orc_semi_stoch_init.argtypes = [vp, C.c_int64, C.c_int]
    lib.orc_semi_stoch_info.restype = C.c_int64
    lib.orc_semi_stoch_info.argtypes = [vp, P(C.c_uint64), P(C.c_int64)]
    return lib


def oracle_report(orc, system, st, nprocs=1, start=None, end=None):
    """The oracle's report table; start / end are restart.RestartBuffers."""
    rows = (N.FqhReportRow * (st.nreports + 1))()
    nr = C.c_int()
    if start is not None or end is not None:
        rc = orc.orc_fciqmc_restart(system.view, C.byref(st), C.byref(start.st) if start else None,
                                    C.byref(end.st) if end else None, rows, st.nreports + 1, C.byref(nr))
        assert rc == 0, rc
    elif nprocs == 1:
        orc.orc_fciqmc(system.view, C.byref(st), rows, st.nreports + 1, C.byref(nr))
    else:
        orc.orc_fciqmc_ranks(system.view, C.byref(st), nprocs, rows, st.nreports + 1, C.byref(nr))
    return api.rows_to_table(rows, nr.value, 2 if st.replica_tricks else 1)


def compare_tables(got, ref, rtol=1e-10, rspawn_atol=6e-5):
    """Golden comparison as testcode does it (test_suite/userconfig:14-15): every
    column to 1e-10 (absolute or relative); R_spawn is printed with 4 decimals."""
    ref = np.asarray(ref)[:, :got.shape[1]]
    n = min(len(got), len(ref))
    assert len(got) == len(ref), f"row count {len(got)} != {len(ref)}"
    g, r = got[:n], ref[:n]
    diff = np.abs(g - r)
    ok = (diff <= rtol) | (diff <= rtol * np.abs(r))
    ok[:, -1] = diff[:, -1] <= rspawn_atol
    bad = np.argwhere(~ok)
    if len(bad):
        i, j = bad[0]
        raise AssertionError(f"row {i} col {j}: got {g[i, j]!r} ref {r[i, j]!r}\n got {g[i]}\n ref {r[i]}")


def engine_params(system_view, st):
    """fq_params as the driver builds them from a qmc block."""
    p = N.FqParams()
    p.nspaces = 2 if st.replica_tricks else 1
    p.initiator = st.initiator
    p.initiator_pop = st.initiator_threshold
    p.quadrature_initiator = st.quadrature_initiator
    p.real_amplitudes = st.real_amplitudes
    p.pop_bit_shift = (11 if st.real_amplitude_force_32 else 31) if st.real_amplitudes else 0
    rf = 1 << p.pop_bit_shift
    p.spawn_cutoff = int(np.ceil(st.spawn_cutoff * rf)) if st.real_amplitudes else 0
    p.excit_gen = st.excit_gen
    p.rng_mode = st.rng_mode
    p.philox_seed = st.rng_seed
    p.main_capacity = 1 << 16
    p.spawn_capacity = 1 << 16
    p.device = 0
    return p


def materialize_data(name, directory):
    """Write a golden entry's integral dump(s) (stored gzipped) into directory."""
    import gzip
    for fname in goldens()[name].get("data_files", []):
        target = os.path.join(directory, fname.split(".", 1)[1][:-3])  # "<name>.INTDUMP.gz" -> INTDUMP
        with gzip.open(os.path.join(GOLDEN, fname), "rb") as src, open(target, "wb") as dst:
            dst.write(src.read())
    return directory


def deck(name, directory, lib=None):
    """(system, fciqmc options) for a golden entry, its data files in directory."""
    e = goldens()[name]
    materialize_data(name, directory)
    return api.load_deck(e["input"], base_dir=directory, lib=lib)


def download(lib, eng, W, S):
    """The engine's walker list (states, pops, K_ii) as numpy arrays."""
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


def odownload(orc, h, W, S):
    """The oracle's walker list (states, pops, K_ii) as numpy arrays."""
    k = orc.orc_nstates(h)
    st = np.zeros(max(k, 1) * W, dtype=np.uint64)
    po = np.zeros(max(k, 1) * S, dtype=np.int64)
    da = np.zeros(max(k, 1))
    orc.orc_download(h, st.ctypes.data_as(C.POINTER(C.c_uint64)), po.ctypes.data_as(C.POINTER(C.c_int64)),
                     da.ctypes.data_as(C.POINTER(C.c_double)))
    return st[:k * W], po[:k * S], da[:k]


def lockstep(system, st, ncycles, init_pop=None, shift=0.0, label=""):
    """Run the engine (replay mode) and the oracle side by side from init_pop
    walkers on the reference determinant, with the same dSFMT stream and shift;
    assert bit-identical cycle outputs and walker lists after every cycle.
    Returns the last cycle's output."""
    lib = N.lib()
    orc = oracle()
    p = engine_params(system.view, st)
    W, S = system.view.nwords, p.nspaces
    eng = C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(eng)), "fq_create")
    h = orc.orc_create(system.view, C.byref(p), st.rng_seed)
    try:
        rf = 1 << p.pop_bit_shift
        f0 = np.array([system.view.f0[w] for w in range(W)], dtype=np.uint64)
        ip = st.init_pop if init_pop is None else init_pop
        pop0 = np.full(S, int(round(ip)) * rf, dtype=np.int64)
        dat0 = np.zeros(1)
        args = (f0.ctypes.data_as(C.POINTER(C.c_uint64)), pop0.ctypes.data_as(C.POINTER(C.c_int64)),
                dat0.ctypes.data_as(C.POINTER(C.c_double)), 1)
        N.check(lib.fq_load(eng, *args), "fq_load")
        orc.orc_load(h, *args)
        stream = np.zeros(1 << 22)
        lib.fqh_dsfmt_fill(st.rng_seed, stream.ctypes.data_as(C.POINTER(C.c_double)), len(stream))
        N.check(lib.fq_push_rng(eng, stream.ctypes.data_as(C.POINTER(C.c_double)), len(stream)), "fq_push_rng")
        sh = (C.c_double * 2)(shift, shift)
        eold = (C.c_double * 2)(0.0, 0.0)
        consumed = 0
        o1 = None
        for cyc in range(ncycles):
            o1, o2 = N.FqCycleOut(), N.FqCycleOut()
            N.check(lib.fq_cycle(eng, sh, eold, st.tau, C.byref(o1)), "fq_cycle")
            orc.orc_cycle(h, sh, eold, st.tau, C.byref(o2))
            consumed += o1.rng_consumed
            assert o1.rng_consumed == o2.rng_consumed, cyc
            assert o1.nspawn_events == o2.nspawn_events and o1.ndeath == o2.ndeath, cyc
            assert o1.nstates == o2.nstates, cyc
            for s in range(S):
                assert o1.nparticles[s] == o2.nparticles[s], cyc
                assert o1.D0_population[s] == o2.D0_population[s], cyc
                assert o1.proj_energy[s] == o2.proj_energy[s], cyc
            a, b = download(lib, eng, W, S), odownload(orc, h, W, S)
            for x, y in zip(a, b):
                assert np.array_equal(x, y), cyc
            if cyc % 50 == 0:
                print(f"  {label} cycle {cyc}: nstates={o1.nstates} N={o1.nparticles[0]}", flush=True)
        assert consumed == orc.orc_rng_consumed(h)
        return o1
    finally:
        lib.fq_destroy(eng)
        orc.orc_destroy(h)


def production_lockstep(system, p, ncycles, init_pop, shift=0.0, tau=0.01, label="", every=50, core=None):
    """Run the engine in production mode (Philox) and the oracle in its
    production mode with the same seed side by side from init_pop walkers on
    the reference determinant.  Per cycle: spawn/death/attempt counts, list
    length, the walker lists (determinants, populations, K_ii) and the
    reference population must be identical; the projected energy and the
    particle totals (summed in a different order on the GPU) agree to 1e-12
    relative.  core = (cycle, size, separate) makes a semi-stochastic core
    before that cycle on both (same determinants, same stored element count).
    Returns the list of cycle outputs."""
    lib = N.lib()
    orc = oracle()
    W, S = system.view.nwords, p.nspaces
    eng = C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(eng)), "fq_create")
    h = orc.orc_create(system.view, C.byref(p), 1)
    orc.orc_set_production(h, p.philox_seed)
    outs = []
    try:
        rf = 1 << p.pop_bit_shift
        f0 = np.array([system.view.f0[w] for w in range(W)], dtype=np.uint64)
        pop0 = np.full(S, int(round(init_pop)) * rf, dtype=np.int64)
        dat0 = np.zeros(1)
        args = (f0.ctypes.data_as(C.POINTER(C.c_uint64)), pop0.ctypes.data_as(C.POINTER(C.c_int64)),
                dat0.ctypes.data_as(C.POINTER(C.c_double)), 1)
        N.check(lib.fq_load(eng, *args), "fq_load")
        orc.orc_load(h, *args)
        sh = (C.c_double * 2)(shift, shift)
        eold = (C.c_double * 2)(0.0, 0.0)
        for cyc in range(ncycles):
            if core is not None and cyc == core[0]:
                N.check(lib.fq_semi_stoch_init(eng, core[1], core[2]), "fq_semi_stoch_init")
                orc.orc_semi_stoch_init(h, core[1], core[2])
                nd, nnz = C.c_int64(), C.c_int64()
                N.check(lib.fq_semi_stoch_info(eng, C.byref(nd), C.byref(nnz), None, 0), "fq_semi_stoch_info")
                g_states = np.zeros(max(nd.value, 1) * W, dtype=np.uint64)
                N.check(lib.fq_semi_stoch_info(eng, C.byref(nd), None,
                                               g_states.ctypes.data_as(C.POINTER(C.c_uint64)), nd.value),
                        "fq_semi_stoch_info")
                o_states = np.zeros(max(nd.value, 1) * W, dtype=np.uint64)
                o_nnz = C.c_int64()
                assert orc.orc_semi_stoch_info(h, o_states.ctypes.data_as(C.POINTER(C.c_uint64)),
                                               C.byref(o_nnz)) == nd.value
                assert nd.value == min(core[1], o1.nstates if cyc else 1) and nnz.value == o_nnz.value
                assert np.array_equal(g_states, o_states)
                print(f"  {label} core: {nd.value} determinants, {nnz.value} stored elements", flush=True)
            o1, o2 = N.FqCycleOut(), N.FqCycleOut()
            N.check(lib.fq_cycle(eng, sh, eold, tau, C.byref(o1)), "fq_cycle")
            orc.orc_cycle(h, sh, eold, tau, C.byref(o2))
            assert (o1.nspawn_events, o1.ndeath, o1.nattempts, o1.nstates) == \
                   (o2.nspawn_events, o2.ndeath, o2.nattempts, o2.nstates), cyc
            assert not o1.spawn_overflow and not o1.main_overflow
            for s in range(S):
                assert o1.D0_population[s] == o2.D0_population[s], cyc
                np.testing.assert_allclose(o1.proj_energy[s], o2.proj_energy[s], rtol=1e-12, atol=1e-12)
                np.testing.assert_allclose(o1.nparticles[s], o2.nparticles[s], rtol=1e-12)
            a, b = download(lib, eng, W, S), odownload(orc, h, W, S)
            for x, y in zip(a, b):
                assert np.array_equal(x, y), cyc
            outs.append(o1)
            if cyc % every == 0:
                print(f"  {label} cycle {cyc}: nstates={o1.nstates} spawned={o1.nspawn_events} "
                      f"N={o1.nparticles[0]:.6g}", flush=True)
        return outs
    finally:
        lib.fq_destroy(eng)
        orc.orc_destroy(h)


def replicate_estimates(system, q, seeds, skip=150):
    """Independent production runs (one per Philox seed) of the same deck:
    each gives the projected energy as the ratio of the production window's
    means of sum H_0j N_j and N_0, and the mean shift; the runs are
    independent, so the spread of these per-run values gives an honest
    standard error however long the population-control correlation time is
    (a single short run often has no reblocking plateau).  Returns
    (E mean, its standard error, shift mean, its standard error), energies
    including H00."""
    es, ss = [], []
    for sd in seeds:
        qq = dict(q)
        qq["rng_seed"] = sd
        tab = api.fciqmc(system, qq, rng_mode="philox")
        prod = tab[skip:]
        es.append(system.H00 + prod[:, 2].mean() / prod[:, 3].mean())
        ss.append(system.H00 + prod[:, 1].mean())
    es, ss = np.array(es), np.array(ss)
    k = len(seeds)
    return es.mean(), es.std(ddof=1) / np.sqrt(k), ss.mean(), ss.std(ddof=1) / np.sqrt(k)
