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
from helpers import goldens, oracle, compare_tables, engine_params, lockstep

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


@pytest.mark.parametrize("name,ncycles", [("hubbard_k_1d_ifciqmc_real_32", 400), ("hubbard_k_1d_renorm", 60)])
def test_replay_walker_lists_match_oracle_every_cycle(name, ncycles):
    """Cycle-level lock-step: same stream, same shift, bit-identical lists."""
    e = goldens()[name]
    print(f"[replay lockstep] {name}", flush=True)
    system, opts = api.load_deck(e["input"])
    st = api.qmc_in(opts["qmc"], opts.get("fciqmc"), rng_mode="replay")
    lockstep(system, st, ncycles, label=name)


MOLECULAR = {"H2O-RHF-cc-pVTZ": 30, "Ne-RHF-aug-cc-pVDZ": 40, "H2-sto-3g-replica": 1500}


@pytest.mark.parametrize("name", sorted(MOLECULAR))
def test_replay_reproduces_molecular_golden(name, tmp_path):
    """Molecular renorm generator, Slater--Condon rules, (Ne) the initiator
    rules and (H2) a replica pair on the GPU, replaying the reference's stream:
    the first report rows of the golden table, bit-exact at 1e-10."""
    from helpers import deck
    e = goldens()[name]
    print(f"[replay molecular] {name}", flush=True)
    system, opts = deck(name, str(tmp_path))
    q = dict(opts["qmc"])
    q["nreports"] = MOLECULAR[name]
    got = api.fciqmc(system, q, opts.get("fciqmc"), rng_mode="replay")
    compare_tables(got, e["report"][:MOLECULAR[name] + 1])


@pytest.mark.parametrize("excit_gen", ["renorm", "heat_bath"])
def test_replay_lockstep_two_word_determinants(tmp_path, excit_gen):
    """68 spin orbitals: determinants span two 64-bit words (orbitals 65-68 in
    the second); real_64 initiator run, walker lists bit-identical each cycle."""
    import os
    from hande_amd import synthetic
    print(f"[replay lockstep] two words, {excit_gen}", flush=True)
    path = synthetic.write_fcidump(os.path.join(str(tmp_path), "FCIDUMP"), 34, 6, 3406)
    system = api.read_in(int_file=path, nel=6, ms=0)
    assert system.view.nwords == 2
    st = api.qmc_in(dict(tau=0.01, rng_seed=19, init_pop=30, real_amplitudes=True, initiator=True,
                         excit_gen=excit_gen), rng_mode="replay")
    out = lockstep(system, st, 80, label=f"W=2 {excit_gen}")
    assert out.nstates > 50


def test_hubbard_4x4_configuration_lockstep():
    """The reference-runnable Hubbard configuration: 4x4 periodic lattice in
    momentum space, 5 up + 5 down electrons, U/t = 4, totally symmetric
    k-sector, integer populations, tau = 0.005, no initiator.  The GPU replays
    the dSFMT stream with walker lists bit-identical to the oracle's after
    every cycle (60 cycles from 1000 walkers on the Hartree--Fock determinant,
    shift held at -18 relative to it), then the production path (Philox) does
    the same against the oracle's production mode."""
    from hande_amd import _native as N
    from helpers import production_lockstep
    system = api.hubbard_k(electrons=10, lattice=[[4, 0], [0, 4]], ms=0, sym="tot_sym", U=4, t=1)
    print("[replay lockstep] Hubbard 4x4, 10 electrons", flush=True)
    st = api.qmc_in(dict(tau=0.005, rng_seed=7, init_pop=1000), rng_mode="replay")
    out = lockstep(system, st, 60, shift=-18.0, label="4x4 replay")
    assert out.nstates > 5000
    p = N.FqParams()
    p.nspaces, p.initiator, p.real_amplitudes, p.pop_bit_shift, p.spawn_cutoff = 1, 0, 0, 0, 0
    p.excit_gen, p.rng_mode, p.philox_seed = N.FQ_EXCIT_RENORM, N.FQ_RNG_PHILOX, 7
    p.main_capacity = p.spawn_capacity = 1 << 18
    outs = production_lockstep(system, p, 60, 1000, shift=-18.0, tau=0.005, label="4x4 production")
    assert outs[-1].nstates > 5000


def test_replay_two_rank_real_64_golden():
    """The reference's two-rank real_64 initiator run
    (test_suite/ifciqmc_real_64/np2/hubbard_k_14s_14e_U2_real_64) replayed on
    two emulated ranks, one engine each on the GPU, spawned walkers exchanged
    by MurmurHash2 ownership: the first 300 of its 600 report rows (the
    oracle's emulation matches all 600) against the golden table at 1e-10."""
    e = goldens()["hubbard_k_14s_real_64_np2"]
    system, opts = api.load_deck(e["input"])
    q = dict(opts["qmc"])
    q["nreports"] = 300
    print("[replay] two emulated ranks, 300 reports", flush=True)
    got = api.fciqmc(system, q, opts.get("fciqmc"), rng_mode="replay", device=0, nprocs=2)
    compare_tables(got, e["report"][:301])


def test_replay_restart_golden_and_dump(tmp_path):
    """The reference's restart test (test_suite/fciqmc_real_64/np1/
    He2-aug-cc-pVDZ_real_64_restart): the engine continues from the
    reference's HDF5 restart file (integer populations read into real_64
    amplitudes) and reproduces all 101 rows of its benchmark table; the state
    it dumps after the last report is bit-identical to the oracle's."""
    from hande_amd import restart as R
    from helpers import deck, oracle_report
    name = "He2-aug-cc-pVDZ_real_64_restart"
    system, opts = deck(name, str(tmp_path))
    path = api.restart_file(opts, str(tmp_path))
    print(f"[replay restart] {name} from {path}", flush=True)
    got, end = api.fciqmc(system, opts["qmc"], opts.get("fciqmc"), rng_mode="replay", restart=path,
                          return_state=True)
    compare_tables(got, goldens()[name]["report"])
    st = api.qmc_in(opts["qmc"], opts.get("fciqmc"))
    rs = R.read_restart(path, system.view.nbasis, 1, api.pop_bit_shift(st))
    oend = R.RestartBuffers(system.view.nwords, 1, 1 << 16)
    oracle_report(oracle(), system, st, start=R.RestartBuffers.from_state(rs), end=oend)
    ref = oend.to_state(1, api.pop_bit_shift(st))
    assert end.nstates == ref.nstates and end.mc_cycles_done == ref.mc_cycles_done == 10000
    assert np.array_equal(end.states, ref.states) and np.array_equal(end.pops, ref.pops)
    assert np.array_equal(end.dat, ref.dat)
    assert np.array_equal(end.shift, ref.shift) and np.array_equal(end.proj_energy, ref.proj_energy)


def test_replay_auto_shift_damping_restart_golden(tmp_path):
    """test_suite/ifciqmc_real_64/np1/He2-aug-cc-pVDZ_real_64_auto_shift_damping_restart
    on the engine: a restart with the initiator approximation, vary_shift on
    from the start, on-the-fly blocking and automatic shift damping; all 501
    first 130 of its 501 rows at 1e-10, past the damping change after report
    100 (which moves every later shift; the oracle matches all 501 rows).
    Replay runs spawning on one GPU thread, so the whole table (2500 cycles at
    ~9000 walkers) would take minutes."""
    import time
    from helpers import deck
    name = "He2-aug-cc-pVDZ_real_64_auto_shift_damping_restart"
    system, opts = deck(name, str(tmp_path))
    path = api.restart_file(opts, str(tmp_path))
    print(f"[replay restart] {name}, 130 of 500 reports", flush=True)
    t0 = time.time()
    got = api.fciqmc(system, opts["qmc"], opts.get("fciqmc"), rng_mode="replay", restart=path,
                     blocking=opts.get("blocking"), max_reports=130)
    print(f"  {time.time() - t0:.1f} s", flush=True)
    compare_tables(got, goldens()[name]["report"][:131])
