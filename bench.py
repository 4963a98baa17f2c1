#!/usr/bin/env python
"""Benchmark: spawn attempts/s of the FCIQMC evolve loop on one MI355X.

Workload (the molecular i-FCIQMC configuration): seeded synthetic molecular
Hamiltonian, 26 spatial orbitals (52 spin orbitals), 10 electrons, C1
symmetry; real_64 fixed-point populations, initiator threshold 3, target
population 1e7, tau 1e-3, renormalised excitation generator, FP64 integrals.

A step is one Monte Carlo cycle (spawning + death + annihilation) of the whole
walker population resident in HBM.  The walker distribution is prepared
untimed: target/1000 walkers on the Hartree--Fock determinant are propagated
at the production time step (1e-3; a larger step moves the stationary state of
these systems, whose diagonal elements spread over tens of Hartree); once the
population passes the target the shift varies with a proportional-integral
update that pulls the population to the target, for 250 cycles.  Then W
warm-up cycles run at tau = 1e-3 with the reference's shift update (damping
0.05) before K timed cycles.  A prepared
list can be saved (--save-state) and reloaded (--load-state), so a profiled
run contains only cycles of the measured shape.

Prints one JSON line (rank 0).  Extra fields:
  roofline      the spawning-attempt kernel (the largest device-time share):
                algorithmic HBM bytes per launch / its average launch time
                (HIP event pair around each launch on the engine's stream)
                vs the 8 TB/s HBM3E peak;
  cpu_baseline  the CPU oracle (a serial restatement of the reference loop)
                timed on the same walker state: one instance per host core on
                a slice of the determinants each (as MPI ranks would hold), and
                one instance on one core over the whole list;
  batched       the batched workload (64 Hubbard-k rings of 10-18 sites) run
                as graph-captured cycles: cycle time, attempts/s and the
                schedule the tuner chose (--no-batched skips it).
Under ``rocprofv3 --selected-regions`` only the timed cycles are profiled
(roctxProfilerResume/Pause bracket them).
"""
import argparse
import concurrent.futures
import ctypes as C
import json
import math
import os
import sys
import tempfile
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from hande_amd import _native as N  # noqa: E402
from hande_amd import api, synthetic  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Vector ALU issue peak of one MI355X: 256 CUs x 4 SIMDs x 16 lanes per clock
# at 2.4 GHz (a wave64 full-rate 32-bit instruction takes 4 cycles on a SIMD).
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class ProfilerRegion:
    """roctxProfilerPause/Resume when the roctx library is present (no-ops otherwise)."""

    def __init__(self):
        self.lib = None
        for name in ("librocprofiler-sdk-roctx.so.1", "/opt/rocm/lib/librocprofiler-sdk-roctx.so.1"):
            try:
                self.lib = C.CDLL(name)
                break
            except OSError:
                continue
        if self.lib is not None:
            for fn in ("roctxProfilerPause", "roctxProfilerResume"):
                getattr(self.lib, fn).argtypes = [C.c_uint64]
                getattr(self.lib, fn).restype = C.c_int

    def pause(self):
        if self.lib is not None:
            self.lib.roctxProfilerPause(0)

    def resume(self):
        if self.lib is not None:
            self.lib.roctxProfilerResume(0)


def molecular_params(seed, main_cap, spawn_cap, excit_gen=N.FQ_EXCIT_RENORM, mix=0.1, min_attempts=0, spaces=1):
    """i-FCIQMC engine options of the molecular workloads: real_64 populations,
    initiator threshold 3, spawn cutoff 0.01, production Philox stream."""
    p = N.FqParams()
    p.nspaces = spaces
    p.initiator = 1
    p.initiator_pop = 3.0
    p.quadrature_initiator = 1
    p.real_amplitudes = 1
    p.pop_bit_shift = 31
    p.spawn_cutoff = int(math.ceil(0.01 * (1 << 31)))
    p.excit_gen = excit_gen
    p.rng_mode = N.FQ_RNG_PHILOX
    p.philox_seed = seed
    p.main_capacity = main_cap
    p.spawn_capacity = spawn_cap
    p.device = 0
    p.proposal_mix = mix
    p.proposal_min_attempts = min_attempts
    return p


def make_engine(lib, system, target, seed, main_cap, spawn_cap, **kw):
    p = molecular_params(seed, main_cap, spawn_cap, **kw)
    eng = C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(eng)), "fq_create")
    return eng, p


class Propagator:
    """Cycle driver with the reference's shift update (src/energy_evaluation.F90:659-711):
    every `update_every` cycles S -= damping/(tau*every) * ln(N/N_old) once the
    population has passed the target.  With pull > 0 (preparation only) an
    integral term -pull*damping^2/(tau*every)*ln(N/target) also pulls the
    population to the target."""

    def __init__(self, lib, eng, target, damping=0.05, update_every=5, pull=0.0, spaces=1):
        self.lib, self.eng = lib, eng
        self.target, self.damping, self.every = target, damping, update_every
        self.pull = pull  # integral gain towards the target (preparation only)
        self.S = spaces   # walker spaces (2 for a replica pair), each with its own shift
        self.shift = (C.c_double * 2)(0.0, 0.0)
        self.eold = (C.c_double * 2)(0.0, 0.0)
        self.vary_s = [False, False]
        self.n_old = [None, None]
        self.count = 0
        self.proj = [0.0, 0.0]
        self.d0 = [0.0, 0.0]
        self.last = None

    @property
    def vary(self):
        return all(self.vary_s[:self.S])

    @vary.setter
    def vary(self, value):
        self.vary_s = [bool(value)] * 2

    def cycle(self, tau):
        out = N.FqCycleOut()
        N.check(self.lib.fq_cycle(self.eng, self.shift, self.eold, tau, C.byref(out)), "fq_cycle")
        if out.spawn_overflow or out.main_overflow:  # a truncated cycle: no number from it is valid
            raise RuntimeError(f"list capacity exceeded (spawned {out.nspawn_events}, states {out.nstates}); "
                               f"raise --spawn-cap / --main-cap")
        self.count += 1
        for s in range(self.S):
            self.proj[s] += out.proj_energy[s]
            self.d0[s] += out.D0_population[s]
        if self.count % self.every == 0:
            dt = tau * self.every
            for s in range(self.S):
                n = out.nparticles[s]
                if self.vary_s[s] and self.n_old[s]:
                    self.shift[s] = self.shift[s] - math.log(n / self.n_old[s]) * self.damping / dt
                    if self.pull:
                        self.shift[s] -= self.pull * self.damping ** 2 / dt * math.log(n / self.target)
                if abs(self.d0[s]) > 0:
                    self.eold[s] = self.proj[s] / self.d0[s]
                if not self.vary_s[s] and n > self.target:
                    self.vary_s[s] = True
                self.n_old[s] = n
                self.proj[s] = self.d0[s] = 0.0
        self.last = out
        return out


def sort_key_words(nbasis, nel, spawn_capacity):
    """64-bit words of the sort key: the combinatorial rank of a determinant
    (ceil(log2 C(nbasis, nel)) bits) above its spawned-list slot
    (ceil(log2 spawn_capacity) bits)."""
    bits = max(1, (math.comb(nbasis, nel) - 1).bit_length()) + max(1, (spawn_capacity - 1).bit_length())
    return 1 if bits <= 64 else 2


def entry_bytes(W, S):
    """Bytes of a main-list entry (determinant words, populations, K_ii) and of
    a spawned entry (determinant words, populations, flag)."""
    return 8 * W + 8 * S + 8, 8 * W + 8 * S + 8


def compulsory_bytes(prof, W=1, S=1):
    """The hot path's compulsory HBM bytes per cycle: the main list read by the
    spawning/death pass and by annihilation and written once, the spawned
    list written by spawning and read by annihilation
    (3 N_d B_m + 2 N_s B_s with the written list N_d' B_m)."""
    bm, bs = entry_bytes(W, S)
    nd, ns, nout = prof.nstates_in, prof.nspawned, prof.nstates_out
    return {"cycle": 2 * nd * bm + nout * bm + 2 * ns * bs,
            "annihilation": nd * bm + nout * bm + ns * bs}


def phase_bytes(prof, W=1, S=1, kw=1):
    """Algorithmic HBM bytes of each phase of one cycle as the kernels move
    them; kw: words of the sort key."""
    nd, ns, nr, nout = prof.nstates_in, prof.nspawned, prof.nruns, prof.nstates_out
    bm, bs = entry_bytes(W, S)
    return {
        # read det/pop/K_ii, write post-death pops, attempt counts, flags
        "det_pass": nd * (bm + 8 * S + 8 * S + 8),
        # parents (det, pop, flag) and attempt offsets once, spawned entries and their sort keys written
        "spawn": nd * (8 * W + 8 * S + 8 + 8 * S) + ns * (bs + 8 * kw),
        # one read of the keys for every pass's digit counts; each pass reads and writes them once
        "sort": ns * 8 * kw + prof.radix_passes * ns * 2 * 8 * kw,
        # sorted keys read (and each key's neighbour), each entry's populations
        # and flag gathered by slot, each run's determinant gathered, runs written
        "reduce": ns * (2 * 8 * kw + 8 * S + 8) + nr * (8 * W) + nr * (8 * W + 8 * S + 8 + 8),
        # main list read twice (merge search ~ touched lines ignored) and written once, runs read
        "merge": nd * (2 * bm + 3 * 8) + nout * bm + nr * (8 * W + 8 * S + 8 + 8 * 3),
    }


def download_list(lib, eng, W, S=1):
    n = C.c_int64()
    lib.fq_nstates(eng, C.byref(n))
    nd = n.value
    st = np.zeros(nd * W, dtype=np.uint64)
    po = np.zeros(nd * S, dtype=np.int64)
    da = np.zeros(nd)
    got = C.c_int64()
    N.check(lib.fq_download(eng, st.ctypes.data_as(C.POINTER(C.c_uint64)), po.ctypes.data_as(C.POINTER(C.c_int64)),
                            da.ctypes.data_as(C.POINTER(C.c_double)), nd, C.byref(got)), "fq_download")
    return st, po, da


def traffic_summary(workload):
    """HBM bytes per launch of the spawning kernel measured by the rocprofv3
    FETCH_SIZE / WRITE_SIZE passes for this workload (profiles/), or None."""
    path = os.path.join(ROOT, "profiles", "k_spawn_mol26_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t if t.get("workload") == workload else None


def issue_summary(workload):
    """VALU issue of the spawning kernel from the committed rocprofv3 counter
    pass (profiles/k_spawn_mol26_valu_raw.json), or None.  busy: fraction of
    each SIMD's cycles with a VALU instruction in flight, SQ_ACTIVE_INST_VALU x
    4 cycles / (1024 SIMDs x GRBM_GUI_ACTIVE / 8), the GRBM counter being summed
    over the 8 XCDs (it equals 8 x kernel time x clock); lanes: active lanes
    per VALU instruction / 64, SQ_THREAD_CYCLES_VALU / (64 x
    SQ_ACTIVE_INST_VALU)."""
    path = os.path.join(ROOT, "profiles", "k_spawn_mol26_valu_raw.json")
    try:
        with open(path) as f:
            v = json.load(f)
    except (OSError, ValueError):
        return None
    if v.get("workload") != workload:
        return None
    c = v["raw_per_launch"]
    return {"valu_busy": c["SQ_ACTIVE_INST_VALU"] * 4 / (1024 * c["GRBM_GUI_ACTIVE"] / 8),
            "valu_lane_utilization": c["SQ_THREAD_CYCLES_VALU"] / (64 * c["SQ_ACTIVE_INST_VALU"]),
            "valu_insts_per_launch": c["SQ_INSTS_VALU"], "lane_ops_per_launch": c["SQ_THREAD_CYCLES_VALU"],
            "source": "profiles/k_spawn_mol26_valu_raw.json"}


def spawn_kernel_bytes(prof, W=1, S=1):
    """Algorithmic HBM bytes of the spawning-attempt kernel in one cycle: every
    parent's determinant words and population read once, every spawned entry
    (determinant words, population per space, initiator flag) written once."""
    return prof.nstates_in * (8 * W + 8 * S) + prof.nspawned * (8 * W + 8 * S + 8)


BOX_CPU_SHARE = 16  # host cores the GPU box grants one GPU's jobs, whatever os.cpu_count() shows


def host_cores():
    """Threads for the CPU baseline: the cores this process may run on, at most
    the box's CPU share for one GPU (16), and the count visible, both reported."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(BOX_CPU_SHARE, n)), n


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_flags():
    """Compiler flags of oracle/liboracle.so (from its Makefile)."""
    try:
        with open(os.path.join(ROOT, "oracle", "Makefile")) as f:
            for line in f:
                if line.startswith("CXXFLAGS"):
                    return "g++ " + line.split("=", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def hash_owner(states, W, nranks):
    """Owner rank of each determinant: a 64-bit mix of its words mod nranks, so
    the walkers spread over the ranks as the reference's hashed ownership
    spreads them over MPI ranks (src/spawning.F90:770-838)."""
    h = np.zeros(len(states) // W, dtype=np.uint64)
    words = states.reshape(-1, W)
    with np.errstate(over="ignore"):
        for w in range(W):
            x = words[:, w] ^ h
            x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            h = x ^ (x >> np.uint64(31))
    return (h % np.uint64(nranks)).astype(np.int64)


def cpu_baseline(lib, system, eng, p, shift, eold, tau, budget_s=10.0):
    """Time the CPU oracle on the GPU's current walker list.

    Parallel form: one oracle instance per host core (threads; the oracle's
    ctypes calls release the GIL), each owning the determinants a hash assigns
    to it (sorted, with its own seed), as the reference's MPI ranks own a
    hashed share of the list; no inter-rank exchange is timed, which favours
    the CPU.  Every instance is created and loaded first; then all start
    together and run whole cycles until a common deadline ``budget_s`` later,
    and the rate is the sum of the attempts over the longest instance's wall
    time.  Serial form: one instance over the whole list on one core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import helpers
    orc = helpers.oracle()
    W = system.view.nwords
    S = p.nspaces
    st, po, da = download_list(lib, eng, W, S)
    nd = len(da)
    walkers = np.abs(po).sum() / 2**31
    P, visible = host_cores()

    def make(idx, seed):
        h = orc.orc_create(system.view, C.byref(p), seed)
        sst = np.ascontiguousarray(st.reshape(nd, W)[idx]).reshape(-1)
        spo = np.ascontiguousarray(po.reshape(nd, S)[idx]).reshape(-1)
        sda = np.ascontiguousarray(da[idx])
        orc.orc_load(h, sst.ctypes.data_as(C.POINTER(C.c_uint64)), spo.ctypes.data_as(C.POINTER(C.c_int64)),
                     sda.ctypes.data_as(C.POINTER(C.c_double)), len(sda))
        return h

    def run(h, barrier, budget):
        out = N.FqCycleOut()
        att, cyc = 0, 0
        if barrier is not None:
            barrier.wait()
        t0 = time.perf_counter()
        stop_at = t0 + budget
        while True:
            orc.orc_cycle(h, shift, eold, tau, C.byref(out))
            att += out.nattempts
            cyc += 1
            now = time.perf_counter()
            if now >= stop_at:
                break
        return att, cyc, now - t0

    owner = hash_owner(st, W, P)
    hs = [make(np.nonzero(owner == k)[0], 7 + k) for k in range(P)]
    barrier = threading.Barrier(P)
    with concurrent.futures.ThreadPoolExecutor(max_workers=P) as ex:
        futs = [ex.submit(run, hs[k], barrier, budget_s) for k in range(P)]
        res = [f.result() for f in futs]
    for h in hs:
        orc.orc_destroy(h)
    wall_p = max(r[2] for r in res)
    att_p = sum(r[0] for r in res)
    cyc_p = min(r[1] for r in res)
    h1 = make(np.arange(nd), 7)
    att_1, cyc_1, wall_1 = run(h1, None, budget_s)
    orc.orc_destroy(h1)
    return {"value": att_p / wall_p, "unit": "spawn attempts/s", "cores": P, "kind": "port",
            "sample": f"{P} oracle instances (serial C++ restatement of the reference cycle), one thread each, "
                      f"each owning a hashed 1/{P} share of the {nd} determinants of the same {walkers:.3g}-walker "
                      f"state, all running whole cycles for {budget_s:g} s from a common start "
                      f"(>= {cyc_p} cycles each), {att_p} attempts in {wall_p:.1f} s",
            "cpu_model": host_cpu_model(), "logical_cpus_visible": visible,
            "cpu_share": f"{BOX_CPU_SHARE} cores (the GPU box's share for one GPU)",
            "compiler": oracle_flags(),
            "single_core": {"value": att_1 / wall_1, "cores": 1,
                            "sample": f"one instance over all {nd} determinants, {cyc_1} cycles, "
                                      f"{att_1} attempts in {wall_1:.1f} s"},
            "parallel_efficiency": (att_p / wall_p) / (P * att_1 / wall_1)}


def copy_probe(mib):
    import torch
    n = mib * (1 << 20) // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    ok = bool(torch.equal(dst[-4:], src[-4:]))
    print(json.dumps({"copy_probe_bytes": 8 * n, "repeats": 3, "ok": ok}), flush=True)


# ---------------------------------------------------------------------------
# E_proj statistical efficiency 1/(sigma^2 t)
# ---------------------------------------------------------------------------
def prepare(lib, system, target, seed, cap, equil_tau=1e-3, hold=300, max_cycles=4000):
    """Untimed preparation (renormalised generator): returns the walker list and shift."""
    eng, _p = make_engine(lib, system, target, seed, cap, cap)
    f0 = system.f0()
    pop0 = np.array([max(1, int(target * 1e-3)) << 31], dtype=np.int64)
    dat0 = np.zeros(1)
    N.check(lib.fq_load(eng, f0.ctypes.data_as(C.POINTER(C.c_uint64)), pop0.ctypes.data_as(C.POINTER(C.c_int64)),
                        dat0.ctypes.data_as(C.POINTER(C.c_double)), 1), "fq_load")
    prop = Propagator(lib, eng, target, damping=0.2, pull=1.0)
    vary_at = None
    t0 = time.perf_counter()
    for i in range(max_cycles):
        prop.cycle(equil_tau)
        if prop.vary and vary_at is None:
            vary_at = i
        if (i + 1) % 200 == 0:
            log(f"[eff] prep cycle {i + 1}: N={prop.last.nparticles[0]:.4g}, {time.perf_counter() - t0:.1f}s")
        if vary_at is not None and i - vary_at >= hold:
            break
    st, po, da = download_list(lib, eng, system.view.nwords)
    log(f"[eff] prepared N={prop.last.nparticles[0]:.4g} on {len(da)} determinants, shift {prop.shift[0]:.4f}")
    lib.fq_destroy(eng)
    return st, po, da, prop.shift[0], prop.eold[0]


def production(lib, eng, target, state, tau, cycles, warmup, report=10, prof=None, pull=0.0):
    """Warm-up then timed production cycles from a prepared state; returns the
    per-report (sum H_0j N_j, N_0) series, the per-report walker counts, wall
    time and attempts.  pull > 0 holds every generator's population at the
    target (an integral term on the shift), so generators are compared at
    matched walker numbers.  prof (a dict) receives per-cycle means of the
    learned proposal's time and rows and of the cycle time."""
    st, po, da, shift, eold = state
    N.check(lib.fq_load(eng, st.ctypes.data_as(C.POINTER(C.c_uint64)), po.ctypes.data_as(C.POINTER(C.c_int64)),
                        da.ctypes.data_as(C.POINTER(C.c_double)), len(da)), "fq_load")
    prop = Propagator(lib, eng, target, pull=pull)
    prop.shift[0], prop.eold[0], prop.vary = shift, eold, True
    for _ in range(warmup):
        prop.cycle(tau)
    proj, d0, walkers, attempts = [], [], [], 0
    accp = accd = 0.0
    t0 = time.perf_counter()
    pr = N.FqProfileOut()
    acc_prof = np.zeros(3)
    for i in range(cycles):
        out = prop.cycle(tau)
        attempts += out.nattempts
        if prof is not None:
            N.check(lib.fq_last_profile(eng, C.byref(pr)), "fq_last_profile")
            acc_prof += (pr.proposal_ms, pr.proposal_rows, pr.total_ms)
        accp += out.proj_energy[0]
        accd += out.D0_population[0]
        if (i + 1) % report == 0:
            proj.append(accp)
            d0.append(accd)
            walkers.append(out.nparticles[0])
            accp = accd = 0.0
        if (i + 1) % 500 == 0:
            log(f"[eff] production cycle {i + 1}: N={out.nparticles[0]:.4g}, {time.perf_counter() - t0:.1f}s")
    wall = time.perf_counter() - t0
    if prof is not None:
        m = acc_prof / max(cycles, 1)
        prof.update(proposal_ms=float(m[0]), proposal_rows=float(m[1]), cycle_ms=float(m[2]))
    return np.array(proj), np.array(d0), np.array(walkers), wall, attempts, prop.last


def attempt_coverage(pops, rf, thresholds):
    """Share of the walkers (so, roughly, of the spawning attempts) on
    determinants holding at least K walkers, and the number of such
    determinants, for each threshold K."""
    a = np.abs(np.asarray(pops, dtype=np.float64).reshape(-1)) / rf
    tot = a.sum()
    return {int(k): {"walker_share": float(a[a >= k].sum() / tot), "determinants": int((a >= k).sum())}
            for k in thresholds}


def efficiency_learned(args):
    """E_proj efficiency 1/(sigma^2 t) of the generators on mol40 at matched
    walker numbers (every run's shift holds the population at the target).
    sigma is the Flyvbjerg--Petersen reblocked error of the ratio estimator
    at the Wolff/Lee optimal block (pyblock/pyhande semantics); a run whose
    reblocking finds no optimal block reports no efficiency, and ratios are
    given only between runs that both converged."""
    from hande_amd import reblock
    lib = N.lib()
    norb, nel, seed = synthetic.MOLECULAR_CONFIGS["mol40"]
    tmp = tempfile.mkdtemp(prefix="fqmc_eff_")
    system = api.read_in(int_file=synthetic.write_fcidump(os.path.join(tmp, "FCIDUMP"), norb, nel, seed), nel=nel, ms=0)
    cap = int(max(6 * args.target, 1 << 20))
    state = prepare(lib, system, args.target, 11, cap)
    k = int(str(args.min_attempts).split(",")[0])
    coverage = attempt_coverage(state[1], float(1 << 31), sorted({1, 4, 16, 64, k}))
    log(f"[eff] walker share on determinants with >= K walkers: "
        + ", ".join(f"K={kk}: {v['walker_share']:.3f} ({v['determinants']})" for kk, v in coverage.items()))
    gens = {"renorm": (N.FQ_EXCIT_RENORM, 0), "heat_bath": (N.FQ_EXCIT_HEAT_BATH, 0),
            "learned": (N.FQ_EXCIT_LEARNED, k)}
    wanted = [g for g in str(args.eff_gens).split(",") if g]
    res, training = {}, {}
    sd = int(str(args.eff_seeds).split(",")[0])
    for gen in wanted:
        code, kk = gens[gen]
        eng, _p = make_engine(lib, system, args.target, sd, cap, cap, excit_gen=code, mix=args.proposal_mix,
                              min_attempts=kk)
        t_train = 0.0
        if code == N.FQ_EXCIT_LEARNED:
            if args.objective == "variance":  # also the objective of the pre-training below
                N.check(lib.fq_mlp_online(eng, 0, 0, 1, 1e-3, 1), "fq_mlp_online")
            st, po, da = state[:3]
            N.check(lib.fq_load(eng, st.ctypes.data_as(C.POINTER(C.c_uint64)),
                                po.ctypes.data_as(C.POINTER(C.c_int64)), da.ctypes.data_as(C.POINTER(C.c_double)),
                                len(da)), "fq_load")
            losses = np.zeros(args.train_steps)
            t1 = time.perf_counter()
            N.check(lib.fq_mlp_train(eng, 16384, args.train_steps, 1e-3, 5,
                                     losses.ctypes.data_as(C.POINTER(C.c_double))), "fq_mlp_train")
            t_train = time.perf_counter() - t1
            training = {"seconds": t_train, "loss_first10": float(losses[:10].mean()),
                        "loss_last10": float(losses[-10:].mean()), "objective": args.objective}
            if args.online_every > 0:
                N.check(lib.fq_mlp_online(eng, args.online_every, args.online_steps, 16384, 1e-3,
                                          1 if args.objective == "variance" else 0), "fq_mlp_online")
                training["online"] = {"every": args.online_every, "steps": args.online_steps, "batch": 16384}
            log(f"[eff] learned: trained {args.train_steps} steps in {t_train:.1f}s: loss "
                f"{losses[:10].mean():.4f} -> {losses[-10:].mean():.4f}")
        prof = {}
        proj, d0, walkers, wall, att, last = production(lib, eng, args.target, state, 1e-3, args.eff_cycles,
                                                        args.eff_warmup, prof=prof, pull=1.0)
        lib.fq_destroy(eng)
        rb = reblock.projected_energy(proj, d0)
        t = wall + t_train
        r = {"E_proj": system.H00 + rb["energy"], "sigma": rb["error"], "converged": rb["converged"],
             "reblock_level": rb["level"], "reports": len(proj), "wall_s": wall, "train_s": t_train,
             "efficiency": reblock.efficiency(rb["error"], t) if rb["converged"] else None,
             "attempts_per_s": att / wall, "walkers_mean": float(walkers.mean()),
             "walkers_sd": float(walkers.std()), "determinants": int(last.nstates),
             "cycle_ms": prof["cycle_ms"], "seed": sd}
        if code == N.FQ_EXCIT_LEARNED:
            r.update(min_attempts=kk, proposal_ms=prof["proposal_ms"], network_rows=prof["proposal_rows"],
                     proposal_mix=args.proposal_mix)
        res[gen] = r
        log(f"[eff] {gen}: E_proj {r['E_proj']:.6f} +- {r['sigma']:.6f} (reblocked, level {rb['level']} of "
            f"{len(proj)} reports, converged {rb['converged']}), walkers {r['walkers_mean']:.4g}, {wall:.1f}s, "
            f"efficiency {r['efficiency']}")

    def ratio(a, b):
        x, y = res.get(a), res.get(b)
        if not x or not y or x["efficiency"] is None or y["efficiency"] is None:
            return None
        return {"ratio": x["efficiency"] / y["efficiency"],
                "energies_agree_2sigma": bool(abs(x["E_proj"] - y["E_proj"]) < 2 * math.hypot(x["sigma"], y["sigma"]))}

    print(json.dumps({"metric": "E_proj efficiency 1/(sigma^2 t) on mol40 at matched walker numbers",
                      "unit": "ratio", "higher_is_better": True,
                      "learned_vs_renorm": ratio("learned", "renorm"),
                      "learned_vs_heat_bath": ratio("learned", "heat_bath"),
                      "heat_bath_vs_renorm": ratio("heat_bath", "renorm"),
                      "walker_coverage": coverage, "n_gpus": 1, "data": "synthetic",
                      "config": {"workload": f"i-FCIQMC mol40: synthetic molecular integrals {norb} spatial orbitals, "
                                             f"{nel} electrons, C1, real_64, initiator 3, tau 1e-3, target "
                                             f"{args.target:g} held by the shift, {args.eff_cycles} production "
                                             f"cycles in reports of 10",
                                 "proposal_mix": args.proposal_mix, "proposal_min_attempts": k,
                                 "train_steps": args.train_steps, "warmup_cycles": args.eff_warmup, "seed": sd,
                                 "sigma": "Flyvbjerg-Petersen reblocking, Wolff/Lee optimal block, ratio with "
                                          "covariance"},
                      "training": training, "runs": res}), flush=True)


def c3_instances(n=64, seed=2024):
    """The batched workload: n Hubbard-k rings of 10-18 sites, fillings near
    half, U/t uniform on [1, 8], targets log-uniform on [1e4, 1e6]."""
    rng = np.random.default_rng(seed)
    decks, targets = [], []
    for k in range(n):
        L = int(rng.choice([10, 12, 14, 16, 18]))
        nel = L + int(rng.choice([-2, 0, 0, 2]))
        U = float(rng.uniform(1.0, 8.0))
        decks.append(dict(electrons=nel, lattice=[[L]], ms=0, sym=1, U=U, t=1.0))
        targets.append(float(10 ** rng.uniform(4, 6)))
    return decks, np.array(targets)


def integer_params(seed, cap):
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


def efficiency_batched(args):
    """64 instances: one batched engine against the same instances propagated
    one engine at a time, both from the same prepared lists at cycle 0 with
    the same Philox keys (instance k: philox_seed + k) and the same shift
    control (a function of each instance's own outputs), so every instance
    follows one chain in both forms; the chains are checked equal report by
    report.  sigma of each instance's E_proj is its reblocked error (as
    pyhande computes it); the efficiency of a form is the sum over instances
    of 1/(sigma_k^2 t) with t the form's wall time."""
    from hande_amd import reblock
    lib = N.lib()
    decks, targets = c3_instances(args.instances)
    K = len(decks)
    systems = [api.hubbard_k(**d) for d in decks]
    tau, every = 0.005, 5
    cap = int(6 * targets.sum()) + (1 << 20)
    arr = (N.FqSystem * K)(*[s.view for s in systems])
    beng = C.c_void_p()
    N.check(lib.fq_create_batch(arr, K, C.byref(integer_params(77, cap)), C.byref(beng)), "fq_create_batch")
    inst = np.arange(K, dtype=np.int32)
    f0 = np.array([int(s.f0()[0]) for s in systems], dtype=np.uint64)
    pops = np.full(K, 10, dtype=np.int64)
    N.check(lib.fq_load_batch(beng, inst.ctypes.data_as(C.POINTER(C.c_int32)), f0.ctypes.data_as(C.POINTER(C.c_uint64)),
                              pops.ctypes.data_as(C.POINTER(C.c_int64)), np.zeros(K).ctypes.data_as(C.POINTER(C.c_double)),
                              K), "load")
    shift = np.zeros(K)
    eold = np.zeros(K)
    taus = np.full(K, tau)
    vary = np.zeros(K, bool)
    n_old = np.zeros(K)
    outs = (N.FqCycleOut * K)()

    def batch_cycle():
        N.check(lib.fq_cycle_batch(beng, shift.ctypes.data_as(C.POINTER(C.c_double)),
                                   eold.ctypes.data_as(C.POINTER(C.c_double)), taus.ctypes.data_as(C.POINTER(C.c_double)),
                                   outs), "fq_cycle_batch")

    # preparation (untimed): stronger damping plus an integral pull to the
    # target, from a twentieth of it, so that strongly correlated rings
    # (growth ~10% per cycle at zero shift) do not overshoot
    c = 0
    t0 = time.perf_counter()
    while c < args.prep_cycles and not (vary.all() and c > 400):
        batch_cycle()
        c += 1
        if c % every == 0:
            n = np.array([outs[k].nparticles[0] for k in range(K)])
            upd = vary & (n_old > 0)
            shift[upd] -= (np.log(n[upd] / n_old[upd]) * 0.2 + 0.04 * np.log(n[upd] / targets[upd])) / (tau * every)
            vary[:] |= n > targets / 20
            n_old[:] = n
        if c % 200 == 0:
            log(f"[batch] prep cycle {c}: {sum(outs[k].nparticles[0] for k in range(K)):.4g} walkers, "
                f"{int(vary.sum())}/{K} instances controlled, {time.perf_counter() - t0:.1f}s")
    st, po, da = download_list(lib, beng, 1)
    lib.fq_destroy(beng)
    shift0, nold0 = shift.copy(), n_old.copy()
    log(f"[batch] prepared {K} instances in {c} cycles ({time.perf_counter() - t0:.1f}s): {po.sum():.4g} walkers")
    cycles = args.eff_cycles

    class Control:
        """The production shift update (damping 0.05 every 5 cycles, the
        reference's rule) and E_proj report sums of one set of instances."""

        def __init__(self, ks):
            self.ks = np.asarray(ks)
            self.shift, self.n_old = shift0[self.ks].copy(), nold0[self.ks].copy()
            self.eold = np.zeros(len(self.ks))
            self.accp, self.accd = np.zeros(len(self.ks)), np.zeros(len(self.ks))
            self.proj, self.d0, self.count = [], [], 0

        def update(self, o):  # o: sequence of FqCycleOut for self.ks
            self.count += 1
            self.accp += [x.proj_energy[0] for x in o]
            self.accd += [x.D0_population[0] for x in o]
            if self.count % every == 0:
                n = np.array([x.nparticles[0] for x in o])
                self.shift -= np.log(n / self.n_old) * 0.05 / (tau * every)
                self.n_old = n
                self.eold = np.where(self.accd != 0, self.accp / np.where(self.accd != 0, self.accd, 1), 0.0)
                self.proj.append(self.accp.copy())
                self.d0.append(self.accd.copy())
                self.accp[:] = 0.0
                self.accd[:] = 0.0

    # batched form, from the prepared lists at cycle 0
    beng = C.c_void_p()
    N.check(lib.fq_create_batch(arr, K, C.byref(integer_params(77, cap)), C.byref(beng)), "fq_create_batch")
    N.check(lib.fq_batch_graph(beng, 1), "fq_batch_graph")  # each batched cycle one graph launch
    binst = (st >> np.uint64(48)).astype(np.int32)
    bst = np.ascontiguousarray(st & np.uint64((1 << 48) - 1))
    N.check(lib.fq_load_batch(beng, binst.ctypes.data_as(C.POINTER(C.c_int32)), bst.ctypes.data_as(C.POINTER(C.c_uint64)),
                              po.ctypes.data_as(C.POINTER(C.c_int64)), da.ctypes.data_as(C.POINTER(C.c_double)),
                              len(da)), "load")
    cb = Control(range(K))
    attempts = 0
    t1 = time.perf_counter()
    for i in range(cycles):
        shift[:], eold[:] = cb.shift, cb.eold
        batch_cycle()
        cb.update([outs[k] for k in range(K)])
        attempts += sum(outs[k].nattempts for k in range(K))
        if (i + 1) % 500 == 0:
            log(f"[batch] batched cycle {i + 1}: {time.perf_counter() - t1:.1f}s")
    t_batch = time.perf_counter() - t1
    info = N.FqScheduleOut()
    N.check(lib.fq_batch_schedule(beng, C.byref(info)), "fq_batch_schedule")
    lib.fq_destroy(beng)
    # the same instances, one engine each
    t_solo, att_solo, same = 0.0, 0, True
    sproj, sd0 = [], []
    for k in range(K):
        sel = binst == k
        sk, pk, dk = (np.ascontiguousarray(x[sel]) for x in (bst, po, da))
        eng = C.c_void_p()
        N.check(lib.fq_create(systems[k].view, C.byref(integer_params(77 + k, cap)), C.byref(eng)), "fq_create")
        N.check(lib.fq_load(eng, sk.ctypes.data_as(C.POINTER(C.c_uint64)), pk.ctypes.data_as(C.POINTER(C.c_int64)),
                            dk.ctypes.data_as(C.POINTER(C.c_double)), len(dk)), "fq_load")
        ck = Control([k])
        o = N.FqCycleOut()
        t2 = time.perf_counter()
        for i in range(cycles):
            N.check(lib.fq_cycle(eng, (C.c_double * 2)(ck.shift[0], 0.0), (C.c_double * 2)(ck.eold[0], 0.0), tau,
                                 C.byref(o)), "fq_cycle")
            ck.update([o])
            att_solo += o.nattempts
        t_solo += time.perf_counter() - t2
        lib.fq_destroy(eng)
        p_k, d_k = np.array(ck.proj)[:, 0], np.array(ck.d0)[:, 0]
        same = same and np.array_equal(p_k, np.array(cb.proj)[:, k]) and np.array_equal(d_k, np.array(cb.d0)[:, k])
        sproj.append(p_k)
        sd0.append(d_k)
        if (k + 1) % 8 == 0:
            log(f"[batch] one engine per instance: {k + 1}/{K} done, {t_solo:.1f}s")
    per, inv_var = [], []
    for k in range(K):
        rb = reblock.projected_energy(np.array(cb.proj)[:, k], np.array(cb.d0)[:, k])
        per.append({"E_proj": systems[k].H00 + rb["energy"], "sigma": rb["error"], "converged": rb["converged"]})
        if rb["converged"] and rb["error"] > 0:
            inv_var.append(1.0 / rb["error"] ** 2)
    s_inv = float(np.sum(inv_var))
    eff_b, eff_s = s_inv / t_batch, s_inv / t_solo
    print(json.dumps({"metric": "E_proj efficiency 1/(sigma^2 t), batched vs one engine per instance",
                      "value": eff_b / eff_s if s_inv > 0 else None, "unit": "ratio", "higher_is_better": True,
                      "n_gpus": 1, "data": "synthetic", "chains_identical": bool(same),
                      "instances_converged": len(inv_var),
                      "efficiency": {"batched": eff_b, "one_engine_per_instance": eff_s,
                                     "definition": "sum over converged instances of 1/(sigma_k^2 t)"},
                      "config": {"workload": f"{K} Hubbard-k rings of 10-18 sites, U/t 1-8, targets 1e4-1e6, "
                                             f"integer populations, tau {tau}, {cycles} cycles, reports of {every}"},
                      "batched": {"wall_s": t_batch, "attempts_per_s": attempts / t_batch,
                                  "graph_cycles": info.graph_cycles, "ordinary_cycles": info.eager_cycles,
                                  "schedule": {"digit_bits": info.digit_bits, "spawn_blocks": info.spawn_blocks,
                                               "tuned_octaves": info.tuned_octaves}},
                      "one_engine_per_instance": {"wall_s": t_solo, "attempts_per_s": att_solo / t_solo},
                      "instances": per}), flush=True)


def batched_leg(cycles=200, prep_cycles=1500, graph=True):
    """The batched workload (64 Hubbard-k rings of 10-18 sites, targets 1e4-1e6,
    integer populations) with graph-captured cycles: prepared as the batched
    efficiency mode prepares it, then `cycles` timed graph cycles (their device
    time from the engine's events around each graph launch); the tuner's
    candidate timings and its choice are reported."""
    lib = N.lib()
    decks, targets = c3_instances(64)
    K = len(decks)
    systems = [api.hubbard_k(**d) for d in decks]
    tau, every = 0.005, 5
    cap = int(6 * targets.sum()) + (1 << 20)
    arr = (N.FqSystem * K)(*[s.view for s in systems])
    eng = C.c_void_p()
    N.check(lib.fq_create_batch(arr, K, C.byref(integer_params(77, cap)), C.byref(eng)), "fq_create_batch")
    try:
        N.check(lib.fq_batch_graph(eng, 1 if graph else 0), "fq_batch_graph")
        inst = np.arange(K, dtype=np.int32)
        f0 = np.array([int(s.f0()[0]) for s in systems], dtype=np.uint64)
        pops = np.full(K, 10, dtype=np.int64)
        N.check(lib.fq_load_batch(eng, inst.ctypes.data_as(C.POINTER(C.c_int32)),
                                  f0.ctypes.data_as(C.POINTER(C.c_uint64)), pops.ctypes.data_as(C.POINTER(C.c_int64)),
                                  np.zeros(K).ctypes.data_as(C.POINTER(C.c_double)), K), "load")
        shift, eold, taus = np.zeros(K), np.zeros(K), np.full(K, tau)
        vary, n_old = np.zeros(K, bool), np.zeros(K)
        outs = (N.FqCycleOut * K)()
        prof = N.FqProfileOut()
        t0 = time.perf_counter()
        dev_ms, wall, attempts = [], 0.0, 0
        for c in range(prep_cycles + cycles):
            t1 = time.perf_counter()
            N.check(lib.fq_cycle_batch(eng, shift.ctypes.data_as(C.POINTER(C.c_double)),
                                       eold.ctypes.data_as(C.POINTER(C.c_double)),
                                       taus.ctypes.data_as(C.POINTER(C.c_double)), outs), "fq_cycle_batch")
            if c >= prep_cycles:
                wall += time.perf_counter() - t1
                lib.fq_last_profile(eng, C.byref(prof))
                dev_ms.append(prof.total_ms)
                attempts += sum(outs[k].nattempts for k in range(K))
            if (c + 1) % every == 0:
                n = np.array([outs[k].nparticles[0] for k in range(K)])
                upd = vary & (n_old > 0)
                damp = 0.2 if c < prep_cycles else 0.05
                pull = 0.04 if c < prep_cycles else 0.0
                shift[upd] -= (np.log(n[upd] / n_old[upd]) * damp + pull * np.log(n[upd] / targets[upd])) / (tau * every)
                vary[:] |= n > targets / 20
                n_old[:] = n
        info = N.FqScheduleOut()
        N.check(lib.fq_batch_schedule(eng, C.byref(info)), "fq_batch_schedule")
        walkers = sum(outs[k].nparticles[0] for k in range(K))
        log(f"[bench] batched: {walkers:.4g} walkers, {np.median(dev_ms):.3f} ms/cycle device "
            f"({'graph' if graph else 'ordinary launches'}), "
            f"{1e3 * wall / cycles:.3f} ms/cycle wall, {time.perf_counter() - t0:.1f}s")
        return {"workload": f"{K} Hubbard-k rings of 10-18 sites, U/t 1-8, targets 1e4-1e6, integer populations, "
                            f"tau {tau}",
                "walkers": walkers, "cycles": cycles, "graph_launches": bool(graph), "attempts_per_s": attempts / wall,
                "ms_per_cycle_wall": 1e3 * wall / cycles, "ms_per_cycle_device": float(np.median(dev_ms)),
                "graph": {"graph_cycles": info.graph_cycles, "ordinary_cycles": info.eager_cycles,
                          "graphs_captured": info.graphs, "tuned": bool(info.tuned),
                          "attempt_octave": info.octave, "tuned_octaves": info.tuned_octaves,
                          "chosen": {"digit_bits": info.digit_bits, "spawn_blocks": info.spawn_blocks},
                          "candidates": [{"digit_bits": info.cand_digit_bits[q],
                                          "spawn_blocks": info.cand_spawn_blocks[q],
                                          "median_ms": info.cand_ms[q]} for q in range(info.ncand)]}}
    finally:
        lib.fq_destroy(eng)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="mol26", choices=sorted(synthetic.MOLECULAR_CONFIGS))
    ap.add_argument("--target", type=float, default=1e7)
    ap.add_argument("--replicas", type=int, default=1, choices=[1, 2], help="walker spaces (2: replica pair)")
    ap.add_argument("--equil-cycles", type=int, default=4000, help="max untimed preparation cycles")
    ap.add_argument("--hold-cycles", type=int, default=250, help="preparation cycles after the shift starts varying")
    ap.add_argument("--equil-tau", type=float, default=1e-3,
                    help="preparation time step (the production one: a larger step changes the stationary state "
                         "of systems with a wide diagonal spread)")
    ap.add_argument("--prep-damping", type=float, default=0.2, help="shift damping while preparing")
    ap.add_argument("--save-state", default=None, help="write the prepared walker list and shift to this .npz")
    ap.add_argument("--load-state", default=None, help="start from a list written by --save-state (no preparation)")
    ap.add_argument("--start-frac", type=float, default=1e-3, help="initial HF population / target")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched graph-cycle leg")
    ap.add_argument("--main-cap", type=int, default=0, help="main-list capacity (default 4 x target)")
    ap.add_argument("--spawn-cap", type=int, default=0, help="spawned-list capacity (default 4 x target)")
    ap.add_argument("--batched-only", action="store_true", help="run only the batched leg and print its JSON")
    ap.add_argument("--batched-graph", type=int, default=1, choices=[0, 1],
                    help="batched leg: graph-captured cycles (1) or ordinary launches (0)")
    ap.add_argument("--efficiency", choices=["learned", "batched"], default=None,
                    help="measure the E_proj statistical efficiency of the learned proposal (against heat-bath, "
                         "mol40) or of batched instances (against one engine per instance) instead")
    ap.add_argument("--eff-cycles", type=int, default=None,
                    help="timed cycles of an efficiency mode (default 20000 learned: 2000 reports of 10 for the "
                         "reblocking; 5000 batched: 1000 reports of 5)")
    ap.add_argument("--train-steps", type=int, default=1000)
    ap.add_argument("--objective", choices=["xent", "variance"], default="variance",
                    help="learned proposal training objective: cross-entropy to |H| or the spawning variance")
    ap.add_argument("--online-every", type=int, default=50,
                    help="learned proposal: train online every this many production cycles (0: pre-training only)")
    ap.add_argument("--online-steps", type=int, default=2, help="learned proposal: Adam steps per online training")
    ap.add_argument("--proposal-mix", type=float, default=0.1,
                    help="learned proposal: fraction of doubles from the uniform fallback")
    ap.add_argument("--eff-seeds", default="23", help="learned efficiency: production seed")
    ap.add_argument("--eff-gens", default="renorm,heat_bath,learned",
                    help="learned efficiency: generators to run (comma-separated; ratios need both runs)")
    ap.add_argument("--eff-warmup", type=int, default=500, help="untimed cycles per generator before measuring")
    ap.add_argument("--min-attempts", default="64",
                    help="learned proposal: attempt threshold for the network; determinants below use flat "
                         "weights with the learned affinity")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--prep-cycles", type=int, default=3000)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB three times (calibrates the HBM "
                         "byte counters FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.eff_cycles is None:
        args.eff_cycles = 20000 if args.efficiency == "learned" else 5000
    if args.batched_only:
        print(json.dumps(batched_leg(graph=bool(args.batched_graph))), flush=True)
        return
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.efficiency == "learned":
        efficiency_learned(args)
        return
    if args.efficiency == "batched":
        efficiency_batched(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the walker-propagation path runs on a single MI355X here "
              f"(--gpus {args.gpus} requested). Multi-GPU sharding is a design, not a measured path.",
              file=sys.stderr)
        sys.exit(2)
    norb, nel, seed = synthetic.MOLECULAR_CONFIGS[args.config]
    tmp = tempfile.mkdtemp(prefix="fqmc_bench_")
    path = synthetic.write_fcidump(os.path.join(tmp, "FCIDUMP"), norb, nel, seed)
    lib = N.lib()
    system = api.read_in(int_file=path, nel=nel, ms=0)
    log(f"[bench] {args.config}: nbasis={system.view.nbasis} nel={nel} H00={system.view.H00:.8f}")
    cap = int(max(4 * args.target, 1 << 20))
    S = args.replicas
    W = system.view.nwords
    mcap = args.main_cap or cap
    scap = args.spawn_cap or cap
    eng, p = make_engine(lib, system, args.target, 7, mcap, scap, spaces=S)
    rf = 1 << 31
    f0 = system.f0()
    region = ProfilerRegion()
    region.pause()
    prop = Propagator(lib, eng, args.target, damping=args.prep_damping, pull=1.0, spaces=S)
    tau = 1e-3
    if args.load_state:
        z = np.load(args.load_state)
        st, po, da = (np.ascontiguousarray(z[k]) for k in ("states", "pops", "kii"))
        N.check(lib.fq_load(eng, st.ctypes.data_as(C.POINTER(C.c_uint64)), po.ctypes.data_as(C.POINTER(C.c_int64)),
                            da.ctypes.data_as(C.POINTER(C.c_double)), len(da)), "fq_load")
        prop.shift[0], prop.eold[0] = float(z["shift"]), float(z["eold"])
        prop.vary = True
        log(f"[bench] loaded {len(da)} determinants from {args.load_state}")
    else:
        pop0 = np.full(S, max(1, int(args.target * args.start_frac)) * rf, dtype=np.int64)
        dat0 = np.zeros(1)
        N.check(lib.fq_load(eng, f0.ctypes.data_as(C.POINTER(C.c_uint64)), pop0.ctypes.data_as(C.POINTER(C.c_int64)),
                            dat0.ctypes.data_as(C.POINTER(C.c_double)), 1), "fq_load")
        t0 = time.perf_counter()
        vary_at = None
        for i in range(args.equil_cycles):
            out = prop.cycle(args.equil_tau)
            if prop.vary and vary_at is None:
                vary_at = i
            if i % 50 == 0:
                log(f"[bench] prepare {i}: N={out.nparticles[0]:.4g} dets={out.nstates} shift={prop.shift[0]:.4f} "
                    f"({time.perf_counter() - t0:.1f}s)")
            if vary_at is not None and i - vary_at >= args.hold_cycles:
                break
        log(f"[bench] prepared: N={prop.last.nparticles[0]:.4g} dets={prop.last.nstates} shift={prop.shift[0]:.4f}")
    if args.save_state:
        st, po, da = download_list(lib, eng, W, S)
        np.savez(args.save_state, states=st, pops=po, kii=da, shift=prop.shift[0], eold=prop.eold[0])
        log(f"[bench] saved {len(da)} determinants to {args.save_state}")
    prop.damping = 0.05  # the workload's shift damping from here on
    prop.pull = 0.0
    for i in range(args.warmup):
        prop.cycle(tau)
    profs = []
    attempts = 0
    wall = 0.0
    region.resume()
    for i in range(args.steps):
        t1 = time.perf_counter()
        out = prop.cycle(tau)
        wall += time.perf_counter() - t1
        attempts += out.nattempts
        pr = N.FqProfileOut()
        lib.fq_last_profile(eng, C.byref(pr))
        profs.append(pr)
    region.pause()
    value = attempts / wall
    last = prop.last
    phases = ["det_pass", "spawn", "sort", "reduce", "merge"]
    avg_ms = {ph: float(np.mean([getattr(pr, ph + "_ms") for pr in profs])) for ph in phases}
    kw = sort_key_words(system.view.nbasis, system.view.nel, cap)
    avg_bytes = {ph: float(np.mean([phase_bytes(pr, W, S, kw)[ph] for pr in profs])) for ph in phases}
    comp = {k: float(np.mean([compulsory_bytes(pr, W, S)[k] for pr in profs])) for k in ("cycle", "annihilation")}
    launches = sum(pr.spawn_launches for pr in profs)
    k_ms = sum(pr.spawn_kernel_ms for pr in profs) / launches
    k_bytes = sum(spawn_kernel_bytes(pr, W, S) for pr in profs) / launches
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    # the spawning kernel dominates when its time exceeds every other phase's total
    dominant = all(k_ms * launches / len(profs) >= avg_ms[ph] for ph in phases if ph != "spawn")
    annih_ms = sum(pr.sort_ms + pr.reduce_ms + pr.merge_ms for pr in profs)
    annih_rate = sum(pr.nspawned for pr in profs) / (annih_ms * 1e-3)
    log("[bench] phase ms: " + ", ".join(f"{ph}={avg_ms[ph]:.3f}" for ph in phases) +
        f", cycle={np.mean([pr.total_ms for pr in profs]):.3f}")
    log("[bench] phase GB/s (algorithmic): " +
        ", ".join(f"{ph}={avg_bytes[ph] / (avg_ms[ph] * 1e-3) / 1e9:.1f}" for ph in phases if avg_ms[ph] > 0))
    log(f"[bench] k_spawn: {launches} launches, {k_ms:.3f} ms/launch, {k_bytes / 1e6:.1f} MB/launch, "
        f"{achieved:.1f} GB/s; annihilated entries/s {annih_rate:.4g}")
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(lib, system, eng, p, prop.shift, prop.eold, tau)
        except Exception as ex:  # the baseline is reported, never the target
            log(f"[bench] cpu baseline failed: {ex}")
    workload = (f"i-FCIQMC {args.config}: synthetic molecular integrals {norb} spatial orbitals, "
                f"{nel} electrons, C1, real_64, initiator 3, tau 1e-3, renorm, target {args.target:g}"
                + (", replica pair" if S == 2 else ""))
    traffic = traffic_summary(workload)
    issue = issue_summary(workload)
    cycle_ms = float(np.mean([pr.total_ms for pr in profs]))
    annih_phase_ms = avg_ms["sort"] + avg_ms["reduce"] + avg_ms["merge"]
    valu = None
    if issue:
        lane_ops = issue["lane_ops_per_launch"] / (k_ms * 1e-3)
        valu = {"achieved": lane_ops / 1e12, "peak": VALU_PEAK_LANE_OPS / 1e12, "unit": "Tlane-op/s",
                "frac": lane_ops / VALU_PEAK_LANE_OPS, "busy": issue["valu_busy"],
                "lane_utilization": issue["valu_lane_utilization"], "source": issue["source"]}
    result = {
        "metric": "spawn attempts/s",
        "value": value,
        "unit": "spawn attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "walkers": sum(last.nparticles[k] for k in range(S)),
            "determinants": int(last.nstates),
            "attempts_per_step": attempts / args.steps,
            "parallelism": "1 GPU",
        },
        "annihilated_entries_per_s": annih_rate,
        # the dominant kernel in HBM terms (achieved / peak / frac / traffic), with
        # what actually limits it: VALU issue (busy >= 0.9 of SIMD cycles, lanes
        # lost to divergence), priced against the VALU lane-op peak in "valu";
        # "annihilation" and "cycle" are the HBM-bound phases on their
        # compulsory bytes (3 N_d B_m + 2 N_s B_s per cycle)
        "roofline": {"bound": "hbm", "kernel": "k_spawn", "dominant": dominant, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "limiter": "valu-issue" if valu else None, "valu": valu,
                     "launch_ms": k_ms, "bytes_per_launch": k_bytes, "launches": launches,
                     "annihilation": {"ms": annih_phase_ms, "compulsory_bytes": comp["annihilation"],
                                      "achieved": comp["annihilation"] / (annih_phase_ms * 1e-3) / 1e9,
                                      "unit": "GB/s",
                                      "frac": comp["annihilation"] / (annih_phase_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
                     "cycle": {"ms": cycle_ms, "compulsory_bytes": comp["cycle"],
                               "achieved": comp["cycle"] / (cycle_ms * 1e-3) / 1e9, "unit": "GB/s",
                               "frac": comp["cycle"] / (cycle_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
                     "phase_ms": avg_ms, "phase_bytes": avg_bytes},
        "cpu_baseline": cpu,
    }
    lib.fq_destroy(eng)
    if not args.no_batched and not args.load_state:
        try:
            result["batched"] = batched_leg()
        except Exception as ex:  # a side measurement; the headline line still prints
            log(f"[bench] batched leg failed: {ex}")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
