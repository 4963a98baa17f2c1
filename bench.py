#!/usr/bin/env python
"""Benchmark: spawn attempts/s of the FCIQMC evolve loop on one MI355X.

Workload (the molecular i-FCIQMC configuration): seeded synthetic molecular
Hamiltonian, 26 spatial orbitals (52 spin orbitals), 10 electrons, C1
symmetry; real_64 fixed-point populations, initiator threshold 3, target
population 1e7, tau 1e-3, renormalised excitation generator, FP64 integrals.

A step is one Monte Carlo cycle (spawning + death + annihilation) of the whole
walker population resident in HBM.  The walker distribution is prepared
untimed: 2e5 walkers on the Hartree--Fock determinant are propagated at a larger time
step (0.01) until the population passes the 1e7 target and the shift has held
it for 150 cycles, then W warm-up cycles run at tau = 1e-3 before K timed
cycles.

Prints one JSON line (rank 0).  Extra fields:
  roofline      dominant kernel phase: algorithmic HBM bytes per launch / its
                average device time (HIP events on the engine's stream) vs the
                8 TB/s HBM3E peak;
  cpu_baseline  the CPU oracle (a serial restatement of the reference loop)
                timed on one cycle of the same state on one host core.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from hande_amd import _native as N  # noqa: E402
from hande_amd import api, synthetic  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_engine(lib, system, target, seed, main_cap, spawn_cap):
    p = N.FqParams()
    p.nspaces = 1
    p.initiator = 1
    p.initiator_pop = 3.0
    p.quadrature_initiator = 1
    p.real_amplitudes = 1
    p.pop_bit_shift = 31
    p.spawn_cutoff = int(math.ceil(0.01 * (1 << 31)))
    p.excit_gen = N.FQ_EXCIT_RENORM
    p.rng_mode = N.FQ_RNG_PHILOX
    p.philox_seed = seed
    p.main_capacity = main_cap
    p.spawn_capacity = spawn_cap
    p.device = 0
    eng = C.c_void_p()
    N.check(lib.fq_create(system.view, C.byref(p), C.byref(eng)), "fq_create")
    return eng, p


class Propagator:
    """Cycle driver with the reference's shift update (src/energy_evaluation.F90:659-711).

    When the population first exceeds the target the shift starts from the
    projected energy of the last update window (vary_shift_from = 'proje',
    src/energy_evaluation.F90:638-648)."""

    def __init__(self, lib, eng, target, damping=0.05, update_every=5):
        self.lib, self.eng = lib, eng
        self.target, self.damping, self.every = target, damping, update_every
        self.shift = (C.c_double * 2)(0.0, 0.0)
        self.eold = (C.c_double * 2)(0.0, 0.0)
        self.vary = False
        self.n_old = None
        self.count = 0
        self.proj = self.d0 = 0.0
        self.last = None

    def cycle(self, tau):
        out = N.FqCycleOut()
        N.check(self.lib.fq_cycle(self.eng, self.shift, self.eold, tau, C.byref(out)), "fq_cycle")
        self.count += 1
        self.proj += out.proj_energy[0]
        self.d0 += out.D0_population[0]
        n = out.nparticles[0]
        if self.count % self.every == 0:
            if self.vary and self.n_old:
                self.shift[0] = self.shift[0] - math.log(n / self.n_old) * self.damping / (tau * self.every)
            if abs(self.d0) > 0:
                self.eold[0] = self.proj / self.d0
            if not self.vary and n > self.target:
                self.vary = True
                self.shift[0] = self.eold[0]
            self.n_old = n
            self.proj = self.d0 = 0.0
        self.last = out
        return out


def phase_bytes(prof, W=1, S=1):
    """Algorithmic HBM bytes of each phase of one cycle (compulsory traffic)."""
    nd, ns, nr, nout = prof.nstates_in, prof.nspawned, prof.nruns, prof.nstates_out
    bm = 8 * W + 8 * S + 8          # main-list entry: det words, pops, K_ii
    bs = 8 * W + 8 * S + 8 + 8      # spawned entry: det words, pops, flag, attempt ordinal
    return {
        # read det/pop/K_ii, write post-death pops, attempt counts, flags
        "det_pass": nd * (bm + 8 * S + 8 * S + 8),
        # parents (det, pop, flag) and attempt offsets once, spawned entries written
        "spawn": nd * (8 * W + 8 * S + 8 + 8 * S) + ns * bs,
        # each 8-bit LSD pass reads keys+indices twice (histogram, scatter) and writes them once
        "sort": prof.radix_passes * ns * (3 * (8 * W + 4)),
        # run starts, gather of pops/flags/ordinals, unique runs written
        "reduce": ns * (8 * W + 8 + 4 + 8 * S + 8 + 8) + nr * (8 * W + 8 * S + 8 + 8),
        # main list read twice (merge search ~ touched lines ignored) and written once, runs read
        "merge": nd * (2 * bm + 3 * 8) + nout * bm + nr * (8 * W + 8 * S + 8 + 8 * 3),
    }


def cpu_baseline(lib, system, eng, p, shift, eold, tau, max_seconds=30.0):
    """Time the CPU oracle on one cycle from the GPU's current walker list."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import helpers
    orc = helpers.oracle()
    n = C.c_int64()
    lib.fq_nstates(eng, C.byref(n))
    nd = n.value
    W = system.view.nwords
    st = np.zeros(nd * W, dtype=np.uint64)
    po = np.zeros(nd, dtype=np.int64)
    da = np.zeros(nd)
    got = C.c_int64()
    N.check(lib.fq_download(eng, st.ctypes.data_as(C.POINTER(C.c_uint64)), po.ctypes.data_as(C.POINTER(C.c_int64)),
                            da.ctypes.data_as(C.POINTER(C.c_double)), nd, C.byref(got)), "fq_download")
    h = orc.orc_create(system.view, C.byref(p), 7)
    orc.orc_load(h, st.ctypes.data_as(C.POINTER(C.c_uint64)), po.ctypes.data_as(C.POINTER(C.c_int64)),
                 da.ctypes.data_as(C.POINTER(C.c_double)), nd)
    out = N.FqCycleOut()
    t0 = time.perf_counter()
    orc.orc_cycle(h, shift, eold, tau, C.byref(out))
    dt = time.perf_counter() - t0
    orc.orc_destroy(h)
    return {"value": out.nattempts / dt, "unit": "spawn attempts/s", "cores": 1, "kind": "port",
            "sample": f"one full cycle ({out.nattempts} attempts, {nd} determinants) of the same "
                      f"{np.abs(po).sum() / 2**31:.3g}-walker state, serial C++ oracle, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="mol26", choices=sorted(synthetic.MOLECULAR_CONFIGS))
    ap.add_argument("--target", type=float, default=1e7)
    ap.add_argument("--equil-cycles", type=int, default=1200, help="max untimed preparation cycles")
    ap.add_argument("--hold-cycles", type=int, default=150, help="preparation cycles after the shift starts varying")
    ap.add_argument("--equil-tau", type=float, default=0.01)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
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
    eng, p = make_engine(lib, system, args.target, 7, cap, cap)
    rf = 1 << 31
    f0 = system.f0()
    pop0 = np.array([int(args.target / 50) * rf], dtype=np.int64)
    dat0 = np.zeros(1)
    N.check(lib.fq_load(eng, f0.ctypes.data_as(C.POINTER(C.c_uint64)), pop0.ctypes.data_as(C.POINTER(C.c_int64)),
                        dat0.ctypes.data_as(C.POINTER(C.c_double)), 1), "fq_load")
    prop = Propagator(lib, eng, args.target)
    tau = 1e-3
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
    for i in range(args.warmup):
        prop.cycle(tau)
    profs = []
    attempts = 0
    wall = 0.0
    for i in range(args.steps):
        t1 = time.perf_counter()
        out = prop.cycle(tau)
        wall += time.perf_counter() - t1
        attempts += out.nattempts
        pr = N.FqProfileOut()
        lib.fq_last_profile(eng, C.byref(pr))
        profs.append(pr)
    value = attempts / wall
    last = prop.last
    phases = ["det_pass", "spawn", "sort", "reduce", "merge"]
    avg_ms = {ph: float(np.mean([getattr(pr, ph + "_ms") for pr in profs])) for ph in phases}
    avg_bytes = {ph: float(np.mean([phase_bytes(pr)[ph] for pr in profs])) for ph in phases}
    dom = max(phases, key=lambda ph: avg_ms[ph])
    achieved = avg_bytes[dom] / (avg_ms[dom] * 1e-3) / 1e9
    log("[bench] phase ms: " + ", ".join(f"{ph}={avg_ms[ph]:.3f}" for ph in phases) +
        f", cycle={np.mean([pr.total_ms for pr in profs]):.3f}")
    log("[bench] phase GB/s (algorithmic): " +
        ", ".join(f"{ph}={avg_bytes[ph] / (avg_ms[ph] * 1e-3) / 1e9:.1f}" for ph in phases if avg_ms[ph] > 0))
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(lib, system, eng, p, prop.shift, prop.eold, tau)
        except Exception as ex:  # the baseline is reported, never the target
            log(f"[bench] cpu baseline failed: {ex}")
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
            "workload": f"i-FCIQMC {args.config}: synthetic molecular integrals {norb} spatial orbitals, "
                        f"{nel} electrons, C1, real_64, initiator 3, tau 1e-3, renorm, target {args.target:g}",
            "walkers": last.nparticles[0],
            "determinants": int(last.nstates),
            "attempts_per_step": attempts / args.steps,
            "parallelism": "1 GPU",
        },
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "phase_ms": avg_ms, "phase_bytes": avg_bytes},
        "cpu_baseline": cpu,
    }
    lib.fq_destroy(eng)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
