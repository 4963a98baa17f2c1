"""Python mirror of the reference's FCIQMC input surface.

``hubbard_k{}``, ``read_in{}`` and ``fciqmc{}`` take the same keys and
defaults as the reference's Lua functions (src/lua_hande_systems.f90:300-380,
560-620; src/lua_hande_calc.f90:1244-1503; defaults src/qmc_data.f90:121-303)
and run on the MI355X engine through the C-ABI in include/fqmc_host.h.
"""
import ctypes as C
import os

import numpy as np

from . import _native as N
from .lua_input import LuaCall, parse

_EXCIT = {"renorm": N.FQ_EXCIT_RENORM, "no_renorm": N.FQ_EXCIT_NO_RENORM, "heat_bath": N.FQ_EXCIT_HEAT_BATH,
          "learned": N.FQ_EXCIT_LEARNED}
_RNG = {"philox": N.FQ_RNG_PHILOX, "replay": N.FQ_RNG_REPLAY}


class System:
    """A system set up on the host; owns the tables the engine uploads."""

    def __init__(self, handle, lib):
        self._h = handle
        self._lib = lib
        self.view = lib.fqh_view(handle).contents

    @property
    def handle(self):
        return self._h

    def __del__(self):
        if getattr(self, "_h", None) and self._lib is not None:
            self._lib.fqh_free(self._h)
            self._h = None

    @property
    def nbasis(self):
        return self.view.nbasis

    @property
    def nel(self):
        return self.view.nel

    @property
    def H00(self):
        return self.view.H00

    def reference_occ(self):
        occ = (C.c_int * self.view.nel)()
        N.check(self._lib.fqh_ref_occ(self._h, occ), "fqh_ref_occ")
        return list(occ)

    def f0(self):
        return np.array([self.view.f0[i] for i in range(self.view.nwords)], dtype=np.uint64)

    def basis(self):
        n = self.view.nbasis
        ms = (C.c_int * n)()
        lab = (C.c_int * n)()
        eps = (C.c_double * n)()
        N.check(self._lib.fqh_basis_info(self._h, ms, lab, eps), "fqh_basis_info")
        return list(ms), list(lab), list(eps)

    def diagonal(self, f):
        arr = np.ascontiguousarray(f, dtype=np.uint64)
        return self._lib.fqh_diagonal(self._h, arr.ctypes.data_as(C.POINTER(C.c_uint64)))


def hubbard_k(electrons=None, nel=None, lattice=None, ms=0, sym=None, U=1.0, t=1.0, lib=None, **_):
    """hubbard_k{...}: Hubbard model in a momentum-space basis."""
    lib = lib or N.lib()
    nel = electrons if electrons is not None else nel
    if nel is None or lattice is None:
        raise ValueError("hubbard_k needs electrons/nel and lattice")
    ndim = len(lattice)
    flat = [int(x) for vec in lattice for x in vec]
    mode, index = N.FQH_SYM_AUFBAU, 0
    if isinstance(sym, str):
        mode = N.FQH_SYM_TOT if sym == "tot_sym" else N.FQH_SYM_AUFBAU
    elif sym is not None:
        mode, index = N.FQH_SYM_INDEX, int(sym)
    h = C.c_void_p()
    arr = (C.c_int * len(flat))(*flat)
    N.check(lib.fqh_hubbard_k(int(nel), int(ms), ndim, arr, float(U), float(t), mode, index, C.byref(h)),
            "hubbard_k")
    return System(h, lib)


def read_in(int_file=None, nel=None, electrons=None, ms=0, sym=None, CAS=None, base_dir=".", lib=None, **_):
    """read_in{...}: molecular Hamiltonian from an FCIDUMP-format integral file (RHF)."""
    lib = lib or N.lib()
    nel = electrons if electrons is not None else nel
    path = int_file if os.path.isabs(int_file) else os.path.join(base_dir, int_file)
    cas = CAS or (-1, -1)
    symv = -1 if (sym is None or isinstance(sym, str)) else int(sym)
    h = C.c_void_p()
    N.check(lib.fqh_read_in(path.encode(), int(nel or 0), int(ms), symv, int(cas[0]), int(cas[1]), C.byref(h)),
            "read_in")
    return System(h, lib)


BLOCKING_DEFAULTS = {  # blocking{} (src/qmc_data.f90:425-460)
    "blocking_on_the_fly": False, "auto_shift_damping": False, "shift_damping_precision": 2.0,
    "force_shift_damping_opt": False, "start_save_frequency": -1, "start_point_number": -1,
    "start_point": -1, "error_limit": 0.0, "min_blocks_used": 10, "blocks_used": 2147483647,
}


def qmc_in(qmc, fciqmc_opts=None, rng_mode="replay", device=0, blocking=None, semi_stoch=None):
    """Build the qmc/fciqmc/blocking/semi_stoch option block with the
    reference's defaults."""
    q = dict(qmc)
    f = dict(fciqmc_opts or {})
    b = dict(blocking or {})
    ss = dict(semi_stoch or {})
    unknown = set(b) - set(BLOCKING_DEFAULTS) - {"filename"}
    if unknown:
        raise ValueError(f"unknown blocking options: {sorted(unknown)}")
    b = {**BLOCKING_DEFAULTS, **b}
    if b["auto_shift_damping"] and not b["blocking_on_the_fly"]:
        raise ValueError("auto_shift_damping needs blocking_on_the_fly")
    st = N.FqhQmcIn()
    st.tau = float(q["tau"])
    st.rng_seed = int(q.get("rng_seed", 7))
    st.init_pop = float(q.get("init_pop", 10))
    st.mc_cycles = int(q.get("mc_cycles", 1))
    st.nreports = int(q.get("nreports", 1))
    st.target_population = float(q.get("target_population", 1.0e300))
    st.shift_damping = float(q.get("shift_damping", 0.05))
    st.initial_shift = float(q.get("initial_shift", 0.0))
    # vary_shift_from defaults to initial_shift; 'proje' starts from the projected
    # energy (src/lua_hande_calc.f90:1483-1491)
    vsf = q.get("vary_shift_from", st.initial_shift)
    st.vary_shift_from_proje = int(isinstance(vsf, str) and vsf == "proje")
    st.vary_shift_from = 0.0 if st.vary_shift_from_proje else float(vsf)
    st.initiator = int(bool(q.get("initiator", False)))
    st.initiator_threshold = float(q.get("initiator_threshold", 3.0))
    st.real_amplitudes = int(bool(q.get("real_amplitudes", False)))
    st.real_amplitude_force_32 = int(bool(q.get("real_amplitude_force_32", False)))
    st.spawn_cutoff = float(q.get("spawn_cutoff", 0.01))
    st.excit_gen = _EXCIT[q.get("excit_gen", "renorm")]
    st.state_size = int(q.get("state_size", -1000))
    st.spawned_state_size = int(q.get("spawned_state_size", -500))
    st.replica_tricks = int(bool(f.get("replica_tricks", False)))
    st.quadrature_initiator = int(bool(f.get("quadrature_initiator", True)))
    st.rng_mode = _RNG[rng_mode]
    st.device = int(device)
    # learned proposal options (no reference counterpart)
    st.proposal_mix = float(q.get("proposal_mix", 0.0))
    st.proposal_min_attempts = int(q.get("proposal_min_attempts", 0))
    st.vary_shift_set = int("vary_shift" in q)
    st.vary_shift = int(bool(q.get("vary_shift", False)))
    st.blocking_on_the_fly = int(bool(b["blocking_on_the_fly"]))
    st.auto_shift_damping = int(bool(b["auto_shift_damping"]))
    st.shift_damping_precision = float(b["shift_damping_precision"])
    st.force_shift_damping_opt = int(bool(b["force_shift_damping_opt"]))
    st.start_save_frequency = int(b["start_save_frequency"])
    st.start_point_number = int(b["start_point_number"])
    st.start_point = int(b["start_point"])
    st.error_limit = float(b["error_limit"])
    st.min_blocks_used = int(b["min_blocks_used"])
    st.blocks_used = int(b["blocks_used"])
    # semi_stoch{} (src/lua_hande_calc.f90:1620-1740): space 'high', separate
    # (the default) or combined annihilation; real amplitudes are switched on
    # as the reference does; shift_start_iteration replaces start_iteration
    unknown = set(ss) - {"space", "size", "start_iteration", "shift_start_iteration", "separate_annihilation"}
    if unknown:
        raise ValueError(f"unsupported semi_stoch options: {sorted(unknown)}")
    if ss:
        if ss.get("space") != "high":
            raise ValueError("semi_stoch: only space = 'high' is implemented")
        if "size" not in ss:
            raise ValueError("semi_stoch: target space size not given")
        st.real_amplitudes = 1
        st.semi_stoch_size = int(ss["size"])
        st.semi_stoch_separate = 1 if ss.get("separate_annihilation", True) else 0
        if "shift_start_iteration" in ss:
            st.semi_stoch_start_iter = 2147483647
            st.semi_stoch_shift_iter = int(ss["shift_start_iteration"])
        else:
            st.semi_stoch_start_iter = int(ss.get("start_iteration", 1))
            st.semi_stoch_shift_iter = -1
    else:
        st.semi_stoch_size, st.semi_stoch_start_iter, st.semi_stoch_shift_iter = 0, 1, -1
        st.semi_stoch_separate = 1
    return st


REPORT_COLUMNS = ("iterations", "Shift", "\\sum H_0j N_j", "N_0", "# H psips", "# states",
                  "# spawn_events", "R_spawn")


def rows_to_table(rows, nrows, nspaces=1):
    """Report rows as a float array with the reference's column order."""
    out = []
    for r in rows[:nrows]:
        line = [float(r.iterations)]
        for s in range(nspaces):
            line += [r.shift[s], r.proj_energy[s], r.D0_population[s], r.nparticles[s]]
        line += [float(r.nstates), float(r.nspawn_events), r.rspawn]
        out.append(line)
    return np.array(out)


def pop_bit_shift(st):
    """Bits of fixed-point fraction in the population encoding (real_bit_shift,
    src/qmc.F90:1462-1479): 31 for real_64 amplitudes, 11 for real_32, 0 for integers."""
    return (11 if st.real_amplitude_force_32 else 31) if st.real_amplitudes else 0


def fciqmc(sys, qmc, fciqmc=None, rng_mode="replay", device=0, lib=None, nprocs=1, restart=None,
           return_state=False, blocking=None, max_reports=None, semi_stoch=None, **_):
    """fciqmc{...}: run FCIQMC on the GPU engine and return the report table.

    nprocs > 1 (replay mode) replays a run of that many MPI ranks: one engine
    per emulated rank on the device, spawned walkers exchanged on the host.
    restart (a ``restart.RestartState`` or the path of a HANDE restart file)
    continues a stored run, as ``restart = { read = ... }`` does; the file's
    shift damping applies unless qmc sets one (src/qmc.F90:186-192).
    return_state=True returns (table, RestartState after the last report), the
    state ``restart = { write = ... }`` stores (``restart.write_restart``).
    max_reports stops the run after that many reports (a prefix of the run
    nreports describes).  semi_stoch is the semi_stoch{} table (space
    "high"; one rank)."""
    from . import restart as R
    lib = lib or sys._lib
    st = qmc_in(qmc, fciqmc, rng_mode=rng_mode, device=device, blocking=blocking, semi_stoch=semi_stoch)
    S = 2 if st.replica_tricks else 1
    pbs = pop_bit_shift(st)
    rs = None
    if restart is not None:
        rs = restart if isinstance(restart, R.RestartState) else R.read_restart(
            restart, sys.view.nbasis, S, pbs)
        if rs.pop_bit_shift != pbs:
            raise R.RestartError("restart state encoded with another population scale")
        if "shift_damping" not in qmc and rs.shift_damping is not None:
            st.shift_damping = rs.shift_damping
    max_rows = min(st.nreports, max_reports if max_reports is not None else st.nreports) + 1
    rows = (N.FqhReportRow * max_rows)()
    nrows = C.c_int()
    end = None
    if nprocs > 1:
        if rs is not None or return_state:
            raise ValueError("restarts are written and read on one rank")
        rc = lib.fqh_fciqmc_ranks(sys.handle, C.byref(st), int(nprocs), rows, max_rows, C.byref(nrows))
    elif rs is None and not return_state:
        rc = lib.fqh_fciqmc(sys.handle, C.byref(st), rows, max_rows, C.byref(nrows))
    else:
        start = R.RestartBuffers.from_state(rs) if rs is not None else None
        cap = N.capacity_from(st.state_size, 8 * sys.view.nwords + 16 * S)
        end = R.RestartBuffers(sys.view.nwords, S, cap) if return_state else None
        rc = lib.fqh_fciqmc_restart(sys.handle, C.byref(st), C.byref(start.st) if start else None,
                                    C.byref(end.st) if end else None, rows, max_rows, C.byref(nrows))
    table = rows_to_table(rows, nrows.value, S)
    if rc == N.FQ_ERR_OVERFLOW:
        raise ListOverflow(table)
    N.check(rc, "fciqmc")
    if return_state:
        return table, end.to_state(S, pbs, like=rs)
    return table


class ListOverflow(RuntimeError):
    """The spawned or main walker list filled up and the report loop stopped
    early, as the reference stops on its spawn%error / psip_list%error flags
    (src/fciqmc.f90:415-422).  ``table`` holds the rows reported until then."""

    def __init__(self, table):
        super().__init__(f"walker list overflow: report loop stopped after {len(table)} rows "
                         "(raise state_size / spawned_state_size)")
        self.table = table


def restart_file(opts, base_dir=".", rank=0):
    """Path of the restart file a deck's ``restart = { read = Y }`` names
    (restart_stem.Y.pX.H5 with stem HANDE.RS, src/restart_hdf5.F90:112-127;
    read = true picks the highest existing Y), or None if it reads none."""
    import glob
    import os
    import re
    r = opts.get("restart") or {}
    rid = r.get("read")
    if rid is None or rid is False:
        return None
    if rid is True or (isinstance(rid, (int, float)) and rid < 0):
        ids = []
        for f in glob.glob(os.path.join(base_dir, f"HANDE.RS.*.p{rank}.H5")):
            m = re.search(r"HANDE\.RS\.(\d+)\.p\d+\.H5$", f)
            if m:
                ids.append(int(m.group(1)))
        if not ids:
            raise FileNotFoundError(f"no restart files in {base_dir}")
        rid = max(ids)
    return os.path.join(base_dir, f"HANDE.RS.{int(rid)}.p{rank}.H5")


def load_deck(text, base_dir=".", lib=None):
    """Parse a deck; return (system, fciqmc options) for its fciqmc{} call."""
    env, calls = parse(text)
    lib = lib or N.lib()
    call = next((c for c in calls if isinstance(c, LuaCall) and c.func == "fciqmc"), None)
    if call is None:
        raise ValueError("deck has no fciqmc{} call")
    sysdef = call.args["sys"]
    if not isinstance(sysdef, LuaCall):
        raise ValueError("fciqmc{} needs sys = hubbard_k{...} or read_in{...}")
    if sysdef.func == "hubbard_k":
        system = hubbard_k(lib=lib, **sysdef.args)
    elif sysdef.func == "read_in":
        system = read_in(base_dir=base_dir, lib=lib, **sysdef.args)
    else:
        raise ValueError(f"system {sysdef.func} is outside the FCIQMC hot path")
    return system, call.args
