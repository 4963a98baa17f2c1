"""HANDE restart files (HANDE.RS.<id>.p<rank>.H5): read and write.

The file layout is the reference's (src/restart_hdf5.F90:24-90): groups
``metadata``, ``qmc/psips``, ``qmc/state``, ``qmc/reference``, ``basis`` and
``rng``.  Reading follows read_restart_hdf5 (src/restart_hdf5.F90:539-927):

* determinant bit strings stored with 32-bit words (``metadata/i0_length`` 32)
  are joined into 64-bit words as bit_string_32_to_64 does
  (src/restart_utils.F90:38-89,460-478);
* populations stored as 32-bit integers are widened (convert_pops_32_to_64,
  src/restart_utils.F90:239-278) and rescaled from the file's
  ``population scale factor`` to the run's 2^pop_bit_shift with integer
  multiplication or truncating division (change_pop_scaling_64,
  src/restart_utils.F90:454-472);
* shift, vary-shift flags, cycle count, shift damping, the projected energy and
  the reference population of the last report come from ``qmc/state`` and
  ``qmc/reference``.

Writing follows dump_restart_hdf5 (src/restart_hdf5.F90:307-537) for the items
the walker-propagation path owns, as restart version 2 (no RNG state: the
replay stream restarts from rng_seed, as the reference does for a file whose
``rng`` group is empty).

Not supported (``RestartError``): restart version 1 files (their projected
energy must be re-estimated from the list), files written on several ranks,
non-blocking-communication files (``received_list``), bit strings carrying
extra information words, and a change of basis size.
"""
from __future__ import annotations

import ctypes as C
import time
import uuid as _uuid
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import hdf5

RESTART_VERSION_WRITTEN = 2
FCIQMC_CALC = 4  # calc type of an FCIQMC calculation (src/calc.F90:35)


class RestartError(ValueError):
    """The restart file cannot be used for this calculation."""


@dataclass
class RestartState:
    """Walker list and report-loop state carried between two runs."""
    states: np.ndarray          # (n, nwords) uint64, main-list order
    pops: np.ndarray            # (n, nspaces) int64, scaled by 2^pop_bit_shift
    dat: np.ndarray             # (n,) float64, K_ii
    mc_cycles_done: int
    shift: np.ndarray           # (nspaces,)
    vary_shift: np.ndarray      # (nspaces,) bool
    proj_energy: np.ndarray     # (nspaces,)
    D0_population: np.ndarray   # (nspaces,)
    pop_bit_shift: int = 0
    shift_damping: float | None = None
    shift_damping_status: int = -2             # sd_no_optimisation unless the file says otherwise
    reference: np.ndarray | None = None        # (nwords,) uint64
    hs_reference: np.ndarray | None = None
    extra: dict = field(default_factory=dict)  # other items read from the file

    @property
    def nstates(self):
        return len(self.dat)

    @property
    def nspaces(self):
        return self.pops.shape[1]

    def total_population(self):
        """sum |N_i| per space, as init_estimators recounts it."""
        rf = float(1 << self.pop_bit_shift)
        return np.array([np.sum(np.abs(self.pops[:, s].astype(np.float64) / rf)) for s in range(self.nspaces)])


def _words64(dets, i0_length, nbasis):
    """(n, nwords_file) ints of i0_length bits -> (n, ceil(nbasis/64)) uint64."""
    nw = (nbasis + 63) // 64
    d = np.asarray(dets)
    if d.ndim == 1:
        d = d[:, None]
    if i0_length == 64:
        out = d.astype(np.int64).view(np.uint64)
    elif i0_length == 32:
        u = d.astype(np.int32).view(np.uint32).astype(np.uint64)
        if u.shape[1] % 2:
            u = np.concatenate([u, np.zeros((u.shape[0], 1), np.uint64)], axis=1)
        out = u[:, 0::2] | (u[:, 1::2] << np.uint64(32))
    else:
        raise RestartError(f"i0_length {i0_length} is not supported")
    if out.shape[1] < nw:
        raise RestartError("bit strings are shorter than the basis")
    return np.ascontiguousarray(out[:, :nw])


def _rescale(pops, old, new):
    """change_pop_scaling_64: multiply or divide (truncating toward zero)."""
    pops = pops.astype(np.int64)
    if old < new:
        return pops * np.int64(new // old)
    if old > new:
        f = np.int64(old // new)
        return np.sign(pops) * (np.abs(pops) // f)
    return pops


def read_restart(path, nbasis, nspaces=1, pop_bit_shift=0):
    """Read a restart file for a run with nbasis spin orbitals, nspaces
    population spaces and populations encoded with 2^pop_bit_shift."""
    t = hdf5.read(path)
    meta = t.get("metadata", {})
    version = int(meta.get("restart version", 1))
    if version < 2:
        raise RestartError("restart version 1 files are not supported (no stored projected energy)")
    if version > 3:
        raise RestartError(f"restart version {version} is newer than this reader")
    if int(meta.get("nprocs", 1)) != 1:
        raise RestartError("restart files written on several ranks are not supported")
    i0_length = int(meta.get("i0_length", 64))
    basis = t.get("basis", {})
    nbasis_file = int(basis.get("nbasis", nbasis))
    if nbasis_file != nbasis:
        raise RestartError(f"restart file has {nbasis_file} basis functions, the system {nbasis}")
    if int(basis.get("info string len", 0)) != 0:
        raise RestartError("bit strings with extra information words are not supported")
    psips = t["qmc"]["psips"]
    if "received_list" in psips:
        raise RestartError("non-blocking-communication restart files are not supported")
    states = _words64(psips["determinants"], i0_length, nbasis)
    pops = np.asarray(psips["populations"])
    if pops.ndim == 1:
        pops = pops[:, None]
    if pops.shape[1] != nspaces:
        raise RestartError(f"restart file has {pops.shape[1]} population spaces, the run {nspaces}")
    scale = int(np.asarray(psips.get("population scale factor", [1 << pop_bit_shift])).ravel()[0])
    pops = _rescale(pops, scale, 1 << pop_bit_shift)
    dat = np.asarray(psips["data"], dtype=np.float64)
    dat = dat[:, 0] if dat.ndim == 2 else dat
    state = t["qmc"]["state"]
    ref = t["qmc"].get("reference", {})

    def spaces(x, dtype=np.float64):
        a = np.asarray(x, dtype=dtype).ravel()
        if len(a) < nspaces:
            raise RestartError("restart file holds fewer spaces than the run")
        return a[:nspaces].copy()

    sd = state.get("shift_damping")
    rs = RestartState(
        states=states,
        pops=np.ascontiguousarray(pops),
        dat=np.ascontiguousarray(dat),
        mc_cycles_done=int(state["ncycles"]),
        shift=spaces(state["shift"]),
        vary_shift=spaces(state.get("vary shift", np.zeros(nspaces)), np.int64).astype(bool),
        proj_energy=spaces(state["projected energy"]),
        D0_population=spaces(ref["reference population @ t-1"]),
        pop_bit_shift=pop_bit_shift,
        shift_damping=None if sd is None else float(np.asarray(sd).ravel()[0]),
        reference=(_words64(np.asarray(ref["reference determinant"])[None, :], i0_length, nbasis)[0]
                   if "reference determinant" in ref else None),
        hs_reference=(_words64(np.asarray(ref["Hilbert space reference determinant"])[None, :], i0_length, nbasis)[0]
                      if "Hilbert space reference determinant" in ref else None),
    )
    for k in ("hash_seed", "move_freq"):
        if k in state:
            rs.extra[k] = int(state[k])
    if "shift_damping_status" in state:
        rs.shift_damping_status = int(state["shift_damping_status"])
    for k in ("uuid", "hande version", "date"):
        if k in meta:
            rs.extra[k] = meta[k]
    return rs


def restart_tree(rs: RestartState, nbasis, shift_damping=0.05, calc_uuid=None):
    """The group/dataset tree dump_restart_hdf5 writes for this state."""
    S = rs.nspaces
    rf = 1 << rs.pop_bit_shift
    ref = rs.reference if rs.reference is not None else rs.states[0]
    hs = rs.hs_reference if rs.hs_reference is not None else ref
    zeros = np.zeros(S)
    return {
        "metadata": {
            "restart version": np.int32(RESTART_VERSION_WRITTEN),
            "hande version": "fqmc-mi355x",
            "date": time.strftime("%H:%M:%S %d/%m/%Y"),
            "uuid": calc_uuid or str(_uuid.uuid4()),
            "calc type": np.int32(FCIQMC_CALC),
            "nprocs": np.int32(1),
            "i0_length": np.int32(64),
        },
        "qmc": {
            "psips": {
                "determinants": rs.states.astype(np.uint64).view(np.int64),
                "populations": rs.pops.astype(np.int64),
                "data": rs.dat.astype(np.float64)[:, None],
                "total population": rs.total_population(),
                "processor map": np.zeros(1, np.int32),
                "population scale factor": np.array([rf], np.int64),
            },
            "state": {
                "ncycles": np.int32(rs.mc_cycles_done),
                "shift": rs.shift.astype(np.float64),
                "projected energy": rs.proj_energy.astype(np.float64),
                "projected energy real": rs.proj_energy.astype(np.float64),
                "projected energy imag": zeros,
                "hash_seed": np.int32(rs.extra.get("hash_seed", 7)),
                "move_freq": np.int32(rs.extra.get("move_freq", 0)),
                "vary shift": rs.vary_shift.astype(np.int32),
                "shift_damping": np.array([shift_damping if rs.shift_damping is None else rs.shift_damping]),
                "shift_damping_status": np.int32(rs.shift_damping_status),
            },
            "reference": {
                "reference determinant": np.asarray(ref, np.uint64).view(np.int64),
                "Hilbert space reference determinant": np.asarray(hs, np.uint64).view(np.int64),
                "reference population @ t-1": rs.D0_population.astype(np.float64),
                "reference population @ t-1 real": rs.D0_population.astype(np.float64),
                "reference population @ t-1 imag": zeros,
            },
        },
        "basis": {"nbasis": np.int32(nbasis), "info string len": np.int32(0)},
        "rng": {},
    }


def write_restart(path, rs: RestartState, nbasis, shift_damping=0.05, calc_uuid=None):
    """Write rs as a restart file; returns the file's bytes."""
    return hdf5.write(path, restart_tree(rs, nbasis, shift_damping, calc_uuid))


# ------------------------------------------------------- the C-ABI structures


class RestartBuffers:
    """An fqh_restart struct whose arrays live as long as this object."""

    def __init__(self, nwords, nspaces, capacity):
        self.states = np.zeros((capacity, nwords), np.uint64)
        self.pops = np.zeros((capacity, nspaces), np.int64)
        self.dat = np.zeros(capacity, np.float64)
        self.st = N.FqhRestart()
        self.st.capacity = capacity
        self.st.states = self.states.ctypes.data_as(C.POINTER(C.c_uint64))
        self.st.pops = self.pops.ctypes.data_as(C.POINTER(C.c_int64))
        self.st.dat = self.dat.ctypes.data_as(C.POINTER(C.c_double))

    @classmethod
    def from_state(cls, rs: RestartState):
        b = cls(rs.states.shape[1], rs.nspaces, max(rs.nstates, 1))
        n = rs.nstates
        b.states[:n] = rs.states
        b.pops[:n] = rs.pops
        b.dat[:n] = rs.dat
        st = b.st
        st.nstates = n
        st.mc_cycles_done = int(rs.mc_cycles_done)
        st.shift_damping = float(rs.shift_damping) if rs.shift_damping is not None else 0.0
        st.shift_damping_status = int(rs.shift_damping_status)
        for s in range(rs.nspaces):
            st.shift[s] = float(rs.shift[s])
            st.vary_shift[s] = int(bool(rs.vary_shift[s]))
            st.proj_energy[s] = float(rs.proj_energy[s])
            st.D0_population[s] = float(rs.D0_population[s])
        return b

    def to_state(self, nspaces, pop_bit_shift, like: RestartState | None = None):
        st = self.st
        n = int(st.nstates)
        if n > self.states.shape[0]:
            raise RestartError("end state larger than its buffers")
        rs = RestartState(
            states=self.states[:n].copy(), pops=self.pops[:n].copy(), dat=self.dat[:n].copy(),
            mc_cycles_done=int(st.mc_cycles_done),
            shift=np.array(st.shift[:nspaces]), vary_shift=np.array(st.vary_shift[:nspaces]).astype(bool),
            proj_energy=np.array(st.proj_energy[:nspaces]), D0_population=np.array(st.D0_population[:nspaces]),
            pop_bit_shift=pop_bit_shift, shift_damping=float(st.shift_damping),
            shift_damping_status=int(st.shift_damping_status))
        if like is not None:
            rs.reference, rs.hs_reference, rs.extra = like.reference, like.hs_reference, dict(like.extra)
        return rs
