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
    lib.orc_fciqmc.argtypes = [P(N.FqSystem), P(N.FqhQmcIn), P(N.FqhReportRow), C.c_int, P(C.c_int)]
    return lib


def oracle_report(orc, system, st):
    rows = (N.FqhReportRow * (st.nreports + 1))()
    nr = C.c_int()
    orc.orc_fciqmc(system.view, C.byref(st), rows, st.nreports + 1, C.byref(nr))
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
