"""ctypes view of the engine and host C-ABI (include/fqmc.h, include/fqmc_host.h).

The shared library ``hande_amd/libfqmc.so`` is built in-tree by
``hande_amd/build.py`` (hipcc, gfx950).  Loading fails loudly when it is
missing: there is no fallback path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfqmc.so")

FQ_OK = 0
FQ_ERR_OVERFLOW = -8
FQ_ERR_SCRATCH = -9
FQ_ERR = {
    -1: "FQ_ERR_ARG", -2: "FQ_ERR_NOMEM", -3: "FQ_ERR_DEVICE", -4: "FQ_ERR_RNG_EXHAUSTED",
    -5: "FQ_ERR_CAPACITY", -6: "FQ_ERR_NO_DEVICE", -7: "FQ_ERR_BIASED", -8: "FQ_ERR_OVERFLOW",
    -9: "FQ_ERR_SCRATCH",
}
FQ_SYS_HUB_K, FQ_SYS_READ_IN = 1, 2
FQ_EXCIT_RENORM, FQ_EXCIT_NO_RENORM, FQ_EXCIT_HEAT_BATH, FQ_EXCIT_LEARNED = 0, 1, 2, 3
FQ_RNG_PHILOX, FQ_RNG_REPLAY = 0, 1
FQH_SYM_AUFBAU, FQH_SYM_TOT, FQH_SYM_INDEX = 0, 1, 2
FQ_MAX_SPACES = 2
FQ_MAX_WORDS = 2

i32, i64, u64, f64 = C.c_int32, C.c_int64, C.c_uint64, C.c_double
P = C.POINTER


class FqSystem(C.Structure):
    _fields_ = [
        ("system", i32), ("nbasis", i32), ("nel", i32), ("nalpha", i32), ("nbeta", i32),
        ("nvirt", i32), ("nvirt_alpha", i32), ("nvirt_beta", i32), ("nwords", i32),
        ("f0", P(u64)), ("H00", f64), ("sp_eigv", P(f64)),
        ("nsym", i32), ("sym_table", P(i32)), ("inv_sym", P(i32)), ("coulomb_k", f64),
        ("norb", i32), ("nsym_pg", i32), ("orb_sym", P(i32)), ("nbasis_sym_spin", P(i32)),
        ("max_sym_spin", i32), ("sym_spin_basis_fns", P(i32)), ("ecore", f64),
        ("h1", P(f64)), ("eri", P(f64)), ("n_eri", i64),
        ("pattempt_single", f64), ("pattempt_double", f64),
    ]


class FqParams(C.Structure):
    _fields_ = [
        ("nspaces", i32), ("initiator", i32), ("initiator_pop", f64), ("quadrature_initiator", i32),
        ("real_amplitudes", i32), ("pop_bit_shift", i32), ("spawn_cutoff", i64), ("excit_gen", i32),
        ("rng_mode", i32), ("philox_seed", u64), ("main_capacity", i64), ("spawn_capacity", i64),
        ("device", i32), ("proposal_mix", f64), ("proposal_min_attempts", i32),
    ]


class FqCycleOut(C.Structure):
    _fields_ = [
        ("proj_energy", f64 * FQ_MAX_SPACES), ("D0_population", f64 * FQ_MAX_SPACES),
        ("nparticles", f64 * FQ_MAX_SPACES), ("nstates", i64), ("nspawn_events", i64),
        ("ndeath", i64), ("nattempts", i64), ("rng_consumed", i64),
        ("spawn_overflow", i32), ("main_overflow", i32),
    ]


class FqScheduleOut(C.Structure):
    _fields_ = [
        ("graph", i32), ("tuned", i32), ("digit_bits", i32), ("spawn_blocks", i32), ("ncand", i32),
        ("chosen", i32), ("octave", i32), ("tuned_octaves", i32), ("graph_cycles", i64), ("graphs", i64), ("eager_cycles", i64),
        ("cand_ms", f64 * 8), ("cand_digit_bits", i32 * 8), ("cand_spawn_blocks", i32 * 8),
    ]


class FqProfileOut(C.Structure):
    _fields_ = [
        ("det_pass_ms", f64), ("spawn_ms", f64), ("sort_ms", f64), ("reduce_ms", f64), ("merge_ms", f64),
        ("total_ms", f64), ("nstates_in", i64), ("nattempts", i64), ("nspawned", i64), ("nruns", i64),
        ("nstates_out", i64), ("radix_passes", i32), ("spawn_launches", i32), ("spawn_kernel_ms", f64),
        ("proposal_ms", f64), ("proposal_rows", i64),
    ]


class FqhQmcIn(C.Structure):
    _fields_ = [
        ("tau", f64), ("rng_seed", i64), ("init_pop", f64), ("mc_cycles", i32), ("nreports", i32),
        ("target_population", f64), ("shift_damping", f64), ("initial_shift", f64),
        ("vary_shift_from", f64), ("initiator", i32), ("initiator_threshold", f64),
        ("real_amplitudes", i32), ("real_amplitude_force_32", i32), ("spawn_cutoff", f64),
        ("excit_gen", i32), ("state_size", i64), ("spawned_state_size", i64),
        ("replica_tricks", i32), ("quadrature_initiator", i32), ("rng_mode", i32), ("device", i32),
        ("vary_shift_from_proje", i32), ("proposal_mix", f64), ("proposal_min_attempts", i32),
        ("vary_shift_set", i32), ("vary_shift", i32),
        ("blocking_on_the_fly", i32), ("auto_shift_damping", i32), ("shift_damping_precision", f64),
        ("force_shift_damping_opt", i32), ("start_save_frequency", i32), ("start_point_number", i32),
        ("start_point", i32), ("error_limit", f64), ("min_blocks_used", i32), ("blocks_used", i32),
        ("semi_stoch_size", i32), ("semi_stoch_start_iter", i32), ("semi_stoch_shift_iter", i32),
        ("semi_stoch_separate", i32),
    ]


class FqhReportRow(C.Structure):
    _fields_ = [
        ("iterations", i64), ("shift", f64 * FQ_MAX_SPACES), ("proj_energy", f64 * FQ_MAX_SPACES),
        ("D0_population", f64 * FQ_MAX_SPACES), ("nparticles", f64 * FQ_MAX_SPACES),
        ("nstates", i64), ("nspawn_events", i64), ("rspawn", f64), ("time_per_cycle", f64),
        ("shift_damping", f64), ("blocked_proj_energy", f64), ("blocked_proj_error", f64),
    ]


class FqhRestart(C.Structure):
    _fields_ = [
        ("nstates", i64), ("capacity", i64), ("states", P(C.c_uint64)), ("pops", P(i64)), ("dat", P(f64)),
        ("mc_cycles_done", i64), ("shift", f64 * FQ_MAX_SPACES), ("vary_shift", C.c_int32 * FQ_MAX_SPACES),
        ("proj_energy", f64 * FQ_MAX_SPACES), ("D0_population", f64 * FQ_MAX_SPACES),
        ("shift_damping", f64), ("shift_damping_status", i32),
    ]


def _declare(lib):
    vp = C.c_void_p
    sig = {
        "fqh_hubbard_k": (C.c_int, [C.c_int, C.c_int, C.c_int, P(C.c_int), f64, f64, C.c_int, C.c_int, P(vp)]),
        "fqh_read_in": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P(vp)]),
        "fqh_view": (P(FqSystem), [vp]),
        "fqh_ref_occ": (C.c_int, [vp, P(C.c_int)]),
        "fqh_basis_info": (C.c_int, [vp, P(C.c_int), P(C.c_int), P(f64)]),
        "fqh_diagonal": (f64, [vp, P(u64)]),
        "fqh_free": (None, [vp]),
        "fqh_fciqmc": (C.c_int, [vp, P(FqhQmcIn), P(FqhReportRow), C.c_int, P(C.c_int)]),
        "fqh_dsfmt_fill": (C.c_int, [i64, P(f64), i64]),
        "fq_create": (C.c_int, [P(FqSystem), P(FqParams), P(vp)]),
        "fq_load": (C.c_int, [vp, P(u64), P(i64), P(f64), i64]),
        "fq_push_rng": (C.c_int, [vp, P(f64), i64]),
        "fq_rng_available": (C.c_int, [vp, P(i64)]),
        "fq_cycle": (C.c_int, [vp, P(f64), P(f64), f64, P(FqCycleOut)]),
        "fq_cycle_noalloc": (C.c_int, [vp, P(f64), P(f64), f64, P(FqCycleOut)]),
        "fq_semi_stoch_init": (C.c_int, [vp, i64, C.c_int]),
        "fq_semi_stoch_info": (C.c_int, [vp, P(i64), P(i64), P(u64), i64]),
        "fq_cycle_spawn": (C.c_int, [vp, P(f64), P(f64), f64, P(FqCycleOut)]),
        "fq_spawned_list": (C.c_int, [vp, P(u64), P(i64), P(i64), i64, P(i64)]),
        "fq_cycle_finish": (C.c_int, [vp, P(u64), P(i64), P(i64), i64, P(FqCycleOut)]),
        "fqh_fciqmc_ranks": (C.c_int, [vp, P(FqhQmcIn), C.c_int, P(FqhReportRow), C.c_int, P(C.c_int)]),
        "fqh_qmc_in_defaults": (None, [P(FqhQmcIn)]),
        "fqh_fciqmc_restart": (C.c_int, [vp, P(FqhQmcIn), P(FqhRestart), P(FqhRestart), P(FqhReportRow), C.c_int,
                                         P(C.c_int)]),
        "fqh_murmur2": (C.c_uint32, [C.c_char_p, C.c_int, C.c_uint32]),
        "fqh_owner_rank": (C.c_int, [P(u64), C.c_int, C.c_int]),
        "fq_annihilate": (C.c_int, [vp, P(u64), P(i64), P(i64), i64, P(FqCycleOut)]),
        "fq_nstates": (C.c_int, [vp, P(i64)]),
        "fq_download": (C.c_int, [vp, P(u64), P(i64), P(f64), i64, P(i64)]),
        "fq_last_timing": (C.c_int, [vp, P(f64), P(f64), P(f64)]),
        "fq_last_profile": (C.c_int, [vp, P(FqProfileOut)]),
        "fq_create_batch": (C.c_int, [P(FqSystem), i32, P(FqParams), P(vp)]),
        "fq_load_batch": (C.c_int, [vp, P(i32), P(u64), P(i64), P(f64), i64]),
        "fq_cycle_batch": (C.c_int, [vp, P(f64), P(f64), P(f64), P(FqCycleOut)]),
        "fq_batch_graph": (C.c_int, [vp, i32]),
        "fq_batch_schedule": (C.c_int, [vp, P(FqScheduleOut)]),
        "fq_mlp_nparams": (C.c_int, [vp, P(i64)]),
        "fq_mlp_set_params": (C.c_int, [vp, P(C.c_float), i64]),
        "fq_mlp_get_params": (C.c_int, [vp, P(C.c_float), i64]),
        "fq_mlp_forward": (C.c_int, [vp, P(u64), i64, P(C.c_float)]),
        "fq_mlp_affinity": (C.c_int, [vp, P(C.c_float), i64]),
        "fq_mlp_loss_grad": (C.c_int, [vp, P(u64), P(i32), P(f64), P(i32), i64, P(f64), P(C.c_float)]),
        "fq_mlp_variance_grad": (C.c_int, [vp, P(u64), P(i32), P(f64), P(f64), P(i32), i64, P(f64), P(C.c_float)]),
        "fq_mlp_online": (C.c_int, [vp, i32, i32, i64, f64, i32]),
        "fq_mlp_train": (C.c_int, [vp, i64, i32, f64, u64, P(f64)]),
        "fq_probe_excitations": (C.c_int, [vp, P(u64), i64, u64, P(i32), P(f64), P(f64)]),
        "fq_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        if hasattr(lib, name):
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
    return lib


_LIB = None


def lib(path=None):
    """Load the engine library (built in-tree); raise if it is missing."""
    global _LIB
    if path is not None:
        return _declare(C.CDLL(path))
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                "(hipcc, gfx950). There is no CPU fallback.")
        _LIB = _declare(C.CDLL(LIB_PATH))
    return _LIB


def capacity_from(size, element_bytes, spawned=False):
    """Slots from a state_size-style option: >= 0 is a count, < 0 is MB per
    core (src/particle_t_utils.f90:84-101), as the report-loop driver reads it."""
    if size >= 0:
        return int(size)
    mb = -float(size) * 1.0e6
    return int(mb / (2.0 * element_bytes) if spawned else mb / element_bytes)


def check(rc, what):
    if rc != FQ_OK:
        raise RuntimeError(f"{what} failed: {FQ_ERR.get(rc, rc)}")
