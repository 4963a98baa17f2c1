/*
 * fqmc_host.h -- host side above the engine C-ABI: system set-up (the
 * read-only tables fq_create consumes) and the FCIQMC report-loop driver.
 *
 * These functions stand in for the reference's Lua-facing set-up
 * (hubbard_k{} -> init_system/init_model_basis_fns/init_momentum_symmetry,
 * src/system.f90:418, src/basis.f90:258, src/momentum_symmetry.f90:17;
 * read_in{} -> read_in_integrals, src/read_in.F90:12) and for do_fciqmc's
 * report loop (src/fciqmc.f90:276-464) with end_report_loop /
 * update_energy_estimators / update_shift (src/qmc_common.F90:1082,
 * src/energy_evaluation.F90:126-711).
 */
#ifndef FQMC_HOST_H
#define FQMC_HOST_H

#include <stdint.h>
#include "fqmc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fqh_system fqh_system;

/* Symmetry selection for the reference determinant (sys%symmetry,
 * sys%tot_sym, sys%aufbau_sym in src/lua_hande_systems.f90:161-173). */
enum fqh_sym_mode { FQH_SYM_AUFBAU = 0, FQH_SYM_TOT = 1, FQH_SYM_INDEX = 2 };

/* hubbard_k{ electrons, lattice, ms, sym, U, t }.
 * lattice: ndim*ndim ints, vector j is lattice[j*ndim .. j*ndim+ndim-1]. */
int fqh_hubbard_k(int nel, int ms, int ndim, const int *lattice, double U, double t, int sym_mode,
                  int sym, fqh_system **out);

/* read_in{ int_file, nel, ms, sym, CAS } for an RHF integral dump.
 * sym < 0 selects the Aufbau symmetry; cas_nel/cas_norb <= 0 disable CAS. */
int fqh_read_in(const char *path, int nel, int ms, int sym, int cas_nel, int cas_norb, fqh_system **out);

/* Tables in engine layout.  Valid while the fqh_system lives. */
const fq_system *fqh_view(const fqh_system *s);
/* Reference determinant occupation list (1-based orbitals, nel entries) and
 * basis-function listing for checks against the reference's printout. */
int fqh_ref_occ(const fqh_system *s, int *occ);
int fqh_basis_info(const fqh_system *s, int *ms, int *sym_or_k, double *sp_eigv);
void fqh_free(fqh_system *s);

/* Slater--Condon rules on the host (for checks): <D|H|D> of a determinant. */
double fqh_diagonal(const fqh_system *s, const uint64_t *f);

/* fciqmc{ qmc = {...} } options used on the hot path (src/qmc_data.f90:121-303). */
typedef struct fqh_qmc_in {
  double tau;
  int64_t rng_seed;
  double init_pop;
  int32_t mc_cycles;
  int32_t nreports;
  double target_population;
  double shift_damping;
  double initial_shift;
  double vary_shift_from;
  int32_t initiator;
  double initiator_threshold;
  int32_t real_amplitudes;
  int32_t real_amplitude_force_32;
  double spawn_cutoff;
  int32_t excit_gen;            /* enum fq_excit_gen */
  int64_t state_size;           /* main list slots (negative: MB, as the reference) */
  int64_t spawned_state_size;   /* spawned list slots (negative: MB) */
  int32_t replica_tricks;
  int32_t quadrature_initiator;
  int32_t rng_mode;             /* enum fq_rng_mode */
  int32_t device;
  int32_t vary_shift_from_proje; /* vary_shift_from = 'proje': the shift starts from the report's projected
                                    energy (src/energy_evaluation.F90:638-648) */
  double proposal_mix;           /* fq_params.proposal_mix (learned proposal; 0: default) */
  int32_t proposal_min_attempts; /* fq_params.proposal_min_attempts */
  int32_t vary_shift_set;        /* qmc{ vary_shift = ... } given: it overrides the restart's flag and, if
                                    false, resets the shift to initial_shift (src/qmc.F90:882-891) */
  int32_t vary_shift;
  /* blocking{} (src/qmc_data.f90:425-460; defaults in fqh_qmc_in_defaults) */
  int32_t blocking_on_the_fly;
  int32_t auto_shift_damping;
  double shift_damping_precision;
  int32_t force_shift_damping_opt;
  int32_t start_save_frequency;
  int32_t start_point_number;
  int32_t start_point;
  double error_limit;
  int32_t min_blocks_used;
  int32_t blocks_used;
  /* semi_stoch{} (src/qmc_data.f90:304-330), space "high": the
   * semi_stoch_size most populated determinants become the deterministic core
   * at iteration semi_stoch_start_iter (or semi_stoch_shift_iter iterations
   * after the shift starts varying, if not -1); size 0: off.
   * semi_stoch_separate: 1 = separate annihilation (the reference's default:
   * the projection is added to the core populations directly), 0 = combined
   * (its result joins the spawned list). */
  int32_t semi_stoch_size;
  int32_t semi_stoch_start_iter;
  int32_t semi_stoch_shift_iter;
  int32_t semi_stoch_separate;
} fqh_qmc_in;

/* Fill the blocking{} fields with the reference's defaults (on-the-fly
 * blocking and auto shift damping off, start points chosen automatically,
 * shift_damping_precision 2, min_blocks_used 10, blocks_used huge). */
void fqh_qmc_in_defaults(fqh_qmc_in *in);

/* One row of the QMC report table (src/qmc_io.f90:412-508). */
typedef struct fqh_report_row {
  int64_t iterations;
  double shift[FQ_MAX_SPACES];
  double proj_energy[FQ_MAX_SPACES];
  double D0_population[FQ_MAX_SPACES];
  double nparticles[FQ_MAX_SPACES];
  int64_t nstates;
  int64_t nspawn_events;
  double rspawn;
  double time_per_cycle;
  double shift_damping;         /* damping in force after this report (auto_shift_damping may change it) */
  double blocked_proj_energy;   /* on-the-fly blocking: optimal-block E_proj - E_HF so far (0 if none) */
  double blocked_proj_error;    /* and its standard error */
} fqh_report_row;

/* Run do_fciqmc on the engine: rows[0] is the iteration-0 status line.
 * Fills at most max_rows rows and sets *nrows: the run stops once the table is
 * full, so a short table runs a prefix of the calculation nreports describes
 * (everything that depends on nreports, such as the blocking start points,
 * is as in the full run).  On error returns the engine's status code. */
int fqh_fciqmc(const fqh_system *s, const fqh_qmc_in *in, fqh_report_row *rows, int max_rows, int *nrows);

/* do_fciqmc on nprocs emulated MPI ranks (replay mode): one engine per rank on
 * the same device, rank r drawing from dSFMT seeded rng_seed + r and owning the
 * determinants that hash to it; each cycle's spawned entries are exchanged on
 * the host as comm_spawn_t does (src/spawn_data.F90:624-739) and the report
 * columns are the ranks' sums, as after the reference's MPI_Allreduce
 * (src/energy_evaluation.F90:126-201).  nprocs = 1 is fqh_fciqmc; a
 * semi-stochastic run (semi_stoch_size > 0) needs nprocs = 1. */
int fqh_fciqmc_ranks(const fqh_system *s, const fqh_qmc_in *in, int nprocs, fqh_report_row *rows, int max_rows,
                     int *nrows);

/* The state a restart file carries between two runs (src/restart_hdf5.F90:30-90:
 * qmc/psips determinants, populations and data; qmc/state ncycles, shift and
 * vary shift; qmc/reference "reference population @ t-1"; qmc/state
 * "projected energy").  Populations are in the run's encoding (times
 * 2^pop_bit_shift for real amplitudes); dat is K_ii = <D|H|D> - H_00 per state. */
typedef struct fqh_restart {
  int64_t nstates;      /* walkers in states / pops / dat */
  int64_t capacity;     /* room in states / pops / dat (written by fqh_fciqmc_restart's end state) */
  uint64_t *states;     /* nwords x nstates, determinant-major, sorted as the main list */
  int64_t *pops;        /* nspaces x nstates, determinant-major */
  double *dat;          /* nstates */
  int64_t mc_cycles_done;
  double shift[FQ_MAX_SPACES];
  int32_t vary_shift[FQ_MAX_SPACES];
  double proj_energy[FQ_MAX_SPACES];    /* sum_j H_0j N_j of the last report */
  double D0_population[FQ_MAX_SPACES];  /* N_0 of the last report */
  double shift_damping;                 /* qmc/state shift_damping (<= 0: none stored) */
  int32_t shift_damping_status;         /* qmc/state shift_damping_status (sd_no_optimisation = -2, ...) */
} fqh_restart;

/* do_fciqmc restarted from `start` (read_restart_hdf5 and the restart branches
 * of init_qmc / init_estimators / do_fciqmc, src/restart_hdf5.F90:539-927,
 * src/qmc.F90:105-172,827-908, src/fciqmc.f90:147-262 with a restart file of
 * version >= 2, so the first row's projected energy and N_0 are the stored
 * ones): the walker list, shift, vary-shift flags and iteration count continue
 * from `start`, the replay stream restarts from rng_seed (a file without an
 * rng group), and the total population is recounted from the populations.
 * start = NULL is fqh_fciqmc.  end (may be NULL) receives the state after the
 * last report (dump_restart_hdf5, src/restart_hdf5.F90:307-537): FQ_ERR_CAPACITY
 * with end->nstates set if end->capacity is too small. */
int fqh_fciqmc_restart(const fqh_system *s, const fqh_qmc_in *in, const fqh_restart *start, fqh_restart *end,
                       fqh_report_row *rows, int max_rows, int *nrows);

/* MurmurHash2 of len bytes (lib/external/MurmurHash2.c:16) and the rank owning
 * a determinant among nprocs (assign_particle_processor with hash seed 7 and
 * no hash shift, src/spawning.F90:770-838; -1 for bad arguments). */
uint32_t fqh_murmur2(const void *key, int len, uint32_t seed);
int fqh_owner_rank(const uint64_t *f, int nbasis, int nprocs);

/* The replay stream: n close-open dSFMT-19937 draws for rng_seed, in the
 * order get_rand_close_open returns them
 * (lib/dSFMT_F03_interface/dSFMT_interface.F90:192-250,387-409). */
int fqh_dsfmt_fill(int64_t seed, double *out, int64_t n);

#ifdef __cplusplus
}
#endif
#endif /* FQMC_HOST_H */
