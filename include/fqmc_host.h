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
} fqh_qmc_in;

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
} fqh_report_row;

/* Run do_fciqmc on the engine: rows[0] is the iteration-0 status line.
 * Fills at most max_rows rows and sets *nrows.  On error returns the engine's
 * status code. */
int fqh_fciqmc(const fqh_system *s, const fqh_qmc_in *in, fqh_report_row *rows, int max_rows, int *nrows);

/* The replay stream: n close-open dSFMT-19937 draws for rng_seed, in the
 * order get_rand_close_open returns them
 * (lib/dSFMT_F03_interface/dSFMT_interface.F90:192-250,387-409). */
int fqh_dsfmt_fill(int64_t seed, double *out, int64_t n);

#ifdef __cplusplus
}
#endif
#endif /* FQMC_HOST_H */
