/*
 * fqmc.h -- C-ABI of the MI355X FCIQMC walker-propagation engine.
 *
 * The engine replaces the body of the reference's Monte Carlo cycle
 * (src/fciqmc.f90:307-399): the loop over occupied determinants (decode,
 * projected-energy accumulation, initiator flag, decide_nattempts, spawning,
 * stochastic death) followed by direct_annihilation (sort, self-annihilation,
 * merge with the main list, rounding of small real populations, insertion of
 * new determinants).  The host keeps the calc/system setup, the report loop
 * and the shift update (src/energy_evaluation.F90:126-711) and talks to the
 * engine only through the functions below: plain C types, no exceptions, an
 * int status code on every call.
 *
 * Determinant layout at the boundary follows the reference's particle_t
 * (src/qmc_data.f90:615-682): states(nwords, n) column-major 64-bit words
 * (DET_SIZE=64), pops(nspaces, n) int64 fixed point (value = pop / 2^shift,
 * src/particle_t_utils.f90:55-82), dat(1, n) = <D|H|D> - H00.  Orbital i
 * (1-based) lives in word (i-1)/64, bit (i-1)%64 (src/basis_types.f90:152-169);
 * odd orbitals are alpha, even are beta.
 */
#ifndef FQMC_H
#define FQMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define FQ_OK 0
#define FQ_ERR_ARG (-1)            /* invalid argument / unsupported combination */
#define FQ_ERR_NOMEM (-2)          /* device allocation failed */
#define FQ_ERR_DEVICE (-3)         /* HIP runtime error */
#define FQ_ERR_RNG_EXHAUSTED (-4)  /* replay stream ran out of draws mid-cycle */
#define FQ_ERR_CAPACITY (-5)       /* initial walker list larger than the main list */
#define FQ_ERR_NO_DEVICE (-6)      /* no usable GPU */
#define FQ_ERR_BIASED (-7)         /* heat-bath tables cannot reach every allowed single excitation
                                      (the stop_all of src/excit_gen_heat_bath_mol.F90:185-202) */
#define FQ_ERR_OVERFLOW (-8)       /* report loop stopped early: the spawned or main list filled up
                                      (spawn%error / psip_list%error, src/fciqmc.f90:415-422) */
#define FQ_ERR_SCRATCH (-9)        /* fq_cycle_noalloc: the cycle needs more than the fixed scratch holds */

/* ---- enumerations ------------------------------------------------------ */
enum fq_system_kind { FQ_SYS_HUB_K = 1, FQ_SYS_READ_IN = 2 };
enum fq_excit_gen {
  FQ_EXCIT_RENORM = 0,     /* excit_gen 'renorm' (Hubbard-k split generator / gen_excit_mol) */
  FQ_EXCIT_NO_RENORM = 1,  /* excit_gen 'no_renorm' (Hubbard-k only) */
  FQ_EXCIT_HEAT_BATH = 2,  /* excit_gen 'heat_bath': gen_excit_mol_heat_bath, src/excit_gen_heat_bath_mol.F90:252-548 (read_in only) */
  FQ_EXCIT_LEARNED = 3     /* learned double proposal mixed with a uniform fallback; singles as 'renorm'
                              (read_in only, production RNG only; no reference counterpart) */
};
enum fq_rng_mode {
  FQ_RNG_PHILOX = 0, /* production: counter-based Philox4x32-10 per attempt */
  FQ_RNG_REPLAY = 1  /* correctness: consume a host-supplied dSFMT stream in reference order */
};

#define FQ_MAX_SPACES 2
#define FQ_MAX_WORDS 2

/* Read-only system tables (host-computed once, uploaded by fq_create).
 * Hubbard-k fields follow src/momentum_symmetry.f90:17-71 and
 * src/system.f90:578; molecular fields follow src/molecular_integrals.F90
 * (RHF store, 8-fold packed two-body index, :736-845) and
 * src/pg_symmetry.f90 (XOR products of abelian irreps). */
typedef struct fq_system {
  int32_t system;      /* enum fq_system_kind */
  int32_t nbasis;      /* spin orbitals */
  int32_t nel, nalpha, nbeta;
  int32_t nvirt, nvirt_alpha, nvirt_beta;
  int32_t nwords;      /* 64-bit words per determinant, ceil(nbasis/64) */
  const uint64_t *f0;  /* reference determinant, nwords words */
  double H00;          /* <D0|H|D0> */
  const double *sp_eigv; /* nbasis single-particle energies */
  /* Hubbard model in momentum space */
  int32_t nsym;              /* number of wavevectors = nbasis/2 */
  const int32_t *sym_table;  /* nsym*nsym, entry [(i-1)*nsym+(j-1)] = label of k_i+k_j (1-based) */
  const int32_t *inv_sym;    /* nsym, label of -k_i (1-based) */
  double coulomb_k;          /* U / nsites */
  /* molecular Hamiltonian from an integral dump (RHF) */
  int32_t norb;                    /* spatial orbitals, nbasis/2 */
  int32_t nsym_pg;                 /* irreps 0..nsym_pg-1, products by XOR */
  const int32_t *orb_sym;          /* nbasis: irrep of each spin orbital */
  const int32_t *nbasis_sym_spin;  /* 2*nsym_pg: [ims*nsym_pg + sym], ims 0 = beta, 1 = alpha */
  int32_t max_sym_spin;            /* max(nbasis_sym_spin) */
  const int32_t *sym_spin_basis_fns; /* 2*nsym_pg*max_sym_spin: [(ims*nsym_pg+sym)*max_sym_spin + k] = orbital (1-based) */
  double ecore;
  const double *h1;                /* norb*norb one-body integrals (spatial, row-major) */
  const double *eri;               /* packed (ij|kl), index tri(tri(i,j),tri(k,l)) over spatial orbitals */
  int64_t n_eri;
  double pattempt_single, pattempt_double;
} fq_system;

/* Per-calculation parameters (src/qmc_data.f90:121-303 defaults). */
typedef struct fq_params {
  int32_t nspaces;              /* 1, or 2 for a replica pair */
  int32_t initiator;            /* initiator approximation (src/ifciqmc.f90) */
  double initiator_pop;         /* initiator threshold, decoded units (default 3.0) */
  int32_t quadrature_initiator; /* combine pairs of spaces (default true) */
  int32_t real_amplitudes;
  int32_t pop_bit_shift;        /* pop_real_factor = 2^shift: 0, 11 (real_32) or 31 (real_64) */
  int64_t spawn_cutoff;         /* encoded: ceiling(cutoff*real_factor); 0 for integer runs */
  int32_t excit_gen;            /* enum fq_excit_gen */
  int32_t rng_mode;             /* enum fq_rng_mode */
  uint64_t philox_seed;         /* production-mode key */
  int64_t main_capacity;        /* main walker list slots (state_size) */
  int64_t spawn_capacity;       /* spawned list slots (spawned_state_size) */
  int32_t device;               /* HIP device ordinal */
  double proposal_mix;          /* FQ_EXCIT_LEARNED: fraction of doubles drawn from the uniform
                                   fallback, in (0, 1]; 0 selects 0.1 */
  int32_t proposal_min_attempts; /* FQ_EXCIT_LEARNED: determinants with fewer spawning attempts in a
                                    cycle skip the network: their doubles use flat hole and particle
                                    weights with the learned affinity, in the same mixture (still
                                    exact); 0 runs the network on every determinant */
} fq_params;

/* Per-cycle results, the quantities the reference accumulates per cycle. */
typedef struct fq_cycle_out {
  double proj_energy[FQ_MAX_SPACES];   /* sum_j H_0j N_j over dets at cycle start */
  double D0_population[FQ_MAX_SPACES]; /* reference population at cycle start */
  double nparticles[FQ_MAX_SPACES];    /* sum |N_i| after annihilation (decoded) */
  int64_t nstates;                     /* occupied determinants after annihilation */
  int64_t nspawn_events;               /* spawned entries created this cycle */
  int64_t ndeath;                      /* sum |kill| (encoded units) */
  int64_t nattempts;                   /* spawning attempts made this cycle */
  int64_t rng_consumed;                /* replay mode: draws taken from the stream */
  int32_t spawn_overflow;              /* spawned list full (spawn%error) */
  int32_t main_overflow;               /* main list full (psip_list%error) */
} fq_cycle_out;

typedef struct fq_engine fq_engine;

/* Create an engine on params->device and upload the system tables.
 * Replaces the allocation in init_psip_list/init_spawn_store
 * (src/qmc.F90:773,1418) and the table set-up read by the procedure pointers
 * bound in init_proc_pointers (src/qmc.F90:217-703).
 * FQ_ERR_ARG for an inconsistent request, including the heat-bath or learned
 * generator on a system other than read_in or with more than 32 electrons,
 * spawned lists whose sort keys would not fit 128 bits (ceil(log2 C(nbasis, nel))
 * plus ceil(log2 spawn_capacity) above 128: half-filled systems near 128 spin orbitals),
 * the heat-bath generator when its per-attempt weights (8 bytes per
 * electron per thread of a 256-thread workgroup) do not fit the device's
 * workgroup LDS, and the learned generator when its affinity table (4 nbasis^2
 * bytes) with the spawning tile's staging does not; FQ_ERR_BIASED when the heat-bath tables cannot reach some
 * single excitation (the reference stops there). */
int fq_create(const fq_system *sys, const fq_params *params, fq_engine **out);

/* Load a sorted walker list (initial_distribution, src/qmc.F90:1507, or a
 * restart).  states: nwords*n, pops: nspaces*n, dat: n (K_ii). */
int fq_load(fq_engine *e, const uint64_t *states, const int64_t *pops, const double *dat, int64_t n);

/* Semi-stochastic projection (init_semi_stoch_t, src/semi_stoch.F90:134-377,
 * space "high"): make the target_size most populated determinants of the
 * current list (sum over spaces of |N|, chosen as find_most_populated_dets
 * does, :1318-1385) the deterministic core, and build the Hamiltonian among
 * them on the device (create_determ_hamil, :533-685).  From the next cycle the
 * core is projected exactly (determ_projection, :859-1094): separate = 1 adds
 * the encoded result to the core populations (separate annihilation, the
 * reference's default), 0 appends it to the spawned list (combined
 * annihilation); core parents are initiators, their spawns onto the core and
 * their deaths are left to the projection, and core determinants stay in the
 * list whatever their population.  target_size 0 switches it off; fq_load
 * also clears the core.  Not for batched engines or during a split cycle. */
int fq_semi_stoch_init(fq_engine *e, int64_t target_size, int separate);
/* The core's size, its Hamiltonian's stored elements (nnz, may be NULL) and,
 * if states is not NULL, its determinants (list order, nwords each). */
int fq_semi_stoch_info(fq_engine *e, int64_t *ndeterm, int64_t *nnz, uint64_t *states, int64_t capacity);

/* Append draws to the device replay stream (FQ_RNG_REPLAY only).  The host
 * generates them with dSFMT exactly as get_rand_close_open hands them out
 * (lib/dSFMT_F03_interface/dSFMT_interface.F90:387-409). */
int fq_push_rng(fq_engine *e, const double *draws, int64_t n);
/* Draws pushed but not yet consumed. */
int fq_rng_available(fq_engine *e, int64_t *n);

/* One Monte Carlo cycle (src/fciqmc.f90:307-399): spawning and death over
 * the main list, then direct_annihilation (src/annihilation.f90:9).
 * shift / eproj_old: per space (eproj_old = proj_energy_old used by death,
 * src/fciqmc.f90:278,370).  In replay mode, returns FQ_ERR_RNG_EXHAUSTED if
 * the stream ran dry; the engine state is then undefined. */
int fq_cycle(fq_engine *e, const double *shift, const double *eproj_old, double tau, fq_cycle_out *out);

/* A replay-mode cycle split at the spawned list, for emulated MPI ranks (a
 * multi-rank reference run replayed on one GPU; the exchange step of a sharded
 * engine).  fq_cycle_spawn runs the loop over determinants (estimators,
 * spawning, death) and keeps the spawned list on the device (out: proj_energy,
 * D0_population, nspawn_events = entries spawned, ndeath, nattempts);
 * fq_spawned_list copies it (creation order: states nwords*n, pops nspaces*n,
 * flags n; capacity entries); fq_cycle_finish annihilates the caller's list
 * (the entries this rank receives) into the post-death walker list
 * (direct_annihilation, src/annihilation.f90:9) and completes out.  fq_cycle
 * and fq_annihilate return FQ_ERR_ARG while a split cycle is pending. */
int fq_cycle_spawn(fq_engine *e, const double *shift, const double *eproj_old, double tau, fq_cycle_out *out);
int fq_spawned_list(fq_engine *e, uint64_t *states, int64_t *pops, int64_t *flags, int64_t capacity, int64_t *n);
int fq_cycle_finish(fq_engine *e, const uint64_t *sstates, const int64_t *spops, const int64_t *sflags, int64_t ns,
                    fq_cycle_out *out);

/* fq_cycle without device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  fq_cycle
 * grows its prefix-sum scratch on demand; this variant takes it from a fixed
 * device array of the library (8 MiB per device, held until the process ends,
 * enough for the longest scan an engine issues) and calls no hipMalloc or
 * hipFree.  Limits: production RNG, generators other than the learned proposal
 * (FQ_ERR_ARG otherwise); FQ_ERR_SCRATCH (walker list unchanged) if a scan
 * should ever exceed the array.  The array is shared by every engine on a
 * device: calls that overlap on one device race.  Results equal fq_cycle's bit
 * for bit. */
int fq_cycle_noalloc(fq_engine *e, const double *shift, const double *eproj_old, double tau, fq_cycle_out *out);

/* Annihilate a caller-supplied spawned list into the main list
 * (direct_annihilation without the spawning loop; src/annihilation.f90:9-214
 * and src/spawn_data.F90:359-1101).  sstates: nwords*ns, spops: nspaces*ns,
 * sflags: ns initiator flag words (ignored without the initiator approximation).
 * Replay mode draws the real-amplitude rounding numbers from the stream. */
int fq_annihilate(fq_engine *e, const uint64_t *sstates, const int64_t *spops, const int64_t *sflags,
                  int64_t ns, fq_cycle_out *out);

/* Current walker count and a copy of the main list to the host (report /
 * restart boundaries only).  Arrays must hold capacity entries. */
int fq_nstates(fq_engine *e, int64_t *n);
int fq_download(fq_engine *e, uint64_t *states, int64_t *pops, double *dat, int64_t capacity, int64_t *n);

/* Device-side timing of the last cycle's kernels (milliseconds). */
int fq_last_timing(fq_engine *e, double *spawn_ms, double *annihilate_ms, double *total_ms);

/* Per-phase device times of the last fq_cycle, from HIP events recorded on
 * the engine's stream around each phase, and the sizes that phase handled. */
typedef struct fq_profile_out {
  double det_pass_ms;    /* per-determinant pass: estimators, nattempts, death */
  double spawn_ms;       /* spawning-attempt kernel */
  double sort_ms;        /* radix sort (or reference quicksort) of the spawned list */
  double reduce_ms;      /* run detection and self-annihilation */
  double merge_ms;       /* merge with the main list, rounding, compaction, insertion */
  double total_ms;       /* whole cycle, first to last kernel */
  int64_t nstates_in;    /* occupied determinants at the start of the cycle */
  int64_t nattempts;     /* spawning attempts */
  int64_t nspawned;      /* spawned entries written */
  int64_t nruns;         /* unique spawned determinants */
  int64_t nstates_out;   /* occupied determinants after annihilation */
  int32_t radix_passes;  /* 8-bit digit passes of the sort */
  int32_t spawn_launches;    /* launches of the spawning-attempt kernel this cycle */
  double spawn_kernel_ms;    /* summed device time of those launches (event pair around each) */
  double proposal_ms;        /* learned proposal: network marking, inference and attempt offsets */
  int64_t proposal_rows;     /* learned proposal: determinants the network ran on */
} fq_profile_out;
int fq_last_profile(fq_engine *e, fq_profile_out *out);

/* ---- Batched independent instances ------------------------------------------
 * Many small systems (or independent replicas) propagated by one engine: one
 * walker list, one set of kernel launches per cycle for all instances.  No
 * counterpart in the reference (each HANDE run is one system).  Instance k's
 * tables come from systems[k]; all share params (integer populations,
 * production RNG, one 64-bit word per determinant, nbasis <= 48).  The
 * instance index occupies bits 48..63 of a determinant's word (fq_download
 * returns it there).  Instance k draws from Philox key philox_seed + k with
 * counters numbered from its own first determinant and attempt, so each
 * instance follows exactly the chain it would follow alone in an engine
 * created with philox_seed + k. */
int fq_create_batch(const fq_system *systems, int32_t ninst, const fq_params *params, fq_engine **engine);
/* n walkers: inst[i] is the instance of states[i] (bits 48..63 clear). */
int fq_load_batch(fq_engine *e, const int32_t *inst, const uint64_t *states, const int64_t *pops, const double *dat,
                  int64_t n);
/* One cycle of every instance: shift and eproj_old are ninst x nspaces, tau is
 * per instance; out[k] receives instance k's results. */
int fq_cycle_batch(fq_engine *e, const double *shift, const double *eproj_old, const double *tau, fq_cycle_out *out);

/* Graph-captured batched cycles (no counterpart in the reference).  With
 * enable = 1, fq_cycle_batch runs each cycle as one hipGraph launch: the
 * cycle's launches are captured once per (list buffer, length step, attempt
 * step) with grids sized by those steps (four per octave) and list lengths
 * read on the device, so the host neither waits nor launches inside a cycle;
 * the cycle's inputs enter through one pinned-memory copy at the graph's
 * start and its outputs leave through copies at its end.  The first graph
 * cycles try each candidate schedule (radix digit widths, spawning grid) in
 * turn and keep the fastest by median cycle time, separately for each octave
 * of the attempt bound (their relative cost changes with the workload's
 * size); every schedule gives the
 * same walker lists, bit for bit, so the chain does not depend on the choice.
 * A cycle whose attempt bound exceeds the spawning staging capacity, and the
 * first cycle after fq_load_batch, run as ordinary launches. */
typedef struct fq_schedule_out {
  int32_t graph;          /* graph cycles enabled */
  int32_t tuned;          /* the tuner has chosen for the octave of the last graph cycle's attempt bound */
  int32_t digit_bits;     /* widest radix digit of the chosen schedule: 9 (8- and 9-bit digits) or 8 */
  int32_t spawn_blocks;   /* spawning grid cap of the chosen schedule (0: one block per 1024-attempt tile) */
  int32_t ncand;          /* candidate schedules */
  int32_t chosen;         /* index of the chosen one (-1 while tuning) */
  int32_t octave;         /* floor(log2) of the last graph cycle's attempt bound (-1: none yet) */
  int32_t tuned_octaves;  /* octaves with a chosen schedule (the tuner works per octave) */
  int64_t graph_cycles;   /* cycles run as graph launches */
  int64_t graphs;         /* graphs captured */
  int64_t eager_cycles;   /* cycles run as ordinary launches */
  double cand_ms[8];      /* median device time of one cycle under each candidate (ms; 0 if untried) */
  int32_t cand_digit_bits[8];
  int32_t cand_spawn_blocks[8];
} fq_schedule_out;
int fq_batch_graph(fq_engine *e, int32_t enable);
int fq_batch_schedule(fq_engine *e, fq_schedule_out *out);

/* ---- Learned double-excitation proposal (excit_gen FQ_EXCIT_LEARNED) ---------
 * No counterpart in the reference: an addition beside its heat-bath generator
 * (src/excit_gen_heat_bath_mol.F90), which produces the training samples.
 * The network maps a determinant's occupations (and occupied orbital
 * energies) through two 256-wide GELU layers to hole and particle weights; a
 * learned determinant-independent hole-particle affinity table weights the
 * particle paired with each hole.  Parameters are float32, laid out
 * W1[2n][256], b1[256], W2[256][256], b2[256], W3[256][2n], b3[2n],
 * theta[n][n] (n = nbasis; affinity exp(clamp(theta, -60, 60))).  fq_create
 * initialises the network from philox_seed and theta to zero; inference runs
 * in bf16 with fp32 accumulation every cycle. */
int fq_mlp_nparams(fq_engine *e, int64_t *n);
int fq_mlp_set_params(fq_engine *e, const float *params, int64_t n);  /* also resets the optimiser */
int fq_mlp_get_params(fq_engine *e, float *params, int64_t n);
/* Proposal weights exp(clamp(logit, -60, 60)) of n determinants (n x W words,
 * determinant-major as in fq_load) into out (n x 2*nbasis: holes, then particles). */
int fq_mlp_forward(fq_engine *e, const uint64_t *dets, int64_t n, float *out);
/* The stored affinity table (n = nbasis * nbasis floats, [hole-1][particle-1]). */
int fq_mlp_affinity(fq_engine *e, float *out, int64_t n);
/* Training loss sum_s w_s (-log q(ex_s | D_s)) / sum_s w_s of n samples
 * (exc: n x 4 = i, j, a, b with i<j, a<b) and its gradient (nparams floats;
 * grad may be NULL); fp32 forward and backward.  flat (n flags, may be NULL):
 * a flagged sample is scored with the weights of a determinant the network
 * does not run on (flat hole and particle weights, the affinity). */
int fq_mlp_loss_grad(fq_engine *e, const uint64_t *dets, const int32_t *exc, const double *w, const int32_t *flat,
                     int64_t n, double *loss, float *grad);
/* steps Adam steps (learning rate lr), each on batch heat-bath samples drawn on
 * the device from the current walker list (determinant chosen with weight
 * 1 + floor|N_i|); a sample whose determinant has 1 + floor|N_i| below
 * proposal_min_attempts is scored flat, as production draws for it.  losses
 * (may be NULL) receives each step's loss. */
int fq_mlp_train(fq_engine *e, int64_t batch, int32_t steps, double lr, uint64_t seed, double *losses);
/* The spawning-variance objective sum_s w_s |H_s| / q(ex_s | D_s) / sum_s w_s
 * (habs: |H_s| per sample; with w_s = |H_s| / p_s for samples drawn with
 * probability p_s an estimate of sum |H|^2 / q) and its gradient, as
 * fq_mlp_loss_grad. */
int fq_mlp_variance_grad(fq_engine *e, const uint64_t *dets, const int32_t *exc, const double *w,
                         const double *habs, const int32_t *flat, int64_t n, double *loss, float *grad);
/* Online training: with every > 0, a production fq_cycle whose cycle index is
 * a multiple of every first runs steps Adam steps (learning rate lr) on batch
 * heat-bath samples of the current walker list, as fq_mlp_train, and the
 * cycle's proposal is the updated one (the sampler stays exact for any
 * parameters).  objective: 0 the weighted cross-entropy to q* ~ |H|
 * (fq_mlp_loss_grad), 1 the spawning variance (fq_mlp_variance_grad); it is
 * also the objective of fq_mlp_train from then on.  every = 0 stops it. */
int fq_mlp_online(fq_engine *e, int32_t every, int32_t steps, int64_t batch, double lr, int32_t objective);

/* ---- Excitation probe (molecular systems) --------------------------------------
 * n independent calls of the engine's excitation generator from one
 * determinant (W words), Philox stream keyed by seed: exc (n x 6: allowed,
 * nexcit, from1, from2, to1, to2), pgen and the matrix element of each.  The
 * generator is the one the spawning kernel uses (gen_excit_mol,
 * gen_excit_mol_heat_bath or the learned proposal), without attempt_to_spawn.
 * The learned proposal uses the network's row for the determinant, or, when
 * proposal_min_attempts > 1 (a determinant with one attempt), the flat weights. */
int fq_probe_excitations(fq_engine *e, const uint64_t *det, int64_t n, uint64_t seed, int32_t *exc, double *pgen,
                         double *hmatel);

void fq_destroy(fq_engine *e);

#ifdef __cplusplus
}
#endif
#endif /* FQMC_H */
