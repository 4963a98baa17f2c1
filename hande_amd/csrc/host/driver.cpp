// FCIQMC report-loop driver on the engine C-ABI.
//
// Host counterpart of do_fciqmc's report loop (src/fciqmc.f90:276-464):
// initial_distribution (src/qmc.F90:1507-1560), initial_ci_projected_energy
// and initial_qmc_status (src/qmc_common.F90:654-797), init_report_loop /
// init_mc_cycle / end_mc_cycle / spawning_rate (src/qmc_common.F90:929-1017,
// 1240-1304), and end_report_loop -> update_energy_estimators ->
// communicated_energy_estimators -> update_shift
// (src/energy_evaluation.F90:126-201,425-711, get_sanitized_projected_energy
// :1433-1454).  The per-cycle work runs in the engine (fq_cycle).  In replay
// mode this driver owns the dSFMT stream (seed rng_seed + iproc with iproc = 0,
// src/fciqmc.f90:206) and keeps the engine's device copy topped up.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "blocking.hpp"
#include "dsfmt.hpp"
#include "fqmc.h"
#include "fqmc_host.h"

namespace {

// Elements from a state_size-style input: >= 0 is a count, < 0 is MB per core
// (src/particle_t_utils.f90:84-101, src/qmc.F90:1445-1448).
int64_t capacity_from(int64_t size, int64_t element_bytes, bool spawned) {
  if (size >= 0) return size;
  const double mb = -static_cast<double>(size) * 1.0e6;
  return static_cast<int64_t>(spawned ? mb / (2.0 * element_bytes) : mb / element_bytes);
}

class StreamFeeder {
 public:
  StreamFeeder(fq_engine* e, uint32_t seed) : e_(e), rng_(seed) {}
  // Ensure at least `need` unconsumed draws are on the device.
  int ensure(int64_t need) {
    int64_t have = 0;
    int rc = fq_rng_available(e_, &have);
    if (rc != FQ_OK) return rc;
    if (have >= need) return FQ_OK;
    const int64_t add = std::max<int64_t>(need - have, 1 << 20);
    buf_.resize(add);
    rng_.fill(buf_.data(), add);
    return fq_push_rng(e_, buf_.data(), add);
  }

 private:
  fq_engine* e_;
  fqmc::Dsfmt19937 rng_;
  std::vector<double> buf_;
};

}  // namespace

extern "C" void fqh_qmc_in_defaults(fqh_qmc_in* in) {
  if (!in) return;
  in->vary_shift_set = 0;
  in->vary_shift = 0;
  in->blocking_on_the_fly = 0;
  in->auto_shift_damping = 0;
  in->shift_damping_precision = 2.0;
  in->force_shift_damping_opt = 0;
  in->start_save_frequency = -1;
  in->start_point_number = -1;
  in->start_point = -1;
  in->error_limit = 0.0;
  in->min_blocks_used = 10;
  in->blocks_used = 2147483647;
  in->semi_stoch_size = 0;
  in->semi_stoch_start_iter = 1;
  in->semi_stoch_shift_iter = -1;
  in->semi_stoch_separate = 1;
}

extern "C" int fqh_fciqmc(const fqh_system* hs, const fqh_qmc_in* in, fqh_report_row* rows, int max_rows,
                          int* nrows) {
  return fqh_fciqmc_restart(hs, in, nullptr, nullptr, rows, max_rows, nrows);
}

extern "C" int fqh_fciqmc_restart(const fqh_system* hs, const fqh_qmc_in* in, const fqh_restart* start,
                                  fqh_restart* end, fqh_report_row* rows, int max_rows, int* nrows) {
  if (!hs || !in || !rows || !nrows) return FQ_ERR_ARG;
  if (start && start->nstates > 0 && (!start->states || !start->pops || !start->dat)) return FQ_ERR_ARG;
  const fq_system* sys = fqh_view(hs);
  fq_params p{};
  p.nspaces = in->replica_tricks ? 2 : 1;
  p.initiator = in->initiator;
  p.initiator_pop = in->initiator_threshold;
  p.quadrature_initiator = in->quadrature_initiator;
  p.real_amplitudes = in->real_amplitudes;
  p.pop_bit_shift = in->real_amplitudes ? (in->real_amplitude_force_32 ? 11 : 31) : 0;
  const int64_t rf = int64_t{1} << p.pop_bit_shift;
  // spawn%cutoff = ceiling(cutoff*real_factor) (src/spawn_data.F90:215); 0 for integers (src/qmc.F90:1485)
  p.spawn_cutoff = in->real_amplitudes ? static_cast<int64_t>(std::ceil(in->spawn_cutoff * static_cast<double>(rf))) : 0;
  p.excit_gen = in->excit_gen;
  p.proposal_mix = in->proposal_mix;
  p.proposal_min_attempts = in->proposal_min_attempts;
  p.rng_mode = in->rng_mode;
  p.philox_seed = static_cast<uint64_t>(in->rng_seed);
  p.device = in->device;
  const int S = p.nspaces;
  const int W = sys->nwords;
  p.main_capacity = capacity_from(in->state_size, 8 * W + 8 * S + 8 * S, false);
  p.spawn_capacity = capacity_from(in->spawned_state_size, 8 * (W + S + (in->initiator ? 1 : 0)), true);
  fq_engine* e = nullptr;
  int rc = fq_create(sys, &p, &e);
  if (rc != FQ_OK) return rc;
  const bool replay = in->rng_mode == FQ_RNG_REPLAY;
  StreamFeeder feeder(e, static_cast<uint32_t>(static_cast<int32_t>(in->rng_seed)));
  double proj[FQ_MAX_SPACES] = {0, 0}, d0[FQ_MAX_SPACES] = {0, 0};
  double nparticles[FQ_MAX_SPACES] = {0, 0};
  int64_t nstates = 1, cycles0 = 0;
  if (!start) {
    // initial_distribution: the reference determinant with nint(D0_pop)*real_factor.
    std::vector<int64_t> pop0(S, static_cast<int64_t>(std::llround(in->init_pop)) * rf);
    double dat0 = 0.0;
    rc = fq_load(e, sys->f0, pop0.data(), &dat0, 1);
    for (int s = 0; s < S; ++s) {
      nparticles[s] = std::fabs(static_cast<double>(pop0[s]) / static_cast<double>(rf));
      // initial_ci_projected_energy over the single reference determinant
      d0[s] = static_cast<double>(pop0[s]) / static_cast<double>(rf);
      proj[s] = 0.0;
    }
  } else {
    // read_restart_hdf5 + init_estimators: the stored list and estimators; the
    // population recounted as sum |pop| / real_factor in list order.
    rc = fq_load(e, start->states, start->pops, start->dat, start->nstates);
    nstates = start->nstates;
    cycles0 = start->mc_cycles_done;
    for (int s = 0; s < S; ++s) {
      proj[s] = start->proj_energy[s];
      d0[s] = start->D0_population[s];
      double t = 0.0;
      for (int64_t i = 0; i < start->nstates; ++i)
        t = t + std::fabs(static_cast<double>(start->pops[i * S + s]) / static_cast<double>(rf));
      nparticles[s] = t;
    }
  }
  if (rc != FQ_OK) {
    fq_destroy(e);
    return rc;
  }
  fqmc::ShiftControl ctl;
  ctl.init(*in, S, nparticles, start);
  const double* shift = ctl.shift;
  int nr = 0;
  auto emit = [&](int64_t iter, int64_t nspawn, double rspawn, double tcyc) {
    if (nr >= max_rows) return;
    fqh_report_row& r = rows[nr++];
    std::memset(&r, 0, sizeof(r));
    r.iterations = iter;
    for (int s = 0; s < S; ++s) {
      r.shift[s] = shift[s];
      r.proj_energy[s] = proj[s];
      r.D0_population[s] = d0[s];
      r.nparticles[s] = ctl.ntot_old[s];
    }
    r.nstates = nstates;
    r.nspawn_events = nspawn;
    r.rspawn = rspawn;
    r.time_per_cycle = tcyc;
    r.shift_damping = ctl.damping;
  };
  emit(cycles0, 0, 0.0, 0.0);
  bool overflowed = false;
  int reports_done = 0;
  // the iteration the semi-stochastic core is built (src/fciqmc.f90:227-235):
  // start_iter, or shift_iter iterations after the shift starts varying (at
  // once if it already varies)
  int64_t ss_iter = std::max<int64_t>(in->semi_stoch_start_iter, cycles0 + 1);
  auto all_vary = [&]() {
    for (int s = 0; s < S; ++s)
      if (!ctl.vary[s]) return false;
    return true;
  };
  if (all_vary() && in->semi_stoch_shift_iter != -1) ss_iter = cycles0 + 1;
  for (int ireport = 1; ireport <= in->nreports && nr < max_rows; ++ireport) {  // a full table ends the run
    double eproj_old[FQ_MAX_SPACES] = {0, 0};
    for (int s = 0; s < S; ++s) eproj_old[s] = (std::fabs(d0[s]) < 2.2250738585072014e-308) ? 0.0 : proj[s] / d0[s];
    double rspawn = 0.0;
    for (int s = 0; s < S; ++s) proj[s] = d0[s] = 0.0;
    int64_t nspawn_events = 0;
    bool error = false;
    const auto t0 = std::chrono::steady_clock::now();
    for (int ic = 0; ic < in->mc_cycles; ++ic) {
      const int64_t iter = cycles0 + static_cast<int64_t>(ireport - 1) * in->mc_cycles + ic + 1;
      if (in->semi_stoch_size > 0 && iter == ss_iter) {
        rc = fq_semi_stoch_init(e, in->semi_stoch_size, in->semi_stoch_separate);
        if (rc != FQ_OK) break;
      }
      const int64_t nattempts = static_cast<int64_t>(std::llround(2 * nparticles[0]));
      if (replay) {
        double wsum = 0.0;
        for (int s = 0; s < S; ++s) wsum += nparticles[s];
        const int64_t bound = 4096 + 8 * S * nstates + static_cast<int64_t>(48.0 * (wsum + S * nstates));
        rc = feeder.ensure(bound);
        if (rc != FQ_OK) break;
      }
      fq_cycle_out out;
      rc = fq_cycle(e, shift, eproj_old, in->tau, &out);
      if (rc != FQ_OK) break;
      for (int s = 0; s < S; ++s) {
        proj[s] = proj[s] + out.proj_energy[s];
        d0[s] = d0[s] + out.D0_population[s];
        nparticles[s] = out.nparticles[s];
      }
      nstates = out.nstates;
      nspawn_events = out.nspawn_events;
      double rate = 0.0;
      if (nattempts > 0) {
        const double ndeath_real = static_cast<double>(out.ndeath) / static_cast<double>(rf);
        rate = (static_cast<double>(nspawn_events) + ndeath_real) / static_cast<double>(nattempts);
      }
      rspawn = rspawn + rate;
      error = error || out.spawn_overflow || out.main_overflow;
    }
    if (rc != FQ_OK) break;
    const double tcyc =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / in->mc_cycles;
    double ntot[FQ_MAX_SPACES];
    for (int s = 0; s < S; ++s) {
      ntot[s] = nparticles[s];
      proj[s] = proj[s] / in->mc_cycles;
      d0[s] = d0[s] / in->mc_cycles;
    }
    rspawn = rspawn / in->mc_cycles;
    const bool vary_before = all_vary();
    const bool soft_exit = ctl.end_report(ntot, proj, d0);
    if (!vary_before && all_vary() && in->semi_stoch_shift_iter != -1)  // end_report_loop, src/qmc_common.F90:1203
      ss_iter = in->semi_stoch_shift_iter + cycles0 + static_cast<int64_t>(ireport) * in->mc_cycles + 1;
    emit(cycles0 + static_cast<int64_t>(ireport) * in->mc_cycles, nspawn_events, rspawn, tcyc);
    ctl.after_report(ireport, proj[0], d0[0]);
    if (nr > 0 && nr <= max_rows) {  // the damping and blocked estimate once this report is processed
      rows[nr - 1].shift_damping = ctl.damping;
      if (ctl.blocking()) {
        rows[nr - 1].blocked_proj_energy = ctl.blocks().proj_energy_mean();
        rows[nr - 1].blocked_proj_error = ctl.blocks().proj_energy_error();
      }
    }
    reports_done = ireport;
    if (error) {
      overflowed = true;
      break;
    }
    if (soft_exit) break;
  }
  if (rc == FQ_OK && end) {
    // dump_restart_hdf5: the list, shift, flags and the last report's estimators
    rc = fq_download(e, end->states, end->pops, end->dat, end->capacity, &end->nstates);
    end->mc_cycles_done = cycles0 + static_cast<int64_t>(reports_done) * in->mc_cycles;
    for (int s = 0; s < FQ_MAX_SPACES; ++s) {
      end->shift[s] = s < S ? shift[s] : 0.0;
      end->vary_shift[s] = s < S && ctl.vary[s] ? 1 : 0;
      end->proj_energy[s] = s < S ? proj[s] : 0.0;
      end->D0_population[s] = s < S ? d0[s] : 0.0;
    }
    end->shift_damping = ctl.damping;
    end->shift_damping_status = ctl.status;
  }
  fq_destroy(e);
  *nrows = nr;
  // A full spawned or main list stops the report loop after the current
  // report, as the reference's sticky spawn%error / psip_list%error flags do
  // (src/fciqmc.f90:415-422); the caller sees it as FQ_ERR_OVERFLOW with the
  // rows emitted so far.
  if (rc == FQ_OK && overflowed) rc = FQ_ERR_OVERFLOW;
  return rc;
}

// MurmurHash2 (Appleby; the reference's copy is lib/external/MurmurHash2.c:16):
// 32-bit hash of len bytes read as little-endian 32-bit words.
extern "C" uint32_t fqh_murmur2(const void* key, int len, uint32_t seed) {
  const unsigned char* data = static_cast<const unsigned char*>(key);
  const uint32_t m = 0x5bd1e995u;
  uint32_t h = seed ^ static_cast<uint32_t>(len);
  for (; len >= 4; data += 4, len -= 4) {
    uint32_t k;
    std::memcpy(&k, data, 4);  // little-endian host
    k *= m;
    k ^= k >> 24;
    k *= m;
    h *= m;
    h ^= k;
  }
  switch (len) {
    case 3: h ^= static_cast<uint32_t>(data[2]) << 16; [[fallthrough]];
    case 2: h ^= static_cast<uint32_t>(data[1]) << 8; [[fallthrough]];
    case 1: h ^= data[0]; h *= m;
  }
  h ^= h >> 13;
  h *= m;
  h ^= h >> 15;
  return h;
}

// The rank owning a determinant (assign_particle_processor with hash_shift 0
// and one processor-map slot per rank, src/spawning.F90:770-838,
// src/load_balancing.F90:141-170): MurmurHash2 with seed 7 (src/qmc.F90:1496)
// of the first ceiling(nbasis/32)*4 bytes of the bit string
// (lib/local/hash.f90), the signed hash modulo nprocs as Fortran's modulo.
extern "C" int fqh_owner_rank(const uint64_t* f, int nbasis, int nprocs) {
  if (!f || nprocs < 1) return -1;
  unsigned char bytes[8 * FQ_MAX_WORDS] = {};
  const int nw = std::min((nbasis + 63) / 64, FQ_MAX_WORDS);
  for (int w = 0; w < nw; ++w)
    for (int b = 0; b < 8; ++b) bytes[8 * w + b] = static_cast<unsigned char>(f[w] >> (8 * b));
  const int32_t hash = static_cast<int32_t>(fqh_murmur2(bytes, (nbasis + 31) / 32 * 4, 7u));
  int r = hash % nprocs;
  return r < 0 ? r + nprocs : r;
}

// The report loop on nprocs emulated MPI ranks, one engine each on the same
// device, replay mode: rank r consumes dSFMT seeded rng_seed + r
// (src/fciqmc.f90:206), owns the determinants that hash to it (the reference
// determinant's initial walkers start on its owner, src/qmc.F90:1574-1578),
// spawns (fq_cycle_spawn), and receives the spawned entries addressed to it in
// rank order, each source's in creation order (the per-destination blocks of
// comm_spawn_t, src/spawn_data.F90:624-739), before annihilating them
// (fq_cycle_finish).  Report quantities are the ranks' sums
// (update_energy_estimators' MPI_Allreduce, src/energy_evaluation.F90:126-201),
// R_spawn their mean (:579).  A host-side exchange: the replay of multi-rank
// reference runs on one GPU, and the exchange step a sharded engine performs.
extern "C" int fqh_fciqmc_ranks(const fqh_system* hs, const fqh_qmc_in* in, int nprocs, fqh_report_row* rows,
                                int max_rows, int* nrows) {
  if (!hs || !in || !rows || !nrows || nprocs < 1) return FQ_ERR_ARG;
  if (nprocs == 1) return fqh_fciqmc(hs, in, rows, max_rows, nrows);
  if (in->rng_mode != FQ_RNG_REPLAY || in->semi_stoch_size > 0) return FQ_ERR_ARG;  // semi-stochastic: one rank
  const fq_system* sys = fqh_view(hs);
  fq_params p{};
  p.nspaces = in->replica_tricks ? 2 : 1;
  p.initiator = in->initiator;
  p.initiator_pop = in->initiator_threshold;
  p.quadrature_initiator = in->quadrature_initiator;
  p.real_amplitudes = in->real_amplitudes;
  p.pop_bit_shift = in->real_amplitudes ? (in->real_amplitude_force_32 ? 11 : 31) : 0;
  const int64_t rf = int64_t{1} << p.pop_bit_shift;
  p.spawn_cutoff = in->real_amplitudes ? static_cast<int64_t>(std::ceil(in->spawn_cutoff * static_cast<double>(rf))) : 0;
  p.excit_gen = in->excit_gen;
  p.rng_mode = FQ_RNG_REPLAY;
  p.device = in->device;
  const int S = p.nspaces;
  const int W = sys->nwords;
  p.main_capacity = capacity_from(in->state_size, 8 * W + 8 * S + 8 * S, false);
  p.spawn_capacity = capacity_from(in->spawned_state_size, 8 * (W + S + (in->initiator ? 1 : 0)), true);
  std::vector<fq_engine*> eng(nprocs, nullptr);
  std::vector<StreamFeeder> feeders;
  int rc = FQ_OK;
  for (int r = 0; r < nprocs && rc == FQ_OK; ++r) rc = fq_create(sys, &p, &eng[r]);
  auto cleanup = [&]() {
    for (fq_engine* e : eng)
      if (e) fq_destroy(e);
  };
  if (rc != FQ_OK) {
    cleanup();
    return rc;
  }
  for (int r = 0; r < nprocs; ++r)
    feeders.emplace_back(eng[r], static_cast<uint32_t>(static_cast<int32_t>(in->rng_seed) + r));
  std::vector<int64_t> pop0(S, static_cast<int64_t>(std::llround(in->init_pop)) * rf);
  double dat0 = 0.0;
  const int d0_rank = fqh_owner_rank(sys->f0, sys->nbasis, nprocs);
  rc = fq_load(eng[d0_rank], sys->f0, pop0.data(), &dat0, 1);
  if (rc != FQ_OK) {
    cleanup();
    return rc;
  }
  std::vector<double> npart(static_cast<size_t>(nprocs) * S, 0.0);  // per-rank particle totals
  std::vector<int64_t> nst(nprocs, 0);
  for (int s = 0; s < S; ++s) npart[d0_rank * S + s] = std::fabs(static_cast<double>(pop0[s]) / static_cast<double>(rf));
  nst[d0_rank] = 1;
  double proj[FQ_MAX_SPACES] = {0, 0}, d0[FQ_MAX_SPACES] = {0, 0};
  auto total = [&](int s) {
    double t = 0.0;
    for (int r = 0; r < nprocs; ++r) t = t + npart[r * S + s];
    return t;
  };
  double ntot0[FQ_MAX_SPACES] = {0, 0};
  for (int s = 0; s < S; ++s) {
    ntot0[s] = total(s);
    d0[s] = static_cast<double>(pop0[s]) / static_cast<double>(rf);
    proj[s] = 0.0;
  }
  fqmc::ShiftControl ctl;  // estimators are the ranks' sums, so the parent's bookkeeping applies
  ctl.init(*in, S, ntot0, nullptr);
  const double* shift = ctl.shift;
  int nr = 0;
  auto emit = [&](int64_t iter, int64_t nspawn, double rspawn, double tcyc) {
    if (nr >= max_rows) return;
    fqh_report_row& row = rows[nr++];
    std::memset(&row, 0, sizeof(row));
    row.iterations = iter;
    for (int s = 0; s < S; ++s) {
      row.shift[s] = shift[s];
      row.proj_energy[s] = proj[s];
      row.D0_population[s] = d0[s];
      row.nparticles[s] = ctl.ntot_old[s];
    }
    int64_t t = 0;
    for (int r = 0; r < nprocs; ++r) t += nst[r];
    row.nstates = t;
    row.nspawn_events = nspawn;
    row.rspawn = rspawn;
    row.time_per_cycle = tcyc;
    row.shift_damping = ctl.damping;
  };
  emit(0, 0, 0.0, 0.0);
  bool overflowed = false;
  std::vector<fq_cycle_out> out(nprocs);
  std::vector<std::vector<uint64_t>> sst(nprocs);
  std::vector<std::vector<int64_t>> spo(nprocs), sfl(nprocs);
  std::vector<uint64_t> rst;
  std::vector<int64_t> rpo, rfl;
  for (int ireport = 1; ireport <= in->nreports && nr < max_rows && rc == FQ_OK; ++ireport) {
    double eproj_old[FQ_MAX_SPACES] = {0, 0};
    for (int s = 0; s < S; ++s) eproj_old[s] = (std::fabs(d0[s]) < 2.2250738585072014e-308) ? 0.0 : proj[s] / d0[s];
    std::vector<double> rproj(static_cast<size_t>(nprocs) * S, 0.0), rd0(static_cast<size_t>(nprocs) * S, 0.0);
    std::vector<double> rsp(nprocs, 0.0);
    std::vector<int64_t> nsp(nprocs, 0);
    bool error = false;
    const auto t0 = std::chrono::steady_clock::now();
    for (int ic = 0; ic < in->mc_cycles && rc == FQ_OK; ++ic) {
      std::vector<int64_t> nattempts(nprocs);
      for (int r = 0; r < nprocs && rc == FQ_OK; ++r) {
        nattempts[r] = static_cast<int64_t>(std::llround(2 * npart[r * S]));
        double wsum = 0.0;
        for (int s = 0; s < S; ++s) wsum += npart[r * S + s];
        rc = feeders[r].ensure(4096 + 8 * S * nst[r] + static_cast<int64_t>(48.0 * (wsum + S * nst[r])));
        if (rc == FQ_OK) rc = fq_cycle_spawn(eng[r], shift, eproj_old, in->tau, &out[r]);
        if (rc != FQ_OK) break;
        const int64_t ns = out[r].nspawn_events;
        sst[r].resize(static_cast<size_t>(W) * ns);
        spo[r].resize(static_cast<size_t>(S) * ns);
        sfl[r].resize(ns);
        int64_t got = 0;
        rc = fq_spawned_list(eng[r], sst[r].data(), spo[r].data(), sfl[r].data(), ns, &got);
      }
      for (int r = 0; r < nprocs && rc == FQ_OK; ++r) {  // the entries rank r receives, in rank order
        rst.clear();
        rpo.clear();
        rfl.clear();
        for (int q = 0; q < nprocs; ++q) {
          const int64_t nq = static_cast<int64_t>(sfl[q].size());
          for (int64_t i = 0; i < nq; ++i) {
            if (fqh_owner_rank(&sst[q][i * W], sys->nbasis, nprocs) != r) continue;
            rst.insert(rst.end(), sst[q].begin() + i * W, sst[q].begin() + (i + 1) * W);
            rpo.insert(rpo.end(), spo[q].begin() + i * S, spo[q].begin() + (i + 1) * S);
            rfl.push_back(sfl[q][i]);
          }
        }
        fq_cycle_out fin;
        rc = fq_cycle_finish(eng[r], rst.data(), rpo.data(), rfl.data(), static_cast<int64_t>(rfl.size()), &fin);
        if (rc != FQ_OK) break;
        for (int s = 0; s < S; ++s) {
          rproj[r * S + s] = rproj[r * S + s] + out[r].proj_energy[s];
          rd0[r * S + s] = rd0[r * S + s] + out[r].D0_population[s];
          npart[r * S + s] = fin.nparticles[s];
        }
        nst[r] = fin.nstates;
        nsp[r] = out[r].nspawn_events;
        double rate = 0.0;
        if (nattempts[r] > 0) {
          const double ndeath_real = static_cast<double>(out[r].ndeath) / static_cast<double>(rf);
          rate = (static_cast<double>(nsp[r]) + ndeath_real) / static_cast<double>(nattempts[r]);
        }
        rsp[r] = rsp[r] + rate;
        error = error || out[r].spawn_overflow || fin.main_overflow;
      }
    }
    if (rc != FQ_OK) break;
    const double tcyc =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / in->mc_cycles;
    double ntot[FQ_MAX_SPACES], rspawn = 0.0;
    int64_t nspawn_events = 0;
    for (int s = 0; s < S; ++s) {
      proj[s] = d0[s] = 0.0;
      for (int r = 0; r < nprocs; ++r) {
        proj[s] = proj[s] + rproj[r * S + s];
        d0[s] = d0[s] + rd0[r * S + s];
      }
      ntot[s] = total(s);
      proj[s] = proj[s] / in->mc_cycles;
      d0[s] = d0[s] / in->mc_cycles;
    }
    for (int r = 0; r < nprocs; ++r) {
      rspawn = rspawn + rsp[r];
      nspawn_events += nsp[r];
    }
    rspawn = rspawn / (in->mc_cycles * nprocs);
    const bool soft_exit = ctl.end_report(ntot, proj, d0);
    emit(static_cast<int64_t>(ireport) * in->mc_cycles, nspawn_events, rspawn, tcyc);
    ctl.after_report(ireport, proj[0], d0[0]);
    if (nr > 0 && nr <= max_rows) {
      rows[nr - 1].shift_damping = ctl.damping;
      if (ctl.blocking()) {
        rows[nr - 1].blocked_proj_energy = ctl.blocks().proj_energy_mean();
        rows[nr - 1].blocked_proj_error = ctl.blocks().proj_energy_error();
      }
    }
    if (error) {
      overflowed = true;
      break;
    }
    if (soft_exit) break;
  }
  cleanup();
  *nrows = nr;
  if (rc == FQ_OK && overflowed) rc = FQ_ERR_OVERFLOW;
  return rc;
}

// Close-open dSFMT-19937 draws as the reference's buffered generator hands
// them out (for checks of the replay stream against golden vectors).
extern "C" int fqh_dsfmt_fill(int64_t seed, double* out, int64_t n) {
  if (!out || n < 0) return FQ_ERR_ARG;
  fqmc::Dsfmt19937 rng(static_cast<uint32_t>(static_cast<int32_t>(seed)));
  rng.fill(out, n);
  return FQ_OK;
}
