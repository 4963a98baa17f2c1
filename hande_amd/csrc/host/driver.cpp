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
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

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

extern "C" int fqh_fciqmc(const fqh_system* hs, const fqh_qmc_in* in, fqh_report_row* rows, int max_rows,
                          int* nrows) {
  if (!hs || !in || !rows || !nrows) return FQ_ERR_ARG;
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
  // initial_distribution: the reference determinant with nint(D0_pop)*real_factor.
  std::vector<int64_t> pop0(S, static_cast<int64_t>(std::llround(in->init_pop)) * rf);
  double dat0 = 0.0;
  rc = fq_load(e, sys->f0, pop0.data(), &dat0, 1);
  if (rc != FQ_OK) {
    fq_destroy(e);
    return rc;
  }
  double shift[FQ_MAX_SPACES] = {0, 0}, proj[FQ_MAX_SPACES] = {0, 0}, d0[FQ_MAX_SPACES] = {0, 0};
  double ntot_old[FQ_MAX_SPACES] = {0, 0}, nparticles[FQ_MAX_SPACES] = {0, 0};
  bool vary[FQ_MAX_SPACES] = {false, false};
  for (int s = 0; s < S; ++s) {
    shift[s] = in->initial_shift;
    nparticles[s] = std::fabs(static_cast<double>(pop0[s]) / static_cast<double>(rf));
    ntot_old[s] = nparticles[s];
    // initial_ci_projected_energy over the single reference determinant
    d0[s] = static_cast<double>(pop0[s]) / static_cast<double>(rf);
    proj[s] = 0.0;
  }
  int64_t nstates = 1;
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
      r.nparticles[s] = ntot_old[s];
    }
    r.nstates = nstates;
    r.nspawn_events = nspawn;
    r.rspawn = rspawn;
    r.time_per_cycle = tcyc;
  };
  emit(0, 0, 0.0, 0.0);
  bool overflowed = false;
  for (int ireport = 1; ireport <= in->nreports; ++ireport) {
    double eproj_old[FQ_MAX_SPACES] = {0, 0};
    for (int s = 0; s < S; ++s) eproj_old[s] = (std::fabs(d0[s]) < 2.2250738585072014e-308) ? 0.0 : proj[s] / d0[s];
    double rspawn = 0.0;
    for (int s = 0; s < S; ++s) proj[s] = d0[s] = 0.0;
    int64_t nspawn_events = 0;
    bool error = false;
    const auto t0 = std::chrono::steady_clock::now();
    for (int ic = 0; ic < in->mc_cycles; ++ic) {
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
    for (int s = 0; s < S; ++s) {
      if (vary[s]) {
        const double denom = 1.0 * in->tau * in->mc_cycles;
        shift[s] = shift[s] - std::log(ntot[s] / ntot_old[s]) * in->shift_damping / denom -
                   std::log(ntot[s] / in->target_population) * 0.0 / denom;
      }
    }
    for (int s = 0; s < S; ++s) ntot_old[s] = ntot[s];
    for (int s = 0; s < S; ++s) {
      if (!vary[s] && ntot[s] > in->target_population) {
        vary[s] = true;
        shift[s] = in->vary_shift_from_proje ? proj[s] / d0[s] : in->vary_shift_from;
      }
    }
    emit(static_cast<int64_t>(ireport) * in->mc_cycles, nspawn_events, rspawn, tcyc);
    if (error) {
      overflowed = true;
      break;
    }
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

// Close-open dSFMT-19937 draws as the reference's buffered generator hands
// them out (for checks of the replay stream against golden vectors).
extern "C" int fqh_dsfmt_fill(int64_t seed, double* out, int64_t n) {
  if (!out || n < 0) return FQ_ERR_ARG;
  fqmc::Dsfmt19937 rng(static_cast<uint32_t>(static_cast<int32_t>(seed)));
  rng.fill(out, n);
  return FQ_OK;
}
