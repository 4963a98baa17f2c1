// CPU ORACLE -- test infrastructure only.
//
// A serial restatement of the reference's FCIQMC evolve loop, used by tests/
// (and the CPU-baseline leg of bench.py) as the checker for the GPU engine.
// Nothing in the product links or calls this file.
//
// It follows, statement by statement and in the same floating-point order:
//   src/fciqmc.f90:276-464,635-769  report loop, MC cycle, spawning attempts
//   src/qmc_common.F90:379-406,950-1017,1240-1304  decide_nattempts,
//       init_mc_cycle, end_mc_cycle, spawning_rate
//   src/ifciqmc.f90:14-57  set_parent_flag
//   src/spawning.F90:215-315,680-766,907-1182  split/standard spawners,
//       set_child_sign, attempt_to_spawn, add_(flagged_)spawned_particle
//   src/stoch_utils.f90:51-136  stochastic rounding
//   src/death.f90:11-130  stochastic_death
//   src/spawn_data.F90:359-418,859-1101  qsort + (initiator) self-annihilation
//   src/sort.f90:213-396  qsort_int_64_list (exact comparison/swap sequence)
//   src/search.f90:47-135  binary_search
//   src/annihilation.f90:214-486,537-675,677-901  merge with the main list,
//       remove_unoccupied_dets, round_low_population_spawns, insert_new_walkers
//   src/excitations.F90:75-200,247-406  get_excitation, permutations, create_excited_det
//   src/excit_gen_hub_k.f90:29-159,469-770  Hubbard-k generators and pgen
//   src/hamiltonian_hub_k.f90:91-232, src/hubbard_k.f90  Hubbard-k matrix elements
//   src/excit_gen_mol.f90:16-101,384-946,1140-1327  molecular renorm generator
//   src/excit_gen_heat_bath_mol.F90:14-548, src/excit_gen_utils.f90:9-66,142-167,
//       lib/local/alias.f90:13-184  heat-bath tables, alias method, generator
//   src/hamiltonian_molecular.f90:73-346  molecular Slater--Condon rules
//   src/energy_evaluation.F90:425-711,795-986,1433-1454  estimators, shift
//   lib/dSFMT_F03_interface/dSFMT_interface.F90:387-409  RNG stream
// The dSFMT stream generator itself is hande_amd/csrc/host/dsfmt.hpp, pinned
// against the reference's C source in tests/golden.
//
// Production mode (orc_set_production) runs the same cycle with the GPU
// engine's counter-based draws instead of one dSFMT stream: every random
// decision reads its own Philox4x32-10 stream (Salmon, Moraes, Dror, Shaw,
// SC'11; restated here and pinned by the Random123 known-answer vectors),
// keyed by the seed and by (index, purpose, cycle):
//   attempt count, death   (det i, space s)      purposes 2, 3
//   spawning attempt       (global attempt t)    purpose 1 (det-major order)
//   main-list rounding     (list position i)     purpose 4
//   new-spawn rounding     (unique-key run r)    purpose 5
// and equal spawned keys keep their spawning order (stable sort).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "blocking.hpp"
#include "dsfmt.hpp"
#include "fqmc.h"
#include "fqmc_host.h"

namespace {

constexpr double kDepsilon = 1.e-12;

inline int popc(uint64_t x) { return __builtin_popcountll(x); }
inline int64_t fsign64(int64_t a, int64_t b) {  // Fortran SIGN(a,b) for integers
  int64_t m = a < 0 ? -a : a;
  return b >= 0 ? m : -m;
}
inline int fmod2(int x) { return x % 2; }        // Fortran MOD (truncating)
inline int fmodulo2(int x) { return ((x % 2) + 2) % 2; }  // Fortran MODULO

struct Excit {
  int nexcit = 0;
  int from[2] = {0, 0};
  int to[2] = {0, 0};
  bool perm = false;
};

struct Det {
  uint64_t f[FQ_MAX_WORDS];
};

struct Spawned {
  uint64_t f[FQ_MAX_WORDS];
  int64_t pop[FQ_MAX_SPACES];
  int64_t flag;
  int64_t run;  // index of its unique key among all keys of the sorted list
};

// Philox4x32-10: counter (c0, c1, c2, block), key (k0, k1); ten rounds of two
// 32x32->64 multiplications with the Weyl key schedule.
void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = uint64_t{0xD2511F53u} * x0;
    const uint64_t p1 = uint64_t{0xCD9E8D57u} * x2;
    const uint32_t y0 = static_cast<uint32_t>(p1 >> 32) ^ x1 ^ k0;
    const uint32_t y2 = static_cast<uint32_t>(p0 >> 32) ^ x3 ^ k1;
    x1 = static_cast<uint32_t>(p1);
    x3 = static_cast<uint32_t>(p0);
    x0 = y0;
    x2 = y2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = x0;
  out[1] = x1;
  out[2] = x2;
  out[3] = x3;
}

// One keyed stream: blocks 0, 1, ... of the counter, two 32-bit words per
// draw, each draw (w0 >> 6) * 2^26 + (w1 >> 6) times 2^-52.
struct PhiloxStream {
  uint32_t key[2] = {0, 0}, ctr[4] = {0, 0, 0, 0}, buf[4] = {0, 0, 0, 0};
  int left = 0;
  void reset(uint64_t k, uint32_t a, uint32_t b, uint32_t c) {
    key[0] = static_cast<uint32_t>(k);
    key[1] = static_cast<uint32_t>(k >> 32);
    ctr[0] = a;
    ctr[1] = b;
    ctr[2] = c;
    ctr[3] = 0;
    left = 0;
  }
  double next() {
    if (left < 2) {
      philox4x32_10(ctr, key, buf);
      ++ctr[3];
      left = 4;
    }
    const uint32_t w0 = buf[4 - left], w1 = buf[5 - left];
    left -= 2;
    const uint64_t v = (static_cast<uint64_t>(w0 >> 6) << 26) | (w1 >> 6);
    return static_cast<double>(v) * (1.0 / 4503599627370496.0);
  }
};
enum : uint32_t {
  kDrawAttempt = 1,
  kDrawNattempts = 2,
  kDrawDeath = 3,
  kDrawRoundMain = 4,
  kDrawRoundSpawn = 5,
  kDrawDeterm = 8  // rounding of the deterministic projection's spawns
};

class Oracle {
 public:
  Oracle(const fq_system* s, const fq_params* p, uint32_t seed) : rng_(seed) {
    sys_ = *s;
    par_ = *p;
    W = s->nwords;
    S = p->nspaces;
    nbasis = s->nbasis;
    nel = s->nel;
    rf = int64_t{1} << p->pop_bit_shift;
    f0.assign(s->f0, s->f0 + W);
    sp_eigv.assign(s->sp_eigv, s->sp_eigv + nbasis);
    if (s->system == FQ_SYS_HUB_K) {
      sym_table.assign(s->sym_table, s->sym_table + s->nsym * s->nsym);
      inv_sym.assign(s->inv_sym, s->inv_sym + s->nsym);
    } else {
      nsym = s->nsym_pg;
      orb_sym.assign(s->orb_sym, s->orb_sym + nbasis);
      nbss.assign(s->nbasis_sym_spin, s->nbasis_sym_spin + 2 * nsym);
      ssbf.assign(s->sym_spin_basis_fns, s->sym_spin_basis_fns + 2 * nsym * s->max_sym_spin);
      h1.assign(s->h1, s->h1 + s->norb * s->norb);
      eri.assign(s->eri, s->eri + s->n_eri);
    }
    // excit_mask(:, orb) = bits of all orbitals after orb (src/excitations.F90:17-45).
    excit_mask.assign(static_cast<size_t>(nbasis + 1) * W, 0);
    for (int i = 1; i <= nbasis; ++i)
      for (int j = i + 1; j <= nbasis; ++j) excit_mask[i * W + (j - 1) / 64] |= uint64_t{1} << ((j - 1) % 64);
    if (s->system == FQ_SYS_READ_IN && p->excit_gen == FQ_EXCIT_HEAT_BATH) build_heat_bath();
  }

  // ---- state ----------------------------------------------------------------
  std::vector<Det> states;
  std::vector<int64_t> pops;  // S per det
  std::vector<double> dat;    // K_ii
  double nparticles[FQ_MAX_SPACES] = {0, 0};
  bool main_error = false, spawn_error = false;
  // Order of equal keys before self-annihilation: the reference's quicksort
  // (default) or a stable sort (the first entry of a run is then the earliest
  // spawned, the rule the GPU's production mode uses).
  bool stable_order = false;

  void load(const uint64_t* st, const int64_t* p, const double* d, int64_t n) {
    states.resize(n);
    pops.assign(p, p + n * S);
    dat.assign(d, d + n);
    for (int64_t i = 0; i < n; ++i)
      for (int w = 0; w < W; ++w) states[i].f[w] = st[i * W + w];
    for (int s = 0; s < S; ++s) {
      // src/qmc.F90:864 nparticles = sum(abs(real(pops)/real_factor))
      double acc = 0.0;
      for (int64_t i = 0; i < n; ++i) acc = acc + std::fabs(static_cast<double>(pops[i * S + s]) / rf);
      nparticles[s] = acc;
    }
    determ_ = DetermSpace{};  // a new list has no semi-stochastic core
  }

  uint64_t rng_consumed() const { return rng_.consumed(); }
  // the core's determinants (list order) and its stored Hamiltonian's size
  int64_t determ_info(uint64_t* st, int64_t* nnz) const {
    const int64_t D = determ_size();
    if (nnz) *nnz = D ? static_cast<int64_t>(determ_.mat.size()) : 0;
    if (st)
      for (int64_t k = 0; k < D; ++k)
        for (int w = 0; w < W; ++w) st[k * W + w] = determ_.dets[k].f[w];
    return D;
  }

  void download(uint64_t* st, int64_t* p, double* d) const {
    for (size_t i = 0; i < states.size(); ++i) {
      for (int w = 0; w < W; ++w) st[i * W + w] = states[i].f[w];
      for (int s = 0; s < S; ++s) p[i * S + s] = pops[i * S + s];
      d[i] = dat[i];
    }
  }

  // ---- estimators over the current list (initial_ci_projected_energy) ------
  void projected_energy(double* proj, double* d0) {
    for (int s = 0; s < S; ++s) proj[s] = d0[s] = 0.0;
    for (size_t i = 0; i < states.size(); ++i) {
      std::vector<int> occ = decode_occ(states[i].f);
      for (int s = 0; s < S; ++s) {
        double pop = static_cast<double>(pops[i * S + s]) / rf;
        Excit c = get_excitation(states[i].f, f0.data());
        update_proj_energy(states[i].f, occ, pop, c, proj[s], d0[s]);
      }
    }
  }

  // ---- one MC cycle (src/fciqmc.f90:307-399) ---------------------------------
  void cycle(const double* shift, const double* eproj_old, double tau, fq_cycle_out* out) {
    cycle_spawn(shift, eproj_old, tau, out);
    cycle_finish(out);
  }

  // The cycle up to the spawned list (estimators, spawning, death): the list
  // is then in `spawned` in creation order, ready to be exchanged between
  // emulated ranks before cycle_finish annihilates it.
  void cycle_spawn(const double* shift, const double* eproj_old, double tau, fq_cycle_out* out) {
    std::memset(out, 0, sizeof(*out));
    rng0_ = rng_.consumed();
    spawned.clear();
    int64_t ndeath = 0;
    int64_t attempts = 0;
    int64_t t = 0;  // production: global attempt index, det-major
    std::vector<double> vec;     // semi-stochastic: the core determinants' populations (determ%vector)
    std::vector<int64_t> dpos;   // and their list positions (determ%indices)
    for (size_t idet = 0; idet < states.size(); ++idet) {
      Decoded dd = decode_full(states[idet].f);
      double real_pop[FQ_MAX_SPACES];
      for (int s = 0; s < S; ++s) real_pop[s] = static_cast<double>(pops[idet * S + s]) / rf;
      // set_determ_info / set_parent_flag: a core determinant is an initiator
      const bool dparent = is_determ(states[idet].f);
      if (dparent) {
        for (int s = 0; s < S; ++s) vec.push_back(real_pop[s]);
        dpos.push_back(static_cast<int64_t>(idet));
      }
      int64_t flag = dparent ? 0 : parent_flag(real_pop);
      for (int s = 0; s < S; ++s) {
        Excit c = get_excitation(states[idet].f, f0.data());
        update_proj_energy(states[idet].f, dd.occ, real_pop[s], c, out->proj_energy[s], out->D0_population[s]);
        const int64_t i64 = static_cast<int64_t>(idet);
        const uint32_t hi = static_cast<uint32_t>(i64 >> 32) ^ (static_cast<uint32_t>(s) << 28);
        draw_stream(static_cast<uint32_t>(i64), hi, kDrawNattempts);
        int nat = decide_nattempts(real_pop[s]);
        attempts += nat;
        const int64_t parent = pops[idet * S + s];
        for (int ip = 0; ip < nat; ++ip, ++t) {
          draw_stream(static_cast<uint32_t>(t), static_cast<uint32_t>(t >> 32), kDrawAttempt);
          Excit conn;
          int64_t nspawn = spawn(dd, states[idet].f, parent, tau, conn);
          if (nspawn != 0 && dparent) {  // core-to-core spawns are the deterministic projection's
            uint64_t child[FQ_MAX_WORDS] = {0, 0};
            excited_det(states[idet].f, conn, child);
            if (is_determ(child)) nspawn = 0;
          }
          if (nspawn != 0) create_spawned(states[idet].f, conn, nspawn, s, flag);
        }
        if (!dparent) {  // a core determinant's death is part of the deterministic projection
          draw_stream(static_cast<uint32_t>(i64), hi, kDrawDeath);
          death(tau, dat[idet], shift[s], eproj_old[s], pops[idet * S + s], nparticles[s], ndeath);
        }
      }
    }
    if (determ_.on) {
      if (determ_.separate)
        determ_separate(vec, dpos, shift, eproj_old, tau);
      else
        determ_projection(vec, shift, eproj_old, tau);
    }
    out->nspawn_events = static_cast<int64_t>(spawned.size());
    out->ndeath = ndeath;
    out->nattempts = attempts;
  }

  // init_semi_stoch_t with space "high" (src/semi_stoch.F90:134-377,
  // create_high_pop_space :1198-1316, find_most_populated_dets :1318-1385,
  // create_determ_hamil :533-685, one rank): the target_size most populated
  // determinants of the current list (sum over spaces of |N|, kept by the
  // reference's replace-the-minimum scan), sorted as the main list, and the
  // Hamiltonian among them as a CSR matrix, H_ij = <D_i|H|D_j> (H_ii - H_00 on
  // the diagonal, elements of magnitude <= 1e-12 dropped).  For separate
  // annihilation the stored matrix is its transpose, T_ij = <D_j|H|D_i>: row i
  // then lists the entries the reference's transposed product
  // (csrpgemv, lib/local/csr.f90:148-204) adds into element i, in the order it
  // adds them.  Every core determinant is in the list, so
  // add_determ_dets_to_psip_list inserts nothing.
  void semi_stoch_init(int target_size, bool separate) {
    const int64_t n = static_cast<int64_t>(states.size());
    const int64_t nd = std::min<int64_t>(target_size, n);
    if (nd <= 0) {  // an empty core: off
      determ_ = DetermSpace{};
      return;
    }
    auto absum = [&](int64_t i) {
      int64_t t = 0;
      for (int s = 0; s < S; ++s) t += std::llabs(pops[i * S + s]);
      return t;
    };
    std::vector<int64_t> idx(nd), pop(nd);
    int64_t min_pop = 0, min_ind = 0;
    for (int64_t i = 0; i < nd; ++i) {
      idx[i] = i;
      pop[i] = absum(i);
      if (i == 0 || pop[i] < min_pop) {
        min_pop = pop[i];
        min_ind = i;
      }
    }
    for (int64_t i = nd; i < n; ++i) {
      const int64_t a = absum(i);
      if (a > min_pop) {
        idx[min_ind] = i;
        pop[min_ind] = a;
        min_pop = pop[0];
        min_ind = 0;
        for (int64_t j = 1; j < nd; ++j)
          if (pop[j] < min_pop) {
            min_pop = pop[j];
            min_ind = j;
          }
      }
    }
    determ_ = DetermSpace{};
    for (int64_t i = 0; i < nd; ++i) determ_.dets.push_back(states[idx[i]]);
    std::sort(determ_.dets.begin(), determ_.dets.end(),
              [&](const Det& a, const Det& b) { return cmp(a.f, b.f) == 1; });
    const int64_t D = nd;
    determ_.row_ptr.assign(D + 1, 0);
    for (int64_t i = 0; i < D; ++i) {
      determ_.row_ptr[i] = static_cast<int64_t>(determ_.mat.size());
      for (int64_t j = 0; j < D; ++j) {
        double h = separate ? hmatel(determ_.dets[j].f, determ_.dets[i].f) : hmatel(determ_.dets[i].f, determ_.dets[j].f);
        if (i == j) h = h - sys_.H00;
        h = 1.0 * h;  // the quasi-Newton weight, 1 here
        if (std::fabs(h) > 1.0e-12) {
          determ_.mat.push_back(h);
          determ_.col.push_back(j);
        }
      }
    }
    determ_.row_ptr[D] = static_cast<int64_t>(determ_.mat.size());
    determ_.on = true;
    determ_.separate = separate;
  }
  int64_t determ_size() const { return determ_.on ? static_cast<int64_t>(determ_.dets.size()) : 0; }

  // direct_annihilation of `spawned` (as received) and the cycle's totals.
  void cycle_finish(fq_cycle_out* out) {
    annihilate();
    for (int s = 0; s < S; ++s) out->nparticles[s] = nparticles[s];
    out->nstates = static_cast<int64_t>(states.size());
    out->rng_consumed = static_cast<int64_t>(rng_.consumed() - rng0_);
    out->spawn_overflow = spawn_error;
    out->main_overflow = main_error;
    ++cyc_;
  }
  std::vector<Spawned>& spawned_list() { return spawned; }

  // Annihilate an externally supplied spawned list (entries in the given order).
  void annihilate_external(const uint64_t* sst, const int64_t* sp, const int64_t* sf, int64_t ns,
                           fq_cycle_out* out) {
    std::memset(out, 0, sizeof(*out));
    uint64_t rng0 = rng_.consumed();
    spawned.resize(ns);
    for (int64_t i = 0; i < ns; ++i) {
      std::memset(&spawned[i], 0, sizeof(Spawned));
      for (int w = 0; w < W; ++w) spawned[i].f[w] = sst[i * W + w];
      for (int s = 0; s < S; ++s) spawned[i].pop[s] = sp[i * S + s];
      spawned[i].flag = sf ? sf[i] : 0;
    }
    out->nspawn_events = ns;
    annihilate();
    for (int s = 0; s < S; ++s) out->nparticles[s] = nparticles[s];
    out->nstates = static_cast<int64_t>(states.size());
    out->rng_consumed = static_cast<int64_t>(rng_.consumed() - rng0);
  }

 private:
  fq_system sys_;
  fq_params par_;
  int W, S, nbasis, nel, nsym = 1;
  int64_t rf;
  fqmc::Dsfmt19937 rng_;
  std::vector<uint64_t> f0;
  std::vector<double> sp_eigv;
  std::vector<int> sym_table, inv_sym, orb_sym, nbss, ssbf;
  std::vector<double> h1, eri;
  std::vector<uint64_t> excit_mask;
  std::vector<Spawned> spawned;

  // the semi-stochastic core space (semi_stoch_t)
  struct DetermSpace {
    bool on = false;
    bool separate = false;       // separate (true) or combined annihilation
    std::vector<Det> dets;        // sorted as the main list
    std::vector<int64_t> row_ptr;  // CSR rows in that order
    std::vector<int64_t> col;
    std::vector<double> mat;
  } determ_;

  bool is_determ(const uint64_t* f) const {  // check_if_determ
    if (!determ_.on) return false;
    size_t lo = 0, hi = determ_.dets.size();
    while (lo < hi) {
      const size_t mid = (lo + hi) / 2;
      const int c = cmp(determ_.dets[mid].f, f);
      if (c == 0) return true;
      if (c == 1)
        lo = mid + 1;
      else
        hi = mid;
    }
    return false;
  }

  // get_hmatel for two determinants (src/hamiltonian_molecular.f90:12-71 and
  // the Hubbard-k rules): the Slater--Condon element by excitation level.
  double hmatel(const uint64_t* f1, const uint64_t* f2) {
    Excit e = get_excitation(f1, f2);
    if (e.nexcit == 0) return sc0(f1);
    if (sys_.system == FQ_SYS_HUB_K) {
      if (e.nexcit != 2) return 0.0;
      double h = two_e_hub_k(e.from[0], e.from[1], e.to[0], e.to[1]);
      return e.perm ? -h : h;
    }
    if (e.nexcit == 1) return sc1_excit(decode_occ(f1), e.from[0], e.to[0], e.perm);
    if (e.nexcit == 2) return sc2_excit(e.from[0], e.from[1], e.to[0], e.to[1], e.perm);
    return 0.0;
  }

  // determ_proj_combined_annihil + create_spawned_particle_determ
  // (src/semi_stoch.F90:897-975,1096-1146): for each core determinant (CSR
  // row) and space, -tau (sum_j H_ij N_j + (E_old*0 - S) N_i) (the quasi-Newton
  // weight difference is 0 without quasi-Newton), encoded with one stochastic
  // rounding draw, appended to the spawned list (a zero included) as from an
  // initiator.
  void determ_projection(const std::vector<double>& vec, const double* shift, const double* eproj_old, double tau) {
    const int64_t D = static_cast<int64_t>(determ_.dets.size());
    for (int64_t row = 0; row < D; ++row) {
      for (int s = 0; s < S; ++s) {
        double out = 0.0;
        for (int64_t k = determ_.row_ptr[row]; k < determ_.row_ptr[row + 1]; ++k)
          out = out + determ_.mat[k] * vec[determ_.col[k] * S + s];
        const double rho = 1.0 - 1.0;  // quasi_newton_pop_control - the row's weight
        out = -tau * (out + (eproj_old[s] * rho - 1.0 * shift[s]) * vec[row * S + s]);
        draw_stream(static_cast<uint32_t>(row), static_cast<uint32_t>(s) << 28, kDrawDeterm);
        const double scaled = out * static_cast<double>(rf);
        const double sgn = std::copysign(1.0, out);
        int64_t nspawn = static_cast<int64_t>(std::fabs(scaled));
        if (std::fabs(scaled) - static_cast<double>(nspawn) > rand()) nspawn = nspawn + 1;
        nspawn = nspawn * static_cast<int64_t>(std::llround(sgn));
        if (static_cast<int64_t>(spawned.size()) + 1 > par_.spawn_capacity) {
          spawn_error = true;
          continue;
        }
        Spawned sp;
        std::memset(&sp, 0, sizeof(sp));
        for (int w = 0; w < W; ++w) sp.f[w] = determ_.dets[row].f[w];
        sp.pop[s] = nspawn;
        sp.flag = 0;
        spawned.push_back(sp);
      }
    }
  }
  // determ_proj_separate_annihil + deterministic_annihilation
  // (src/semi_stoch.F90:1009-1094, src/annihilation.f90:488-535): per core
  // determinant and space, y_i = (-tau (E_old*0 - S)) v_i, then
  // y_i += (-tau T_ik) v_k over row i of the stored transpose (the
  // reference's y(icol) += a*m*x(irow) loop, in its order); once every y is
  // formed, each is encoded with one stochastic rounding draw per space (row
  // by row, spaces inner) and added to the core determinant's population in
  // the main list (these are not spawning events).
  void determ_separate(const std::vector<double>& vec, const std::vector<int64_t>& dpos, const double* shift,
                       const double* eproj_old, double tau) {
    const int64_t D = static_cast<int64_t>(determ_.dets.size());
    std::vector<double> y(static_cast<size_t>(D) * S);
    const double rho = 1.0 - 1.0;  // quasi_newton_pop_control - the row's weight
    const double a = -1.0 * tau;
    for (int s = 0; s < S; ++s) {
      const double c = -(tau * (eproj_old[s] * rho - shift[s] * 1.0));
      for (int64_t i = 0; i < D; ++i) {
        double v = c * vec[i * S + s];
        for (int64_t k = determ_.row_ptr[i]; k < determ_.row_ptr[i + 1]; ++k)
          v = v + a * determ_.mat[k] * vec[determ_.col[k] * S + s];
        y[i * S + s] = v;
      }
    }
    for (int64_t i = 0; i < D; ++i) {
      int64_t* p = &pops[dpos[i] * S];
      for (int s = 0; s < S; ++s) {
        draw_stream(static_cast<uint32_t>(i), static_cast<uint32_t>(s) << 28, kDrawDeterm);
        double scaled = y[i * S + s] * static_cast<double>(rf);
        const int64_t sgn = scaled < 0.0 ? -1 : 1;
        scaled = std::fabs(scaled);
        int64_t nspawn = static_cast<int64_t>(scaled);
        if (scaled - static_cast<double>(nspawn) > rand()) nspawn = nspawn + 1;
        const int64_t old = p[s];
        p[s] = p[s] + sgn * nspawn;
        nparticles[s] = nparticles[s] + static_cast<double>(std::llabs(p[s]) - std::llabs(old)) / rf;
      }
    }
  }



 public:
  // Draws come from the dSFMT stream, or from a caller-supplied list (the
  // generator probe used by the tests).
  const double* forced_ = nullptr;
  int64_t nforced_ = 0, iforced_ = 0;
  double rand() {
    if (forced_) return iforced_ < nforced_ ? forced_[iforced_++] : (++iforced_, 0.5);
    if (prod_) return ph_.next();
    return rng_.next();
  }
  // Production mode: counter-based Philox streams (see the file header).
  bool prod_ = false;
  uint64_t pseed_ = 0;
  uint32_t cyc_ = 0;  // cycles run so far
  uint64_t rng0_ = 0;  // draws consumed before the current cycle
  PhiloxStream ph_;
  void set_production(uint64_t seed) {
    prod_ = true;
    pseed_ = seed;
    stable_order = true;
  }
  void draw_stream(uint32_t a, uint32_t b, uint32_t purpose) {
    if (prod_) ph_.reset(pseed_, a, b, cyc_ ^ (purpose << 24));
  }
  bool btest(const uint64_t* f, int orb) const { return (f[(orb - 1) / 64] >> ((orb - 1) % 64)) & 1; }
  int ms_of(int orb) const { return (orb & 1) ? 1 : -1; }
  int ims_of(int orb) const { return (ms_of(orb) + 3) / 2; }  // 1 = beta, 2 = alpha

  struct Decoded {
    std::vector<int> occ, occ_a, occ_b, unocc_a, unocc_b;
    std::vector<int> symunocc;  // [(ims-1)*nsym + sym]
  };

  std::vector<int> decode_occ(const uint64_t* f) const {
    std::vector<int> occ;
    for (int i = 1; i <= nbasis; ++i)
      if (btest(f, i)) occ.push_back(i);
    return occ;
  }

  Decoded decode_full(const uint64_t* f) const {
    Decoded d;
    d.occ = decode_occ(f);
    if (sys_.system == FQ_SYS_HUB_K) {
      for (int i = 1; i <= nbasis; ++i) {
        bool o = btest(f, i);
        if (i & 1)
          (o ? d.occ_a : d.unocc_a).push_back(i);
        else
          (o ? d.occ_b : d.unocc_b).push_back(i);
      }
    } else {
      d.symunocc = nbss;
      for (int o : d.occ) d.symunocc[(ims_of(o) - 1) * nsym + orb_sym[o - 1]] -= 1;
    }
    return d;
  }
  int symunocc(const Decoded& d, int ims, int sym) const { return d.symunocc[(ims - 1) * nsym + sym]; }

  // ---- excitations ---------------------------------------------------------
  Excit get_excitation(const uint64_t* f1, const uint64_t* f2) const {
    Excit e;
    bool differ = false;
    for (int w = 0; w < W; ++w) differ = differ || f1[w] != f2[w];
    if (!differ) return e;
    int nx = 0;
    for (int w = 0; w < W; ++w) nx += popc(f1[w] ^ f2[w]);
    e.nexcit = nx / 2;
    const int shift = nel - e.nexcit;
    if (e.nexcit <= 2) {
      int ie1 = 0, ie2 = 0, iel1 = 0, iel2 = 0, perm = 0;
      for (int w = 0; w < W; ++w) {
        if (f1[w] == f2[w]) {
          if (fmodulo2(ie1 - ie2) == 1) {
            int n = popc(f1[w]);
            iel1 += n;
            iel2 += n;
          }
          continue;
        }
        for (int j = 0; j < 64; ++j) {
          bool t1 = (f1[w] >> j) & 1, t2 = (f2[w] >> j) & 1;
          const int orb = w * 64 + j + 1;
          if (t2) ++iel2;
          if (t1) {
            ++iel1;
            if (!t2) {
              ++ie1;
              e.from[ie1 - 1] = orb;
              perm += shift - iel1 + ie1;
            }
          } else if (t2) {
            ++ie2;
            e.to[ie2 - 1] = orb;
            perm += shift - iel2 + ie2;
          }
        }
      }
      e.perm = fmod2(perm) == 1;
    }
    return e;
  }
  void perm1(const uint64_t* f, Excit& e) const {
    int perm = 0;
    for (int w = 0; w < W; ++w)
      perm += popc(f[w] & (excit_mask[e.from[0] * W + w] ^ excit_mask[e.to[0] * W + w]));
    if (e.from[0] > e.to[0]) perm -= 1;
    e.perm = fmod2(perm) == 1;
  }
  void perm2(const uint64_t* f, Excit& e) const {
    int perm = 0;
    for (int w = 0; w < W; ++w) {
      uint64_t ia = excit_mask[e.from[0] * W + w] ^ excit_mask[e.to[0] * W + w];
      uint64_t jb = excit_mask[e.from[1] * W + w] ^ excit_mask[e.to[1] * W + w];
      perm += popc((f[w] & ia) ^ (f[w] & jb));
    }
    if (e.from[0] > e.to[0]) perm += 1;
    if (e.from[0] > e.to[1]) perm += 1;
    if (e.from[1] > e.to[1] || e.from[1] < e.to[0]) perm += 1;
    e.perm = fmod2(perm) == 1;
  }
  void excited_det(const uint64_t* f, const Excit& e, uint64_t* out) const {
    for (int w = 0; w < W; ++w) out[w] = f[w];
    for (int i = 0; i < e.nexcit; ++i) {
      int o = e.from[i];
      out[(o - 1) / 64] &= ~(uint64_t{1} << ((o - 1) % 64));
      o = e.to[i];
      out[(o - 1) / 64] |= uint64_t{1} << ((o - 1) % 64);
    }
  }

  // ---- Hubbard-k -------------------------------------------------------------
  int st(int a, int b) const { return sym_table[(a - 1) * sys_.nsym + (b - 1)]; }
  double two_e_hub_k(int p1, int p2, int p3, int p4) const {
    bool s12 = ms_of(p1) == ms_of(p2), s34 = ms_of(p3) == ms_of(p4);
    int dk = st((p1 + 1) / 2, (p2 + 1) / 2);
    dk = st(dk, inv_sym[(p3 + 1) / 2 - 1]);
    bool conserved = dk == (p4 + 1) / 2;
    if (!s12 && (s12 == s34) && conserved) return ms_of(p1) == ms_of(p3) ? sys_.coulomb_k : -sys_.coulomb_k;
    return 0.0;
  }

  // ---- molecular integrals ---------------------------------------------------
  static int64_t tri(int64_t i, int64_t j) { return (i * (i - 1)) / 2 + j; }
  int sp(int i) const { return (i + 1) / 2; }
  double h1g(int i, int j) const { return h1[(sp(i) - 1) * sys_.norb + (sp(j) - 1)]; }
  double eri_nz(int i, int j, int a, int b) const {
    int ii = std::max(i, a), aa = std::min(i, a), jj = std::max(j, b), bb = std::min(j, b);
    int64_t ia = tri(sp(ii), sp(aa)), jb = tri(sp(jj), sp(bb));
    return eri[(ia < jb ? tri(jb, ia) : tri(ia, jb)) - 1];
  }
  double sc1_excit(const std::vector<int>& occ, int i, int a, bool perm) const {
    double h = h1g(i, a);
    for (int o : occ) {
      if (o != i) {
        h = h + eri_nz(i, o, a, o);
        if (ms_of(o) == ms_of(i)) h = h - eri_nz(i, o, o, a);
      }
    }
    if (perm) h = -h;
    return h;
  }
  double sc2_excit(int i, int j, int a, int b, bool perm) const {
    double h = 0.0;
    if (ms_of(i) == ms_of(a)) h = eri_nz(i, j, a, b);
    if (ms_of(i) == ms_of(b)) h = h - eri_nz(i, j, b, a);
    if (perm) h = -h;
    return h;
  }
  double sc0(const uint64_t* f) const {
    std::vector<int> occ = decode_occ(f);
    if (sys_.system == FQ_SYS_HUB_K) {
      double h = static_cast<double>(sys_.nalpha * sys_.nbeta) * sys_.coulomb_k;
      for (int o : occ) h = h + sp_eigv[o - 1];
      return h;
    }
    double h = sys_.ecore;
    for (size_t ie = 0; ie < occ.size(); ++ie) {
      int i = occ[ie];
      h = h + h1g(i, i);
      for (size_t je = ie + 1; je < occ.size(); ++je) {
        int j = occ[je];
        h = h + eri_nz(i, j, i, j);
        if (ms_of(i) == ms_of(j)) h = h - eri_nz(i, j, j, i);
      }
    }
    return h;
  }

  void update_proj_energy(const uint64_t* f, const std::vector<int>& occ, double pop, Excit& c, double& proj,
                          double& d0) const {
    if (c.nexcit == 0) {
      d0 = d0 + pop;
      return;
    }
    if (sys_.system == FQ_SYS_HUB_K) {
      if (c.nexcit == 2) {
        double h = two_e_hub_k(c.from[0], c.from[1], c.to[0], c.to[1]);
        if (c.perm) h = -h;
        proj = proj + h * pop;
      }
      return;
    }
    if (c.nexcit == 1) {
      if (ms_of(c.from[0]) == ms_of(c.to[0]) && orb_sym[c.from[0] - 1] == orb_sym[c.to[0] - 1]) {
        double h = sc1_excit(occ, c.from[0], c.to[0], c.perm);
        proj = proj + h * pop;
      }
    } else if (c.nexcit == 2) {
      if (ms_of(c.from[0]) + ms_of(c.from[1]) == ms_of(c.to[0]) + ms_of(c.to[1])) {
        int ij = orb_sym[c.from[0] - 1] ^ orb_sym[c.from[1] - 1];
        int ab = orb_sym[c.to[0] - 1] ^ orb_sym[c.to[1] - 1];
        if (ij == ab) {
          double h = sc2_excit(c.from[0], c.from[1], c.to[0], c.to[1], c.perm);
          proj = proj + h * pop;
        }
      }
    }
  }

  int64_t parent_flag(const double* pop) const {
    int64_t flag = 0;
    if (!par_.initiator) return 0;
    const double ip = par_.initiator_pop;
    if (par_.quadrature_initiator && S % 2 == 0) {
      for (int i = 0; i < S; i += 2) {
        double mag = pop[i] * pop[i] + pop[i + 1] * pop[i + 1];
        if (!(mag > ip * ip)) flag |= (int64_t{1} << i) | (int64_t{1} << (i + 1));
      }
    } else {
      for (int i = 0; i < S; ++i)
        if (!(std::fabs(pop[i]) > ip)) flag |= int64_t{1} << i;
    }
    return flag;
  }

  int decide_nattempts(double population) {
    int n = std::abs(static_cast<int>(population));
    double pextra = std::fabs(population) - n;
    if (std::fabs(pextra) > kDepsilon) {
      if (pextra > rand()) n = n + 1;
    }
    return n;
  }

  int64_t stoch_round_spawn(double pspawn) {
    const int64_t cutoff = par_.spawn_cutoff;
    int64_t nspawn;
    if (pspawn < static_cast<double>(cutoff)) {
      nspawn = (pspawn > rand() * static_cast<double>(cutoff)) ? cutoff : 0;
    } else {
      nspawn = static_cast<int64_t>(pspawn);
      double padd = pspawn - static_cast<double>(nspawn);
      if (padd > rand()) nspawn = nspawn + 1;
    }
    return nspawn;
  }

  // ---- spawners --------------------------------------------------------------
  int64_t spawn(const Decoded& d, const uint64_t* f, int64_t parent, double tau, Excit& conn) {
    if (sys_.system == FQ_SYS_HUB_K) return spawn_hub_k(d, f, parent, tau, conn);
    if (par_.excit_gen == FQ_EXCIT_HEAT_BATH) return spawn_heat_bath(d, f, parent, tau, conn);
    return spawn_mol(d, f, parent, tau, conn);
  }

  void choose_ij_hub_k(const Decoded& d, int& i, int& j, int& ij_sym) {
    double r = rand();
    const int na = sys_.nalpha, nb = sys_.nbeta;
    int ind = static_cast<int>(r * na * nb) + 1;
    i = static_cast<int>((ind - 1.0) / nb) + 1;
    j = ind - (i - 1) * nb;
    i = d.occ_a[i - 1];
    j = d.occ_b[j - 1];
    ij_sym = st((i + 1) / 2, (j + 1) / 2);
  }
  bool find_ab_hub_k(const Decoded& d, const uint64_t* f, int ij_sym, int& a, int& b) {
    int r = static_cast<int>(rand() * sys_.nvirt_alpha) + 1;
    a = d.unocc_a[r - 1];
    int ka = (a + 1) / 2;
    b = 2 * st(ij_sym, inv_sym[ka - 1]);
    return !btest(f, b);
  }
  double pgen_hub_k(const Decoded& d, const uint64_t* f, int ij_sym) const {
    int forbidden = 0;
    for (int a = 0; a < sys_.nvirt_alpha; ++a) {
      int ka = (d.unocc_a[a] + 1) / 2;
      int b = 2 * st(ij_sym, inv_sym[ka - 1]);
      if (btest(f, b)) ++forbidden;
    }
    return 1.0 / static_cast<double>(sys_.nalpha * sys_.nbeta * (sys_.nvirt_alpha - forbidden));
  }
  double sc2_hub_k_excit(const uint64_t* f, Excit& c) const {
    if (c.from[0] > c.from[1]) std::swap(c.from[0], c.from[1]);
    if (c.to[0] > c.to[1]) std::swap(c.to[0], c.to[1]);
    double h = sys_.coulomb_k;
    perm2(f, c);
    if (c.perm) h = -h;
    if (fmod2(c.from[0] - c.to[0]) != 0) h = -h;
    return h;
  }

  int64_t spawn_hub_k(const Decoded& d, const uint64_t* f, int64_t parent, double tau, Excit& conn) {
    int i, j, ij_sym;
    double pgen, absh;
    conn = Excit();
    choose_ij_hub_k(d, i, j, ij_sym);
    conn.from[0] = i;
    conn.from[1] = j;
    if (par_.excit_gen == FQ_EXCIT_NO_RENORM) {
      int a, b;
      bool allowed = find_ab_hub_k(d, f, ij_sym, a, b);
      conn.to[0] = a;
      conn.to[1] = b;
      conn.nexcit = 2;
      if (allowed) {
        pgen = 1.0 / static_cast<double>(sys_.nalpha * sys_.nbeta * sys_.nvirt_alpha);
        absh = std::fabs(sys_.coulomb_k);
      } else {
        pgen = 1.0;
        absh = 0.0;
      }
    } else {
      pgen = pgen_hub_k(d, f, ij_sym);
      absh = std::fabs(sys_.coulomb_k);
    }
    int64_t nspawn = stoch_round_spawn(static_cast<double>(rf) * tau * absh / pgen);
    if (nspawn != 0) {
      double h;
      if (par_.excit_gen == FQ_EXCIT_NO_RENORM) {
        h = sc2_hub_k_excit(f, conn);
      } else {
        int sym2 = st((conn.from[0] + 1) / 2, (conn.from[1] + 1) / 2);
        conn.nexcit = 2;
        int a, b;
        while (!find_ab_hub_k(d, f, sym2, a, b)) {
        }
        conn.to[0] = a;
        conn.to[1] = b;
        h = sc2_hub_k_excit(f, conn);
      }
      nspawn = (h > 0.0) ? -fsign64(nspawn, parent) : fsign64(nspawn, parent);
    }
    return nspawn;
  }

  // Molecular renormalised generator (src/excit_gen_mol.f90:16-101).
  int64_t spawn_mol(const Decoded& d, const uint64_t* f, int64_t parent, double tau, Excit& conn) {
    double pgen = 1.0, h = 0.0;
    bool allowed = false;
    conn = Excit();
    if (rand() < sys_.pattempt_single) {
      int i = 0, a = 0;
      allowed = choose_ia(d, f, i, a);
      conn.nexcit = 1;
      conn.from[0] = i;
      conn.to[0] = a;
      if (allowed) {
        pgen = sys_.pattempt_single * pgen_single(d, a);
        perm1(f, conn);
        h = sc1_excit(d.occ, i, a, conn.perm);
      } else {
        h = 0.0;
        pgen = 1.0;
      }
    } else {
      int i, j, ij_sym, ij_spin;
      double pgen_ij;
      choose_ij_mol(d, i, j, ij_sym, ij_spin, pgen_ij);
      conn.from[0] = i;
      conn.from[1] = j;
      int a = 0, b = 0;
      allowed = choose_ab_mol(d, f, ij_sym, ij_spin, a, b);
      conn.to[0] = a;
      conn.to[1] = b;
      conn.nexcit = 2;
      if (allowed) {
        pgen = sys_.pattempt_double * pgen_ij * pgen_double(d, ij_sym, a, b, ij_spin);
        perm2(f, conn);
        h = sc2_excit(conn.from[0], conn.from[1], conn.to[0], conn.to[1], conn.perm);
      } else {
        h = 0.0;
        pgen = 1.0;
      }
    }
    // attempt_to_spawn (src/spawning.F90:711-766)
    double pspawn = tau * std::fabs(h) / pgen;
    pspawn = pspawn * static_cast<double>(rf);
    int64_t nspawn = stoch_round_spawn(pspawn);
    if (nspawn > 0) nspawn = (h > 0.0) ? -fsign64(nspawn, parent) : fsign64(nspawn, parent);
    return nspawn;
  }

  // ---- heat-bath generator ---------------------------------------------------
  // Tables of init_excit_mol_heat_bath (src/excit_gen_heat_bath_mol.F90:14-248)
  // with the loops in serial order (one OpenMP thread): i outermost, then j,
  // a, b.  Stored i-major: ija(i,j,a) = hb_ija%weights(a,j,i), etc.
 public:
  struct HeatBath {
    int nb = 0;
    bool biased = false;
    std::vector<double> i_w, ij_w, ija_w, ija_tot, ija_u, ijab_w, ijab_tot, ijab_u;
    std::vector<int> ija_k, ijab_k;
    size_t ij(int i, int j) const { return static_cast<size_t>(i - 1) * nb + (j - 1); }
    size_t ija(int i, int j, int a) const { return ij(i, j) * nb + (a - 1); }
    size_t ijab(int i, int j, int a, int b) const { return ija(i, j, a) * nb + (b - 1); }
  } hb;

  // generate_alias_tables (lib/local/alias.f90:101-184), 1-based outcomes.
  static void alias_tables(int n, const double* w, double tot, double* u, int* k) {
    std::vector<int> under(n), over(n);
    int nunder = 0, nover = 0;
    const double scale = static_cast<double>(n) / tot;
    for (int i = 1; i <= n; ++i) {
      u[i - 1] = w[i - 1] * scale;
      if (u[i - 1] <= 1.0)
        under[nunder++] = i;
      else
        over[nover++] = i;
      k[i - 1] = i;
    }
    while (nover > 0 && nunder > 0) {
      const int ov = over[nover - 1], un = under[nunder - 1];
      k[un - 1] = ov;
      nunder = nunder - 1;
      u[ov - 1] = u[ov - 1] - (1.0 - u[un - 1]);
      if (u[ov - 1] < 1.0) {
        under[nunder++] = over[nover - 1];
        nover = nover - 1;
      }
    }
  }
  // select_weighted_value_precalc (lib/local/alias.f90:13-62).
  int alias_draw(int n, const double* u, const int* k) {
    double x = rand() * n;
    int kk = static_cast<int>(std::floor(x));
    x = x - kk;
    kk = kk + 1;
    return x < u[kk - 1] ? kk : k[kk - 1];
  }
  // select_weighted_value (lib/local/alias.f90:64-99).
  int weighted_draw(int n, const double* w, double tot) {
    std::vector<double> u(n);
    std::vector<int> k(n);
    alias_tables(n, w, tot, u.data(), k.data());
    return alias_draw(n, u.data(), k.data());
  }

 private:
  void build_heat_bath() {
    const int nb = nbasis;
    hb.nb = nb;
    const size_t n2 = static_cast<size_t>(nb) * nb, n3 = n2 * nb, n4 = n3 * nb;
    hb.i_w.assign(nb, 0.0);
    hb.ij_w.assign(n2, 0.0);
    hb.ija_tot.assign(n2, 0.0);
    hb.ija_w.assign(n3, 0.0);
    hb.ijab_tot.assign(n3, 0.0);
    hb.ijab_w.assign(n4, 0.0);
    hb.ija_u.assign(n3, 0.0);
    hb.ija_k.assign(n3, 0);
    hb.ijab_u.assign(n4, 0.0);
    hb.ijab_k.assign(n4, 0);
    std::vector<int> j_nonzero(n2, 0);  // (i, a)
    for (int i = 1; i <= nb; ++i) {
      double i_weight = 0.0;
      for (int j = 1; j <= nb; ++j) {
        double ij_weight = 0.0;
        if (i != j) {
          const int it = std::min(i, j), jt = std::max(i, j);
          const int ij_sym = orb_sym[it - 1] ^ orb_sym[jt - 1];
          double ija_weights_tot = 0.0, i_weight_extra = 0.0;
          for (int a = 1; a <= nb; ++a) {
            double ija_weight = 0.0;
            if (a != i && a != j) {
              const int isymb = ij_sym ^ orb_sym[a - 1];
              for (int b = 1; b <= nb; ++b) {
                double ijab_weight = 0.0;
                const bool spin_ok = (ms_of(it) == ms_of(a) && ms_of(jt) == ms_of(b)) ||
                                     (ms_of(it) == ms_of(b) && ms_of(jt) == ms_of(a));
                if (spin_ok && orb_sym[b - 1] == isymb && b != a && b != i && b != j) {
                  const int at = std::min(a, b), bt = std::max(a, b);
                  const double h = std::fabs(sc2_excit(it, jt, at, bt, false));
                  i_weight_extra = i_weight_extra + h;
                  ij_weight = ij_weight + h;
                  ija_weight = ija_weight + h;
                  ijab_weight = ijab_weight + h;
                }
                hb.ijab_w[hb.ijab(i, j, a, b)] = ijab_weight;
                hb.ijab_tot[hb.ija(i, j, a)] = hb.ijab_tot[hb.ija(i, j, a)] + ijab_weight;
              }
            }
            hb.ija_w[hb.ija(i, j, a)] = ija_weight;
            ija_weights_tot = ija_weights_tot + ija_weight;
            if (ija_weight > kDepsilon) j_nonzero[hb.ij(i, a)] += 1;
          }
          i_weight = i_weight + i_weight_extra;
          hb.ija_tot[hb.ij(i, j)] = hb.ija_tot[hb.ij(i, j)] + ija_weights_tot;
        }
        hb.ij_w[hb.ij(i, j)] = ij_weight;
      }
      hb.i_w[i - 1] = i_weight;
      // bias test (src/excit_gen_heat_bath_mol.F90:185-202), gamma_sym = 0
      for (int a = 1; a <= nb; ++a)
        if (j_nonzero[hb.ij(i, a)] < nb - nel && i != a && orb_sym[a - 1] == orb_sym[i - 1] &&
            ms_of(a) == ms_of(i))
          hb.biased = true;
    }
    for (int i = 1; i <= nb; ++i)
      for (int j = 1; j <= nb; ++j) {
        if (!(std::fabs(hb.ija_tot[hb.ij(i, j)]) > 0.0)) continue;
        alias_tables(nb, &hb.ija_w[hb.ija(i, j, 1)], hb.ija_tot[hb.ij(i, j)], &hb.ija_u[hb.ija(i, j, 1)],
                     &hb.ija_k[hb.ija(i, j, 1)]);
        for (int a = 1; a <= nb; ++a)
          if (std::fabs(hb.ijab_tot[hb.ija(i, j, a)]) > 0.0)
            alias_tables(nb, &hb.ijab_w[hb.ijab(i, j, a, 1)], hb.ijab_tot[hb.ija(i, j, a)],
                         &hb.ijab_u[hb.ijab(i, j, a, 1)], &hb.ijab_k[hb.ijab(i, j, a, 1)]);
      }
  }

  // psingle of the single/double decision for i -> a given j
  // (src/excit_gen_heat_bath_mol.F90:367-380).
  double hb_psingle(const std::vector<int>& occ, int i, int j, int a) const {
    if (!(orb_sym[a - 1] == orb_sym[i - 1] && ms_of(a) == ms_of(i))) return 0.0;
    const double hm = std::fabs(sc1_excit(occ, i, a, false));
    const double t = hb.ijab_tot[hb.ija(i, j, a)];
    return hm < t ? hm / (t + hm) : 0.5;
  }

 public:
  // Result of one call of the heat-bath generator (for the generator probe).
  struct HbResult {
    int allowed = 0, nexcit = 0, from[2] = {0, 0}, to[2] = {0, 0};
    double pgen = 1.0, hmatel = 0.0;
  };

  // gen_excit_mol_heat_bath (src/excit_gen_heat_bath_mol.F90:252-548) with
  // select_ij_heat_bath / find_i_d_weights (src/excit_gen_utils.f90:9-66,142-167).
  HbResult gen_heat_bath(const std::vector<int>& occ, const uint64_t* f, Excit& conn) {
    HbResult r;
    conn = Excit();
    std::vector<double> iw(nel), ijw(nel), jiw(nel);
    double iw_tot = 0.0;
    for (int k = 0; k < nel; ++k) {
      iw[k] = hb.i_w[occ[k] - 1];
      iw_tot = iw_tot + iw[k];
    }
    const int i_ind = weighted_draw(nel, iw.data(), iw_tot);
    const int i = occ[i_ind - 1];
    double ij_tot = 0.0;
    for (int k = 0; k < nel; ++k) {
      ijw[k] = hb.ij_w[hb.ij(i, occ[k])];
      ij_tot = ij_tot + ijw[k];
    }
    bool allowed = false;
    int j = 0, j_ind = 0;
    double ji_tot = 0.0;
    if (ij_tot > 0.0) {
      j_ind = weighted_draw(nel, ijw.data(), ij_tot);
      j = occ[j_ind - 1];
      for (int k = 0; k < nel; ++k) {
        jiw[k] = hb.ij_w[hb.ij(j, occ[k])];
        ji_tot = ji_tot + jiw[k];
      }
      allowed = true;
    }
    allowed = allowed && std::fabs(hb.ija_tot[hb.ij(i, j)]) > 0.0;
    int a = 0;
    bool dbl = false;
    double psingle = 0.0, hm_ia = 0.0;
    if (allowed) {
      a = alias_draw(nbasis, &hb.ija_u[hb.ija(i, j, 1)], &hb.ija_k[hb.ija(i, j, 1)]);
      if (!btest(f, a)) {
        if (orb_sym[a - 1] == orb_sym[i - 1] && ms_of(a) == ms_of(i)) {
          conn.nexcit = 1;
          conn.from[0] = i;
          conn.to[0] = a;
          perm1(f, conn);
          hm_ia = std::fabs(sc1_excit(occ, i, a, conn.perm));
          const double x = rand();
          const double t = hb.ijab_tot[hb.ija(i, j, a)];
          psingle = hm_ia < t ? hm_ia / (t + hm_ia) : 0.5;
          dbl = !(x < psingle);
        } else {
          dbl = true;
          psingle = 0.0;
        }
      } else {
        allowed = false;
      }
    }
    double pgen = 1.0, h = 0.0;
    if (allowed) {
      if (dbl) {
        const int b = alias_draw(nbasis, &hb.ijab_u[hb.ijab(i, j, a, 1)], &hb.ijab_k[hb.ijab(i, j, a, 1)]);
        if (!btest(f, b)) {
          const double pi = iw[i_ind - 1] / iw_tot, pj = ijw[j_ind - 1] / ij_tot;
          const double pj2 = iw[j_ind - 1] / iw_tot, pi2 = jiw[i_ind - 1] / ji_tot;
          const double pgen_ija = (pi * pj) * (hb.ija_w[hb.ija(i, j, a)] / hb.ija_tot[hb.ij(i, j)]) *
                                  (1.0 - psingle) * (hb.ijab_w[hb.ijab(i, j, a, b)] / hb.ijab_tot[hb.ija(i, j, a)]);
          double ps = hb_psingle(occ, i, j, b);
          const double pgen_ijb = (pi * pj) * (hb.ija_w[hb.ija(i, j, b)] / hb.ija_tot[hb.ij(i, j)]) * (1.0 - ps) *
                                  (hb.ijab_w[hb.ijab(i, j, b, a)] / hb.ijab_tot[hb.ija(i, j, b)]);
          ps = hb_psingle(occ, j, i, a);
          const double pgen_jia = (pj2 * pi2) * (hb.ija_w[hb.ija(j, i, a)] / hb.ija_tot[hb.ij(j, i)]) * (1.0 - ps) *
                                  (hb.ijab_w[hb.ijab(j, i, a, b)] / hb.ijab_tot[hb.ija(j, i, a)]);
          ps = hb_psingle(occ, j, i, b);
          const double pgen_jib = (pj2 * pi2) * (hb.ija_w[hb.ija(j, i, b)] / hb.ija_tot[hb.ij(j, i)]) * (1.0 - ps) *
                                  (hb.ijab_w[hb.ijab(j, i, b, a)] / hb.ijab_tot[hb.ija(j, i, b)]);
          pgen = pgen_ija + pgen_ijb + pgen_jia + pgen_jib;
          conn = Excit();
          conn.nexcit = 2;
          conn.from[0] = std::min(i, j);
          conn.from[1] = std::max(i, j);
          conn.to[0] = std::min(a, b);
          conn.to[1] = std::max(a, b);
          perm2(f, conn);
          h = sc2_excit(conn.from[0], conn.from[1], conn.to[0], conn.to[1], conn.perm);
        } else {
          allowed = false;
        }
      } else {
        h = sc1_excit(occ, i, a, conn.perm);
        pgen = 0.0;
        for (int q = 0; q < nel; ++q) {
          const int oq = occ[q];
          if (i != oq && a != oq) {
            const double t = hb.ijab_tot[hb.ija(i, oq, a)];
            const double psq = hm_ia < t ? hm_ia / (t + hm_ia) : 0.5;
            pgen = pgen + (psq * (ijw[q] / ij_tot) * (hb.ija_w[hb.ija(i, oq, a)] / hb.ija_tot[hb.ij(i, oq)]));
          }
        }
        pgen = pgen * (iw[i_ind - 1] / iw_tot);
      }
    }
    if (!allowed) {
      h = 0.0;
      pgen = 1.0;
    }
    r.allowed = allowed;
    r.nexcit = allowed ? conn.nexcit : 0;
    for (int k = 0; k < 2; ++k) {
      r.from[k] = conn.from[k];
      r.to[k] = conn.to[k];
    }
    r.pgen = pgen;
    r.hmatel = h;
    return r;
  }

  // spawn_standard with the heat-bath generator (src/spawning.F90:34-124).
  int64_t spawn_heat_bath(const Decoded& d, const uint64_t* f, int64_t parent, double tau, Excit& conn) {
    const HbResult r = gen_heat_bath(d.occ, f, conn);
    double pspawn = tau * std::fabs(r.hmatel) / r.pgen;
    pspawn = pspawn * static_cast<double>(rf);
    int64_t nspawn = stoch_round_spawn(pspawn);
    if (nspawn > 0) nspawn = (r.hmatel > 0.0) ? -fsign64(nspawn, parent) : fsign64(nspawn, parent);
    return nspawn;
  }

  bool choose_ia(const Decoded& d, const uint64_t* f, int& i, int& a) {
    bool allowed = false;
    for (int k = 0; k < nel; ++k) {
      int o = d.occ[k];
      if (symunocc(d, ims_of(o), orb_sym[o - 1]) != 0) {
        allowed = true;
        break;
      }
    }
    if (!allowed) return false;
    while (true) {
      i = d.occ[static_cast<int>(rand() * nel)];
      int imsa = ims_of(i), isyma = orb_sym[i - 1];
      if (symunocc(d, imsa, isyma) != 0) {
        const int g = (imsa - 1) * nsym + isyma;
        while (true) {
          int ind = static_cast<int>(nbss[g] * rand()) + 1;
          a = ssbf[g * sys_.max_sym_spin + ind - 1];
          if (!btest(f, a)) break;
        }
        break;
      }
    }
    return true;
  }
  double pgen_single(const Decoded& d, int a) const {
    int ni = nel;
    for (int k = 0; k < nel; ++k) {
      int o = d.occ[k];
      if (symunocc(d, ims_of(o), orb_sym[o - 1]) == 0) ni = ni - 1;
    }
    return 1.0 / static_cast<double>(ni * symunocc(d, ims_of(a), orb_sym[a - 1]));
  }
  void choose_ij_mol(const Decoded& d, int& i, int& j, int& ij_sym, int& ij_spin, double& pgen_ij) {
    int ind = static_cast<int>(rand() * nel * (nel - 1) / 2) + 1;
    int j_ind = static_cast<int>(1.50 + std::sqrt(2 * ind - 1.750));
    int i_ind = ind - ((j_ind - 1) * (j_ind - 2)) / 2;
    i = d.occ[i_ind - 1];
    j = d.occ[j_ind - 1];
    ij_sym = orb_sym[i - 1] ^ orb_sym[j - 1];
    ij_spin = ms_of(i) + ms_of(j);
    pgen_ij = 2.0 / static_cast<double>(nel * (nel - 1));
  }
  bool choose_ab_mol(const Decoded& d, const uint64_t* f, int sym, int spin, int& a, int& b) {
    bool allowed = false;
    int fac = 1, shift = 0, na = nbasis;
    for (int isyma = 0; isyma < nsym && !allowed; ++isyma) {
      int isymb = isyma ^ sym;
      if (spin == -2) {
        if (symunocc(d, 1, isyma) > 0 &&
            (symunocc(d, 1, isymb) > 1 || (symunocc(d, 1, isymb) == 1 && isyma != isymb)))
          allowed = true;
      } else if (spin == 0) {
        if ((symunocc(d, 1, isyma) > 0 && symunocc(d, 2, isymb) > 0) ||
            (symunocc(d, 2, isyma) > 0 && symunocc(d, 1, isymb) > 0))
          allowed = true;
      } else {
        if (symunocc(d, 2, isyma) > 0 &&
            (symunocc(d, 2, isymb) > 1 || (symunocc(d, 2, isymb) == 1 && isyma != isymb)))
          allowed = true;
      }
    }
    if (spin == -2) {
      fac = 2;
      shift = 0;
      na = nbasis / 2;
    } else if (spin == 0) {
      fac = 1;
      shift = 0;
      na = nbasis;
    } else {
      fac = 2;
      shift = 1;
      na = nbasis / 2;
    }
    if (!allowed) return false;
    while (true) {
      a = static_cast<int>(rand() * na) + 1;
      a = fac * a - shift;
      if (!btest(f, a)) {
        int imsb = (spin - ms_of(a) + 3) / 2;
        int isymb = sym ^ orb_sym[a - 1];
        int nu = symunocc(d, imsb, isymb);
        if (nu > 1 || (nu == 1 && (isymb != orb_sym[a - 1] || spin == 0))) {
          const int g = (imsb - 1) * nsym + isymb;
          while (true) {
            int ind = static_cast<int>(nbss[g] * rand()) + 1;
            b = ssbf[g * sys_.max_sym_spin + ind - 1];
            if (b != a && !btest(f, b)) break;
          }
          break;
        }
      }
    }
    if (a > b) std::swap(a, b);
    return true;
  }
  double pgen_double(const Decoded& d, int ij_sym, int a, int b, int spin) const {
    int imsa = ims_of(a), imsb = ims_of(b);
    int n_aij;
    double p_bija, p_aijb;
    if (spin == -2 || spin == 2) {
      const int ims = spin == -2 ? 1 : 2;
      n_aij = spin == -2 ? sys_.nvirt_beta : sys_.nvirt_alpha;
      for (int isyma = 0; isyma < nsym; ++isyma) {
        int isymb = isyma ^ ij_sym;
        if (symunocc(d, ims, isymb) == 0)
          n_aij = n_aij - symunocc(d, ims, isyma);
        else if (isyma == isymb && symunocc(d, ims, isymb) == 1)
          n_aij = n_aij - symunocc(d, ims, isyma);
      }
      if (orb_sym[a - 1] == orb_sym[b - 1]) {
        p_aijb = 1.0 / static_cast<double>(symunocc(d, imsa, orb_sym[a - 1]) - 1);
        p_bija = 1.0 / static_cast<double>(symunocc(d, imsb, orb_sym[b - 1]) - 1);
      } else {
        p_aijb = 1.0 / static_cast<double>(symunocc(d, imsa, orb_sym[a - 1]));
        p_bija = 1.0 / static_cast<double>(symunocc(d, imsb, orb_sym[b - 1]));
      }
    } else {
      n_aij = sys_.nvirt;
      for (int isyma = 0; isyma < nsym; ++isyma) {
        int isymb = isyma ^ ij_sym;
        if (symunocc(d, 1, isymb) == 0) n_aij = n_aij - symunocc(d, 2, isyma);
        if (symunocc(d, 2, isymb) == 0) n_aij = n_aij - symunocc(d, 1, isyma);
      }
      p_aijb = 1.0 / static_cast<double>(symunocc(d, imsa, orb_sym[a - 1]));
      p_bija = 1.0 / static_cast<double>(symunocc(d, imsb, orb_sym[b - 1]));
    }
    return (1.0 / static_cast<double>(n_aij)) * (p_bija + p_aijb);
  }

  void create_spawned(const uint64_t* f, const Excit& conn, int64_t nspawn, int space, int64_t flag) {
    if (static_cast<int64_t>(spawned.size()) + 1 > par_.spawn_capacity) {
      spawn_error = true;
      return;
    }
    Spawned sp;
    std::memset(&sp, 0, sizeof(sp));
    excited_det(f, conn, sp.f);
    sp.pop[space] = nspawn;
    sp.flag = par_.initiator ? flag : 0;
    spawned.push_back(sp);
  }

  void death(double tau, double Kii, double shift, double E, int64_t& population, double& tot, int64_t& ndeath) {
    const double weight = 1.0, qnpc = 1.0, dmqmc = 1.0, cheby = 1.0;
    double pd = tau * ((Kii - E) * weight + (E - shift) * qnpc) * dmqmc;
    pd = pd * cheby;
    pd = pd * static_cast<double>(population < 0 ? -population : population);
    int64_t kill = static_cast<int64_t>(pd);
    pd = pd - static_cast<double>(kill);
    double r = rand();
    if (std::fabs(pd) > r) kill = pd > 0.0 ? kill + 1 : kill - 1;
    int64_t old = population;
    if (population < 0)
      population = population + kill;
    else
      population = population - kill;
    tot = tot + static_cast<double>(std::llabs(population) - std::llabs(old)) / static_cast<double>(rf);
    ndeath = ndeath + std::llabs(kill);
  }

  // ---- annihilation ----------------------------------------------------------
  // Unsigned multi-word comparison, last word most significant
  // (src/bit_utils.F90:341-480).
  int cmp(const uint64_t* a, const uint64_t* b) const {  // bit_str_cmp: 1 if a<b, -1 if a>b
    for (int w = W - 1; w >= 0; --w) {
      if (a[w] < b[w]) return 1;
      if (a[w] > b[w]) return -1;
    }
    return 0;
  }
  bool gt(const uint64_t* a, const uint64_t* b) const { return cmp(a, b) == -1; }
  bool ge(const uint64_t* a, const uint64_t* b) const { return cmp(a, b) != 1; }

  // qsort_int_64_list (src/sort.f90:213-396), element = Spawned.
  void qsort_spawned(std::vector<Spawned>& list) const {
    const int head = static_cast<int>(list.size());
    auto L = [&](int k) -> Spawned& { return list[k - 1]; };  // 1-based view
    int stack[2][64];
    int nstack = 0, lo = 1, hi = head;
    while (true) {
      if (hi - lo < 7) {
        for (int j = lo + 1; j <= hi; ++j) {
          Spawned tmp = L(j);
          int i;
          for (i = j - 1; i >= 1; --i) {
            if (ge(tmp.f, L(i).f)) break;
            L(i + 1) = L(i);
          }
          L(i + 1) = tmp;
        }
        if (nstack == 0) break;
        hi = stack[1][nstack - 1];
        lo = stack[0][nstack - 1];
        --nstack;
      } else {
        int pivot = (lo + hi) / 2;
        std::swap(L(pivot), L(lo + 1));
        if (gt(L(lo).f, L(hi).f)) std::swap(L(lo), L(hi));
        if (gt(L(lo + 1).f, L(hi).f)) std::swap(L(lo + 1), L(hi));
        if (gt(L(lo).f, L(lo + 1).f)) std::swap(L(lo), L(lo + 1));
        int i = lo + 1, j = hi;
        Spawned tmp = L(lo + 1);
        while (true) {
          ++i;
          while (gt(tmp.f, L(i).f)) ++i;
          --j;
          while (gt(L(j).f, tmp.f)) --j;
          if (j < i) break;
          std::swap(L(i), L(j));
        }
        L(lo + 1) = L(j);
        L(j) = tmp;
        ++nstack;
        if (hi - i + 1 >= j - lo) {
          stack[1][nstack - 1] = hi;
          stack[0][nstack - 1] = i;
          hi = j - 1;
        } else {
          stack[1][nstack - 1] = j - 1;
          stack[0][nstack - 1] = lo;
          lo = i;
        }
      }
    }
  }

  bool same(const uint64_t* a, const uint64_t* b) const {
    for (int w = 0; w < W; ++w)
      if (a[w] != b[w]) return false;
    return true;
  }
  bool allzero(const int64_t* p) const {
    for (int s = 0; s < S; ++s)
      if (p[s] != 0) return false;
    return true;
  }

  // annihilate_spawn_t (src/spawn_data.F90:859-935)
  void self_annihilate_plain() {
    int n = static_cast<int>(spawned.size());
    if (n == 0) return;
    std::vector<Spawned> out;
    int64_t nrun = 0;
    int k = 0;
    while (k < n) {
      Spawned acc = spawned[k];
      acc.run = nrun++;
      int m = k + 1;
      while (m < n && same(spawned[m].f, acc.f)) {
        for (int s = 0; s < S; ++s) acc.pop[s] += spawned[m].pop[s];
        ++m;
      }
      if (!allzero(acc.pop)) out.push_back(acc);
      k = m;
    }
    spawned.swap(out);
  }

  // annihilate_spawn_t_initiator (src/spawn_data.F90:937-1101).  The first
  // entry of each run sets the event count from the sign of its population
  // converted to a default (32-bit) integer, as the reference does.
  void self_annihilate_initiator() {
    int n = static_cast<int>(spawned.size());
    if (n == 0) return;
    std::vector<Spawned> out;
    int64_t nrun = 0;
    int k = 0;
    while (k < n) {
      Spawned acc = spawned[k];
      acc.run = nrun++;
      int64_t init_pop[FQ_MAX_SPACES];
      int events[FQ_MAX_SPACES];
      for (int s = 0; s < S; ++s) {
        if (!((spawned[k].flag >> s) & 1)) {
          init_pop[s] = spawned[k].pop[s];
          events[s] = 0;
        } else {
          init_pop[s] = 0;
          int32_t p32 = static_cast<int32_t>(static_cast<uint32_t>(spawned[k].pop[s] & 0xffffffffULL));
          events[s] = p32 >= 0 ? 1 : -1;
        }
      }
      int m = k + 1;
      while (m < n && same(spawned[m].f, acc.f)) {
        for (int s = 0; s < S; ++s) {
          if (!((spawned[m].flag >> s) & 1))
            init_pop[s] += spawned[m].pop[s];
          else if (spawned[m].pop[s] < 0)
            events[s] -= 1;
          else if (spawned[m].pop[s] > 0)
            events[s] += 1;
        }
        for (int s = 0; s < S; ++s) acc.pop[s] += spawned[m].pop[s];
        ++m;
      }
      acc.flag = 0;
      for (int s = 0; s < S; ++s) {
        int sgn_tot = acc.pop[s] >= 0 ? 1 : -1;
        int sgn_init = init_pop[s] >= 0 ? 1 : -1;
        if (init_pop[s] != 0 && sgn_tot == sgn_init) {
        } else if (std::abs(events[s]) > 1) {
        } else {
          acc.flag += int64_t{1} << s;
        }
      }
      if (!allzero(acc.pop)) out.push_back(acc);
      k = m;
    }
    spawned.swap(out);
  }

  // binary_search (src/search.f90:47-135), 0-based range [istart, iend].
  void bsearch(const uint64_t* item, int istart, int iend, bool& hit, int& pos) const {
    hit = false;
    if (istart > iend) {
      pos = istart;
      return;
    }
    int lo = istart, hi = iend;
    while (hi != lo) {
      pos = (hi + lo) / 2;
      int c = cmp(states[pos].f, item);
      if (c == 0) {
        hit = true;
        break;
      } else if (c == 1) {
        lo = pos + 1;
      } else {
        hi = pos;
      }
    }
    if (hi == lo) {
      int c = cmp(states[hi].f, item);
      if (c == 0) {
        hit = true;
        pos = hi;
      } else if (c == 1) {
        pos = hi + 1;
      } else {
        pos = hi;
      }
    }
  }

  void stoch_round(int64_t* p) {
    for (int s = 0; s < S; ++s) {
      if (std::llabs(p[s]) < rf && p[s] != 0) {
        double r = rand() * static_cast<double>(rf);
        p[s] = (static_cast<double>(std::llabs(p[s])) > r) ? fsign64(rf, p[s]) : 0;
      }
    }
  }

  void annihilate() {
    if (!spawned.empty()) {
      if (stable_order)
        std::stable_sort(spawned.begin(), spawned.end(),
                         [&](const Spawned& a, const Spawned& b) { return cmp(a.f, b.f) == 1; });
      else
        qsort_spawned(spawned);
      if (par_.initiator)
        self_annihilate_initiator();
      else
        self_annihilate_plain();
    }
    if (!spawned.empty()) {
      annihilate_main_list();
      remove_unoccupied();
      if (par_.real_amplitudes) round_low_spawns();
      insert_new();
    } else {
      remove_unoccupied();
    }
  }

  void add_np(const int64_t* now, const int64_t* old) {
    for (int s = 0; s < S; ++s)
      nparticles[s] = nparticles[s] + static_cast<double>(std::llabs(now[s]) - std::llabs(old[s])) / rf;
  }

  void annihilate_main_list() {
    int istart = 0, iend = static_cast<int>(states.size()) - 1;
    std::vector<Spawned> keep;
    for (size_t i = 0; i < spawned.size(); ++i) {
      bool hit;
      int pos;
      bsearch(spawned[i].f, istart, iend, hit, pos);
      if (hit) {
        int64_t old[FQ_MAX_SPACES];
        for (int s = 0; s < S; ++s) old[s] = pops[pos * S + s];
        if (par_.initiator) {
          for (int s = 0; s < S; ++s) {
            if (pops[pos * S + s] != 0)
              pops[pos * S + s] += spawned[i].pop[s];
            else if (!((spawned[i].flag >> s) & 1))
              pops[pos * S + s] = spawned[i].pop[s];
          }
        } else {
          for (int s = 0; s < S; ++s) pops[pos * S + s] += spawned[i].pop[s];
        }
        add_np(&pops[pos * S], old);
        istart = pos + 1;
      } else {
        if (par_.initiator) {
          Spawned sp = spawned[i];
          bool discard = true;
          for (int s = 0; s < S; ++s) {
            if ((spawned[i].flag >> s) & 1)
              sp.pop[s] = 0;
            else
              discard = false;
          }
          if (!discard) keep.push_back(sp);
        } else {
          keep.push_back(spawned[i]);
        }
      }
    }
    spawned.swap(keep);
  }

  void remove_unoccupied() {
    size_t k = 0;
    for (size_t i = 0; i < states.size(); ++i) {
      const bool core = is_determ(states[i].f);
      if (par_.real_amplitudes && !core) {
        int64_t old[FQ_MAX_SPACES];
        for (int s = 0; s < S; ++s) old[s] = pops[i * S + s];
        draw_stream(static_cast<uint32_t>(i), static_cast<uint32_t>(static_cast<uint64_t>(i) >> 32), kDrawRoundMain);
        stoch_round(&pops[i * S]);
        add_np(&pops[i * S], old);
      }
      if (core || !allzero(&pops[i * S])) {
        states[k] = states[i];
        for (int s = 0; s < S; ++s) pops[k * S + s] = pops[i * S + s];
        dat[k] = dat[i];
        ++k;
      }
    }
    states.resize(k);
    pops.resize(k * S);
    dat.resize(k);
  }

  void round_low_spawns() {
    std::vector<Spawned> keep;
    for (Spawned& sp : spawned) {
      draw_stream(static_cast<uint32_t>(sp.run), static_cast<uint32_t>(sp.run >> 32), kDrawRoundSpawn);
      stoch_round(sp.pop);
      if (!allzero(sp.pop)) keep.push_back(sp);
    }
    spawned.swap(keep);
  }

  // insert_new_walkers: backward in-place merge (src/annihilation.f90:677-901).
  void insert_new() {
    const int64_t nnew = static_cast<int64_t>(spawned.size());
    if (static_cast<int64_t>(states.size()) + nnew > par_.main_capacity) {
      main_error = true;
      return;
    }
    const int n0 = static_cast<int>(states.size());
    states.resize(n0 + nnew);
    pops.resize((n0 + nnew) * S);
    dat.resize(n0 + nnew);
    int iend = n0 - 1;
    for (int64_t i = nnew - 1; i >= 0; --i) {
      bool hit;
      int pos;
      bsearch(spawned[i].f, 0, iend, hit, pos);
      for (int j = iend; j >= pos; --j) {
        int k = j + static_cast<int>(i) + 1;
        states[k] = states[j];
        for (int s = 0; s < S; ++s) pops[k * S + s] = pops[j * S + s];
        dat[k] = dat[j];
      }
      int k = pos + static_cast<int>(i);
      for (int w = 0; w < W; ++w) states[k].f[w] = spawned[i].f[w];
      for (int s = 0; s < S; ++s) pops[k * S + s] = spawned[i].pop[s];
      dat[k] = sc0(spawned[i].f) - sys_.H00;
      for (int s = 0; s < S; ++s)
        nparticles[s] = nparticles[s] + std::fabs(static_cast<double>(spawned[i].pop[s]) / rf);
      iend = pos - 1;
    }
  }
};

// Capacity in elements from a state_size / spawned_state_size input: a
// positive value is an element count, a negative one is MB per core
// (src/particle_t_utils.f90:84-101, src/qmc.F90:1445-1448).
int64_t capacity_from(int64_t size, int64_t element_bytes, bool spawned) {
  if (size >= 0) return size;
  double mb = -static_cast<double>(size) * 1.0e6;
  return static_cast<int64_t>(spawned ? mb / (2.0 * element_bytes) : mb / element_bytes);
}

// MurmurHash2 (Appleby; lib/external/MurmurHash2.c:16, the reference's
// copy): 32-bit hash of len bytes read as little-endian 32-bit words.
uint32_t murmur2(const unsigned char* data, int len, uint32_t seed) {
  const uint32_t m = 0x5bd1e995u;
  uint32_t h = seed ^ static_cast<uint32_t>(len);
  for (; len >= 4; data += 4, len -= 4) {
    uint32_t k = static_cast<uint32_t>(data[0]) | (static_cast<uint32_t>(data[1]) << 8) |
                 (static_cast<uint32_t>(data[2]) << 16) | (static_cast<uint32_t>(data[3]) << 24);
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

// assign_particle_processor with hash_shift 0 and one slot per rank
// (src/spawning.F90:770-838, proc_map(i) = mod(i, nprocs),
// src/load_balancing.F90:141-170): murmurhash_bit_string over
// ceiling(nbasis/32)*4 bytes of the bit string, seed 7 (src/qmc.F90:1496,
// lib/local/hash.f90), the signed hash taken modulo nprocs (Fortran modulo).
int owner_rank(const uint64_t* f, int nbasis, int nprocs) {
  unsigned char bytes[8 * FQ_MAX_WORDS];
  for (int w = 0; w < FQ_MAX_WORDS; ++w)
    for (int b = 0; b < 8; ++b) bytes[8 * w + b] = static_cast<unsigned char>(f[w] >> (8 * b));
  const int nbytes = (nbasis + 31) / 32 * 4;
  const int32_t hash = static_cast<int32_t>(murmur2(bytes, nbytes, 7u));
  int r = hash % nprocs;
  if (r < 0) r += nprocs;
  return r;
}

// Report loop of do_fciqmc (src/fciqmc.f90:276-464) with the estimator and
// shift bookkeeping of src/energy_evaluation.F90:425-711 and
// src/qmc_common.F90:929-1017,1240-1304, on nprocs emulated MPI ranks: rank r
// draws from dSFMT seeded rng_seed + r, owns the determinants that hash to it
// (the reference determinant's initial walkers start there), spawns into a
// list of per-destination blocks in creation order, and receives the blocks
// addressed to it in rank order (comm_spawn_t, src/spawn_data.F90:624-739)
// before annihilating them.  Report quantities are the ranks' sums
// (MPI_Allreduce, src/energy_evaluation.F90:126-201), R_spawn their mean.
//
// start (may be NULL) restarts from a stored state instead of the initial
// distribution, as read_restart_hdf5 and the restart branches of init_qmc /
// init_estimators / do_fciqmc do for a version-2 file (src/restart_hdf5.F90:
// 539-927, src/qmc.F90:827-908, src/fciqmc.f90:147-262): each stored walker
// goes to the rank owning it, shift / vary flags / N_0 / projected energy and
// the iteration count continue, the population is recounted.  end (may be
// NULL, nprocs = 1 only) receives the state after the last report
// (dump_restart_hdf5, src/restart_hdf5.F90:307-537).
int run_fciqmc(const fq_system* sys, const fqh_qmc_in* in, fqh_report_row* rows, int max_rows, int* nrows,
               int nprocs, const fqh_restart* start = nullptr, fqh_restart* end = nullptr) {
  if (nprocs < 1 || (end && nprocs != 1) || (in->semi_stoch_size > 0 && nprocs != 1)) return -1;
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
  const int S = p.nspaces;
  const int W = sys->nwords;
  p.main_capacity = capacity_from(in->state_size, 8 * W + 8 * S + 8 * S, false);
  p.spawn_capacity = capacity_from(in->spawned_state_size, 8 * (W + S + (in->initiator ? 1 : 0)), true);
  std::vector<std::unique_ptr<Oracle>> ranks;
  for (int r = 0; r < nprocs; ++r)
    ranks.emplace_back(new Oracle(sys, &p, static_cast<uint32_t>(static_cast<int32_t>(in->rng_seed) + r)));
  if (!start) {
    // initial_distribution (src/qmc.F90:1507-1580): on the reference determinant's rank
    std::vector<int64_t> pop0(S, static_cast<int64_t>(std::llround(in->init_pop)) * rf);
    std::vector<double> dat0(1, 0.0);
    const int d0_rank = owner_rank(sys->f0, sys->nbasis, nprocs);
    ranks[d0_rank]->load(sys->f0, pop0.data(), dat0.data(), 1);
  } else {
    for (int r = 0; r < nprocs; ++r) {
      std::vector<uint64_t> st;
      std::vector<int64_t> pp;
      std::vector<double> dd;
      for (int64_t i = 0; i < start->nstates; ++i) {
        if (owner_rank(start->states + i * W, sys->nbasis, nprocs) != r) continue;
        st.insert(st.end(), start->states + i * W, start->states + (i + 1) * W);
        pp.insert(pp.end(), start->pops + i * S, start->pops + (i + 1) * S);
        dd.push_back(start->dat[i]);
      }
      ranks[r]->load(st.data(), pp.data(), dd.data(), static_cast<int64_t>(dd.size()));
    }
  }
  const int64_t cycles0 = start ? start->mc_cycles_done : 0;
  double proj[FQ_MAX_SPACES], d0[FQ_MAX_SPACES], ntot0[FQ_MAX_SPACES];
  auto total_particles = [&](int s) {
    double t = 0.0;
    for (auto& o : ranks) t = t + o->nparticles[s];
    return t;
  };
  auto total_states = [&]() {
    int64_t t = 0;
    for (auto& o : ranks) t += static_cast<int64_t>(o->states.size());
    return t;
  };
  for (int s = 0; s < S; ++s) {
    ntot0[s] = total_particles(s);
    proj[s] = d0[s] = 0.0;
  }
  // shift, vary flags, damping and on-the-fly blocking between reports
  // (the report-loop bookkeeping the drivers share; pinned by the goldens)
  fqmc::ShiftControl ctl;
  ctl.init(*in, S, ntot0, start);
  const double* shift = ctl.shift;
  if (start) {
    for (int s = 0; s < S; ++s) {
      proj[s] = start->proj_energy[s];
      d0[s] = start->D0_population[s];
    }
  } else {
    for (auto& o : ranks) {
      double pr[FQ_MAX_SPACES], dr[FQ_MAX_SPACES];
      o->projected_energy(pr, dr);
      for (int s = 0; s < S; ++s) {
        proj[s] = proj[s] + pr[s];
        d0[s] = d0[s] + dr[s];
      }
    }
  }
  int nr = 0;
  auto emit = [&](int64_t iter, int64_t nspawn, double rspawn) {
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
    r.nstates = total_states();
    r.shift_damping = ctl.damping;
    r.nspawn_events = nspawn;
    r.rspawn = rspawn;
  };
  emit(cycles0, 0, 0.0);
  int reports_done = 0;
  // the iteration the semi-stochastic core is built (src/fciqmc.f90:229-235)
  int64_t ss_iter = std::max<int64_t>(in->semi_stoch_start_iter, cycles0 + 1);
  auto all_vary = [&]() {
    for (int s = 0; s < S; ++s)
      if (!ctl.vary[s]) return false;
    return true;
  };
  if (all_vary() && in->semi_stoch_shift_iter != -1) ss_iter = cycles0 + 1;
  std::vector<fq_cycle_out> out(nprocs);
  std::vector<std::vector<Spawned>> sent(nprocs);
  for (int ireport = 1; ireport <= in->nreports && nr < max_rows; ++ireport) {  // a full table ends the run
    double eproj_old[FQ_MAX_SPACES];
    for (int s = 0; s < S; ++s) eproj_old[s] = (std::fabs(d0[s]) < 2.2250738585072014e-308) ? 0.0 : proj[s] / d0[s];
    // per-rank report accumulators
    std::vector<double> rproj(nprocs * S, 0.0), rd0(nprocs * S, 0.0), rspawn(nprocs, 0.0);
    std::vector<int64_t> nspawn_events(nprocs, 0);
    bool error = false;
    for (int ic = 0; ic < in->mc_cycles; ++ic) {
      const int64_t iter = cycles0 + static_cast<int64_t>(ireport - 1) * in->mc_cycles + ic + 1;
      if (in->semi_stoch_size > 0 && iter == ss_iter)
        ranks[0]->semi_stoch_init(in->semi_stoch_size, in->semi_stoch_separate != 0);
      std::vector<int64_t> nattempts(nprocs);
      for (int r = 0; r < nprocs; ++r) {
        nattempts[r] = static_cast<int64_t>(std::llround(2 * ranks[r]->nparticles[0]));
        ranks[r]->cycle_spawn(shift, eproj_old, in->tau, &out[r]);
      }
      if (nprocs > 1) {  // the exchange: rank r receives, in rank order, the blocks addressed to it
        for (int r = 0; r < nprocs; ++r) sent[r].swap(ranks[r]->spawned_list());
        for (int r = 0; r < nprocs; ++r) {
          std::vector<Spawned>& recv = ranks[r]->spawned_list();
          recv.clear();
          for (int q = 0; q < nprocs; ++q)
            for (const Spawned& e : sent[q])
              if (owner_rank(e.f, sys->nbasis, nprocs) == r) recv.push_back(e);
        }
      }
      for (int r = 0; r < nprocs; ++r) {
        ranks[r]->cycle_finish(&out[r]);
        for (int s = 0; s < S; ++s) {
          rproj[r * S + s] = rproj[r * S + s] + out[r].proj_energy[s];
          rd0[r * S + s] = rd0[r * S + s] + out[r].D0_population[s];
        }
        nspawn_events[r] = out[r].nspawn_events;
        double rate = 0.0;
        if (nattempts[r] > 0) {
          double ndeath_real = static_cast<double>(out[r].ndeath) / static_cast<double>(rf);
          rate = (static_cast<double>(nspawn_events[r]) + ndeath_real) / static_cast<double>(nattempts[r]);
        }
        rspawn[r] = rspawn[r] + rate;
        error = error || out[r].spawn_overflow || out[r].main_overflow;
      }
    }
    double ntot[FQ_MAX_SPACES], rs = 0.0;
    int64_t nsp = 0;
    for (int s = 0; s < S; ++s) {
      proj[s] = d0[s] = 0.0;
      for (int r = 0; r < nprocs; ++r) {
        proj[s] = proj[s] + rproj[r * S + s];
        d0[s] = d0[s] + rd0[r * S + s];
      }
      ntot[s] = total_particles(s);
      proj[s] = proj[s] / in->mc_cycles;
      d0[s] = d0[s] / in->mc_cycles;
    }
    for (int r = 0; r < nprocs; ++r) {
      rs = rs + rspawn[r];
      nsp += nspawn_events[r];
    }
    rs = rs / (in->mc_cycles * nprocs);
    const bool vary_before = all_vary();
    const bool soft_exit = ctl.end_report(ntot, proj, d0);
    if (!vary_before && all_vary() && in->semi_stoch_shift_iter != -1)  // end_report_loop, src/qmc_common.F90:1203
      ss_iter = in->semi_stoch_shift_iter + cycles0 + static_cast<int64_t>(ireport) * in->mc_cycles + 1;
    emit(cycles0 + static_cast<int64_t>(ireport) * in->mc_cycles, nsp, rs);
    ctl.after_report(ireport, proj[0], d0[0]);
    if (nr > 0 && nr <= max_rows) {
      rows[nr - 1].shift_damping = ctl.damping;
      if (ctl.blocking()) {
        rows[nr - 1].blocked_proj_energy = ctl.blocks().proj_energy_mean();
        rows[nr - 1].blocked_proj_error = ctl.blocks().proj_energy_error();
      }
    }
    reports_done = ireport;
    if (error || soft_exit) break;
  }
  *nrows = nr;
  if (end) {
    Oracle& o = *ranks[0];
    end->nstates = static_cast<int64_t>(o.states.size());
    if (end->nstates > end->capacity) return -5;
    o.download(end->states, end->pops, end->dat);
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
  return 0;
}

}  // namespace

extern "C" {

// Oracle handle with the engine's cycle-level interface.
void* orc_create(const fq_system* s, const fq_params* p, int64_t seed) {
  return new Oracle(s, p, static_cast<uint32_t>(seed));
}
void orc_destroy(void* h) { delete static_cast<Oracle*>(h); }
// Switch an oracle to production mode (Philox streams keyed by seed; cycle 0 next).
void orc_set_production(void* h, uint64_t philox_seed) { static_cast<Oracle*>(h)->set_production(philox_seed); }

// One Philox4x32-10 block (for the known-answer vectors).
void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox4x32_10(ctr, key, out); }

int orc_load(void* h, const uint64_t* st, const int64_t* p, const double* d, int64_t n) {
  static_cast<Oracle*>(h)->load(st, p, d, n);
  return 0;
}
int orc_cycle(void* h, const double* shift, const double* eproj_old, double tau, fq_cycle_out* out) {
  static_cast<Oracle*>(h)->cycle(shift, eproj_old, tau, out);
  return 0;
}
int orc_annihilate(void* h, const uint64_t* sst, const int64_t* sp, const int64_t* sf, int64_t ns,
                   fq_cycle_out* out) {
  static_cast<Oracle*>(h)->annihilate_external(sst, sp, sf, ns, out);
  return 0;
}
int64_t orc_nstates(void* h) { return static_cast<int64_t>(static_cast<Oracle*>(h)->states.size()); }
int orc_download(void* h, uint64_t* st, int64_t* p, double* d) {
  static_cast<Oracle*>(h)->download(st, p, d);
  return 0;
}
uint64_t orc_rng_consumed(void* h) { return static_cast<Oracle*>(h)->rng_consumed(); }
void orc_projected_energy(void* h, double* proj, double* d0) { static_cast<Oracle*>(h)->projected_energy(proj, d0); }
double orc_nparticles(void* h, int s) { return static_cast<Oracle*>(h)->nparticles[s]; }
void orc_set_stable_order(void* h, int on) { static_cast<Oracle*>(h)->stable_order = on != 0; }
// Semi-stochastic core from the current list (fq_semi_stoch_init) and its description.
int orc_semi_stoch_init(void* h, int64_t target_size, int separate) {
  static_cast<Oracle*>(h)->semi_stoch_init(static_cast<int>(target_size), separate != 0);
  return 0;
}
int64_t orc_semi_stoch_info(void* h, uint64_t* states, int64_t* nnz) {
  return static_cast<Oracle*>(h)->determ_info(states, nnz);
}

// Heat-bath probes for the tests.
int orc_hb_biased(void* h) { return static_cast<Oracle*>(h)->hb.biased ? 1 : 0; }
// Copy a heat-bath table ("i_w", "ij_w", "ija_w", "ija_tot", "ija_u", "ijab_w", "ijab_tot", "ijab_u")
// into out[0..n); returns its length (or -1 for an unknown name).
int64_t orc_hb_table(void* h, const char* name, double* out, int64_t n) {
  const auto& hb = static_cast<Oracle*>(h)->hb;
  const std::vector<double>* t = nullptr;
  const std::string nm(name);
  if (nm == "i_w") t = &hb.i_w;
  if (nm == "ij_w") t = &hb.ij_w;
  if (nm == "ija_w") t = &hb.ija_w;
  if (nm == "ija_tot") t = &hb.ija_tot;
  if (nm == "ija_u") t = &hb.ija_u;
  if (nm == "ijab_w") t = &hb.ijab_w;
  if (nm == "ijab_tot") t = &hb.ijab_tot;
  if (nm == "ijab_u") t = &hb.ijab_u;
  if (!t) return -1;
  const int64_t len = static_cast<int64_t>(t->size());
  if (out) std::memcpy(out, t->data(), sizeof(double) * static_cast<size_t>(std::min(n, len)));
  return len;
}
void orc_alias_tables(int n, const double* w, double tot, double* u, int* k) { Oracle::alias_tables(n, w, tot, u, k); }
// One heat-bath generator call from determinant f with the given uniform draws.
// ires = {allowed, nexcit, from1, from2, to1, to2}, dres = {pgen, hmatel};
// returns the number of draws the call consumed.
int64_t orc_hb_generate(void* h, const uint64_t* f, const double* draws, int64_t ndraws, int* ires, double* dres) {
  Oracle* o = static_cast<Oracle*>(h);
  o->forced_ = draws;
  o->nforced_ = ndraws;
  o->iforced_ = 0;
  Excit conn;
  const Oracle::HbResult r = o->gen_heat_bath(o->decode_occ(f), f, conn);
  const int64_t used = o->iforced_;
  o->forced_ = nullptr;
  ires[0] = r.allowed;
  ires[1] = r.nexcit;
  ires[2] = r.from[0];
  ires[3] = r.from[1];
  ires[4] = r.to[0];
  ires[5] = r.to[1];
  dres[0] = r.pgen;
  dres[1] = r.hmatel;
  return used;
}

int orc_fciqmc(const fq_system* sys, const fqh_qmc_in* in, fqh_report_row* rows, int max_rows, int* nrows) {
  return run_fciqmc(sys, in, rows, max_rows, nrows, 1);
}
// The report loop continued from a stored state (see run_fciqmc).
int orc_fciqmc_restart(const fq_system* sys, const fqh_qmc_in* in, const fqh_restart* start, fqh_restart* end,
                       fqh_report_row* rows, int max_rows, int* nrows) {
  return run_fciqmc(sys, in, rows, max_rows, nrows, 1, start, end);
}
// The report loop on nprocs emulated MPI ranks (see run_fciqmc).
int orc_fciqmc_ranks(const fq_system* sys, const fqh_qmc_in* in, int nprocs, fqh_report_row* rows, int max_rows,
                     int* nrows) {
  return run_fciqmc(sys, in, rows, max_rows, nrows, nprocs);
}
// MurmurHash2 of len bytes (the oracle's restatement, for the vector test).
uint32_t orc_murmur2(const unsigned char* data, int len, uint32_t seed) { return murmur2(data, len, seed); }
// Owner rank of a determinant among nprocs ranks.
int orc_owner_rank(const uint64_t* f, int nbasis, int nprocs) {
  uint64_t g[FQ_MAX_WORDS] = {0, 0};
  for (int w = 0; w < (nbasis + 63) / 64 && w < FQ_MAX_WORDS; ++w) g[w] = f[w];
  return owner_rank(g, nbasis, nprocs);
}

}  // extern "C"
