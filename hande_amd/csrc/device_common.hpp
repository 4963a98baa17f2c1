// Device-side arithmetic of one FCIQMC cycle for gfx950.
//
// Every routine here restates a reference routine in the same floating-point
// expression order (the HIP sources are compiled with -ffp-contract=off), so
// that a replayed random stream reproduces the reference's Markov chain:
//   src/excitations.F90:75-200,247-406   excitation level, parity, child det
//   src/hubbard_k.f90, src/hamiltonian_hub_k.f90:91-232   Hubbard-k elements
//   src/excit_gen_hub_k.f90:29-159,469-770   Hubbard-k generators, pgen
//   src/excit_gen_mol.f90:16-101,384-946,1140-1327   molecular renorm generator
//   src/excit_gen_heat_bath_mol.F90:252-548, src/excit_gen_utils.f90:9-66,142-167,
//       lib/local/alias.f90:13-184   heat-bath generator and the alias method
//   src/hamiltonian_molecular.f90:73-346   molecular Slater--Condon rules
//   src/molecular_integrals.F90:736-845   8-fold packed two-body index
//   src/spawning.F90:215-315,680-766, src/stoch_utils.f90:51-136
//   src/death.f90:11-130, src/qmc_common.F90:379-406, src/ifciqmc.f90:14-57
// Determinants are bit strings of up to FQ_MAX_WORDS 64-bit words; orbital o
// (1-based) is bit (o-1)%64 of word (o-1)/64; odd orbitals are alpha.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fqmc.h"

namespace fqd {

constexpr double kDepsilon = 1.e-12;
constexpr uint64_t kAlphaBits = 0x5555555555555555ULL;  // odd orbitals
constexpr uint64_t kBetaBits = 0xAAAAAAAAAAAAAAAAULL;   // even orbitals
constexpr int kMaxRejection = 1 << 20;                  // safety cap on rejection loops

// System tables and run parameters, passed by value to every kernel.
struct Sys {
  int system, nbasis, nel, nalpha, nbeta, nvirt, nvirt_alpha, nvirt_beta, W, S;
  uint64_t f0[FQ_MAX_WORDS];
  uint64_t valid[FQ_MAX_WORDS];  // bits of existing orbitals
  double H00, coulomb_k;
  int nsym;
  const int* sym_table;
  const int* inv_sym;
  const double* sp_eigv;
  int norb, nsym_pg, max_sym_spin;
  const int* orb_sym;
  const int* nbss;
  const int* ssbf;
  double ecore;
  const double* h1;
  const double* eri;
  const double* jmat;  // norb*norb: (pp|qq) = eri(i,j,i,j) for spatial p,q
  const double* kmat;  // norb*norb: (pq|qp) = eri(i,j,j,i)
  const double* s1j;   // norb^3: [p][a][o] = (pa|oo), the Coulomb terms of a single p -> a
  const double* s1k;   // norb^3: [p][a][o] = (po|oa), its exchange terms
  double psingle, pdouble;
  // recip[k] = 1.0 / k for 0 < k < nrecip, and jpair[ind] the second electron
  // index int(1.5 + sqrt(2*ind - 1.75)) of pair index ind of choose_ij_mol,
  // both evaluated on the host (the same correctly rounded values)
  const double* recip;
  int nrecip;
  const int* jpair;
  // Heat-bath tables (excit_gen FQ_EXCIT_HEAT_BATH), i-major over 1-based spin
  // orbitals: hb_ija[hb3(i,j,a)] = hb_ija%weights(a,j,i), hb_ija_tot[hb2(i,j)] =
  // hb_ija%weights_tot(j,i), hb_ijab[hb4(i,j,a,b)] = hb_ijab%weights(b,a,j,i), ...
  const double* hb_i;
  const double* hb_ij;
  const double* hb_ija;
  const double* hb_ija_tot;
  const double* hb_ija_u;
  const int* hb_ija_k;
  const double* hb_ijab;
  const double* hb_ijab_tot;
  const double* hb_ijab_u;
  const int* hb_ijab_k;
  // Learned proposal (excit_gen FQ_EXCIT_LEARNED): row d of mlp_tab holds the
  // hole weights (nbasis) then the particle weights (nbasis) of determinant d
  // of the current main list; mlp_aff[(h-1)*nbasis + p-1] is the learned
  // hole-particle affinity of hole h and particle p; mix is the
  // uniform-fallback fraction of doubles.
  const float* mlp_tab;
  const float* mlp_aff;
  const float* mlp_flat;  // 2*nbasis ones: the weights of determinants without a row
  // affinity row sums: mlp_rs[h-1] over all orbitals, mlp_crs[(h-1)*ncls + c]
  // over spin/irrep class c = (ims-1)*nsym_pg + sym (ncls = 2*nsym_pg), whose
  // members are the bits of mlp_cmask[c*FQ_MAX_WORDS ..]
  const double* mlp_rs;
  const double* mlp_crs;
  const uint64_t* mlp_cmask;
  // alias tables of the rows: mlp_au/ak[(h-1)*nbasis + p-1]; per class
  // mlp_cu/ck[((h-1)*ncls + c)*max_sym_spin + k] over the class's members
  const double* mlp_au;
  const int* mlp_ak;
  const double* mlp_cu;
  const int* mlp_ck;
  int64_t mlp_ld;
  const int64_t* mlp_on;   // per determinant: 1 if the network ran on it this cycle
  const int64_t* mlp_row;  // its row of mlp_tab (exclusive prefix of mlp_on)
  double mix;
  int initiator, quadrature, real_amps, excit_gen;
  double initiator_pop;
  int64_t rf, spawn_cutoff;
  // Semi-stochastic core (combined annihilation): its ndeterm determinants,
  // sorted as the main list, word-major (determ[w * ndeterm + k]); 0: off.
  const uint64_t* determ;
  int64_t ndeterm;
};

// ---------------------------------------------------------------------------
// Random number sources
// ---------------------------------------------------------------------------
// Replay: consume a host-uploaded close-open stream in order.
struct ReplayRng {
  const double* buf;
  int64_t pos;
  int64_t avail;
  bool failed;
  __device__ double next() {
    if (pos >= avail) {
      failed = true;
      return 0.5;
    }
    return buf[pos++];
  }
};

// Philox4x32-10 (Salmon et al., SC'11), counter = (c0, c1, c2, draw-block).
// The four words of a block are held in named registers and consumed in two
// halves (a private array indexed by a run-time position would live in
// scratch memory).
struct Philox {
  uint32_t k0, k1, c0, c1, c2;
  uint32_t block;
  uint32_t o0, o1, o2, o3;
  bool second;  // the block's second half is next
  bool failed;
  __device__ Philox(uint64_t key, uint32_t a, uint32_t b, uint32_t c)
      : k0(static_cast<uint32_t>(key)), k1(static_cast<uint32_t>(key >> 32)), c0(a), c1(b), c2(c), block(0),
        o0(0), o1(0), o2(0), o3(0), second(false), failed(false) {}
  // Each product's high and low words come from one 32x32->64 multiply
  // (v_mad_u64_u32) rather than separate high and low multiplies.
  __device__ static void round(uint32_t* x, uint32_t ka, uint32_t kb) {
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * x[0];
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * x[2];
    const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
    const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
    uint32_t y0 = hi1 ^ x[1] ^ ka, y1 = lo1, y2 = hi0 ^ x[3] ^ kb, y3 = lo0;
    x[0] = y0;
    x[1] = y1;
    x[2] = y2;
    x[3] = y3;
  }
  __device__ void generate() {
    uint32_t x[4] = {c0, c1, c2, block++};
    uint32_t ka = k0, kb = k1;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
      round(x, ka, kb);
      ka += 0x9E3779B9u;
      kb += 0xBB67AE85u;
    }
    o0 = x[0];
    o1 = x[1];
    o2 = x[2];
    o3 = x[3];
  }
  // Uniform double in [0,1) on the 2^-52 grid (as dSFMT's close-open values),
  // from two 32-bit words.  On that grid r*n < n rounds below n for every
  // integer n < 2^52, so int(r*n) is always a valid index in [0, n).
  __device__ double next() {
    uint32_t a, b;
    if (second) {
      a = o2;
      b = o3;
    } else {
      generate();
      a = o0;
      b = o1;
    }
    second = !second;
    uint64_t v = (static_cast<uint64_t>(a >> 6) << 26) | (b >> 6);
    return static_cast<double>(v) * (1.0 / 4503599627370496.0);
  }
};

// The same stream as Philox, with its first K blocks generated at
// construction, where a wavefront's lanes are still converged: the rejection
// loops of the generators then consume stored words (a move of the remaining
// words per draw) instead of generating blocks in divergent code, where a
// block costs the whole wavefront whenever any lane needs one.  Draws past
// the K blocks are generated one block at a time, as Philox does.
template <int K>
struct PhiloxPre {
  uint32_t k0, k1, c0, c1, c2;
  uint32_t block;
  uint32_t w[4 * K];  // unconsumed words, the next draw's pair first
  int left;           // draws (word pairs) left in w
  bool failed;
  __device__ void gen(uint32_t* o) {
    uint32_t x[4] = {c0, c1, c2, block++};
    uint32_t ka = k0, kb = k1;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
      Philox::round(x, ka, kb);
      ka += 0x9E3779B9u;
      kb += 0xBB67AE85u;
    }
    o[0] = x[0];
    o[1] = x[1];
    o[2] = x[2];
    o[3] = x[3];
  }
  __device__ PhiloxPre(uint64_t key, uint32_t a, uint32_t b, uint32_t c)
      : k0(static_cast<uint32_t>(key)), k1(static_cast<uint32_t>(key >> 32)), c0(a), c1(b), c2(c), block(0),
        left(2 * K), failed(false) {
#pragma unroll
    for (int q = 0; q < K; ++q) gen(w + 4 * q);
  }
  __device__ double next() {
    if (left == 0) {
      gen(w);
      left = 2;
    }
    const uint32_t a = w[0], b = w[1];
#pragma unroll
    for (int q = 0; q + 2 < 4 * K; ++q) w[q] = w[q + 2];
    --left;
    uint64_t v = (static_cast<uint64_t>(a >> 6) << 26) | (b >> 6);
    return static_cast<double>(v) * (1.0 / 4503599627370496.0);
  }
};

// Draw purposes: the third counter word of a Philox stream.
enum : uint32_t {
  kDrawAttempt = 1,
  kDrawNattempts = 2,
  kDrawDeath = 3,
  kDrawRoundMain = 4,
  kDrawRoundSpawn = 5,
  kDrawTrain = 6,
  kDrawProbe = 7,
  kDrawDeterm = 8  // rounding of the deterministic projection's spawns
};

// ---------------------------------------------------------------------------
// Bit-string helpers
// ---------------------------------------------------------------------------
// Determinant words are addressed with compile-time indices only (selects
// and loops unrolled to FQ_MAX_WORDS), so they stay in registers.
static_assert(FQ_MAX_WORDS == 2, "bit helpers select between two words");
#define FQ_FOR_WORDS(w, W) _Pragma("unroll") for (int w = 0; w < FQ_MAX_WORDS; ++w) if (w < (W))
__device__ __forceinline__ bool btest(const uint64_t* f, int orb) {
  const int b = orb - 1;
  const uint64_t x = (b >> 6) == 0 ? f[0] : f[1];
  return (x >> (b & 63)) & 1;
}
__device__ __forceinline__ void bset(uint64_t* f, int orb) {
  const int b = orb - 1;
  const uint64_t m = uint64_t{1} << (b & 63);
  if ((b >> 6) == 0)
    f[0] |= m;
  else
    f[1] |= m;
}
__device__ __forceinline__ void bclr(uint64_t* f, int orb) {
  const int b = orb - 1;
  const uint64_t m = ~(uint64_t{1} << (b & 63));
  if ((b >> 6) == 0)
    f[0] &= m;
  else
    f[1] &= m;
}
// Word w of excit_mask(:, orb): bits of all orbitals after orb
// (src/excitations.F90:17-45).
__device__ __forceinline__ uint64_t after_mask(int orb, int w) {
  const int b = orb - 1, bw = b >> 6, bb = b & 63;
  if (w < bw) return 0;
  if (w > bw) return ~uint64_t{0};
  return bb == 63 ? 0 : (~uint64_t{0} << (bb + 1));
}
// k-th (0-based) set bit of x (k < popcount(x)): a fixed descent over the
// halves of the word by popcount (no loop whose trip count differs by lane).
__host__ __device__ __forceinline__ int kth_bit(uint64_t x, int k) {
  int pos = 0;
  int c = __builtin_popcount(static_cast<uint32_t>(x));
  if (k >= c) {
    k -= c;
    x >>= 32;
    pos = 32;
  }
  uint32_t v = static_cast<uint32_t>(x);
  c = __builtin_popcount(v & 0xffffu);
  if (k >= c) {
    k -= c;
    v >>= 16;
    pos += 16;
  }
  c = __builtin_popcount(v & 0xffu);
  if (k >= c) {
    k -= c;
    v >>= 8;
    pos += 8;
  }
  c = __builtin_popcount(v & 0xfu);
  if (k >= c) {
    k -= c;
    v >>= 4;
    pos += 4;
  }
  c = __builtin_popcount(v & 0x3u);
  if (k >= c) {
    k -= c;
    v >>= 2;
    pos += 2;
  }
  if (k >= static_cast<int>(v & 1u)) pos += 1;
  return pos;
}
// Orbital (1-based) of the k-th set bit of (words & pattern) across W words.
__device__ __forceinline__ int select_orb(const uint64_t* x, int W, int k) {
  FQ_FOR_WORDS(w, W) {
    const int c = __popcll(x[w]);
    if (k < c) return (w << 6) + kth_bit(x[w], k) + 1;
    k -= c;
  }
  return 0;
}
__device__ __forceinline__ int64_t fsign64(int64_t a, int64_t b) {
  const int64_t m = a < 0 ? -a : a;
  return b >= 0 ? m : -m;
}
__device__ __forceinline__ int ms_of(int orb) { return (orb & 1) ? 1 : -1; }
__device__ __forceinline__ int ims_of(int orb) { return (orb & 1) ? 2 : 1; }  // 1 = beta, 2 = alpha

struct Excit {
  int nexcit;
  int from[2];
  int to[2];
  bool perm;
};

// get_excitation(f1, f2) (src/excitations.F90:75-200), bit-parallel form:
// the running electron counts at a differing bit are prefix popcounts.
__device__ inline Excit get_excitation(const uint64_t* f1, const uint64_t* f2, int W, int nel) {
  Excit e;
  e.nexcit = 0;
  e.from[0] = e.from[1] = e.to[0] = e.to[1] = 0;
  e.perm = false;
  int nx = 0;
  FQ_FOR_WORDS(w, W) nx += __popcll(f1[w] ^ f2[w]);
  if (nx == 0) return e;
  e.nexcit = nx / 2;
  if (e.nexcit > 2) return e;
  const int shift = nel - e.nexcit;
  int ie1 = 0, ie2 = 0, iel1 = 0, iel2 = 0, perm = 0;
  FQ_FOR_WORDS(w, W) {
    const uint64_t a = f1[w], b = f2[w];
    if (a == b) {
      if ((((ie1 - ie2) % 2) + 2) % 2 == 1) {
        const int n = __popcll(a);
        iel1 += n;
        iel2 += n;
      }
      continue;
    }
    uint64_t d = a ^ b;
    while (d) {
      const int j = __builtin_ctzll(d);
      d &= d - 1;
      const uint64_t le = (j == 63) ? ~uint64_t{0} : ((uint64_t{2} << j) - 1);
      const int orb = (w << 6) + j + 1;
      if ((a >> j) & 1) {
        const int c1 = iel1 + __popcll(a & le);
        ++ie1;
        e.from[ie1 - 1] = orb;
        perm += shift - c1 + ie1;
      } else {
        const int c2 = iel2 + __popcll(b & le);
        ++ie2;
        e.to[ie2 - 1] = orb;
        perm += shift - c2 + ie2;
      }
    }
    iel1 += __popcll(a);
    iel2 += __popcll(b);
  }
  e.perm = (perm % 2) == 1;
  return e;
}

__device__ inline bool perm1(const uint64_t* f, int W, int from, int to) {
  int perm = 0;
  FQ_FOR_WORDS(w, W) perm += __popcll(f[w] & (after_mask(from, w) ^ after_mask(to, w)));
  if (from > to) perm -= 1;
  return (perm % 2) == 1;
}
__device__ inline bool perm2(const uint64_t* f, int W, const Excit& e) {
  int perm = 0;
  FQ_FOR_WORDS(w, W) {
    const uint64_t ia = after_mask(e.from[0], w) ^ after_mask(e.to[0], w);
    const uint64_t jb = after_mask(e.from[1], w) ^ after_mask(e.to[1], w);
    perm += __popcll((f[w] & ia) ^ (f[w] & jb));
  }
  if (e.from[0] > e.to[0]) perm += 1;
  if (e.from[0] > e.to[1]) perm += 1;
  if (e.from[1] > e.to[1] || e.from[1] < e.to[0]) perm += 1;
  return (perm % 2) == 1;
}

// ---------------------------------------------------------------------------
// Hubbard model, momentum space
// ---------------------------------------------------------------------------
__device__ __forceinline__ int st(const Sys& s, int a, int b) { return s.sym_table[(a - 1) * s.nsym + (b - 1)]; }

__device__ inline double two_e_hub_k(const Sys& s, int p1, int p2, int p3, int p4) {
  const bool s12 = ms_of(p1) == ms_of(p2), s34 = ms_of(p3) == ms_of(p4);
  int dk = st(s, (p1 + 1) / 2, (p2 + 1) / 2);
  dk = st(s, dk, s.inv_sym[(p3 + 1) / 2 - 1]);
  if (!s12 && (s12 == s34) && dk == (p4 + 1) / 2) return ms_of(p1) == ms_of(p3) ? s.coulomb_k : -s.coulomb_k;
  return 0.0;
}

// ---------------------------------------------------------------------------
// Molecular integrals (RHF; spatial index (o+1)/2)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double h1g(const Sys& s, int i, int j) {
  return s.h1[((i + 1) / 2 - 1) * s.norb + ((j + 1) / 2 - 1)];
}
// Packed 8-fold index in 32-bit arithmetic: with at most 128 spin orbitals
// (64 spatial) pair indices stay below 2081 and packed indices below 2.2e6.
__device__ __forceinline__ uint32_t tri32(uint32_t i, uint32_t j) { return ((i * (i - 1)) >> 1) + j; }
__device__ __forceinline__ double eri_nz(const Sys& s, int i, int j, int a, int b) {
  const int ii = i > a ? i : a, aa = i > a ? a : i, jj = j > b ? j : b, bb = j > b ? b : j;
  const uint32_t ia = tri32((ii + 1) >> 1, (aa + 1) >> 1), jb = tri32((jj + 1) >> 1, (bb + 1) >> 1);
  return s.eri[(ia < jb ? tri32(jb, ia) : tri32(ia, jb)) - 1];
}
// slater_condon1_mol_excit: the terms <io|ao> and <io|oa> are read from the
// (i, a) rows of the single-excitation tables (the packed store's values).
__device__ inline double sc1_excit(const Sys& s, const uint64_t* f, int i, int a, bool perm) {
  double h = h1g(s, i, a);
  const int64_t row = (static_cast<int64_t>((i + 1) / 2 - 1) * s.norb + ((a + 1) / 2 - 1)) * s.norb;
  const double* __restrict__ rj = s.s1j + row;
  const double* __restrict__ rk = s.s1k + row;
  FQ_FOR_WORDS(w, s.W) {
    uint64_t x = f[w];
    while (x) {
      const int o = (w << 6) + __builtin_ctzll(x) + 1;
      x &= x - 1;
      if (o != i) {
        const int po = (o + 1) / 2 - 1;
        h = h + rj[po];
        if (ms_of(o) == ms_of(i)) h = h - rk[po];
      }
    }
  }
  if (perm) h = -h;
  return h;
}
__device__ inline double sc2_excit(const Sys& s, int i, int j, int a, int b, bool perm) {
  double h = 0.0;
  if (ms_of(i) == ms_of(a)) h = eri_nz(s, i, j, a, b);
  if (ms_of(i) == ms_of(b)) h = h - eri_nz(s, i, j, b, a);
  if (perm) h = -h;
  return h;
}

// <D|H|D> in the reference's summation order (slater_condon0_*).
__device__ inline double diagonal_mol(const Sys& s, const uint64_t* f, const double* J, const double* K);
__device__ inline double diagonal(const Sys& s, const uint64_t* f);

// Index of f among the semi-stochastic core determinants, -1 if it is not one
// (check_if_determ, src/semi_stoch.F90:795-824, as a binary search of the
// sorted core).
__device__ inline int64_t core_index(const Sys& s, const uint64_t* f) {
  int64_t lo = 0, hi = s.ndeterm;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    int c = 0;
#pragma unroll
    for (int w = FQ_MAX_WORDS - 1; w >= 0; --w)
      if (w < s.W && c == 0) {
        const uint64_t x = s.determ[w * s.ndeterm + mid];
        c = x < f[w] ? -1 : (x > f[w] ? 1 : 0);
      }
    if (c == 0) return mid;
    if (c < 0)
      lo = mid + 1;
    else
      hi = mid;
  }
  return -1;
}

// get_hmatel(f1, f2) (src/hamiltonian_molecular.f90:12-71; Hubbard-k in
// momentum space: only double excitations couple): the Slater--Condon element
// by excitation level, single excitations over f1's occupied orbitals.
__device__ inline double hmatel(const Sys& s, const uint64_t* f1, const uint64_t* f2) {
  const Excit e = get_excitation(f1, f2, s.W, s.nel);
  if (e.nexcit == 0) return diagonal(s, f1);
  if (s.system == FQ_SYS_HUB_K) {
    if (e.nexcit != 2) return 0.0;
    const double h = two_e_hub_k(s, e.from[0], e.from[1], e.to[0], e.to[1]);
    return e.perm ? -h : h;
  }
  if (e.nexcit == 1) return sc1_excit(s, f1, e.from[0], e.to[0], e.perm);
  if (e.nexcit == 2) return sc2_excit(s, e.from[0], e.from[1], e.to[0], e.to[1], e.perm);
  return 0.0;
}
__device__ inline double diagonal(const Sys& s, const uint64_t* f) {
  if (s.system == FQ_SYS_HUB_K) {
    double h = static_cast<double>(s.nalpha * s.nbeta) * s.coulomb_k;
    FQ_FOR_WORDS(w, s.W) {
      uint64_t x = f[w];
      while (x) {
        const int o = (w << 6) + __builtin_ctzll(x) + 1;
        x &= x - 1;
        h = h + s.sp_eigv[o - 1];
      }
    }
    return h;
  }
  return diagonal_mol(s, f, s.jmat, s.kmat);
}

// Molecular <D|H|D> with the dense Coulomb/exchange pair tables J, K (the
// packed store's values) wherever they are held (global memory or LDS).
__device__ inline double diagonal_mol(const Sys& s, const uint64_t* f, const double* J, const double* K) {
  // Coulomb and exchange elements come from dense per-pair tables holding the
  // same values as the packed store, in the reference's summation order.
  double h = s.ecore;
  FQ_FOR_WORDS(w, s.W) {
    uint64_t x = f[w];
    while (x) {
      const int i = (w << 6) + __builtin_ctzll(x) + 1;
      x &= x - 1;
      h = h + h1g(s, i, i);
      const int pi = ((i + 1) / 2 - 1) * s.norb;
      // pairs (i, j) with j after i
      FQ_FOR_WORDS(w2, s.W) {
        if (w2 < w) continue;
        uint64_t y = (w2 == w) ? x : f[w2];
        while (y) {
          const int j = (w2 << 6) + __builtin_ctzll(y) + 1;
          y &= y - 1;
          const int pj = (j + 1) / 2 - 1;
          h = h + J[pi + pj];
          if (ms_of(i) == ms_of(j)) h = h - K[pi + pj];
        }
      }
    }
  }
  return h;
}

// Projected-energy contribution of one determinant (update_proj_energy_*).
__device__ inline void proj_energy_term(const Sys& s, const uint64_t* f, double pop, double& proj, double& d0) {
  Excit c = get_excitation(f, s.f0, s.W, s.nel);
  if (c.nexcit == 0) {
    d0 = d0 + pop;
    return;
  }
  if (s.system == FQ_SYS_HUB_K) {
    if (c.nexcit == 2) {
      double h = two_e_hub_k(s, c.from[0], c.from[1], c.to[0], c.to[1]);
      if (c.perm) h = -h;
      proj = proj + h * pop;
    }
    return;
  }
  if (c.nexcit == 1) {
    if (ms_of(c.from[0]) == ms_of(c.to[0]) && s.orb_sym[c.from[0] - 1] == s.orb_sym[c.to[0] - 1]) {
      const double h = sc1_excit(s, f, c.from[0], c.to[0], c.perm);
      proj = proj + h * pop;
    }
  } else if (c.nexcit == 2) {
    if (ms_of(c.from[0]) + ms_of(c.from[1]) == ms_of(c.to[0]) + ms_of(c.to[1]) &&
        (s.orb_sym[c.from[0] - 1] ^ s.orb_sym[c.from[1] - 1]) == (s.orb_sym[c.to[0] - 1] ^ s.orb_sym[c.to[1] - 1])) {
      const double h = sc2_excit(s, c.from[0], c.from[1], c.to[0], c.to[1], c.perm);
      proj = proj + h * pop;
    }
  }
}

__device__ inline int64_t parent_flag(const Sys& s, const double* pop) {
  int64_t flag = 0;
  if (!s.initiator) return 0;
  const double ip = s.initiator_pop;
  if (s.quadrature && (s.S % 2) == 0) {
    for (int i = 0; i < s.S; i += 2) {
      const double mag = pop[i] * pop[i] + pop[i + 1] * pop[i + 1];
      if (!(mag > ip * ip)) flag |= (int64_t{1} << i) | (int64_t{1} << (i + 1));
    }
  } else {
    for (int i = 0; i < s.S; ++i)
      if (!(fabs(pop[i]) > ip)) flag |= int64_t{1} << i;
  }
  return flag;
}

template <class Rng>
__device__ inline int decide_nattempts(Rng& rng, double population) {
  int n = abs(static_cast<int>(population));
  const double pextra = fabs(population) - n;
  if (fabs(pextra) > kDepsilon) {
    if (pextra > rng.next()) n = n + 1;
  }
  return n;
}

template <class Rng>
__device__ inline int64_t stoch_round_spawn(const Sys& s, Rng& rng, double pspawn) {
  const int64_t cutoff = s.spawn_cutoff;
  int64_t nspawn;
  if (pspawn < static_cast<double>(cutoff)) {
    nspawn = (pspawn > rng.next() * static_cast<double>(cutoff)) ? cutoff : 0;
  } else {
    nspawn = static_cast<int64_t>(pspawn);
    const double padd = pspawn - static_cast<double>(nspawn);
    if (padd > rng.next()) nspawn = nspawn + 1;
  }
  return nspawn;
}

// stochastic_death (src/death.f90:11-130) with the FCIQMC defaults
// (quasi-Newton weight 1, pop control 1, dmqmc factor 1, Chebyshev weight 1).
template <class Rng>
__device__ inline int64_t death(Rng& rng, double tau, double Kii, double shift, double E, int64_t& population) {
  const double weight = 1.0, qnpc = 1.0, dmqmc = 1.0, cheby = 1.0;
  double pd = tau * ((Kii - E) * weight + (E - shift) * qnpc) * dmqmc;
  pd = pd * cheby;
  pd = pd * static_cast<double>(population < 0 ? -population : population);
  int64_t kill = static_cast<int64_t>(pd);
  pd = pd - static_cast<double>(kill);
  const double r = rng.next();
  if (fabs(pd) > r) kill = pd > 0.0 ? kill + 1 : kill - 1;
  if (population < 0)
    population = population + kill;
  else
    population = population - kill;
  return kill < 0 ? -kill : kill;
}

// stochastic_round (src/stoch_utils.f90:51-85) of one population.
template <class Rng>
__device__ inline int64_t stoch_round_pop(Rng& rng, int64_t p, int64_t rf) {
  const int64_t ap = p < 0 ? -p : p;
  if (ap < rf && p != 0) {
    const double r = rng.next() * static_cast<double>(rf);
    return (static_cast<double>(ap) > r) ? fsign64(rf, p) : 0;
  }
  return p;
}

// A deterministic-projection amplitude encoded with one stochastic rounding
// draw (create_spawned_particle_determ, src/semi_stoch.F90:1096-1146;
// deterministic_annihilation, src/annihilation.f90:488-535): the integer part
// of |amp| * rf, one more with probability its fraction, with amp's sign.
template <class Rng>
__device__ inline int64_t stoch_encode(Rng& rng, double amp, int64_t rf) {
  const double scaled = amp * static_cast<double>(rf);
  const double a = fabs(scaled);
  int64_t k = static_cast<int64_t>(a);
  if (a - static_cast<double>(k) > rng.next()) k = k + 1;
  return scaled < 0.0 ? -k : k;
}

// ---------------------------------------------------------------------------
// Spawning
// ---------------------------------------------------------------------------
// Result of one spawning attempt.
struct SpawnResult {
  int64_t nspawn;  // encoded, signed
  uint64_t child[FQ_MAX_WORDS];
};

__device__ __forceinline__ void excited_det(const uint64_t* f, int W, const Excit& c, uint64_t* out) {
  FQ_FOR_WORDS(w, W) out[w] = f[w];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    if (i < c.nexcit) {
      bclr(out, c.from[i]);
      bset(out, c.to[i]);
    }
  }
}

// Hubbard-k split generator (spawn_lattice_split_gen with gen_excit_init_hub_k /
// gen_excit_finalise_hub_k, or the no_renorm pair).
template <class Rng>
__device__ inline int64_t spawn_hub_k(const Sys& s, Rng& rng, const uint64_t* f, int64_t parent, double tau,
                                      Excit& conn) {
  uint64_t occa[FQ_MAX_WORDS], occb[FQ_MAX_WORDS], unoa[FQ_MAX_WORDS];
  FQ_FOR_WORDS(w, s.W) {
    occa[w] = f[w] & kAlphaBits;
    occb[w] = f[w] & kBetaBits;
    unoa[w] = ~f[w] & kAlphaBits & s.valid[w];
  }
  // choose_ij_hub_k
  const double r = rng.next();
  const int na = s.nalpha, nb = s.nbeta;
  const int ind = static_cast<int>(r * na * nb) + 1;
  int i = static_cast<int>((ind - 1.0) / nb) + 1;
  int j = ind - (i - 1) * nb;
  i = select_orb(occa, s.W, i - 1);
  j = select_orb(occb, s.W, j - 1);
  const int ij_sym = st(s, (i + 1) / 2, (j + 1) / 2);
  conn.nexcit = 2;
  conn.from[0] = i;
  conn.from[1] = j;
  conn.to[0] = conn.to[1] = 0;
  conn.perm = false;
  double pgen, absh;
  if (s.excit_gen == FQ_EXCIT_NO_RENORM) {
    const int rr = static_cast<int>(rng.next() * s.nvirt_alpha) + 1;
    const int a = select_orb(unoa, s.W, rr - 1);
    const int b = 2 * st(s, ij_sym, s.inv_sym[(a + 1) / 2 - 1]);
    conn.to[0] = a;
    conn.to[1] = b;
    if (!btest(f, b)) {
      pgen = 1.0 / static_cast<double>(s.nalpha * s.nbeta * s.nvirt_alpha);
      absh = fabs(s.coulomb_k);
    } else {
      pgen = 1.0;
      absh = 0.0;
    }
  } else {
    int forbidden = 0;
    FQ_FOR_WORDS(w, s.W) {
      uint64_t x = unoa[w];
      while (x) {
        const int a = (w << 6) + __builtin_ctzll(x) + 1;
        x &= x - 1;
        const int b = 2 * st(s, ij_sym, s.inv_sym[(a + 1) / 2 - 1]);
        if (btest(f, b)) ++forbidden;
      }
    }
    pgen = 1.0 / static_cast<double>(s.nalpha * s.nbeta * (s.nvirt_alpha - forbidden));
    absh = fabs(s.coulomb_k);
  }
  int64_t nspawn = stoch_round_spawn(s, rng, static_cast<double>(s.rf) * tau * absh / pgen);
  if (nspawn != 0) {
    if (s.excit_gen != FQ_EXCIT_NO_RENORM) {
      const int sym2 = st(s, (conn.from[0] + 1) / 2, (conn.from[1] + 1) / 2);
      int a = 0, b = 0;
      for (int tries = 0; tries < kMaxRejection; ++tries) {
        const int rr = static_cast<int>(rng.next() * s.nvirt_alpha) + 1;
        a = select_orb(unoa, s.W, rr - 1);
        b = 2 * st(s, sym2, s.inv_sym[(a + 1) / 2 - 1]);
        if (!btest(f, b) || rng.failed) break;
      }
      conn.to[0] = a;
      conn.to[1] = b;
    }
    // slater_condon2_hub_k_excit
    if (conn.from[0] > conn.from[1]) {
      const int t = conn.from[0];
      conn.from[0] = conn.from[1];
      conn.from[1] = t;
    }
    if (conn.to[0] > conn.to[1]) {
      const int t = conn.to[0];
      conn.to[0] = conn.to[1];
      conn.to[1] = t;
    }
    double h = s.coulomb_k;
    conn.perm = perm2(f, s.W, conn);
    if (conn.perm) h = -h;
    if (((conn.from[0] - conn.to[0]) % 2) != 0) h = -h;
    nspawn = (h > 0.0) ? -fsign64(nspawn, parent) : fsign64(nspawn, parent);
  }
  return nspawn;
}

// Unoccupied counts per (spin, irrep), packed 8 bits per irrep (<= 8 irreps).
struct SymUnocc {
  uint64_t pk[2];  // [0] beta, [1] alpha
  __device__ int get(int ims, int sym) const {
    return static_cast<int>(((ims == 1 ? pk[0] : pk[1]) >> (8 * sym)) & 0xff);
  }
};
__device__ inline SymUnocc sym_unocc(const Sys& s, const uint64_t* f) {
  SymUnocc u;
  u.pk[0] = u.pk[1] = 0;
  if (s.nsym_pg == 1) {  // one irrep: unoccupied counts per spin from popcounts
    int na = 0, nb = 0;
    FQ_FOR_WORDS(w, s.W) {
      na += __popcll(f[w] & kAlphaBits);
      nb += __popcll(f[w] & kBetaBits);
    }
    u.pk[0] = static_cast<uint64_t>(s.nbss[0] - nb);
    u.pk[1] = static_cast<uint64_t>(s.nbss[1] - na);
    return u;
  }
  for (int g = 0; g < s.nsym_pg; ++g) {
    u.pk[0] |= static_cast<uint64_t>(s.nbss[g]) << (8 * g);
    u.pk[1] |= static_cast<uint64_t>(s.nbss[s.nsym_pg + g]) << (8 * g);
  }
  FQ_FOR_WORDS(w, s.W) {
    uint64_t x = f[w];
    while (x) {
      const int o = (w << 6) + __builtin_ctzll(x) + 1;
      x &= x - 1;
      const uint64_t d = uint64_t{1} << (8 * s.orb_sym[o - 1]);
      if (ims_of(o) == 1)
        u.pk[0] -= d;
      else
        u.pk[1] -= d;
    }
  }
  return u;
}
// 1.0 / k: the host-made table's entry when it holds k, else the division
__device__ __forceinline__ double recip_of(const Sys& s, int k) {
  return k > 0 && k < s.nrecip ? s.recip[k] : 1.0 / static_cast<double>(k);
}
__device__ __forceinline__ int ssbf(const Sys& s, int ims, int sym, int k) {
  return s.ssbf[((ims - 1) * s.nsym_pg + sym) * s.max_sym_spin + k];
}
__device__ __forceinline__ int nbss(const Sys& s, int ims, int sym) { return s.nbss[(ims - 1) * s.nsym_pg + sym]; }

// gen_single_excit_mol / choose_ia_mol (src/excit_gen_mol.f90:384-448,521-608):
// renormalised single excitation; sets h and pgen (pgen includes psingle).
template <class Rng>
__device__ inline bool single_mol(const Sys& s, Rng& rng, const uint64_t* f, const SymUnocc u, Excit& conn,
                                  double& h, double& pgen) {
  const int nel = s.nel;
  conn.nexcit = 1;
  bool allowed = false;
  int ni = nel;
  FQ_FOR_WORDS(w, s.W) {
    uint64_t x = f[w];
    while (x) {
      const int o = (w << 6) + __builtin_ctzll(x) + 1;
      x &= x - 1;
      if (u.get(ims_of(o), s.orb_sym[o - 1]) != 0)
        allowed = true;
      else
        ni = ni - 1;
    }
  }
  if (allowed) {
    // i until its spin/irrep has an unoccupied orbital, then a within it (two
    // loops in sequence: the same draws as the nested form, one pass each)
    int i = 0, a = 0;
    bool found = false;
    for (int tries = 0; tries < kMaxRejection; ++tries) {
      i = select_orb(f, s.W, static_cast<int>(rng.next() * nel));
      if (u.get(ims_of(i), s.orb_sym[i - 1]) != 0) {
        found = true;
        break;
      }
      if (rng.failed) break;
    }
    if (found) {
      const int imsa = ims_of(i), isyma = s.orb_sym[i - 1];
      for (int t2 = 0; t2 < kMaxRejection; ++t2) {
        const int ind = static_cast<int>(nbss(s, imsa, isyma) * rng.next()) + 1;
        a = ssbf(s, imsa, isyma, ind - 1);
        if (!btest(f, a) || rng.failed) break;
      }
      conn.from[0] = i;
      conn.to[0] = a;
      pgen = s.psingle * recip_of(s, ni * u.get(ims_of(a), s.orb_sym[a - 1]));
      conn.perm = perm1(f, s.W, i, a);
      h = sc1_excit(s, f, i, a, conn.perm);
    } else {  // draws ran out (replay stream exhausted) before an i was found: no excitation
      h = 0.0;
      pgen = 1.0;
      allowed = false;
    }
  } else {
    h = 0.0;
    pgen = 1.0;
  }
  return allowed;
}

// attempt_to_spawn (src/spawning.F90:711-766): one draw, sign opposite to the
// parent's when h > 0.
template <class Rng>
__device__ inline int64_t attempt_to_spawn(const Sys& s, Rng& rng, double tau, double h, double pgen,
                                           int64_t parent) {
  double pspawn = tau * fabs(h) / pgen;
  pspawn = pspawn * static_cast<double>(s.rf);
  int64_t nspawn = stoch_round_spawn(s, rng, pspawn);
  if (nspawn > 0) nspawn = (h > 0.0) ? -fsign64(nspawn, parent) : fsign64(nspawn, parent);
  return nspawn;
}

// gen_excit_mol (renormalised molecular generator, src/excit_gen_mol.f90:16-101):
// the excitation, h and pgen; returns whether an allowed excitation was made.
template <class Rng>
__device__ inline bool gen_mol(const Sys& s, Rng& rng, const uint64_t* f, Excit& conn, double& h, double& pgen) {
  const SymUnocc u = sym_unocc(s, f);
  pgen = 1.0;
  h = 0.0;
  bool ok = false;
  conn.from[0] = conn.from[1] = conn.to[0] = conn.to[1] = 0;
  conn.perm = false;
  const int nel = s.nel;
  if (rng.next() < s.psingle) {
    ok = single_mol(s, rng, f, u, conn, h, pgen);
  } else {
    // choose_ij_mol
    const int ind = static_cast<int>(rng.next() * nel * (nel - 1) / 2) + 1;
    const int j_ind = s.jpair[ind];
    const int i_ind = ind - ((j_ind - 1) * (j_ind - 2)) / 2;
    const int i = select_orb(f, s.W, i_ind - 1);
    const int j = select_orb(f, s.W, j_ind - 1);
    const int ij_sym = s.orb_sym[i - 1] ^ s.orb_sym[j - 1];
    const int spin = ms_of(i) + ms_of(j);
    const double pgen_ij = 2.0 / static_cast<double>(nel * (nel - 1));
    conn.nexcit = 2;
    conn.from[0] = i;
    conn.from[1] = j;
    // choose_ab_mol
    bool allowed = false;
    for (int isyma = 0; isyma < s.nsym_pg && !allowed; ++isyma) {
      const int isymb = isyma ^ ij_sym;
      if (spin == -2) {
        allowed = u.get(1, isyma) > 0 && (u.get(1, isymb) > 1 || (u.get(1, isymb) == 1 && isyma != isymb));
      } else if (spin == 0) {
        allowed = (u.get(1, isyma) > 0 && u.get(2, isymb) > 0) || (u.get(2, isyma) > 0 && u.get(1, isymb) > 0);
      } else {
        allowed = u.get(2, isyma) > 0 && (u.get(2, isymb) > 1 || (u.get(2, isymb) == 1 && isyma != isymb));
      }
    }
    const int fac = spin == 0 ? 1 : 2, shift = spin == 2 ? 1 : 0, na = spin == 0 ? s.nbasis : s.nbasis / 2;
    int a = 0, b = 0;
    bool found = false;
    if (allowed) {
      // a until it is unoccupied and has a possible partner (a free a without
      // one is redrawn), then b: two loops in sequence drawing exactly as the
      // reference's nested form
      for (int tries = 0; tries < kMaxRejection; ++tries) {
        a = static_cast<int>(rng.next() * na) + 1;
        a = fac * a - shift;
        if (!btest(f, a)) {
          const int imsb = (spin - ms_of(a) + 3) / 2;
          const int isymb = ij_sym ^ s.orb_sym[a - 1];
          const int nu = u.get(imsb, isymb);
          if (nu > 1 || (nu == 1 && (isymb != s.orb_sym[a - 1] || spin == 0))) {
            found = true;
            break;
          }
        }
        if (rng.failed) break;
      }
      if (found) {
        const int imsb = (spin - ms_of(a) + 3) / 2;
        const int isymb = ij_sym ^ s.orb_sym[a - 1];
        for (int t2 = 0; t2 < kMaxRejection; ++t2) {
          const int indb = static_cast<int>(nbss(s, imsb, isymb) * rng.next()) + 1;
          b = ssbf(s, imsb, isymb, indb - 1);
          if ((b != a && !btest(f, b)) || rng.failed) break;
        }
      }
    }
    if (allowed && !found) allowed = false;  // draws ran out before an a was found: no excitation
    if (allowed) {
      if (a > b) {
        const int t = a;
        a = b;
        b = t;
      }
      conn.to[0] = a;
      conn.to[1] = b;
      // calc_pgen_double_mol
      const int imsa = ims_of(a), imsb = ims_of(b);
      int n_aij;
      double p_bija, p_aijb;
      if (spin != 0) {
        const int ims = spin == -2 ? 1 : 2;
        n_aij = spin == -2 ? s.nvirt_beta : s.nvirt_alpha;
        for (int isyma = 0; isyma < s.nsym_pg; ++isyma) {
          const int isymb = isyma ^ ij_sym;
          if (u.get(ims, isymb) == 0)
            n_aij = n_aij - u.get(ims, isyma);
          else if (isyma == isymb && u.get(ims, isymb) == 1)
            n_aij = n_aij - u.get(ims, isyma);
        }
        if (s.orb_sym[a - 1] == s.orb_sym[b - 1]) {
          p_aijb = recip_of(s, u.get(imsa, s.orb_sym[a - 1]) - 1);
          p_bija = recip_of(s, u.get(imsb, s.orb_sym[b - 1]) - 1);
        } else {
          p_aijb = recip_of(s, u.get(imsa, s.orb_sym[a - 1]));
          p_bija = recip_of(s, u.get(imsb, s.orb_sym[b - 1]));
        }
      } else {
        n_aij = s.nvirt;
        for (int isyma = 0; isyma < s.nsym_pg; ++isyma) {
          const int isymb = isyma ^ ij_sym;
          if (u.get(1, isymb) == 0) n_aij = n_aij - u.get(2, isyma);
          if (u.get(2, isymb) == 0) n_aij = n_aij - u.get(1, isyma);
        }
        p_aijb = recip_of(s, u.get(imsa, s.orb_sym[a - 1]));
        p_bija = recip_of(s, u.get(imsb, s.orb_sym[b - 1]));
      }
      const double pd = recip_of(s, n_aij) * (p_bija + p_aijb);
      pgen = s.pdouble * pgen_ij * pd;
      conn.perm = perm2(f, s.W, conn);
      h = sc2_excit(s, conn.from[0], conn.from[1], conn.to[0], conn.to[1], conn.perm);
      ok = true;
    } else {
      h = 0.0;
      pgen = 1.0;
    }
  }
  return ok;
}

// spawn_standard with the renormalised generator (src/spawning.F90:34-124).
template <class Rng>
__device__ inline int64_t spawn_mol(const Sys& s, Rng& rng, const uint64_t* f, int64_t parent, double tau,
                                    Excit& conn) {
  double h = 0.0, pgen = 1.0;
  gen_mol(s, rng, f, conn, h, pgen);
  return attempt_to_spawn(s, rng, tau, h, pgen, parent);
}


// ---------------------------------------------------------------------------
// Heat-bath generator (Holmes et al.; src/excit_gen_heat_bath_mol.F90:252-548)
// ---------------------------------------------------------------------------
constexpr int kHbMaxNel = 32;  // electrons a heat-bath determinant may hold here

__device__ __forceinline__ int64_t hb2(const Sys& s, int i, int j) {
  return static_cast<int64_t>(i - 1) * s.nbasis + (j - 1);
}
__device__ __forceinline__ int64_t hb3(const Sys& s, int i, int j, int a) { return hb2(s, i, j) * s.nbasis + (a - 1); }
__device__ __forceinline__ int64_t hb4(const Sys& s, int i, int j, int a, int b) {
  return hb3(s, i, j, a) * s.nbasis + (b - 1);
}

// Per-thread storage for the on-the-fly weights of a heat-bath attempt: entry
// k at w[k * stride] (a column of LDS in the production kernel).
struct HbBuf {
  double* w = nullptr;
  int stride = 1;
  __host__ __device__ double& operator[](int k) const { return w[k * stride]; }
};

// select_weighted_value (lib/local/alias.f90:64-99): generate_alias_tables
// over the n <= 32 weights wt[0..n), then one draw r.  The construction is
// replayed without its tables: the "over" stack only ever loses its top (the
// highest remaining index whose scaled weight exceeds 1), and an index moved
// from it to the "under" stack is popped again by the very next step, so two
// index masks, that one pending index and the running value of the current
// top reproduce every step (same operations, same order).  The drawn column
// kk keeps itself as alias unless it is popped from "under", when its alias
// is the top of "over" at that step and its threshold its value then.
__host__ __device__ inline int alias_pick(int n, const HbBuf& wt, double tot, double r) {
  const double scale = static_cast<double>(n) / tot;
  double x = r * n;
  int kk = static_cast<int>(floor(x));
  x = x - kk;
  kk = kk + 1;
  uint32_t under = 0, over = 0;
  for (int i = 1; i <= n; ++i) {
    if (wt[i - 1] * scale <= 1.0)
      under |= 1u << (i - 1);
    else
      over |= 1u << (i - 1);
  }
  int ov = 0, pend = 0;
  double u_ov = 0.0, u_pend = 0.0;
  if (over) {
    ov = 32 - __builtin_clz(over);
    u_ov = wt[ov - 1] * scale;
  }
  while (over && (pend || under)) {
    int un;
    double u_un;
    if (pend) {
      un = pend;
      u_un = u_pend;
      pend = 0;
    } else {
      un = 32 - __builtin_clz(under);
      under &= ~(1u << (un - 1));
      u_un = wt[un - 1] * scale;
    }
    if (un == kk) return x < u_un ? kk : ov;
    u_ov = u_ov - (1.0 - u_un);
    if (u_ov < 1.0) {
      pend = ov;
      u_pend = u_ov;
      over &= ~(1u << (ov - 1));
      if (over) {
        ov = 32 - __builtin_clz(over);
        u_ov = wt[ov - 1] * scale;
      }
    }
  }
  return kk;
}
// generate_alias_tables (lib/local/alias.f90:101-184) for n <= 128 weights w
// (1-based results: u scaled probabilities, k aliases in 1..n).
__host__ __device__ inline void build_alias(int n, const double* w, double tot, double* u, int* k) {
  uint8_t under[128], over[128];
  int nunder = 0, nover = 0;
  const double scale = static_cast<double>(n) / tot;
  for (int i = 1; i <= n; ++i) {
    u[i - 1] = w[i - 1] * scale;
    if (u[i - 1] <= 1.0)
      under[nunder++] = static_cast<uint8_t>(i);
    else
      over[nover++] = static_cast<uint8_t>(i);
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

// select_weighted_value_precalc (lib/local/alias.f90:13-62) on a stored table.
__device__ __forceinline__ int alias_draw(int n, const double* u, const int* k, double r) {
  double x = r * n;
  int kk = static_cast<int>(floor(x));
  x = x - kk;
  kk = kk + 1;
  return x < u[kk - 1] ? kk : k[kk - 1];
}

// Probability of a single i -> a given j (src/excit_gen_heat_bath_mol.F90:367-380).
__device__ inline double hb_psingle(const Sys& s, const uint64_t* f, int i, int j, int a) {
  if (!(s.orb_sym[a - 1] == s.orb_sym[i - 1] && ms_of(a) == ms_of(i))) return 0.0;
  const double hm = fabs(sc1_excit(s, f, i, a, false));
  const double t = s.hb_ijab_tot[hb3(s, i, j, a)];
  return hm < t ? hm / (t + hm) : 0.5;
}

// gen_excit_mol_heat_bath: the excitation, its matrix element h and pgen.
// The nel on-the-fly weights (first P(i), then P(j|i)) live in wb; occupied
// orbitals are visited in ascending order straight from the bit string.
template <class Rng>
__device__ inline bool gen_heat_bath(const Sys& s, Rng& rng, const uint64_t* f, Excit& conn, double& h,
                                     double& pgen, const HbBuf& wb) {
  const int nel = s.nel;
  conn.nexcit = 0;
  conn.from[0] = conn.from[1] = conn.to[0] = conn.to[1] = 0;
  conn.perm = false;
  // find_i_d_weights + select_ij_heat_bath
  double iw_tot = 0.0;
  {
    int k = 0;
    FQ_FOR_WORDS(w, s.W) {
      for (uint64_t x = f[w]; x; x &= x - 1) {
        const double v = s.hb_i[(w << 6) + __builtin_ctzll(x)];
        wb[k++] = v;
        iw_tot = iw_tot + v;
      }
    }
  }
  const int i_ind = alias_pick(nel, wb, iw_tot, rng.next());
  const int i = select_orb(f, s.W, i_ind - 1);
  const double iw_i = wb[i_ind - 1];
  double ij_tot = 0.0;
  {
    int k = 0;
    FQ_FOR_WORDS(w, s.W) {
      for (uint64_t x = f[w]; x; x &= x - 1) {
        const double v = s.hb_ij[hb2(s, i, (w << 6) + __builtin_ctzll(x) + 1)];
        wb[k++] = v;
        ij_tot = ij_tot + v;
      }
    }
  }
  bool allowed = false;
  int j = 0, j_ind = 0;
  if (ij_tot > 0.0) {
    j_ind = alias_pick(nel, wb, ij_tot, rng.next());
    j = select_orb(f, s.W, j_ind - 1);
    allowed = true;
  }
  allowed = allowed && fabs(s.hb_ija_tot[hb2(s, i, j)]) > 0.0;
  int a = 0;
  bool dbl = false;
  double psingle = 0.0, hm_ia = 0.0;
  if (allowed) {
    a = alias_draw(s.nbasis, s.hb_ija_u + hb3(s, i, j, 1), s.hb_ija_k + hb3(s, i, j, 1), rng.next());
    if (!btest(f, a)) {
      if (s.orb_sym[a - 1] == s.orb_sym[i - 1] && ms_of(a) == ms_of(i)) {
        conn.nexcit = 1;
        conn.from[0] = i;
        conn.to[0] = a;
        conn.perm = perm1(f, s.W, i, a);
        hm_ia = fabs(sc1_excit(s, f, i, a, conn.perm));
        const double x = rng.next();
        const double t = s.hb_ijab_tot[hb3(s, i, j, a)];
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
  pgen = 1.0;
  h = 0.0;
  if (allowed) {
    if (dbl) {
      const int b = alias_draw(s.nbasis, s.hb_ijab_u + hb4(s, i, j, a, 1), s.hb_ijab_k + hb4(s, i, j, a, 1),
                               rng.next());
      if (!btest(f, b)) {
        double ji_tot = 0.0;
        FQ_FOR_WORDS(w, s.W) {
          for (uint64_t x = f[w]; x; x &= x - 1) ji_tot = ji_tot + s.hb_ij[hb2(s, j, (w << 6) + __builtin_ctzll(x) + 1)];
        }
        // P(i) and P(j|i) of the drawn path; P(j) from the i weights (hb_i)
        const double pi = iw_i / iw_tot, pj = wb[j_ind - 1] / ij_tot;
        const double pj2 = s.hb_i[j - 1] / iw_tot, pi2 = s.hb_ij[hb2(s, j, i)] / ji_tot;
        const double pgen_ija = (pi * pj) * (s.hb_ija[hb3(s, i, j, a)] / s.hb_ija_tot[hb2(s, i, j)]) *
                                (1.0 - psingle) * (s.hb_ijab[hb4(s, i, j, a, b)] / s.hb_ijab_tot[hb3(s, i, j, a)]);
        double ps = hb_psingle(s, f, i, j, b);
        const double pgen_ijb = (pi * pj) * (s.hb_ija[hb3(s, i, j, b)] / s.hb_ija_tot[hb2(s, i, j)]) * (1.0 - ps) *
                                (s.hb_ijab[hb4(s, i, j, b, a)] / s.hb_ijab_tot[hb3(s, i, j, b)]);
        ps = hb_psingle(s, f, j, i, a);
        const double pgen_jia = (pj2 * pi2) * (s.hb_ija[hb3(s, j, i, a)] / s.hb_ija_tot[hb2(s, j, i)]) *
                                (1.0 - ps) * (s.hb_ijab[hb4(s, j, i, a, b)] / s.hb_ijab_tot[hb3(s, j, i, a)]);
        ps = hb_psingle(s, f, j, i, b);
        const double pgen_jib = (pj2 * pi2) * (s.hb_ija[hb3(s, j, i, b)] / s.hb_ija_tot[hb2(s, j, i)]) *
                                (1.0 - ps) * (s.hb_ijab[hb4(s, j, i, b, a)] / s.hb_ijab_tot[hb3(s, j, i, b)]);
        pgen = pgen_ija + pgen_ijb + pgen_jia + pgen_jib;
        conn.nexcit = 2;
        conn.from[0] = i < j ? i : j;
        conn.from[1] = i < j ? j : i;
        conn.to[0] = a < b ? a : b;
        conn.to[1] = a < b ? b : a;
        conn.perm = perm2(f, s.W, conn);
        h = sc2_excit(s, conn.from[0], conn.from[1], conn.to[0], conn.to[1], conn.perm);
      } else {
        allowed = false;
      }
    } else {
      h = sc1_excit(s, f, i, a, conn.perm);
      pgen = 0.0;
      int q = 0;  // wb still holds P(j|i) weights, in occupied order
      FQ_FOR_WORDS(w, s.W) {
        for (uint64_t x = f[w]; x; x &= x - 1, ++q) {
          const int oq = (w << 6) + __builtin_ctzll(x) + 1;
          if (i != oq && a != oq) {
            const double t = s.hb_ijab_tot[hb3(s, i, oq, a)];
            const double psq = hm_ia < t ? hm_ia / (t + hm_ia) : 0.5;
            pgen = pgen + (psq * (wb[q] / ij_tot) * (s.hb_ija[hb3(s, i, oq, a)] / s.hb_ija_tot[hb2(s, i, oq)]));
          }
        }
      }
      pgen = pgen * (iw_i / iw_tot);
    }
  }
  if (!allowed) {
    h = 0.0;
    pgen = 1.0;
    conn.nexcit = 0;
  }
  return allowed;
}

// spawn_standard with the heat-bath generator (src/spawning.F90:34-124).
template <class Rng>
__device__ inline int64_t spawn_heat_bath(const Sys& s, Rng& rng, const uint64_t* f, int64_t parent, double tau,
                                          Excit& conn, const HbBuf& wb) {
  double h = 0.0, pgen = 1.0;
  gen_heat_bath(s, rng, f, conn, h, pgen, wb);
  return attempt_to_spawn(s, rng, tau, h, pgen, parent);
}


// ---------------------------------------------------------------------------
// Learned double-excitation proposal (no reference counterpart; an addition
// beside the heat-bath generator).  Doubles come from a mixture: with
// probability 1 - mix the network component, with probability mix a uniform
// choice (pair, particle, partner).  The network component draws an ordered
// hole pair (h1, h2) by the determinant's hole weights e (h1 by e, then h2 != h1
// by e), the first particle p1 among the unoccupied orbitals by f_p * A[h1][p]
// and its partner p2 within p1's spin/irrep partner class by f_p * A[h2][p]
// (f: the determinant's particle weights, A: the learned hole-particle
// affinity).  pgen is the exact mixture probability of the unordered
// excitation, summed over the four paths (h1, p1) in {i, j} x {a, b}, so the
// sampler is unbiased whatever the weights.  Determinants the network did not
// run on (fewer attempts than the threshold) use flat weights e = f = 1 with
// the learned affinity.  Singles use the renormalised generator.
// ---------------------------------------------------------------------------
struct PartnerClass {
  int ims, sym;
  bool ok;
};
// Class the partner of particle x must come from for the hole pair (i, j).
__device__ inline PartnerClass partner_class(const Sys& s, int i, int j, int x) {
  const int msb = ms_of(i) + ms_of(j) - ms_of(x);
  PartnerClass c;
  c.ok = (msb == 1 || msb == -1);
  c.ims = (msb + 3) / 2;
  c.sym = s.orb_sym[i - 1] ^ s.orb_sym[j - 1] ^ s.orb_sym[x - 1];
  return c;
}
// Unoccupied members of class c other than x: their count and, when fw is
// given, the sums of f_o * A[hi][o] (zi) and f_o * A[hj][o] (zj).
__device__ inline void partner_stats(const Sys& s, const uint64_t* f, const float* fw, const float* ai,
                                     const float* aj, const PartnerClass& c, int x, int& count, double& zi,
                                     double& zj) {
  count = 0;
  zi = zj = 0.0;
  if (!c.ok) return;
  const int nm = nbss(s, c.ims, c.sym);
  for (int k = 0; k < nm; ++k) {
    const int o = ssbf(s, c.ims, c.sym, k);
    if (o != x && !btest(f, o)) {
      ++count;
      if (fw) {
        const double v = static_cast<double>(fw[o - 1]);
        zi = zi + v * static_cast<double>(ai[o - 1]);
        zj = zj + v * static_cast<double>(aj[o - 1]);
      }
    }
  }
}
// Member of class c (unoccupied, not x) by index (uniform) or by the weight
// f_o * arow[o] reaching target.
__device__ inline int partner_pick(const Sys& s, const uint64_t* f, const float* fw, const float* arow,
                                   const PartnerClass& c, int x, int kth, double target, bool by_weight) {
  const int nm = nbss(s, c.ims, c.sym);
  int last = 0, seen = 0;
  double cum = 0.0;
  for (int k = 0; k < nm; ++k) {
    const int o = ssbf(s, c.ims, c.sym, k);
    if (o == x || btest(f, o)) continue;
    last = o;
    if (by_weight) {
      cum = cum + static_cast<double>(fw[o - 1]) * static_cast<double>(arow[o - 1]);
      if (target < cum) return o;
    } else if (seen++ == kth) {
      return o;
    }
  }
  return last;
}

// aff_in / flat_in: copies of s.mlp_aff / s.mlp_flat (in LDS in the spawning
// kernel); nullptr reads the device tables.
constexpr int kFlatTries = 16;  // alias redraws before the exact walk takes over

// Flat weights (e = f = 1, determinants without a network row): normalisers
// as the affinity row sums less their occupied terms, O(nel) instead of
// O(nvirt).
__device__ inline double occ_aff_sum(const Sys& s, const uint64_t* f, const uint64_t* mask, const float* arow) {
  double z = 0.0;
  FQ_FOR_WORDS(w, s.W) {
    for (uint64_t y = mask ? (f[w] & mask[w]) : f[w]; y; y &= y - 1)
      z = z + static_cast<double>(arow[(w << 6) + __builtin_ctzll(y)]);
  }
  return z;
}
// partner_stats with flat weights: count of class c's unoccupied members
// other than x and the sums of A[hi][.] and A[hj][.] over them.
__device__ inline void flat_partner_stats(const Sys& s, const uint64_t* f, int hi, int hj, const float* ai,
                                          const float* aj, const PartnerClass& c, int x, int& count, double& zi,
                                          double& zj) {
  count = 0;
  zi = zj = 0.0;
  if (!c.ok) return;
  const int ci = (c.ims - 1) * s.nsym_pg + c.sym, ncls = 2 * s.nsym_pg;
  const uint64_t* m = s.mlp_cmask + ci * FQ_MAX_WORDS;
  int occ_in = 0;
  FQ_FOR_WORDS(w, s.W) occ_in += __popcll(m[w] & f[w]);
  const bool x_in = btest(m, x);
  count = nbss(s, c.ims, c.sym) - occ_in - (x_in ? 1 : 0);
  zi = s.mlp_crs[(hi - 1) * ncls + ci] - occ_aff_sum(s, f, m, ai) - (x_in ? static_cast<double>(ai[x - 1]) : 0.0);
  zj = s.mlp_crs[(hj - 1) * ncls + ci] - occ_aff_sum(s, f, m, aj) - (x_in ? static_cast<double>(aj[x - 1]) : 0.0);
}

template <class Rng>
__device__ inline bool gen_learned(const Sys& s, Rng& rng, const uint64_t* f, const float* wrow, Excit& conn,
                                   double& h, double& pgen, const float* aff_in = nullptr,
                                   const float* flat_in = nullptr) {
  const float* aff = aff_in ? aff_in : s.mlp_aff;
  const SymUnocc u = sym_unocc(s, f);
  pgen = 1.0;
  h = 0.0;
  bool allowed = false;
  conn.nexcit = 0;
  conn.from[0] = conn.from[1] = conn.to[0] = conn.to[1] = 0;
  conn.perm = false;
  const int nel = s.nel, nb = s.nbasis;
  if (rng.next() < s.psingle) {
    allowed = single_mol(s, rng, f, u, conn, h, pgen);
  } else {
    const bool flat = wrow == nullptr;  // e = f = 1
    const float* row = wrow ? wrow : (flat_in ? flat_in : s.mlp_flat);
    const double mix = s.mix;
    const float* ew = row;
    const float* fw = row + nb;
    uint64_t un[FQ_MAX_WORDS];
    FQ_FOR_WORDS(w, s.W) un[w] = ~f[w] & s.valid[w];
    double zh = 0.0;
    FQ_FOR_WORDS(w, s.W) {
      for (uint64_t x = f[w]; x; x &= x - 1) zh = zh + static_cast<double>(ew[(w << 6) + __builtin_ctzll(x)]);
    }
    const int nvirt = nb - nel;
    int i = 0, j = 0, a = 0, b = 0;
    bool ok = true;
    // normalisers of the network component: Z1 of holes i, j; Z2 of holes i, j
    // over b's partner class given a (zbi, zbj); known after a network draw
    double z1i = 0.0, z1j = 0.0, zbi = 0.0, zbj = 0.0;
    int nbc = 0;
    bool have_z = false;
    if (rng.next() < mix) {
      // uniform component: unordered hole pair (choose_ij_mol's index map), a, partner
      const int ind = static_cast<int>(rng.next() * nel * (nel - 1) / 2) + 1;
      const int j_ind = s.jpair[ind];
      const int i_ind = ind - ((j_ind - 1) * (j_ind - 2)) / 2;
      i = select_orb(f, s.W, i_ind - 1);
      j = select_orb(f, s.W, j_ind - 1);
      a = select_orb(un, s.W, static_cast<int>(rng.next() * nvirt));
      const PartnerClass cb = partner_class(s, i, j, a);
      double zi = 0.0, zj = 0.0;
      partner_stats(s, f, nullptr, nullptr, nullptr, cb, a, nbc, zi, zj);
      if (nbc == 0)
        ok = false;
      else
        b = partner_pick(s, f, nullptr, nullptr, cb, a, static_cast<int>(rng.next() * nbc), 0.0, false);
    } else {
      // network component: holes by weight (i, then j != i), a by f A[i], partner by f A[j]
      double x = rng.next() * zh, cum = 0.0;
      FQ_FOR_WORDS(w, s.W) if (!i)
        for (uint64_t y = f[w]; y; y &= y - 1) {
          const int o = (w << 6) + __builtin_ctzll(y) + 1;
          cum = cum + static_cast<double>(ew[o - 1]);
          if (x < cum) {
            i = o;
            break;
          }
        }
      if (!i) i = select_orb(f, s.W, nel - 1);
      x = rng.next() * (zh - static_cast<double>(ew[i - 1]));
      cum = 0.0;
      int lastj = 0;
      FQ_FOR_WORDS(w, s.W) if (!j)
        for (uint64_t y = f[w]; y; y &= y - 1) {
          const int o = (w << 6) + __builtin_ctzll(y) + 1;
          if (o == i) continue;
          lastj = o;
          cum = cum + static_cast<double>(ew[o - 1]);
          if (x < cum) {
            j = o;
            break;
          }
        }
      if (!j) j = lastj;
      const float* ai = aff + (i - 1) * nb;
      const float* aj = aff + (j - 1) * nb;
      if (flat) {
        z1i = s.mlp_rs[i - 1] - occ_aff_sum(s, f, nullptr, ai);
        z1j = s.mlp_rs[j - 1] - occ_aff_sum(s, f, nullptr, aj);
      } else {
        FQ_FOR_WORDS(w, s.W) {
          for (uint64_t y = un[w]; y; y &= y - 1) {
            const int o = (w << 6) + __builtin_ctzll(y);
            const double v = static_cast<double>(fw[o]);
            z1i = z1i + v * static_cast<double>(ai[o]);
            z1j = z1j + v * static_cast<double>(aj[o]);
          }
        }
      }
      if (flat) {  // A[i][.] over all orbitals by its alias table, occupied ones redrawn
        for (int t = 0; t < kFlatTries && !a; ++t) {
          const int p = alias_draw(nb, s.mlp_au + (i - 1) * nb, s.mlp_ak + (i - 1) * nb, rng.next());
          if (!btest(f, p)) a = p;
        }
      }
      x = a ? 0.0 : rng.next() * z1i;  // the walk below, exact, when the redraws ran out
      cum = 0.0;
      int lasta = 0;
      FQ_FOR_WORDS(w, s.W) if (!a)
        for (uint64_t y = un[w]; y; y &= y - 1) {
          const int o = (w << 6) + __builtin_ctzll(y) + 1;
          lasta = o;
          cum = cum + static_cast<double>(fw[o - 1]) * static_cast<double>(ai[o - 1]);
          if (x < cum) {
            a = o;
            break;
          }
        }
      if (!a) a = lasta;
      const PartnerClass cb = partner_class(s, i, j, a);
      if (flat)
        flat_partner_stats(s, f, i, j, ai, aj, cb, a, nbc, zbi, zbj);
      else
        partner_stats(s, f, fw, ai, aj, cb, a, nbc, zbi, zbj);
      have_z = true;
      if (nbc == 0) {
        ok = false;
      } else {
        if (flat) {  // A[j][.] over the class by its alias table, occupied members and a redrawn
          const int ci = (cb.ims - 1) * s.nsym_pg + cb.sym, n = nbss(s, cb.ims, cb.sym);
          const int64_t o = (static_cast<int64_t>(j - 1) * 2 * s.nsym_pg + ci) * s.max_sym_spin;
          for (int t = 0; t < kFlatTries && !b; ++t) {
            const int q = alias_draw(n, s.mlp_cu + o, s.mlp_ck + o, rng.next());
            const int p = ssbf(s, cb.ims, cb.sym, q - 1);
            if (p != a && !btest(f, p)) b = p;
          }
        }
        if (!b) b = partner_pick(s, f, fw, aj, cb, a, 0, rng.next() * zbj, true);
      }
    }
    if (ok) {
      const float* ai = aff + (i - 1) * nb;
      const float* aj = aff + (j - 1) * nb;
      int nac = 0;
      double zai = 0.0, zaj = 0.0;  // Z2 of holes i, j over a's partner class given b
      if (!have_z && flat) {  // a uniform draw: the network normalisers of (i, j, a)
        flat_partner_stats(s, f, i, j, ai, aj, partner_class(s, i, j, a), a, nbc, zbi, zbj);
        z1i = s.mlp_rs[i - 1] - occ_aff_sum(s, f, nullptr, ai);
        z1j = s.mlp_rs[j - 1] - occ_aff_sum(s, f, nullptr, aj);
      } else if (!have_z) {
        partner_stats(s, f, fw, ai, aj, partner_class(s, i, j, a), a, nbc, zbi, zbj);
        FQ_FOR_WORDS(w, s.W) {
          for (uint64_t y = un[w]; y; y &= y - 1) {
            const int o = (w << 6) + __builtin_ctzll(y);
            const double v = static_cast<double>(fw[o]);
            z1i = z1i + v * static_cast<double>(ai[o]);
            z1j = z1j + v * static_cast<double>(aj[o]);
          }
        }
      }
      if (flat)
        flat_partner_stats(s, f, i, j, ai, aj, partner_class(s, i, j, b), b, nac, zai, zaj);
      else
        partner_stats(s, f, fw, ai, aj, partner_class(s, i, j, b), b, nac, zai, zaj);
      const double q_unif = 2.0 / static_cast<double>(nel * (nel - 1)) / nvirt * (1.0 / nbc + 1.0 / nac);
      {
        const double ei = ew[i - 1], ej = ew[j - 1], fa = fw[a - 1], fb = fw[b - 1];
        const double aia = ai[a - 1], aib = ai[b - 1], aja = aj[a - 1], ajb = aj[b - 1];
        const double pij = ei / zh * ej / (zh - ei), pji = ej / zh * ei / (zh - ej);
        const double q_mlp = pij * (fa * aia / z1i) * (fb * ajb / zbj) + pij * (fb * aib / z1i) * (fa * aja / zaj) +
                             pji * (fa * aja / z1j) * (fb * aib / zbi) + pji * (fb * ajb / z1j) * (fa * aia / zai);
        pgen = s.pdouble * ((1.0 - mix) * q_mlp + mix * q_unif);
      }
      conn.nexcit = 2;
      conn.from[0] = i < j ? i : j;
      conn.from[1] = i < j ? j : i;
      conn.to[0] = a < b ? a : b;
      conn.to[1] = a < b ? b : a;
      conn.perm = perm2(f, s.W, conn);
      h = sc2_excit(s, conn.from[0], conn.from[1], conn.to[0], conn.to[1], conn.perm);
      allowed = true;
    } else {
      h = 0.0;
      pgen = 1.0;
    }
  }
  return allowed;
}

template <class Rng>
__device__ inline int64_t spawn_learned(const Sys& s, Rng& rng, const uint64_t* f, const float* wrow, int64_t parent,
                                        double tau, Excit& conn, const float* aff, const float* flat) {
  double h = 0.0, pgen = 1.0;
  gen_learned(s, rng, f, wrow, conn, h, pgen, aff, flat);
  return attempt_to_spawn(s, rng, tau, h, pgen, parent);
}

// One generator call without the spawning step (the excitation probe): the
// family is chosen at run time; returns whether an allowed excitation was made.
template <class Rng>
__device__ inline bool generate_any(const Sys& s, Rng& rng, const uint64_t* f, const float* wrow, Excit& conn,
                                    double& h, double& pgen) {
  if (s.excit_gen == FQ_EXCIT_HEAT_BATH) {
    double wloc[kHbMaxNel];
    return gen_heat_bath(s, rng, f, conn, h, pgen, HbBuf{wloc, 1});
  }
  if (s.excit_gen == FQ_EXCIT_LEARNED) return gen_learned(s, rng, f, wrow, conn, h, pgen);
  return gen_mol(s, rng, f, conn, h, pgen);
}

// Generator families; the production spawning kernel is compiled once per
// family so each carries only its own generator's registers.
enum Gen : int { kGenHubK = 0, kGenMolRenorm = 1, kGenHeatBath = 2, kGenLearned = 3, kGenAny = 4 };
__host__ __device__ inline int gen_of(int system, int excit_gen) {
  if (system == FQ_SYS_HUB_K) return kGenHubK;
  if (excit_gen == FQ_EXCIT_LEARNED) return kGenLearned;
  return excit_gen == FQ_EXCIT_HEAT_BATH ? kGenHeatBath : kGenMolRenorm;
}

// One spawning attempt from determinant f; wrow is the determinant's row of
// the learned proposal's weights (read only by that generator); hb is the
// heat-bath generator's weight storage (a local array when not given).
template <int G = kGenAny, class Rng>
__device__ inline int64_t spawn_attempt(const Sys& s, Rng& rng, const uint64_t* f, const float* wrow,
                                        int64_t parent, double tau, uint64_t* child, HbBuf hb = HbBuf{},
                                        const float* aff = nullptr, const float* flat = nullptr) {
  Excit conn;
  const int g = G == kGenAny ? gen_of(s.system, s.excit_gen) : G;
  int64_t n;
  if ((G == kGenAny || G == kGenHubK) && g == kGenHubK) {
    n = spawn_hub_k(s, rng, f, parent, tau, conn);
  } else if ((G == kGenAny || G == kGenHeatBath) && g == kGenHeatBath) {
    if constexpr (G == kGenAny) {
      double wloc[kHbMaxNel];
      n = spawn_heat_bath(s, rng, f, parent, tau, conn, HbBuf{wloc, 1});
    } else {
      n = spawn_heat_bath(s, rng, f, parent, tau, conn, hb);
    }
  }
  else if ((G == kGenAny || G == kGenLearned) && g == kGenLearned) {
    n = spawn_learned(s, rng, f, wrow, parent, tau, conn, aff, flat);
  } else {
    n = spawn_mol(s, rng, f, parent, tau, conn);
  }
  if (n != 0) excited_det(f, s.W, conn, child);
  return n;
}

}  // namespace fqd
