// MI355X (gfx950) FCIQMC walker-propagation engine: implementation of the
// C-ABI in include/fqmc.h.
//
// One call to fq_cycle runs the body of the reference's Monte Carlo cycle
// (src/fciqmc.f90:307-399) on the device:
//   spawning + death      production: k_det_pass (one thread per determinant:
//                         projected energy, initiator flag, decide_nattempts,
//                         stochastic death) then k_spawn (one thread per
//                         spawning attempt; parents found in a per-tile window
//                         of the attempt prefix sum; results written to the
//                         attempt's staging slot, then packed in attempt order
//                         by a prefix-sum compaction: no atomics).
//                         replay: k_replay_walk, one thread walking the list
//                         in the reference's order and consuming the uploaded
//                         dSFMT stream exactly as get_rand_close_open does.
//   sort                  production: LSD radix sort (8-bit digits, stable
//                         per-round wavefront ranking) of the spawned list's
//                         combinatorial ranks, which order determinants as
//                         their bit strings do in fewer bits.
//                         replay: the reference's quicksort (src/sort.f90:213-396)
//                         on one thread, so equal keys keep its order.
//   self-annihilation     k_reduce_runs: one thread per run of equal keys,
//                         plain and initiator rules (src/spawn_data.F90:859-1101).
//   merge with main list  k_merge_runs: per block, the main-list window of its
//                         runs' key range, then a search within that window
//                         (src/annihilation.f90:294-486).
//   rounding, compaction  k_round_main / k_round_new: stochastic rounding of small
//                         real populations (src/annihilation.f90:537-675), draws
//                         from Philox or from the replay stream at scan offsets.
//   insertion             k_scatter_main / k_scatter_new: merge-path placement
//                         of surviving main entries and new determinants into the
//                         other buffer, new K_ii from the Slater--Condon diagonal
//                         (src/annihilation.f90:677-901).
// fq_cycle_batch runs the same kernels over many independent instances held in
// one list; the heat-bath tables and the learned proposal are built here and
// in mlp.hip.
// Memory layout (HBM, structure of arrays): determinant word w of entry i at
// det[w*cap + i]; population of space s at pop[s*cap + i]; K_ii at kii[i].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <queue>
#include <type_traits>
#include <vector>

#include "device_common.hpp"
#include "fqmc.h"
#include "mlp.hpp"

using fqd::Sys;

// ---------------------------------------------------------------------------
// Fixed scratch of fq_cycle_noalloc, for callers that cannot allocate device
// memory.  One region per scratch slot (int64 elements): the prefix sums'
// tile sums (slot 0) and tile offsets (slot 1), sized for every scan an
// engine can issue (at most 2^31 entries, tiles of 4096).  Every device holds
// its own copy until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returns with that device current (the array's
// name is a host-side stand-in).  Calls that overlap on one device share it
// and race.
// ---------------------------------------------------------------------------
constexpr int kScratchSlots = 2;
constexpr int64_t kScratchSlotElems[kScratchSlots] = {(int64_t{1} << 19) + 64, (int64_t{1} << 19) + 64};
__host__ __device__ constexpr int64_t scratch_slot_offset(int slot) {
  int64_t off = 0;
  for (int k = 0; k < slot; ++k) off += (kScratchSlotElems[k] + 31) / 32 * 32;  // 256-byte aligned regions
  return off;
}
constexpr int64_t kScratchElems = scratch_slot_offset(kScratchSlots);
static __device__ __align__(256) int64_t scratch[kScratchElems];

namespace {

constexpr int kBlock = 256;
// Batched instances share one list: the instance index sits in bits 48..63 of
// word 0 above the orbital bits (systems of at most 48 spin orbitals).
constexpr int kInstShift = 48;
constexpr uint64_t kOrbMask = (uint64_t{1} << kInstShift) - 1;
constexpr int kMaxInst = 1 << 16;
constexpr int kRunTile = 4 * kBlock;  // sorted entries per run-detection tile
constexpr int kMergeWin = 1024;       // main-list window entries a merge block stages in LDS
// Words per spawned-list record (populations, flag), padded to 16-byte
// alignment so one record is one aligned read after the sort's permutation.
__host__ __device__ constexpr int rec_stride(int S) { return S == 1 ? 2 : 4; }

#define FQ_HIP(call)                                  \
  do {                                                \
    hipError_t err_ = (call);                         \
    if (err_ != hipSuccess) {                         \
      fprintf(stderr, "fqmc: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(err_), __FILE__, __LINE__); \
      return FQ_ERR_DEVICE;                           \
    }                                                 \
  } while (0)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline unsigned grid_for(int64_t n, int64_t per_block = kBlock) {
  int64_t g = ceil_div(n, per_block);
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

// Per-cycle device counters.
struct Stats {
  double proj[FQ_MAX_SPACES];
  double d0[FQ_MAX_SPACES];
  int64_t ndeath;
  int64_t nspawn;          // spawned entries requested (may exceed capacity)
  int64_t spawn_overflow;
  int64_t total_attempts;
  int64_t nruns;
  int64_t new_n;
  int64_t main_overflow;
  int64_t rng_pos;
  int64_t rng_failed;
  unsigned long long abs_pop[FQ_MAX_SPACES];  // sum |pop| after annihilation (encoded)
};

// The semi-stochastic core as the projection kernels see it: CSR rows of the
// stored Hamiltonian (row i: H_ij over j in combined mode, the transpose
// H_ji in separate mode), the cycle's core populations vec[i * S + s] and
// the main-list position of each core determinant.
struct DetermDev {
  const int64_t* row;
  const int32_t* col;
  const double* mat;
  double* vec;
  int64_t* pos;
  int64_t nd;
  int separate;
};
// A determinant-pass flag bit above the initiator flags: the parent is a core
// determinant (its spawns onto the core are the deterministic projection's).
constexpr int64_t kCoreFlag = int64_t{1} << 40;
constexpr int64_t kInitFlagMask = 0xff;

// ---------------------------------------------------------------------------
// Prefix sums.  Up to kSmallScanMax entries: one 1024-thread block (one
// launch).  Up to kPrefixTiles
// tiles: per-tile sums, then each tile adds the sums of the tiles before it
// itself and scans its entries (two launches).  More: the tile sums are
// scanned by the one-block scan in between (three launches).  out[n] = total.
// ---------------------------------------------------------------------------
constexpr int kScanItems = 16;
constexpr int64_t kScanTile = static_cast<int64_t>(kBlock) * kScanItems;
constexpr int kSmallBlock = 1024, kSmallItems = 8;
constexpr int64_t kSmallScanMax = kSmallBlock * kSmallItems;  // one chunk: a single block is latency-bound
constexpr int64_t kPrefixTiles = 2048;

// Inclusive scan of x over the 64 lanes of a wavefront.
__device__ __forceinline__ int64_t wave_inclusive_scan(int64_t x, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}
__device__ __forceinline__ int64_t wave_sum(int64_t x) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}
// Sum of x over a kBlock-thread block (every thread receives it).
__device__ __forceinline__ int64_t block_sum(int64_t x, int64_t* wsh) {
  x = wave_sum(x);
  if ((threadIdx.x & 63) == 0) wsh[threadIdx.x >> 6] = x;
  __syncthreads();
  int64_t t = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) t += wsh[w];
  __syncthreads();
  return t;
}

// A list length read on the device when pn is given (launches sized by a
// bound n: production sorts, graph-captured cycles), clamped to the bound.
__device__ __forceinline__ int64_t list_len(const int64_t* __restrict__ pn, int64_t n) {
  if (!pn) return n;
  const int64_t v = *pn;
  return v < n ? v : n;
}
// The cycle index: read on the device in graph-captured cycles.
__device__ __forceinline__ uint32_t cycle_of(const uint32_t* __restrict__ pc, uint32_t cyc) { return pc ? *pc : cyc; }

__global__ __launch_bounds__(kBlock) void k_tile_sums(const int64_t* __restrict__ in, int64_t n,
                                                      int64_t* __restrict__ sums) {
  __shared__ int64_t wsh[kBlock / 64];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  int64_t acc = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k * kBlock + threadIdx.x;
    if (i < n) acc += in[i];
  }
  acc = block_sum(acc, wsh);
  if (threadIdx.x == 0) sums[blockIdx.x] = acc;
}

// Inclusive Hillis-Steele scan of sh[0..kBlock) in place.
__device__ __forceinline__ void block_inclusive_scan(int64_t* sh) {
  for (int h = 1; h < kBlock; h <<= 1) {
    const int64_t add = threadIdx.x >= h ? sh[threadIdx.x - h] : 0;
    __syncthreads();
    sh[threadIdx.x] += add;
    __syncthreads();
  }
}

// Exclusive scan of one tile (kScanItems consecutive entries per thread) plus
// the tile's offset: with prefix, the sum of part[0..tile) (part: tile sums);
// otherwise part[tile] (part: exclusive tile offsets).  The tile holding
// element n-1 also writes out[n] = total.
__global__ __launch_bounds__(kBlock) void k_tile_scan(const int64_t* __restrict__ in, int64_t n,
                                                      const int64_t* __restrict__ part, int prefix,
                                                      int64_t* __restrict__ out) {
  __shared__ int64_t wsh[kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t off;
  if (prefix) {
    int64_t pre = 0;
    for (int64_t k = threadIdx.x; k < blockIdx.x; k += kBlock) pre += part[k];
    off = block_sum(pre, wsh);
  } else {
    off = part[blockIdx.x];
  }
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t local = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    v[k] = i < n ? in[i] : 0;
    local += v[k];
  }
  const int64_t x = wave_inclusive_scan(local, lane);
  if (lane == 63) wsh[wave] = x;
  __syncthreads();
  int64_t wpre = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) wpre += w < wave ? wsh[w] : 0;
  int64_t run = off + wpre + x - local;
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
    if (i == n - 1) out[n] = run;
  }
}

// Exclusive scan of n entries on one block of kSmallBlock threads: chunks of
// kSmallBlock * kSmallItems with a running carry (one chunk for the small
// scans; many for the tile sums of a very long array).
__global__ __launch_bounds__(kSmallBlock) void k_scan_small(const int64_t* __restrict__ in, int64_t n,
                                                            int64_t* __restrict__ out) {
  __shared__ int64_t wsh[kSmallBlock / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int64_t carry = 0;
  for (int64_t c0 = 0; c0 < n; c0 += static_cast<int64_t>(kSmallBlock) * kSmallItems) {
    const int64_t base = c0 + static_cast<int64_t>(tid) * kSmallItems;
    int64_t v[kSmallItems];
    int64_t local = 0;
#pragma unroll
    for (int k = 0; k < kSmallItems; ++k) {
      const int64_t i = base + k;
      v[k] = i < n ? in[i] : 0;
      local += v[k];
    }
    const int64_t x = wave_inclusive_scan(local, lane);
    if (lane == 63) wsh[wave] = x;
    __syncthreads();
    int64_t wpre = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kSmallBlock / 64; ++w) {
      const int64_t t = wsh[w];
      wpre += w < wave ? t : 0;
      total += t;
    }
    int64_t run = carry + wpre + x - local;
#pragma unroll
    for (int k = 0; k < kSmallItems; ++k) {
      const int64_t i = base + k;
      if (i < n) out[i] = run;
      run += v[k];
    }
    carry += total;
    __syncthreads();  // wsh is rewritten by the next chunk
  }
  if (tid == 0) out[n] = carry;
}

__global__ void k_set_zero_total(int64_t* __restrict__ out) { out[0] = 0; }

}  // namespace

struct fq_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  Sys sys{};
  Sys* d_sysp = nullptr;    // device copy of sys (read by the spawning kernel)
  bool sys_dirty = true;    // sys changed since the last upload
  fq_params par{};
  int W = 1, S = 1;
  // device tables
  int *d_sym_table = nullptr, *d_inv_sym = nullptr, *d_orb_sym = nullptr, *d_nbss = nullptr, *d_ssbf = nullptr;
  double *d_sp_eigv = nullptr, *d_h1 = nullptr, *d_eri = nullptr, *d_jmat = nullptr, *d_kmat = nullptr;
  double *d_s1j = nullptr, *d_s1k = nullptr;
  // Graph-captured batched cycles: while one is captured (capturing), the
  // main list's length is read on the device from dn and the cycle index from
  // dcyc (the graph's first copy fills both from pinned host memory, h_in),
  // and launches are sized by bounds.
  // online training of the learned proposal (fq_mlp_online): steps of batch
  // samples at the start of every online_every-th cycle
  int32_t online_every = 0, online_steps = 0;
  int64_t online_batch = 0;
  double online_lr = 0.0;
  int64_t online_runs = 0;
  bool capturing = false;
  const int64_t* dn = nullptr;
  const uint32_t* dcyc = nullptr;
  bool graph_on = false;
  char* h_in = nullptr;       // pinned: [n][cycle index][tau, shift, E_proj per instance]
  char* d_in = nullptr;
  Stats* h_gstats = nullptr;  // pinned copies the graph writes at its end
  double* h_god = nullptr;
  int64_t* h_goi = nullptr;
  struct GraphEntry {
    int c;
    int64_t nb, tb;
    int sched;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;
  // the tuner, per octave of the attempt bound: the chosen schedule (-1:
  // still trying candidates), cycles tried and each candidate's timings
  int sched[64];
  int tune_iter[64];
  std::vector<float> tune_ms[64][8];
  int last_oct = -1;          // octave of the last graph cycle
  int64_t graph_cycles = 0, eager_cycles = 0;
  hipEvent_t gev[2] = {nullptr, nullptr};
  // schedule knobs (the tuner's candidates): 8-bit radix digits only, and a
  // cap on the spawning grid (0: one block per tile)
  int digit8 = 0;
  int spawn_blocks = 0;
  double* d_recip = nullptr;  // 1.0 / k, k < kRecipN
  int* d_jpair = nullptr;     // choose_ij_mol's j index per pair index
  // heat-bath tables (excit_gen FQ_EXCIT_HEAT_BATH)
  double *d_hb_i = nullptr, *d_hb_ij = nullptr, *d_hb_ija = nullptr, *d_hb_ija_tot = nullptr, *d_hb_ija_u = nullptr;
  double *d_hb_ijab = nullptr, *d_hb_ijab_tot = nullptr, *d_hb_ijab_u = nullptr;
  int *d_hb_ija_k = nullptr, *d_hb_ijab_k = nullptr, *d_hb_bias = nullptr;
  // batched independent instances (fq_create_batch): instance index in bits
  // 48..63 of word 0; per-instance tables held by sub-engines
  bool batched = false;
  int ninst = 1;
  int key_bits = 0;                  // key bits the radix sort orders (nbasis, or 48 + instance bits)
  int rank_words = 0, rank_bits = 0;  // production sort on combinatorial ranks (0: on the bit strings)
  uint64_t* d_binom = nullptr;        // C(p, i), p < nbasis, 1 <= i <= nel, as (low, high) halves
  std::vector<fq_engine*> subs;
  Sys* d_sys_arr = nullptr;          // [ninst]
  double* d_ipar = nullptr;          // tau[ninst], shift[S][ninst], eproj_old[S][ninst]
  int64_t* d_ibound = nullptr;       // [2][ninst + 1]: instance bounds of the old and the new list
  int64_t* d_isb = nullptr;          // [ninst + 1]: instance bounds of the spawned list
  int64_t* d_iatt = nullptr;         // [ninst]: first attempt of each instance
  double* d_cterm = nullptr;         // [2S][cap]: per-determinant projected-energy and N_0 terms
  int64_t* d_cdeath = nullptr;       // [cap]
  // per-instance 256-entry block counts and offsets (old list, new list) and
  // the block partials of the estimator terms (2S), deaths and |N| (1 + S)
  int64_t *d_iblk_cnt = nullptr, *d_iblk_off = nullptr, *d_nblk_off = nullptr;
  double* d_ipart_d = nullptr;
  int64_t* d_ipart_i = nullptr;
  int64_t ipb = 0;  // partial stride: a bound of the blocks of either list
  double* d_iout_d = nullptr;        // [ninst][2S]
  int64_t* d_iout_i = nullptr;       // [ninst][4 + S]
  // learned proposal (excit_gen FQ_EXCIT_LEARNED): network, per-determinant weights
  fqm::Mlp mlp;
  float* mlp_tab = nullptr;  // rows x 2*nbasis, one row per determinant the network runs on
  float* mlp_flat = nullptr;  // 2*nbasis ones: weights of determinants without a row
  const int64_t* p_nruns = nullptr;  // device address of this cycle's run count
  int64_t att_bound = -1;            // bound of the next cycle's attempts (-1: unknown)
  bool keys_ready = false;           // the spawning pass made this cycle's sort keys
  int64_t run_bound = 0;             // host bound of it (the count itself in replay mode)
  int64_t mlp_tab_rows = 0;
  int64_t* mlp_on = nullptr;   // [cap] network runs on this determinant this cycle
  int64_t* mlp_row = nullptr;  // [cap + 1] exclusive prefix of mlp_on
  int64_t* mlp_idx = nullptr;  // [cap] determinants the network runs on
  // main list (double-buffered)
  int64_t cap = 0, n = 0;
  int cur = 0;
  uint64_t* det[2] = {nullptr, nullptr};
  int64_t* pop[2] = {nullptr, nullptr};
  double* kii[2] = {nullptr, nullptr};
  int64_t* pop_new = nullptr;   // S*cap, populations after death
  int64_t* nat = nullptr;       // S*cap attempts per (det, space), then scanned
  int64_t* att_off = nullptr;   // S*cap+1
  int64_t* flags = nullptr;     // cap parent initiator flags
  uint64_t* mmask = nullptr;    // cap/64+2: surviving main entries, one bit each
  int64_t* mcnt = nullptr;      // survivors per 64-entry word
  int64_t* moff = nullptr;      // their exclusive prefix
  int64_t* mdraw = nullptr;     // cap+1 (replay rounding draws)
  int64_t* mdraw_off = nullptr;
  double* part_d = nullptr;     // 4 * (cap/256+1) block partials
  int64_t* part_i = nullptr;    // 2 * (cap/256+1)
  // spawning-attempt staging (one slot per attempt of a batch)
  int64_t bcap = 0;
  uint64_t* st_det = nullptr;   // W*bcap
  int64_t* st_pop = nullptr;    // bcap
  int64_t* st_ent = nullptr;    // bcap
  int64_t* tile_cnt = nullptr;  // successes per spawning tile (bcap / kSpawnTile + 2)
  int64_t* tile_off = nullptr;  // their exclusive prefix
  int64_t* tile_win = nullptr;  // parents of each tile's first and last attempt
  // spawned list
  int64_t scap = 0;
  uint64_t* sdet = nullptr;  // W*scap (in kbuf[1])
  int64_t* srec = nullptr;   // scap records: S populations then the flag, rec_stride(S) words each (in pbuf[1])
  // sort: keys (value << flag_bits | flags) and payload (populations), ping-pong
  int sort_kw = 1, flag_bits = 1, value_bits = 1;
  uint64_t* kbuf[2] = {nullptr, nullptr};  // sort_kw * scap
  int64_t* pbuf[2] = {nullptr, nullptr};   // S * scap
  uint32_t* perm = nullptr;                // scap: replay quicksort order
  int64_t* pass_cnt = nullptr;             // 512 * kSortBlocks digit counts of a pass's blocks
  int64_t* pass_off = nullptr;             // their exclusive prefix (+ total)
  int64_t* run_tcnt = nullptr;             // run heads per run tile
  int64_t* run_toff = nullptr;             // their exclusive prefix; [ntiles] = the run count
  int64_t* run_tfirst = nullptr;           // first head of each run tile (-1: none)
  int64_t* run_tpre = nullptr;             // sums of each run tile's entries before its first head
  uint64_t* rdet = nullptr;      // W*scap
  int64_t* rpop = nullptr;       // S*scap
  int64_t* rflag = nullptr;      // scap
  int64_t* rpos = nullptr;       // scap
  int64_t* rstate = nullptr;     // scap: 0 dead, 1 hit, 2 new candidate
  int64_t* nlist = nullptr;      // scap: surviving new runs in order
  int64_t* nrpos = nullptr;      // scap: their insertion positions (non-decreasing)
  int64_t* bstart = nullptr;     // cap/256 + 2: new runs before each 256-entry block of the main list
  int64_t* mwin = nullptr;       // main-list bounds of each merge block's first and last run
  uint64_t* nmask = nullptr;     // scap/64+2: surviving new runs, one bit each
  int64_t* ncnt = nullptr;       // survivors per 64-run word
  int64_t* noff = nullptr;       // their exclusive prefix
  int64_t* ndraw = nullptr;      // scap+1
  int64_t* ndraw_off = nullptr;
  // scan scratch
  std::vector<int64_t*> scan_tmp;
  std::vector<int64_t> scan_tmp_cap;
  // fq_cycle_spawn .. fq_cycle_finish: a cycle split at the spawned list (replay mode)
  bool pending = false;
  uint32_t pend_cyc = 0;
  int64_t pend_rng_pos = 0;   // stream position after the spawning pass
  int64_t pend_ns = 0;        // entries this engine spawned
  Stats pend_stats{};         // the spawning pass's counters
  // fq_cycle_noalloc: scratch slots come from the device's fixed array
  int64_t* fixed_scratch = nullptr;  // device address of `scratch` on this engine's device (null: unavailable)
  bool noalloc = false;              // inside fq_cycle_noalloc
  bool scratch_short = false;        // a request exceeded its fixed slot
  // replay stream
  double* rng_buf = nullptr;
  int64_t rng_cap = 0, rng_pos = 0, rng_end = 0;
  // per-cycle
  Stats* d_stats = nullptr;
  Stats h_stats{};
  uint64_t cycle_index = 0;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // the spawned-list length, spawn overflow and attempt total, copied behind
  // the spawning kernels into pinned memory (h_len) and marked by ev_len
  int64_t* h_len = nullptr;
  hipEvent_t ev_len = nullptr;
  int64_t spawn_bound = 0;  // attempts the cycle's spawning launches covered
  // semi-stochastic core (fq_semi_stoch_init): its determinants (word-major,
  // sorted as the main list; also in sys.determ), the Hamiltonian among them
  // as CSR rows (the transpose for separate annihilation), and per cycle the
  // core populations and their main-list positions
  int64_t ndeterm = 0, determ_nnz = 0;
  int determ_separate = 0;
  uint64_t* d_determ = nullptr;
  int64_t* d_drow = nullptr;
  int32_t* d_dcol = nullptr;
  double* d_dmat = nullptr;
  double* d_dvec = nullptr;
  int64_t* d_dpos = nullptr;
  float t_spawn = 0, t_annih = 0, t_total = 0;
  // phase events: 0 cycle start, 1 det pass end, 2 spawn start, 3 spawn end,
  // 4 sort start, 5 sort end, 6 reduce end, 7 merge start, 8 merge end
  hipEvent_t pev[9] = {};
  // an event pair around each spawning-kernel launch of a cycle (the first kSpawnEvents)
  static constexpr int kSpawnEvents = 16;
  hipEvent_t sev[2 * kSpawnEvents] = {};
  int spawn_launches = 0;
  fq_profile_out prof{};
  int radix_passes = 0;
};

namespace {

// ---- device allocation helpers ---------------------------------------------
template <class T>
int dalloc(T** p, int64_t count) {
  if (count <= 0) count = 1;
  if (hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * static_cast<size_t>(count)) != hipSuccess) {
    *p = nullptr;
    return FQ_ERR_NOMEM;
  }
  return FQ_OK;
}
template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// Device address of the fixed scratch array on `device` (current), looked up
// once per device; null if the lookup fails.
int64_t* fixed_scratch_address(int device) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static int64_t* addr[kMaxDevices] = {};
  static bool looked[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!looked[device]) {
    void* p = nullptr;
    const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess) {
      fprintf(stderr, "fqmc: hipGetSymbolAddress(scratch) on device %d failed: %s\n", device,
              hipGetErrorString(err));
      p = nullptr;
    }
    addr[device] = static_cast<int64_t*>(p);
    looked[device] = true;
  }
  return addr[device];
}

// Scratch buffer `slot` of at least `need` elements (grown on demand; inside
// fq_cycle_noalloc, the slot's region of the fixed array, or null when the
// request exceeds it).
int64_t* scan_scratch(fq_engine* e, int slot, int64_t need) {
  if (e->noalloc) {
    if (slot >= kScratchSlots || need > kScratchSlotElems[slot]) {
      e->scratch_short = true;
      return nullptr;
    }
    return e->fixed_scratch + scratch_slot_offset(slot);
  }
  while (static_cast<int>(e->scan_tmp.size()) <= slot) {
    e->scan_tmp.push_back(nullptr);
    e->scan_tmp_cap.push_back(0);
  }
  if (e->scan_tmp_cap[slot] < need) {
    dfree(e->scan_tmp[slot]);
    if (dalloc(&e->scan_tmp[slot], need) != FQ_OK) {
      e->scan_tmp_cap[slot] = 0;
      return nullptr;
    }
    e->scan_tmp_cap[slot] = need;
  }
  return e->scan_tmp[slot];
}

// Exclusive scan: out[0..n] with out[n] = total (in and out distinct).
int scan_exclusive(fq_engine* e, const int64_t* in, int64_t* out, int64_t n) {
  if (n == 0) {
    hipLaunchKernelGGL(k_set_zero_total, dim3(1), dim3(1), 0, e->stream, out);
    return FQ_OK;
  }
  if (n <= kSmallScanMax) {
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kSmallBlock), 0, e->stream, in, n, out);
    return FQ_OK;
  }
  const int64_t nb = ceil_div(n, kScanTile);
  int64_t* sums = scan_scratch(e, 0, nb);
  if (!sums) return FQ_ERR_NOMEM;
  hipLaunchKernelGGL(k_tile_sums, dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, e->stream, in, n, sums);
  if (nb <= kPrefixTiles) {
    hipLaunchKernelGGL(k_tile_scan, dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, e->stream, in, n, sums, 1,
                       out);
    return FQ_OK;
  }
  int64_t* offs = scan_scratch(e, 1, nb + 1);
  if (!offs) return FQ_ERR_NOMEM;
  hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kSmallBlock), 0, e->stream, sums, nb, offs);
  hipLaunchKernelGGL(k_tile_scan, dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, e->stream, in, n, offs, 0, out);
  return FQ_OK;
}

// ---------------------------------------------------------------------------
// Kernels: spawning and death
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_det(const uint64_t* det, int64_t cap, int W, int64_t i, uint64_t* f) {
#pragma unroll
  for (int w = 0; w < FQ_MAX_WORDS; ++w) f[w] = w < W ? det[w * cap + i] : 0;
}

// Production: per determinant (loop body of src/fciqmc.f90:315-374 minus the
// spawning attempts).  Block partial sums (projected energy, reference
// population, deaths) are reduced in a fixed order by k_sum_partials, so the
// estimators are deterministic and no global atomics are issued.  A
// semi-stochastic core determinant (set_determ_info) records its population
// and position for the projection, is flagged an initiator with kCoreFlag,
// and skips death.
__global__ __launch_bounds__(kBlock) void k_det_pass(Sys s, int64_t n, int64_t cap, const uint64_t* __restrict__ det,
                                                     const int64_t* __restrict__ pop, const double* __restrict__ kii,
                                                     int64_t* __restrict__ pop_new, int64_t* __restrict__ nat,
                                                     int64_t* __restrict__ flags, double tau, double sh0, double sh1,
                                                     double e0, double e1, uint64_t seed, uint32_t cyc,
                                                     double* __restrict__ part_d, int64_t* __restrict__ part_i,
                                                     double* __restrict__ dvec, int64_t* __restrict__ dpos) {
  __shared__ double sh_d[2 * FQ_MAX_SPACES][kBlock];
  __shared__ int64_t sh_death[kBlock];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  double proj[FQ_MAX_SPACES] = {0.0, 0.0}, d0[FQ_MAX_SPACES] = {0.0, 0.0};
  int64_t ndeath = 0;
  if (i < n) {
    uint64_t f[FQ_MAX_WORDS];
    load_det(det, cap, s.W, i, f);
    double rp[FQ_MAX_SPACES] = {0.0, 0.0};
    int64_t p[FQ_MAX_SPACES] = {0, 0};
    for (int sp = 0; sp < s.S; ++sp) {
      p[sp] = pop[sp * cap + i];
      rp[sp] = static_cast<double>(p[sp]) / static_cast<double>(s.rf);
    }
    const int64_t ci = s.ndeterm > 0 ? fqd::core_index(s, f) : -1;
    if (ci >= 0) {
      for (int sp = 0; sp < s.S; ++sp) dvec[ci * s.S + sp] = rp[sp];
      dpos[ci] = i;
    }
    flags[i] = ci >= 0 ? kCoreFlag : fqd::parent_flag(s, rp);
    const double K = kii[i];
    for (int sp = 0; sp < s.S; ++sp) {
      fqd::proj_energy_term(s, f, rp[sp], proj[sp], d0[sp]);
      const uint32_t hi = static_cast<uint32_t>(i >> 32) ^ (static_cast<uint32_t>(sp) << 28);
      fqd::Philox rng(seed, static_cast<uint32_t>(i), hi, cyc ^ (fqd::kDrawNattempts << 24));
      nat[i * s.S + sp] = fqd::decide_nattempts(rng, rp[sp]);
      int64_t q = p[sp];
      if (ci < 0) {
        fqd::Philox rd(seed, static_cast<uint32_t>(i), hi, cyc ^ (fqd::kDrawDeath << 24));
        ndeath += fqd::death(rd, tau, K, sp == 0 ? sh0 : sh1, sp == 0 ? e0 : e1, q);
      }
      pop_new[sp * cap + i] = q;
    }
  }
  for (int sp = 0; sp < FQ_MAX_SPACES; ++sp) {
    sh_d[2 * sp][threadIdx.x] = proj[sp];
    sh_d[2 * sp + 1][threadIdx.x] = d0[sp];
  }
  sh_death[threadIdx.x] = ndeath;
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      for (int q = 0; q < 2 * FQ_MAX_SPACES; ++q) sh_d[q][threadIdx.x] += sh_d[q][threadIdx.x + k];
      sh_death[threadIdx.x] += sh_death[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int q = 0; q < 2 * FQ_MAX_SPACES; ++q) part_d[q * gridDim.x + blockIdx.x] = sh_d[q][0];
    part_i[blockIdx.x] = sh_death[0];
  }
}

// Fixed-order reduction of the k_det_pass block partials: thread t sums
// blocks t, t+256, ... in order, then a tree in LDS.
__global__ __launch_bounds__(kBlock) void k_sum_partials(const double* __restrict__ part_d,
                                                         const int64_t* __restrict__ part_i, int nb,
                                                         Stats* __restrict__ st) {
  __shared__ double sh[2 * FQ_MAX_SPACES][kBlock];
  __shared__ int64_t shi[kBlock];
  double acc[2 * FQ_MAX_SPACES] = {0.0, 0.0, 0.0, 0.0};
  int64_t ai = 0;
  // blocks t, t + 256, ... added in that order; eight are loaded ahead at a time
  constexpr int kAhead = 8;
  for (int b0 = threadIdx.x; b0 < nb; b0 += kAhead * kBlock) {
    double v[kAhead][2 * FQ_MAX_SPACES];
    int64_t vi[kAhead];
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
      const int b = b0 + k * kBlock;
#pragma unroll
      for (int q = 0; q < 2 * FQ_MAX_SPACES; ++q) v[k][q] = b < nb ? part_d[q * nb + b] : 0.0;
      vi[k] = b < nb ? part_i[b] : 0;
    }
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
      if (b0 + k * kBlock < nb) {
#pragma unroll
        for (int q = 0; q < 2 * FQ_MAX_SPACES; ++q) acc[q] = acc[q] + v[k][q];
        ai += vi[k];
      }
    }
  }
  for (int q = 0; q < 2 * FQ_MAX_SPACES; ++q) sh[q][threadIdx.x] = acc[q];
  shi[threadIdx.x] = ai;
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      for (int q = 0; q < 2 * FQ_MAX_SPACES; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + k];
      shi[threadIdx.x] += shi[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int sp = 0; sp < FQ_MAX_SPACES; ++sp) {
      st->proj[sp] = sh[2 * sp][0];
      st->d0[sp] = sh[2 * sp + 1][0];
    }
    st->ndeath = shi[0];
  }
}

// Last entry e in [lo, hi] with att_off[e] <= t.
__device__ __forceinline__ int64_t parent_of(const int64_t* __restrict__ att_off, int64_t t, int64_t lo, int64_t hi) {
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (att_off[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

constexpr int kSpawnTile = 4 * kBlock;  // attempts per block tile
// Philox blocks of an attempt's stream generated before its generator runs
// (two draws each: the single/double choice, the orbital choices with their
// first rejection redraws, the spawn rounding).  On mol26 (renormalised
// generator) k_spawn took 474, 465, 454 and 425 us with 0, 2, 3 and 4 blocks;
// 4 keeps its 80 VGPRs (6 waves per SIMD).
constexpr int kSpawnPre = 4;
constexpr int kRecipN = 4096;  // entries of the host-made reciprocal table

// Parents of each spawning tile's first and last attempt (one thread per
// tile, so the searches' latencies overlap instead of opening every block).
// Parent windows of the tiles of the batch [t_begin, t_end), t_end clamped to
// the attempt total att_off[nent] (the host may only know a bound of it);
// tiles of the grid past the total get a zero count.
__global__ void k_tile_bounds(int64_t nent, const int64_t* __restrict__ att_off, int64_t t_begin, int64_t t_end,
                              int64_t* __restrict__ tile_win, int64_t* __restrict__ tile_cnt) {
  const int64_t nt_batch = (t_end - t_begin + kSpawnTile - 1) / kSpawnTile;  // tiles the host sized for
  const int64_t tot = att_off[nent];
  t_end = t_end < tot ? t_end : tot;
  const int64_t ntiles = t_end > t_begin ? (t_end - t_begin + kSpawnTile - 1) / kSpawnTile : 0;
  const int64_t tile = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (tile >= ntiles) {
    if (tile < nt_batch) tile_cnt[tile] = 0;
    return;
  }
  const int64_t t0 = t_begin + tile * kSpawnTile;
  const int64_t t1 = (t0 + kSpawnTile < t_end) ? t0 + kSpawnTile : t_end;
  const int64_t p0 = parent_of(att_off, t0, 0, nent - 1);
  tile_win[2 * tile] = p0;
  tile_win[2 * tile + 1] = parent_of(att_off, t1 - 1, p0, nent - 1);
}
constexpr int kSpawnWin = 1024;          // parent-window entries staged in LDS

// ---------------------------------------------------------------------------
// Sort keys.  A spawned entry's key is v << flag_bits | flags over kw 64-bit
// words (kw = 1 or 2): v is the determinant's combinatorial rank (or, for
// batched instances, its encoded word 0, instance index included), flags its
// initiator flag bits (one per space).  Ordering the v fields orders
// determinants as the main list does (last word most significant,
// src/bit_utils.F90:452-480); the LSD radix passes order only the digits of
// v and are stable, so equal determinants keep spawned (attempt) order.  The
// entry's populations travel with the key through the passes (the payload),
// so the run reduction reads everything contiguously, and the determinant is
// recovered from v.
//
// Combinatorial rank: for a bit string with set bits p_1 < ... < p_k
// (0-based), rank = sum_i C(p_i, i).  For a fixed number of set bits this is
// a bijection onto [0, C(nbasis, k)) that preserves the strings' numeric
// order, so the sort needs ceil(log2 C(nbasis, nel)) bits instead of nbasis
// (34 instead of 52 for 10 electrons in 52 spin orbitals).  binom holds
// C(p, i) for p < nbasis, 1 <= i <= nel as (low, high) 64-bit halves.
// ---------------------------------------------------------------------------
struct KeyPlan {
  int kw;          // key words
  int flag_bits;   // width of the flag field (the number of spaces)
  int rank;        // 1: v = combinatorial rank; 0: v = determinant word 0
  int nel, nbasis;
  const uint64_t* binom;
  int binom_lds;   // stage the table in dynamic LDS (k_pack_tiles, k_runs)
};
constexpr int kMaxPasses = 16;
struct DigitPlan {
  int np;                  // passes
  int shift[kMaxPasses];   // first key bit of each pass's digit
  int bits[kMaxPasses];    // digit width (8 or 9)
};
// Keys per thread of a sort tile: the tile's keys and payload sit in LDS.
__host__ __device__ constexpr int sort_items(int kw, int S) { return kw + S <= 2 ? 16 : 8; }
constexpr int kSortBlocks = 512;  // blocks of a radix pass (each owns a contiguous range of tiles)

__device__ __forceinline__ void make_key(const KeyPlan& kp, const uint64_t* f, int W, uint64_t flags, uint64_t* key) {
  unsigned __int128 v = 0;
  if (kp.rank) {
    int k = 0;
    FQ_FOR_WORDS(w, W) {
      for (uint64_t x = f[w]; x && k < kp.nel; x &= x - 1) {
        const int p = (w << 6) + __builtin_ctzll(x);
        if (p >= kp.nbasis) break;  // not a determinant of this system (caller-supplied list)
        const uint64_t* c = kp.binom + 2 * (static_cast<int64_t>(p) * kp.nel + k);
        v += (static_cast<unsigned __int128>(c[1]) << 64) | c[0];
        ++k;
      }
    }
  } else {
    v = f[0];
  }
  const unsigned __int128 k = (v << kp.flag_bits) | (flags & ((uint64_t{1} << kp.flag_bits) - 1));
  key[0] = static_cast<uint64_t>(k);
  key[1] = static_cast<uint64_t>(k >> 64);
}

// Production: one thread per spawning attempt.  Each block takes tiles of
// kSpawnTile consecutive attempts of the batch [t_begin, t_end); the parents
// of a tile's first and last attempts bound a short window of the attempt
// prefix sum (n*S (det, space) entries, det-major) in which each thread
// searches its own parent.  The successful attempts of a tile are ranked in
// attempt order (wavefront ballots plus per-wave totals) and written
// contiguously from the tile's first staging slot, with the tile's count in
// tile_cnt; k_pack_tiles then copies the tiles, in order, into the spawned list.
template <int G, bool B>
__global__ __launch_bounds__(kBlock) void k_spawn(const Sys* __restrict__ sdev, int64_t nent, int64_t cap, int64_t t_begin, int64_t t_end,
                                                  const int64_t* __restrict__ att_off,
                                                  const uint64_t* __restrict__ det, const int64_t* __restrict__ pop,
                                                  double tau, uint64_t seed, uint32_t cyc, int64_t bcap,
                                                  uint64_t* __restrict__ st_det, int64_t* __restrict__ st_pop,
                                                  int64_t* __restrict__ st_ent, int64_t* __restrict__ tile_cnt,
                                                  const int64_t* __restrict__ tile_win,
                                                  const int64_t* __restrict__ pflags,
                                                  const Sys* __restrict__ sys_arr, const int64_t* __restrict__ iatt,
                                                  const double* __restrict__ itau, const uint32_t* __restrict__ pcyc) {
  cyc = cycle_of(pcyc, cyc);
  // the system descriptor is read from device memory (scalar loads on use)
  // rather than passed by value: its many table pointers would otherwise sit
  // in scalar registers for the whole kernel and spill into vector lanes
  const Sys& s = *sdev;
  __shared__ int32_t wcount[kBlock / 64];
  // att_off[lo + k] - t0 (clamped below at -1: every such entry starts at or
  // before the tile's first attempt) when the window fits
  __shared__ int32_t watt[kSpawnWin];
  // heat-bath: each thread's nel on-the-fly weights in a column of dynamic LDS
  fqd::HbBuf hb;
  if constexpr (G == fqd::kGenHeatBath) {
    extern __shared__ double hb_lds[];
    hb = fqd::HbBuf{hb_lds + threadIdx.x, kBlock};
  }
  // Renormalised and learned generators decide single vs double on an
  // attempt's first draw: those tiles are processed doubles first and singles
  // last (their code paths differ), results parked in LDS by attempt, so a
  // wavefront seldom runs both paths; successes are still compacted in
  // attempt order.
  constexpr bool kPart = (G == fqd::kGenMolRenorm || G == fqd::kGenLearned) && !B;
  __shared__ int16_t order[kPart ? kSpawnTile : 1];   // processing slot -> attempt in tile
  __shared__ int64_t rpop[kPart ? kSpawnTile : 1];    // results by attempt in tile
  __shared__ int32_t rent[kPart ? kSpawnTile : 1];    // parent entry - lo
  __shared__ int32_t wcs[kBlock / 64];
  uint64_t* rdet = nullptr;                            // [w * kSpawnTile + r], dynamic LDS
  const float* aff_lds = nullptr;  // learned: affinity table then flat row, after rdet
  const float* flat_lds = nullptr;
  if constexpr (kPart) {
    extern __shared__ uint64_t part_lds[];
    rdet = part_lds;
    if constexpr (G == fqd::kGenLearned) {
      float* a = reinterpret_cast<float*>(part_lds + s.W * kSpawnTile);
      const int na = s.nbasis * s.nbasis;
      for (int k = threadIdx.x; k < na + 2 * s.nbasis; k += kBlock)
        a[k] = k < na ? s.mlp_aff[k] : s.mlp_flat[k - na];
      aff_lds = a;
      flat_lds = a + na;
      __syncthreads();
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1;
  if (const int64_t tot = att_off[nent]; tot < t_end) t_end = tot;  // the host may pass a bound
  const int64_t ntiles = t_end > t_begin ? (t_end - t_begin + kSpawnTile - 1) / kSpawnTile : 0;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t t0 = t_begin + tile * kSpawnTile;
    const int64_t t1 = (t0 + kSpawnTile < t_end) ? t0 + kSpawnTile : t_end;
    const int tn = static_cast<int>(t1 - t0);
    const int64_t lo = tile_win[2 * tile], hi = tile_win[2 * tile + 1];
    const bool staged = hi - lo < kSpawnWin;
    if (staged)
      for (int64_t k = threadIdx.x; k <= hi - lo; k += kBlock) {
        const int64_t v = att_off[lo + k] - t0;
        watt[k] = static_cast<int32_t>(v < -1 ? -1 : v);
      }
    __syncthreads();
    // one spawning attempt t (tile-relative r): its parent entry e, the
    // spawned population and the child determinant
    auto attempt = [&](int64_t t, int r, int64_t& e, int64_t& nspawn, uint64_t* child) {
      if (staged) {  // last k in [0, hi - lo] with att_off[lo + k] <= t
        int a = 0, b = static_cast<int>(hi - lo);
        while (a < b) {
          const int mid = (a + b + 1) >> 1;
          if (watt[mid] <= r)
            a = mid;
          else
            b = mid - 1;
        }
        e = lo + a;
      } else {
        e = parent_of(att_off, t, lo, hi);
      }
      const int sp = static_cast<int>(e % s.S);
      const int64_t i = e / s.S;
      uint64_t f[FQ_MAX_WORDS];
      load_det(det, cap, s.W, i, f);
      const int64_t parent = pop[sp * cap + i];
      if constexpr (B) {
        // batched instances: the instance's tables, time step, key and attempt numbering
        const int inst = static_cast<int>(f[0] >> kInstShift);
        f[0] &= kOrbMask;
        const int64_t trel = t - iatt[inst];
        fqd::PhiloxPre<kSpawnPre> rng(seed + static_cast<uint64_t>(inst), static_cast<uint32_t>(trel),
                                      static_cast<uint32_t>(trel >> 32), cyc ^ (fqd::kDrawAttempt << 24));
        nspawn = fqd::spawn_attempt<G>(sys_arr[inst], rng, f, nullptr, parent, itau[inst], child);
        if (nspawn != 0) child[0] |= static_cast<uint64_t>(inst) << kInstShift;
      } else {
        fqd::PhiloxPre<kSpawnPre> rng(seed, static_cast<uint32_t>(t), static_cast<uint32_t>(t >> 32),
                                      cyc ^ (fqd::kDrawAttempt << 24));
        const float* wrow = (G == fqd::kGenLearned && s.mlp_on[i]) ? s.mlp_tab + s.mlp_row[i] * s.mlp_ld : nullptr;
        nspawn = fqd::spawn_attempt<G>(s, rng, f, wrow, parent, tau, child, hb, aff_lds, flat_lds);
        // a core parent's spawn onto the core is the deterministic projection's
        if (s.ndeterm > 0 && nspawn != 0 && (pflags[i] & kCoreFlag) && fqd::core_index(s, child) >= 0) nspawn = 0;
      }
    };
    if constexpr (kPart) {
      // processing order: doubles from the front, singles from the back
      int32_t nd = 0, ns = 0;
      for (int k = 0; k < kSpawnTile / kBlock; ++k) {
        const int r = k * kBlock + threadIdx.x;
        bool single = false;
        if (r < tn) {
          const int64_t t = t0 + r;
          fqd::Philox rng(seed, static_cast<uint32_t>(t), static_cast<uint32_t>(t >> 32),
                          cyc ^ (fqd::kDrawAttempt << 24));
          single = rng.next() < s.psingle;
        }
        const unsigned long long bd = __ballot(r < tn && !single), bs = __ballot(r < tn && single);
        if (lane == 0) {
          wcount[wave] = __popcll(bd);
          wcs[wave] = __popcll(bs);
        }
        __syncthreads();
        int32_t pd = nd + __popcll(bd & below), ps = ns + __popcll(bs & below);
        for (int w = 0; w < kBlock / 64; ++w) {
          if (w < wave) {
            pd += wcount[w];
            ps += wcs[w];
          }
          nd += wcount[w];
          ns += wcs[w];
        }
        if (r < tn) order[single ? tn - 1 - ps : pd] = static_cast<int16_t>(r);
        __syncthreads();
      }
      for (int k = 0; k < kSpawnTile / kBlock; ++k) {
        const int q = k * kBlock + threadIdx.x;
        if (q < tn) {
          const int r = order[q];
          int64_t e = 0, nspawn = 0;
          uint64_t child[FQ_MAX_WORDS] = {0, 0};
          attempt(t0 + r, r, e, nspawn, child);
          FQ_FOR_WORDS(w, s.W) rdet[w * kSpawnTile + r] = child[w];
          rpop[r] = nspawn;
          rent[r] = static_cast<int32_t>(e - lo);
        }
      }
      __syncthreads();
    }
    int32_t run = 0;  // successes of this tile so far (all threads hold the same value)
    for (int k = 0; k < kSpawnTile / kBlock; ++k) {
      const int r = k * kBlock + threadIdx.x;
      const bool valid = r < tn;
      int64_t e = 0, nspawn = 0;
      uint64_t child[FQ_MAX_WORDS] = {0, 0};
      if (valid) {
        if constexpr (kPart) {
          nspawn = rpop[r];
          e = lo + rent[r];
          FQ_FOR_WORDS(w, s.W) child[w] = rdet[w * kSpawnTile + r];
        } else {
          attempt(t0 + r, r, e, nspawn, child);
        }
      }
      // rank this round's successes in attempt order (wave ballot + per-wave totals)
      const bool ok = valid && nspawn != 0;
      const unsigned long long bal = __ballot(ok);
      if (lane == 0) wcount[wave] = __popcll(bal);
      __syncthreads();
      int32_t pos = run + __popcll(bal & below);
      int32_t round_total = 0;
      for (int w = 0; w < kBlock / 64; ++w) {
        if (w < wave) pos += wcount[w];
        round_total += wcount[w];
      }
      if (ok) {  // compacted into the tile's region of the staging arrays
        const int64_t u = (t0 - t_begin) + pos;
        FQ_FOR_WORDS(w, s.W) st_det[w * bcap + u] = child[w];
        st_pop[u] = nspawn;
        // the spawning space and the parent's initiator flags (the sort key's flag bits)
        st_ent[u] = (e % s.S) | (s.initiator ? (pflags[e / s.S] & kInitFlagMask) << 8 : 0);
      }
      run += round_total;
      __syncthreads();
    }
    if (threadIdx.x == 0) tile_cnt[tile] = run;
    __syncthreads();
  }
}

// Pack each tile's compacted successes into the spawned list after the entries
// of earlier batches (st->nspawn), in attempt order, as sort keys (made here,
// the binomial table staged in LDS) and payload (the population in its
// space's word, 0 in the other); blocks loop over tiles.  st_ent holds the
// spawning space and, from bit 8, the parent's initiator flags.
__global__ __launch_bounds__(kBlock) void k_pack_tiles(int W, int S, int64_t ntiles,
                                                       const int64_t* __restrict__ tile_cnt,
                                                       const int64_t* __restrict__ tile_off,
                                                       const uint64_t* __restrict__ st_det,
                                                       const int64_t* __restrict__ st_pop,
                                                       const int64_t* __restrict__ st_ent, int64_t bcap,
                                                       int64_t scap, KeyPlan kp, uint64_t* __restrict__ keys,
                                                       int64_t* __restrict__ pay, Stats* __restrict__ st) {
  extern __shared__ uint64_t binom_lds[];
  if (kp.rank && kp.binom_lds) {  // blocks loop over tiles, so each stages the table once
    const int nw = 2 * kp.nbasis * kp.nel;
    for (int k = threadIdx.x; k < nw; k += kBlock) binom_lds[k] = kp.binom[k];
    __syncthreads();
    kp.binom = binom_lds;
  }
  const int64_t nspawn0 = st->nspawn;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t cnt = tile_cnt[tile];
    const int64_t base = nspawn0 + tile_off[tile];
    for (int64_t j = threadIdx.x; j < cnt; j += kBlock) {
      const int64_t u = tile * kSpawnTile + j;
      const int64_t slot = base + j;
      if (slot >= scap) {
        st->spawn_overflow = 1;
        continue;
      }
      const int64_t e = st_ent[u];
      const int sp = static_cast<int>(e & 0xff);
      uint64_t f[FQ_MAX_WORDS] = {0, 0};
      FQ_FOR_WORDS(w, W) f[w] = st_det[w * bcap + u];
      const int64_t v = st_pop[u];
      for (int q = 0; q < S; ++q) pay[q * scap + slot] = (q == sp) ? v : 0;
      uint64_t k[2];
      make_key(kp, f, W, static_cast<uint64_t>(e) >> 8, k);
      for (int w = 0; w < kp.kw; ++w) keys[w * scap + slot] = k[w];
    }
  }
}

__global__ void k_add_spawned(const int64_t* __restrict__ pos, int64_t nb, const int64_t* __restrict__ ptot,
                              Stats* __restrict__ st, int64_t tlim) {
  st->nspawn += pos[nb];
  st->total_attempts = *ptot;
  if (tlim >= 0 && *ptot > tlim) st->spawn_overflow = 1;  // attempts past the launches' bound were not made
}

// ---------------------------------------------------------------------------
// Batched instances
// ---------------------------------------------------------------------------
// bound[k] = first position in the sorted list whose instance index is >= k.
__global__ void k_inst_bounds(const uint64_t* __restrict__ w0, int64_t n, int ninst, int shift,
                              int64_t* __restrict__ bound, const int64_t* __restrict__ pn) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > ninst) return;
  n = list_len(pn, n);
  if (k == ninst) {
    bound[k] = n;
    return;
  }
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (static_cast<int>(w0[mid] >> shift) < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  bound[k] = lo;
}

// k_det_pass for a batched list: each determinant uses its instance's tables,
// time step, shift and key; Philox counters count determinants from the
// instance's first, so an instance follows the same chain as alone.  The
// estimator terms go to per-determinant arrays for per-instance sums.
__global__ __launch_bounds__(kBlock) void k_det_pass_batch(Sys s, const Sys* __restrict__ sys_arr,
                                                           const int64_t* __restrict__ bound,
                                                           const double* __restrict__ ipar, int ninst, int64_t n,
                                                           int64_t cap, const uint64_t* __restrict__ det,
                                                           const int64_t* __restrict__ pop,
                                                           const double* __restrict__ kii,
                                                           int64_t* __restrict__ pop_new, int64_t* __restrict__ nat,
                                                           int64_t* __restrict__ flags, uint64_t seed, uint32_t cyc,
                                                           double* __restrict__ cterm,
                                                           int64_t* __restrict__ cdeath,
                                                           const int64_t* __restrict__ pn,
                                                           const uint32_t* __restrict__ pcyc) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  if (i >= list_len(pn, n)) {  // past the list in a launch sized by a bound: no attempts
    for (int sp = 0; sp < s.S; ++sp) nat[i * s.S + sp] = 0;
    return;
  }
  cyc = cycle_of(pcyc, cyc);
  uint64_t f[FQ_MAX_WORDS];
  load_det(det, cap, s.W, i, f);
  const int inst = static_cast<int>(f[0] >> kInstShift);
  f[0] &= kOrbMask;
  const Sys& si = sys_arr[inst];
  const int64_t irel = i - bound[inst];
  const double tau = ipar[inst];
  double rp[FQ_MAX_SPACES] = {0.0, 0.0};
  int64_t p[FQ_MAX_SPACES] = {0, 0};
  for (int sp = 0; sp < s.S; ++sp) {
    p[sp] = pop[sp * cap + i];
    rp[sp] = static_cast<double>(p[sp]) / static_cast<double>(s.rf);
  }
  flags[i] = fqd::parent_flag(s, rp);
  const double K = kii[i];
  int64_t ndeath = 0;
  for (int sp = 0; sp < s.S; ++sp) {
    double proj = 0.0, d0 = 0.0;
    fqd::proj_energy_term(si, f, rp[sp], proj, d0);
    cterm[(2 * sp) * cap + i] = proj;
    cterm[(2 * sp + 1) * cap + i] = d0;
    const uint32_t hi = static_cast<uint32_t>(irel >> 32) ^ (static_cast<uint32_t>(sp) << 28);
    const uint64_t key = seed + static_cast<uint64_t>(inst);
    fqd::Philox rng(key, static_cast<uint32_t>(irel), hi, cyc ^ (fqd::kDrawNattempts << 24));
    nat[i * s.S + sp] = fqd::decide_nattempts(rng, rp[sp]);
    fqd::Philox rd(key, static_cast<uint32_t>(irel), hi, cyc ^ (fqd::kDrawDeath << 24));
    int64_t q = p[sp];
    const double sh = ipar[ninst + sp * ninst + inst], eo = ipar[(1 + s.S) * ninst + sp * ninst + inst];
    ndeath += fqd::death(rd, tau, K, sh, eo, q);
    pop_new[sp * cap + i] = q;
  }
  cdeath[i] = ndeath;
}

// First attempt index of each instance: attempts are numbered over the
// (determinant, space) entries in list order.
__global__ void k_inst_att(int ninst, int S, const int64_t* __restrict__ bound, const int64_t* __restrict__ att_off,
                           int64_t* __restrict__ iatt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ninst) iatt[k] = att_off[S * bound[k]];
}

// Per-instance cycle results: projected-energy terms, N_0 and deaths over the
// old list, attempts, spawned entries, occupied determinants and sum |N| over
// the new list.  The floating-point sums repeat the single-instance reduction
// exactly: blocks of 256 determinants counted from the instance's first, each
// reduced by the k_det_pass tree (k_inst_part, one block per such block of
// every instance), then block b added to partial b mod 256 in order and the
// k_sum_partials tree (k_inst_stats, one block per instance), so a batched
// instance reports the same estimators, bit for bit, as it alone.
// Blocks of 256 entries per instance of a list with instance bounds b.
__global__ void k_inst_nblk(int ninst, const int64_t* __restrict__ b, int64_t* __restrict__ cnt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ninst) cnt[k] = (b[k + 1] - b[k] + kBlock - 1) / kBlock;
}
// The instance owning global block blk: the last k with off[k] <= blk.
__device__ __forceinline__ int inst_of_block(const int64_t* __restrict__ off, int ninst, int64_t blk) {
  int lo = 0, hi = ninst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= blk)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}
// One block per 256-entry block of every instance (grid: a bound of their
// number).  Old list (ob, off): the tree sums of the 2S estimator terms and the
// deaths; new list (nbd, noff): sum |N| per space.
__global__ __launch_bounds__(kBlock) void k_inst_part(int ninst, int S, int64_t cap, const int64_t* __restrict__ ob,
                                                      const int64_t* __restrict__ off, const double* __restrict__ cterm,
                                                      const int64_t* __restrict__ cdeath,
                                                      const int64_t* __restrict__ nbd, const int64_t* __restrict__ noff,
                                                      const int64_t* __restrict__ pop, int64_t pb,
                                                      double* __restrict__ part_d, int64_t* __restrict__ part_i) {
  __shared__ double shd[2 * FQ_MAX_SPACES][kBlock];
  __shared__ int64_t wsh[kBlock / 64];
  const int tid = threadIdx.x;
  const int64_t blk = blockIdx.x;
  const int64_t nold = off[ninst], nnew = noff[ninst];
  if (blk < nold) {
    const int k = inst_of_block(off, ninst, blk);
    const int64_t i = ob[k] + (blk - off[k]) * kBlock + tid;
    const bool in = i < ob[k + 1];
    for (int q = 0; q < 2 * FQ_MAX_SPACES; ++q) shd[q][tid] = (q < 2 * S && in) ? cterm[q * cap + i] : 0.0;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
      if (tid < h)
        for (int q = 0; q < 2 * FQ_MAX_SPACES; ++q) shd[q][tid] += shd[q][tid + h];
      __syncthreads();
    }
    const int64_t d = block_sum(in ? cdeath[i] : 0, wsh);
    if (tid == 0) {
      for (int q = 0; q < 2 * S; ++q) part_d[q * pb + blk] = shd[q][0];
      part_i[blk] = d;
    }
  }
  if (blk < nnew) {
    const int k = inst_of_block(noff, ninst, blk);
    const int64_t i = nbd[k] + (blk - noff[k]) * kBlock + tid;
    const bool in = i < nbd[k + 1];
    for (int sp = 0; sp < S; ++sp) {
      int64_t p = in ? pop[sp * cap + i] : 0;
      p = p < 0 ? -p : p;
      const int64_t t = block_sum(p, wsh);
      if (tid == 0) part_i[(1 + sp) * pb + blk] = t;
    }
  }
}
// One block per instance: its block partials added in the k_sum_partials
// order, then the outputs.
__global__ __launch_bounds__(kBlock) void k_inst_stats(int ninst, int S, const int64_t* __restrict__ ob,
                                                       const int64_t* __restrict__ nbd, const int64_t* __restrict__ sb,
                                                       const int64_t* __restrict__ off,
                                                       const int64_t* __restrict__ noff, int64_t pb,
                                                       const double* __restrict__ part_d,
                                                       const int64_t* __restrict__ part_i,
                                                       const int64_t* __restrict__ att_off,
                                                       double* __restrict__ out_d, int64_t* __restrict__ out_i) {
  __shared__ double shd[2 * FQ_MAX_SPACES][kBlock];
  __shared__ int64_t shi[1 + FQ_MAX_SPACES][kBlock];
  const int k = blockIdx.x, tid = threadIdx.x;
  const int nq = 2 * S;
  double acc[2 * FQ_MAX_SPACES] = {0.0, 0.0, 0.0, 0.0};
  int64_t acci[1 + FQ_MAX_SPACES] = {0, 0, 0};
  for (int64_t vb = tid; vb < off[k + 1] - off[k]; vb += kBlock) {
    for (int q = 0; q < nq; ++q) acc[q] = acc[q] + part_d[q * pb + off[k] + vb];
    acci[0] += part_i[off[k] + vb];
  }
  for (int64_t vb = tid; vb < noff[k + 1] - noff[k]; vb += kBlock)
    for (int sp = 0; sp < S; ++sp) acci[1 + sp] += part_i[(1 + sp) * pb + noff[k] + vb];
  for (int q = 0; q < 2 * FQ_MAX_SPACES; ++q) shd[q][tid] = acc[q];
  for (int q = 0; q < 1 + FQ_MAX_SPACES; ++q) shi[q][tid] = acci[q];
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if (tid < h) {
      for (int q = 0; q < 2 * FQ_MAX_SPACES; ++q) shd[q][tid] += shd[q][tid + h];
      for (int q = 0; q < 1 + FQ_MAX_SPACES; ++q) shi[q][tid] += shi[q][tid + h];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const int64_t a = ob[k], b = ob[k + 1];
    for (int q = 0; q < nq; ++q) out_d[k * nq + q] = shd[q][0];
    int64_t* o = out_i + k * (4 + S);
    o[0] = shi[0][0];                          // deaths
    o[1] = att_off[S * b] - att_off[S * a];    // attempts
    o[2] = sb[k + 1] - sb[k];                  // spawned entries
    o[3] = nbd[k + 1] - nbd[k];                // occupied determinants
    for (int sp = 0; sp < S; ++sp) o[4 + sp] = shi[1 + sp][0];
  }
}

// Replay: the whole spawning/death loop on one thread, in reference order.
// determ_proj_combined_annihil (src/semi_stoch.F90:897-975) for one core
// determinant and space: -tau (sum_j H_ij v_j + (E_old * 0 - S) v_i), the
// row summed in column order (csrpgemv_single_row, lib/local/csr.f90:206-235);
// 0 is rho - w, the quasi-Newton weight difference without quasi-Newton.
__device__ __forceinline__ double determ_combined(const DetermDev& dd, int S, int64_t row, int sp, double tau,
                                                  double shift, double eold) {
  double out = 0.0;
  for (int64_t k = dd.row[row]; k < dd.row[row + 1]; ++k) out = out + dd.mat[k] * dd.vec[dd.col[k] * S + sp];
  return -tau * (out + (eold * (1.0 - 1.0) - 1.0 * shift) * dd.vec[row * S + sp]);
}
// determ_proj_separate_annihil (src/semi_stoch.F90:1009-1094): the shift term
// -(tau (E_old * 0 - S)) v_i, then (-tau T_ik) v_k added over the stored
// transpose's row i in order, as the reference's transposed product
// (csrpgemv, lib/local/csr.f90:148-204) accumulates element i.
__device__ __forceinline__ double determ_separate(const DetermDev& dd, int S, int64_t row, int sp, double tau,
                                                  double shift, double eold) {
  const double a = -1.0 * tau;
  double y = -(tau * (eold * (1.0 - 1.0) - shift * 1.0)) * dd.vec[row * S + sp];
  for (int64_t k = dd.row[row]; k < dd.row[row + 1]; ++k) y = y + a * dd.mat[k] * dd.vec[dd.col[k] * S + sp];
  return y;
}

// Semi-stochastic core (dd.nd > 0, set_determ_info and the determ_parent
// branches of src/fciqmc.f90:330-372): a core parent is an initiator, its
// spawns onto the core are cancelled and its death is left to the
// deterministic projection, which follows the loop (determ_projection,
// src/semi_stoch.F90:859-1094) with one rounding draw per core determinant
// and space.
__global__ void k_replay_walk(Sys s, int64_t n, int64_t cap, const uint64_t* __restrict__ det,
                              const int64_t* __restrict__ pop, const double* __restrict__ kii,
                              int64_t* __restrict__ pop_new, double tau, double sh0, double sh1, double e0,
                              double e1, const double* __restrict__ rbuf, int64_t rpos, int64_t rend, int64_t scap,
                              uint64_t* __restrict__ sdet, int64_t* __restrict__ srec, Stats* __restrict__ st,
                              DetermDev dd) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  fqd::ReplayRng rng{rbuf, rpos, rend, false};
  double proj[FQ_MAX_SPACES] = {0.0, 0.0}, d0[FQ_MAX_SPACES] = {0.0, 0.0};
  int64_t ndeath = 0, nspawn_total = 0, attempts = 0;
  bool overflow = false;
  for (int64_t i = 0; i < n && !rng.failed; ++i) {
    uint64_t f[FQ_MAX_WORDS];
    load_det(det, cap, s.W, i, f);
    double rp[FQ_MAX_SPACES] = {0.0, 0.0};
    int64_t p[FQ_MAX_SPACES] = {0, 0};
    for (int sp = 0; sp < s.S; ++sp) {
      p[sp] = pop[sp * cap + i];
      rp[sp] = static_cast<double>(p[sp]) / static_cast<double>(s.rf);
    }
    const int64_t ci = dd.nd > 0 ? fqd::core_index(s, f) : -1;
    if (ci >= 0) {
      for (int sp = 0; sp < s.S; ++sp) dd.vec[ci * s.S + sp] = rp[sp];
      dd.pos[ci] = i;
    }
    const int64_t flag = ci >= 0 ? 0 : fqd::parent_flag(s, rp);
    for (int sp = 0; sp < s.S; ++sp) {
      fqd::proj_energy_term(s, f, rp[sp], proj[sp], d0[sp]);
      const int nat = fqd::decide_nattempts(rng, rp[sp]);
      attempts += nat;
      for (int a = 0; a < nat && !rng.failed; ++a) {
        uint64_t child[FQ_MAX_WORDS] = {0, 0};
        int64_t ns = fqd::spawn_attempt(s, rng, f, nullptr, p[sp], tau, child);
        if (ns != 0 && ci >= 0 && fqd::core_index(s, child) >= 0) ns = 0;
        if (ns != 0) {
          if (nspawn_total + 1 > scap) {
            overflow = true;
          } else {
            const int64_t slot = nspawn_total++;
            FQ_FOR_WORDS(w, s.W) sdet[w * scap + slot] = child[w];
            int64_t* rec = srec + slot * rec_stride(s.S);
            for (int q = 0; q < s.S; ++q) rec[q] = (q == sp) ? ns : 0;
            rec[s.S] = s.initiator ? flag : 0;
          }
        }
      }
      int64_t q = p[sp];
      if (ci < 0) ndeath += fqd::death(rng, tau, kii[i], sp == 0 ? sh0 : sh1, sp == 0 ? e0 : e1, q);
      p[sp] = q;
      pop_new[sp * cap + i] = q;
    }
  }
  if (dd.nd > 0 && !rng.failed) {
    const double sh[FQ_MAX_SPACES] = {sh0, sh1}, eo[FQ_MAX_SPACES] = {e0, e1};
    if (!dd.separate) {
      for (int64_t row = 0; row < dd.nd && !rng.failed; ++row)
        for (int sp = 0; sp < s.S; ++sp) {
          const int64_t ns = fqd::stoch_encode(rng, determ_combined(dd, s.S, row, sp, tau, sh[sp], eo[sp]), s.rf);
          if (nspawn_total + 1 > scap) {
            overflow = true;
            continue;
          }
          const int64_t slot = nspawn_total++;
          FQ_FOR_WORDS(w, s.W) sdet[w * scap + slot] = s.determ[w * dd.nd + row];
          int64_t* rec = srec + slot * rec_stride(s.S);
          for (int q = 0; q < s.S; ++q) rec[q] = (q == sp) ? ns : 0;
          rec[s.S] = 0;
        }
    } else {
      // every row's projection first, then the rounding draws row by row
      for (int64_t row = 0; row < dd.nd; ++row)
        for (int sp = 0; sp < s.S; ++sp)
          dd.vec[(dd.nd + row) * s.S + sp] = determ_separate(dd, s.S, row, sp, tau, sh[sp], eo[sp]);
      for (int64_t row = 0; row < dd.nd && !rng.failed; ++row)
        for (int sp = 0; sp < s.S; ++sp) {
          const int64_t ns = fqd::stoch_encode(rng, dd.vec[(dd.nd + row) * s.S + sp], s.rf);
          pop_new[sp * cap + dd.pos[row]] += ns;
        }
    }
  }
  for (int sp = 0; sp < FQ_MAX_SPACES; ++sp) {
    st->proj[sp] = proj[sp];
    st->d0[sp] = d0[sp];
  }
  st->ndeath = ndeath;
  st->nspawn = nspawn_total;
  st->spawn_overflow = overflow ? 1 : 0;
  st->total_attempts = attempts;
  st->rng_pos = rng.pos;
  st->rng_failed = rng.failed ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Kernels: sorting
// ---------------------------------------------------------------------------
// Unsigned multi-word comparison, last word most significant (bit_str_cmp).
__device__ __forceinline__ int key_cmp(const uint64_t* keys, int64_t stride, int W, int64_t a, int64_t b) {
  for (int w = W - 1; w >= 0; --w) {
    const uint64_t x = keys[w * stride + a], y = keys[w * stride + b];
    if (x < y) return -1;
    if (x > y) return 1;
  }
  return 0;
}

// Sort passes use DB-bit digits (8 or 9; 1 << DB buckets) of the key's value
// field: the host mixes the two widths to cover it in the fewest passes
// (34 bits: 9+9+8+8).  Digit of a KW-word key starting at bit `shift`.
template <int KW>
__device__ __forceinline__ uint32_t key_digit(const uint64_t* k, int shift, uint32_t mask) {
  if constexpr (KW == 1) {
    return static_cast<uint32_t>(k[0] >> shift) & mask;
  } else {
    const uint64_t x = shift >= 64 ? (k[1] >> (shift - 64))
                       : shift == 0 ? k[0]
                                    : ((k[0] >> shift) | (k[1] << (64 - shift)));
    return static_cast<uint32_t>(x) & mask;
  }
}

// Lanes of the wavefront holding the same DB-bit digit as this lane (among
// the `valid` ones): DB ballots.
template <int DB>
__device__ __forceinline__ unsigned long long match_digit(uint32_t d, bool valid) {
  unsigned long long peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < DB; ++b) {
    const bool bit = (d >> b) & 1;
    const unsigned long long m = __ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

// Exclusive scan of x over the block (every thread receives its prefix);
// tot receives the block total.
__device__ __forceinline__ uint32_t block_exclusive_u32(uint32_t x, uint32_t* wsh, uint32_t& tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inc = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
  if (lane == 63) wsh[wave] = inc;
  __syncthreads();
  uint32_t pre = 0;
  tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    pre += w < wave ? wsh[w] : 0;
    tot += wsh[w];
  }
  __syncthreads();
  return pre + inc - x;
}

// Length of the spawned list: the device count (clamped to the host bound n,
// the capacity at most) when pn is given, else n itself.

// Radix pass, reduce-then-scan over a fixed grid: block b owns the sort tiles
// [b * tpb, (b + 1) * tpb) (tiles of kBlock * sort_items keys; tpb from the
// list length on the device and the grid, so the host need not know the
// length, only a bound of it).  k_pass_count writes the
// block's digit counts digit-major, cnt[d * nblk + b]; their exclusive scan
// gives every block the global start of each digit's keys from its range, and
// k_pass_scatter walks its tiles in order, ranks each tile's keys by digit in
// input order (stable), and writes keys and payload (S population words) to
// their sorted positions.  No block waits on another.
template <int KW, int S, int DB>
__global__ __launch_bounds__(kBlock) void k_pass_count(const uint64_t* __restrict__ keys,
                                                       const int64_t* __restrict__ pn, int64_t nb, int64_t stride,
                                                       int shift, int64_t* __restrict__ cnt) {
  constexpr int NB = 1 << DB, R = sort_items(KW, S), T = kBlock * R;
  __shared__ uint32_t h[NB];
  for (int d = threadIdx.x; d < NB; d += kBlock) h[d] = 0;
  __syncthreads();
  const int64_t n = list_len(pn, nb);
  const int64_t tpb = ((n + T - 1) / T + gridDim.x - 1) / gridDim.x;
  const int lane = threadIdx.x & 63;
  const int64_t lo = static_cast<int64_t>(blockIdx.x) * tpb * T;
  const int64_t hi = lo + tpb * T < n ? lo + tpb * T : n;
  for (int64_t base = lo; base < hi; base += T) {  // block-uniform rounds (wavefront ballots)
    uint32_t d[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {  // every load of the tile in flight before the ballots
      const int64_t i = base + r * kBlock + threadIdx.x;
      uint64_t k[KW];
#pragma unroll
      for (int w = 0; w < KW; ++w) k[w] = i < hi ? keys[w * stride + i] : 0;
      d[r] = key_digit<KW>(k, shift, NB - 1);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const bool valid = base + r * kBlock + threadIdx.x < hi;
      const unsigned long long peers = match_digit<DB>(d[r], valid);
      if (valid && (peers & ((1ull << lane) - 1)) == 0)
        atomicAdd(&h[d[r]], static_cast<uint32_t>(__popcll(peers)));
    }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < NB; d += kBlock) cnt[static_cast<int64_t>(d) * gridDim.x + blockIdx.x] = h[d];
}

// Per digit d (one block each, nb <= kSortBlocks block counts in its row):
// the row's exclusive prefix, in place, and the digit total in tot[d].
__global__ __launch_bounds__(kSortBlocks) void k_pass_rowscan(int nb, int64_t* __restrict__ cnt,
                                                              int64_t* __restrict__ tot) {
  __shared__ int64_t wsh[kSortBlocks / 64];
  const int d = blockIdx.x, b = threadIdx.x, lane = b & 63, wave = b >> 6;
  const int64_t v = b < nb ? cnt[static_cast<int64_t>(d) * nb + b] : 0;
  const int64_t x = wave_inclusive_scan(v, lane);
  if (lane == 63) wsh[wave] = x;
  __syncthreads();
  int64_t pre = 0, t = 0;
#pragma unroll
  for (int w = 0; w < kSortBlocks / 64; ++w) {
    pre += w < wave ? wsh[w] : 0;
    t += wsh[w];
  }
  if (b < nb) cnt[static_cast<int64_t>(d) * nb + b] = pre + x - v;
  if (b == 0) tot[d] = t;
}

template <int KW, int S, int DB>
__global__ __launch_bounds__(kBlock) void k_pass_scatter(const uint64_t* __restrict__ kin,
                                                         const int64_t* __restrict__ pin, uint64_t* __restrict__ kout,
                                                         int64_t* __restrict__ pout, const int64_t* __restrict__ pn,
                                                         int64_t nb, int64_t stride, int shift,
                                                         const int64_t* __restrict__ off,
                                                         const int64_t* __restrict__ dtot) {
  constexpr int NB = 1 << DB, R = sort_items(KW, S), T = kBlock * R, NW = kBlock / 64;
  const int64_t n = list_len(pn, nb);
  const int64_t tpb = ((n + T - 1) / T + gridDim.x - 1) / gridDim.x;
  constexpr uint32_t MASK = NB - 1;
  __shared__ uint64_t sk[KW][T];
  __shared__ int64_t sp[S][T];
  __shared__ uint32_t wc[NW][NB];  // per-wave running digit counts, then per-wave exclusive prefixes
  __shared__ uint32_t toff[NB];    // tile-local start of each digit
  __shared__ int64_t run[NB];      // global position of the next key of each digit from this block
  __shared__ uint32_t wsh[NW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  {
    // digit starts: exclusive scan of the digit totals (every block), plus
    // the keys of this digit in the blocks before this one
    constexpr int DPT = NB / kBlock;
    uint32_t t[DPT], sum = 0;
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
      t[j] = static_cast<uint32_t>(dtot[tid * DPT + j]);
      sum += t[j];
    }
    uint32_t all;
    uint32_t st = block_exclusive_u32(sum, wsh, all);
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
      const int d = tid * DPT + j;
      run[d] = static_cast<int64_t>(st) + off[static_cast<int64_t>(d) * gridDim.x + blockIdx.x];
      st += t[j];
    }
  }
  const int64_t lo = static_cast<int64_t>(blockIdx.x) * tpb * T;
  const int64_t hi = lo + tpb * T < n ? lo + tpb * T : n;
  for (int64_t base = lo; base < hi; base += T) {
    for (int k = tid; k < NW * NB; k += kBlock) (&wc[0][0])[k] = 0;
    __syncthreads();
    // wavefront w ranks keys [base + w*64*R, base + (w+1)*64*R) round by round
    const int64_t wbase = base + static_cast<int64_t>(wave) * 64 * R;
    uint64_t key[R][KW];
    int64_t pay[R][S];
    uint32_t dig[R], rnk[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t i = wbase + r * 64 + lane;
      const bool v = i < hi;
#pragma unroll
      for (int w = 0; w < KW; ++w) key[r][w] = v ? kin[w * stride + i] : 0;
#pragma unroll
      for (int q = 0; q < S; ++q) pay[r][q] = v ? pin[q * stride + i] : 0;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const bool valid = wbase + r * 64 + lane < hi;
      const uint32_t d = key_digit<KW>(key[r], shift, MASK);
      dig[r] = d;
      const unsigned long long peers = match_digit<DB>(d, valid);
      const uint32_t below = static_cast<uint32_t>(__popcll(peers & ((1ull << lane) - 1)));
      const uint32_t old = valid ? wc[wave][d] : 0;
      __builtin_amdgcn_wave_barrier();
      if (valid && below == 0) wc[wave][d] = old + static_cast<uint32_t>(__popcll(peers));
      __builtin_amdgcn_wave_barrier();
      rnk[r] = old + below;
    }
    __syncthreads();
    // tile digit counts (wc becomes the waves' exclusive prefixes per digit)
    constexpr int DPT = NB / kBlock;
    uint32_t cnt[DPT], tsum = 0;
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
      const int d = tid * DPT + j;
      uint32_t c = 0;
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        const uint32_t x = wc[w][d];
        wc[w][d] = c;
        c += x;
      }
      cnt[j] = c;
      tsum += c;
    }
    uint32_t ttot;
    uint32_t to = block_exclusive_u32(tsum, wsh, ttot);
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
      toff[tid * DPT + j] = to;
      to += cnt[j];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (wbase + r * 64 + lane < hi) {
        const uint32_t d = dig[r];
        const uint32_t pos = toff[d] + wc[wave][d] + rnk[r];
#pragma unroll
        for (int w = 0; w < KW; ++w) sk[w][pos] = key[r][w];
#pragma unroll
        for (int q = 0; q < S; ++q) sp[q][pos] = pay[r][q];
      }
    }
    __syncthreads();
    const int tn = static_cast<int>(hi - base < T ? hi - base : T);
    for (int p = tid; p < tn; p += kBlock) {
      uint64_t k[KW];
#pragma unroll
      for (int w = 0; w < KW; ++w) k[w] = sk[w][p];
      const uint32_t d = key_digit<KW>(k, shift, MASK);
      const int64_t g = run[d] + (p - static_cast<int>(toff[d]));
#pragma unroll
      for (int w = 0; w < KW; ++w) kout[w * stride + g] = k[w];
#pragma unroll
      for (int q = 0; q < S; ++q) pout[q * stride + g] = sp[q][p];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < DPT; ++j) run[tid * DPT + j] += cnt[j];
    // (the next tile's first barrier orders these updates before their use)
  }
}

// Sort keys and payload from the spawned list (determinant words, records of
// populations and flag): entry i gets those of entry perm[i] (perm may be
// null: the identity).
__global__ void k_make_keys(KeyPlan kp, int S, const uint64_t* __restrict__ sdet, const int64_t* __restrict__ srec,
                            int W, int64_t ns, int64_t scap, const uint32_t* __restrict__ perm,
                            uint64_t* __restrict__ keys, int64_t* __restrict__ pay) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const int64_t src = perm ? static_cast<int64_t>(perm[i]) : i;
  uint64_t f[FQ_MAX_WORDS] = {0, 0};
  FQ_FOR_WORDS(w, W) f[w] = sdet[w * scap + src];
  const int64_t* rec = srec + src * rec_stride(S);
  uint64_t k[2];
  make_key(kp, f, W, static_cast<uint64_t>(rec[S]), k);
  for (int w = 0; w < kp.kw; ++w) keys[w * scap + i] = k[w];
  for (int q = 0; q < S; ++q) pay[q * scap + i] = rec[q];
}

// Replay: the reference's quicksort (src/sort.f90:213-396) on an index array,
// one thread; comparisons read the spawned bit strings.
__global__ void k_replay_qsort(const uint64_t* __restrict__ sdet, int64_t ns, int W, int64_t scap,
                               uint32_t* __restrict__ idx) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int64_t i = 0; i < ns; ++i) idx[i] = static_cast<uint32_t>(i);
  auto gt = [&](uint32_t a, uint32_t b) { return key_cmp(sdet, scap, W, a, b) > 0; };
  auto ge = [&](uint32_t a, uint32_t b) { return key_cmp(sdet, scap, W, a, b) >= 0; };
  // 1-based view of idx
  auto L = [&](int64_t k) -> uint32_t& { return idx[k - 1]; };
  int64_t stack_lo[64], stack_hi[64];
  int nstack = 0;
  int64_t lo = 1, hi = ns;
  while (true) {
    if (hi - lo < 7) {
      for (int64_t j = lo + 1; j <= hi; ++j) {
        const uint32_t tmp = L(j);
        int64_t i;
        for (i = j - 1; i >= 1; --i) {
          if (ge(tmp, L(i))) break;
          L(i + 1) = L(i);
        }
        L(i + 1) = tmp;
      }
      if (nstack == 0) break;
      hi = stack_hi[nstack - 1];
      lo = stack_lo[nstack - 1];
      --nstack;
    } else {
      const int64_t pivot = (lo + hi) / 2;
      uint32_t t = L(pivot);
      L(pivot) = L(lo + 1);
      L(lo + 1) = t;
      if (gt(L(lo), L(hi))) {
        t = L(lo);
        L(lo) = L(hi);
        L(hi) = t;
      }
      if (gt(L(lo + 1), L(hi))) {
        t = L(lo + 1);
        L(lo + 1) = L(hi);
        L(hi) = t;
      }
      if (gt(L(lo), L(lo + 1))) {
        t = L(lo);
        L(lo) = L(lo + 1);
        L(lo + 1) = t;
      }
      int64_t i = lo + 1, j = hi;
      const uint32_t piv = L(lo + 1);
      while (true) {
        ++i;
        while (gt(piv, L(i))) ++i;
        --j;
        while (gt(L(j), piv)) --j;
        if (j < i) break;
        t = L(i);
        L(i) = L(j);
        L(j) = t;
      }
      L(lo + 1) = L(j);
      L(j) = piv;
      ++nstack;
      if (hi - i + 1 >= j - lo) {
        stack_hi[nstack - 1] = hi;
        stack_lo[nstack - 1] = i;
        hi = j - 1;
      } else {
        stack_hi[nstack - 1] = j - 1;
        stack_lo[nstack - 1] = lo;
        lo = i;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Kernels: self-annihilation of the sorted spawned list
// ---------------------------------------------------------------------------
// Equal determinants are adjacent in the sorted keys (equal value fields).
template <int KW>
__device__ __forceinline__ bool same_value(const uint64_t* a, const uint64_t* b, int low_bits) {
  if constexpr (KW == 1) return (a[0] >> low_bits) == (b[0] >> low_bits);
  else return a[1] == b[1] && (a[0] >> low_bits) == (b[0] >> low_bits);
}

// The determinant of a combinatorial rank, in one pass down the orbitals from
// the highest occupied one (a binary search; 7 steps cover 128 orbitals): at
// orbital p with i electrons left, p is occupied iff C(p, i) <= the remaining
// rank, which then drops by C(p, i).  One table read per orbital and no
// nested loop keep a wavefront's lanes together (stepping electron by
// electron lets each lane's gaps serialise the wavefront).  binom: the table
// (in LDS when staged); sat: it holds C(p, i) saturated at 2^64 - 1, one word
// each (ranks below 2^64), instead of (low, high) pairs.
template <bool WIDE>
__device__ __forceinline__ void unrank(const KeyPlan& kp, const uint64_t* __restrict__ binom, bool sat,
                                       unsigned __int128 rank, uint64_t* f) {
  using V = typename std::conditional<WIDE, unsigned __int128, uint64_t>::type;
  const int nel = kp.nel;
  auto C = [&](int p, int i) -> V {
    if constexpr (WIDE) {
      const uint64_t* t = binom + 2 * (p * nel + (i - 1));
      return (static_cast<unsigned __int128>(t[1]) << 64) | t[0];
    } else {
      if (sat) return binom[p * nel + (i - 1)];
      const uint64_t* t = binom + 2 * (p * nel + (i - 1));
      return t[1] ? ~uint64_t{0} : t[0];  // above the 64-bit ranks: larger than any of them
    }
  };
  V r = static_cast<V>(rank);
  int lo = nel - 1, hi = kp.nbasis - 1;
#pragma unroll
  for (int step = 0; step < 7; ++step) {
    const int mid = (lo + hi + 1) >> 1;
    const bool ok = C(mid, nel) <= r;
    const bool live = lo < hi;
    lo = live && ok ? mid : lo;
    hi = live && !ok ? mid - 1 : hi;
  }
  uint64_t f0 = 0, f1 = 0;
  int i = nel;
  for (int p = lo; i > 0 && p >= 0; --p) {
    const V c = C(p, i);
    if (c <= r) {
      if (p < 64)
        f0 |= uint64_t{1} << p;
      else
        f1 |= uint64_t{1} << (p - 64);
      r -= c;
      --i;
    }
  }
  f[0] = f0;
  f[1] = f1;
}

// One sorted entry's share of its run's sums in annihilate_spawn_t_initiator
// (src/spawn_data.F90:937-1101), per space: c[sp] the population; c[S + sp]
// the population if it was spawned from an initiator (flag clear), else 0;
// c[2S + sp] for a flagged (non-initiator) spawn its event count, the sign of
// the population (0 for 0) -- except a run's first entry, which counts the
// sign of its population's low 32 bits read as a default integer.  Every sum
// is additive, so a run's totals may be formed in any grouping.
template <int S>
__device__ __forceinline__ void entry_share(int64_t flags, const int64_t* pop, bool head, int64_t* c) {
#pragma unroll
  for (int sp = 0; sp < S; ++sp) {
    const int64_t v = pop[sp];
    const bool fl = (flags >> sp) & 1;
    c[sp] = v;
    c[S + sp] = fl ? 0 : v;
    int64_t ev = 0;
    if (fl) {
      if (head) {
        const int32_t v32 = static_cast<int32_t>(static_cast<uint32_t>(static_cast<uint64_t>(v) & 0xffffffffULL));
        ev = v32 >= 0 ? 1 : -1;
      } else {
        ev = v < 0 ? -1 : (v > 0 ? 1 : 0);
      }
    }
    c[2 * S + sp] = ev;
  }
}

// Per tile of kRunTile sorted keys: the number of run heads (first key, or
// value differing from the previous key), the position of the first head (-1
// if none), and the sums (entry_share) of the entries before it -- the part of
// a run that began in an earlier tile.
template <int KW, int S>
__global__ __launch_bounds__(kBlock) void k_run_count(const uint64_t* __restrict__ keys,
                                                      const int64_t* __restrict__ pay, const int64_t* __restrict__ pn,
                                                      int64_t nsb, int64_t stride, int fb,
                                                      int64_t* __restrict__ tcnt, int64_t* __restrict__ tfirst,
                                                      int64_t* __restrict__ tpre) {
  constexpr int R = kRunTile / kBlock, NV = 3 * S;
  const int64_t ns = list_len(pn, nsb);
  __shared__ int64_t wsh[kBlock / 64];
  __shared__ int32_t wfirst[kBlock / 64];
  const int64_t tile = blockIdx.x, t0 = tile * kRunTile;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t c = 0;
  int first_here = kRunTile;  // this wavefront's first head in the tile
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t i = t0 + r * kBlock + threadIdx.x;
    bool h = false;
    if (i < ns) {
      uint64_t k[KW], prev[KW];
#pragma unroll
      for (int w = 0; w < KW; ++w) {
        k[w] = keys[w * stride + i];
        prev[w] = i > 0 ? keys[w * stride + i - 1] : 0;
      }
      h = i == 0 || !same_value<KW>(k, prev, fb);
    }
    c += h ? 1 : 0;
    const unsigned long long b = __ballot(h);
    if (b && first_here == kRunTile) first_here = r * kBlock + wave * 64 + __builtin_ctzll(b);
  }
  if (lane == 0) wfirst[wave] = first_here;
  const int64_t t = block_sum(c, wsh);  // its barriers also publish wfirst
  int first = kRunTile;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) first = wfirst[w] < first ? wfirst[w] : first;
  const uint64_t fmask = (uint64_t{1} << fb) - 1;
  int64_t acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0;
  if (first > 0) {  // block-uniform
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int j = r * kBlock + static_cast<int>(threadIdx.x);
      const int64_t i = t0 + j;
      if (j < first && i < ns) {
        int64_t pop[S], cv[NV];
#pragma unroll
        for (int q = 0; q < S; ++q) pop[q] = pay[q * stride + i];
        entry_share<S>(static_cast<int64_t>(keys[i] & fmask), pop, false, cv);
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] += cv[k];
      }
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = block_sum(acc[k], wsh);
  }
  if (threadIdx.x == 0) {
    tcnt[tile] = t;
    tfirst[tile] = first < kRunTile ? first : -1;
#pragma unroll
    for (int k = 0; k < NV; ++k) tpre[tile * NV + k] = acc[k];
  }
}

// Runs longer than this are reduced by the whole block (a queue in LDS), so a
// determinant spawned onto thousands of times does not hold one lane while
// its wavefront idles; shorter ones by one thread each.
constexpr int kShortRun = 32;

// annihilate_spawn_t(_initiator) (src/spawn_data.F90:859-1101) over the
// sorted keys and populations, one block per tile of kRunTile keys.  The
// tile's run heads are listed in order (wavefront ballots); thread h takes run
// h, sums its entries' shares (entry_share) over consecutive keys and writes
// the run -- its index is the tile's run offset (toff) + h, its determinant is
// recovered from the key by unranking -- so every lane has a run whatever the
// run lengths.  Runs longer than kShortRun are queued and reduced by the whole
// block afterwards.  The tile's last run adds the head-free prefixes of the
// following tiles (k_run_count) up to the next head.  The run's first entry
// is its first in sorted order: the reference's quicksort order in replay
// mode, the earliest spawned in production (the stable sort of an
// attempt-ordered list).
template <int KW, int S, bool WIDE>
__global__ __launch_bounds__(kBlock) void k_runs(Sys s, KeyPlan kp, const uint64_t* __restrict__ keys,
                                                 const int64_t* __restrict__ pay, int64_t ns, int64_t stride,
                                                 const int64_t* __restrict__ toff, const int64_t* __restrict__ tfirst,
                                                 const int64_t* __restrict__ tpre, int64_t ntiles,
                                                 uint64_t* __restrict__ rdet, int64_t* __restrict__ rpop,
                                                 int64_t* __restrict__ rflag, int64_t* __restrict__ rstate,
                                                 const int64_t* __restrict__ pns) {
  ns = list_len(pns, ns);
  constexpr int R = kRunTile / kBlock, NV = 3 * S, NW = kBlock / 64;
  __shared__ int32_t wcount[R][NW];
  __shared__ int32_t hpos[kRunTile + 1];  // tile-relative index of each head, in order; then the tile length
  __shared__ int32_t longq[kRunTile];     // runs (by h) left to the whole block
  __shared__ int32_t nlong;
  __shared__ int64_t wsh[NW];
  extern __shared__ uint64_t binom_lds[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t tile = blockIdx.x, t0 = tile * kRunTile;
  if (t0 >= ns) return;  // block-uniform
  const bool staged = kp.rank && kp.binom_lds;
  if (staged) {  // one-word ranks: the table saturated at 2^64 - 1, one word per entry
    const int nv = kp.nbasis * kp.nel;
    for (int k = tid; k < nv; k += kBlock) {
      if constexpr (WIDE) {
        binom_lds[2 * k] = kp.binom[2 * k];
        binom_lds[2 * k + 1] = kp.binom[2 * k + 1];
      } else {
        binom_lds[k] = kp.binom[2 * k + 1] ? ~uint64_t{0} : kp.binom[2 * k];
      }
    }
  }
  const int n = static_cast<int>(ns - t0 < kRunTile ? ns - t0 : kRunTile);
  const int fb = kp.flag_bits;
  const uint64_t fmask = (uint64_t{1} << fb) - 1;
  bool head[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = r * kBlock + tid;
    const int64_t i = t0 + j;
    head[r] = false;
    if (j < n) {
      uint64_t k[KW], prev[KW];
#pragma unroll
      for (int w = 0; w < KW; ++w) {
        k[w] = keys[w * stride + i];
        prev[w] = i > 0 ? keys[w * stride + i - 1] : 0;
      }
      head[r] = i == 0 || !same_value<KW>(k, prev, fb);
    }
  }
  unsigned long long bal[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    bal[r] = __ballot(head[r]);
    if (lane == 0) wcount[r][wave] = __popcll(bal[r]);
  }
  if (tid == 0) nlong = 0;
  __syncthreads();
  int32_t total = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    int32_t p = total + __popcll(bal[r] & ((1ull << lane) - 1));
    for (int w = 0; w < NW; ++w) {
      if (w < wave) p += wcount[r][w];
      total += wcount[r][w];
    }
    if (head[r]) hpos[p] = r * kBlock + tid;
  }
  if (tid == 0) hpos[total] = n;
  __syncthreads();  // also orders the table's staging
  const int64_t rbase = toff[tile];
  // the shares of run h's entries [a, b) of the tile, from index `from` (stride `step`)
  auto shares = [&](int a, int b, int from, int step, int64_t* acc) {
    for (int j = from; j < b; j += step) {
      int64_t pop[S], cv[NV];
#pragma unroll
      for (int q = 0; q < S; ++q) pop[q] = pay[q * stride + t0 + j];
      entry_share<S>(static_cast<int64_t>(keys[t0 + j] & fmask), pop, j == a, cv);
#pragma unroll
      for (int k = 0; k < NV; ++k) acc[k] += cv[k];
    }
  };
  // run h's totals complete: the continuation of the tile's last run, the
  // initiator rules, and the run's record
  auto finish = [&](int h, int64_t* acc) {
    if (h == total - 1 && t0 + n < ns && tfirst[tile + 1] != 0) {
      for (int64_t u = tile + 1; u < ntiles; ++u) {
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] += tpre[u * NV + k];
        if (tfirst[u] >= 0) break;
      }
    }
    const int64_t run = rbase + h;
    int64_t fl = 0;
    bool nz = false;
#pragma unroll
    for (int sp = 0; sp < S; ++sp) {
      const int64_t tot = acc[sp], init = acc[S + sp], evs = acc[2 * S + sp];
      nz = nz || tot != 0;
      if (s.initiator) {
        const int st = tot >= 0 ? 1 : -1, si = init >= 0 ? 1 : -1;
        if (init != 0 && st == si) {
        } else if (evs > 1 || evs < -1) {
        } else {
          fl += int64_t{1} << sp;
        }
      }
      rpop[sp * stride + run] = tot;
    }
    rflag[run] = fl;
    rstate[run] = nz ? 1 : 0;
    const int64_t i = t0 + hpos[h];
    uint64_t kk[KW];
#pragma unroll
    for (int w = 0; w < KW; ++w) kk[w] = keys[w * stride + i];
    uint64_t d[2] = {0, 0};
    const unsigned __int128 val =
        (KW == 1 ? static_cast<unsigned __int128>(kk[0])
                 : (static_cast<unsigned __int128>(kk[KW - 1]) << 64) | kk[0]) >> fb;
    if (!kp.rank) {
      d[0] = static_cast<uint64_t>(val);  // batched instances: the encoded word itself
    } else if (staged) {
      unrank<WIDE>(kp, binom_lds, !WIDE, val, d);
    } else {
      unrank<WIDE>(kp, kp.binom, false, val, d);
    }
    FQ_FOR_WORDS(w, s.W) rdet[w * stride + run] = d[w];
  };
  for (int h = tid; h < total; h += kBlock) {
    const int a = hpos[h], b = hpos[h + 1];
    if (b - a > kShortRun) {
      longq[atomicAdd(&nlong, 1)] = h;
      continue;
    }
    int64_t acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0;
    shares(a, b, a, 1, acc);
    finish(h, acc);
  }
  __syncthreads();
  const int nq = nlong;
  for (int q = 0; q < nq; ++q) {  // block-uniform
    const int h = longq[q];
    const int a = hpos[h], b = hpos[h + 1];
    int64_t acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0;
    shares(a, b, a + tid, kBlock, acc);
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = block_sum(acc[k], wsh);
    if (tid == 0) finish(h, acc);
  }
}

// ---------------------------------------------------------------------------
// Kernels: semi-stochastic projection
// ---------------------------------------------------------------------------
// Element (i, j) of the core Hamiltonian as create_determ_hamil stores it
// (src/semi_stoch.F90:533-685): <D_i|H|D_j>, H_ii - H_00 on the diagonal,
// times the quasi-Newton weight (1 here); the transpose (<D_j|H|D_i>) for
// separate annihilation.
__device__ __forceinline__ double determ_elem(const Sys& s, const uint64_t* fi, int64_t i, int64_t j, int transp) {
  uint64_t fj[FQ_MAX_WORDS];
#pragma unroll
  for (int w = 0; w < FQ_MAX_WORDS; ++w) fj[w] = w < s.W ? s.determ[w * s.ndeterm + j] : 0;
  double h = transp ? fqd::hmatel(s, fj, fi) : fqd::hmatel(s, fi, fj);
  if (i == j) h = h - s.H00;
  return 1.0 * h;
}
constexpr double kDetermEps = 1.0e-12;  // depsilon: smaller elements are not stored

// Non-zero elements per core row: one block per row, threads over columns.
__global__ __launch_bounds__(kBlock) void k_determ_count(Sys s, int transp, int64_t* __restrict__ cnt) {
  __shared__ int64_t wsh[kBlock / 64];
  const int64_t i = blockIdx.x;
  uint64_t fi[FQ_MAX_WORDS];
#pragma unroll
  for (int w = 0; w < FQ_MAX_WORDS; ++w) fi[w] = w < s.W ? s.determ[w * s.ndeterm + i] : 0;
  int64_t c = 0;
  for (int64_t j = threadIdx.x; j < s.ndeterm; j += kBlock) c += fabs(determ_elem(s, fi, i, j, transp)) > kDetermEps;
  c = block_sum(c, wsh);
  if (threadIdx.x == 0) cnt[i] = c;
}

// The row's elements in column order from row[i] (block-wide ranks per chunk
// of kBlock columns).
__global__ __launch_bounds__(kBlock) void k_determ_fill(Sys s, int transp, const int64_t* __restrict__ row,
                                                        int32_t* __restrict__ col, double* __restrict__ mat) {
  __shared__ uint32_t wsh[kBlock / 64];
  const int64_t i = blockIdx.x;
  uint64_t fi[FQ_MAX_WORDS];
#pragma unroll
  for (int w = 0; w < FQ_MAX_WORDS; ++w) fi[w] = w < s.W ? s.determ[w * s.ndeterm + i] : 0;
  int64_t base = row[i];
  for (int64_t j0 = 0; j0 < s.ndeterm; j0 += kBlock) {  // block-uniform trip count
    const int64_t j = j0 + threadIdx.x;
    double h = 0.0;
    bool keep = false;
    if (j < s.ndeterm) {
      h = determ_elem(s, fi, i, j, transp);
      keep = fabs(h) > kDetermEps;
    }
    uint32_t tot;
    const uint32_t off = block_exclusive_u32(keep ? 1u : 0u, wsh, tot);
    if (keep) {
      col[base + off] = static_cast<int32_t>(j);
      mat[base + off] = h;
    }
    base += tot;
  }
}

// Production deterministic projection, one thread per core determinant and
// space, each with its own draw keyed (row, space, cycle).  Combined
// annihilation writes the encoded amplitude to the spawned list after the
// stochastic spawns (key and payload, as an initiator's spawn; the count is
// added by k_add_nspawn); separate annihilation adds it to the core
// determinant's post-death population.
__global__ __launch_bounds__(kBlock) void k_determ_proj(Sys s, KeyPlan kp, DetermDev dd, double tau, double sh0,
                                                        double sh1, double e0, double e1, uint64_t seed, uint32_t cyc,
                                                        int64_t cap, int64_t* __restrict__ pop_new,
                                                        uint64_t* __restrict__ keys, int64_t* __restrict__ pay,
                                                        int64_t scap, Stats* __restrict__ st) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= dd.nd * s.S) return;
  const int64_t row = t / s.S;
  const int sp = static_cast<int>(t - row * s.S);
  const double shift = sp == 0 ? sh0 : sh1, eold = sp == 0 ? e0 : e1;
  fqd::Philox rng(seed, static_cast<uint32_t>(row), static_cast<uint32_t>(sp) << 28, cyc ^ (fqd::kDrawDeterm << 24));
  if (dd.separate) {
    pop_new[sp * cap + dd.pos[row]] += fqd::stoch_encode(rng, determ_separate(dd, s.S, row, sp, tau, shift, eold), s.rf);
    return;
  }
  const int64_t k = fqd::stoch_encode(rng, determ_combined(dd, s.S, row, sp, tau, shift, eold), s.rf);
  const int64_t slot = st->nspawn + t;
  if (slot >= scap) {
    st->spawn_overflow = 1;
    return;
  }
  for (int q = 0; q < s.S; ++q) pay[q * scap + slot] = q == sp ? k : 0;
  uint64_t f[FQ_MAX_WORDS];
#pragma unroll
  for (int w = 0; w < FQ_MAX_WORDS; ++w) f[w] = w < s.W ? s.determ[w * dd.nd + row] : 0;
  uint64_t key[2];
  make_key(kp, f, s.W, 0, key);
  for (int w = 0; w < kp.kw; ++w) keys[w * scap + slot] = key[w];
}
__global__ void k_add_nspawn(Stats* __restrict__ st, int64_t m) { st->nspawn += m; }

// ---------------------------------------------------------------------------
// Kernels: merge with the main list, rounding, insertion
// ---------------------------------------------------------------------------
// Per surviving run: binary search of its key in the (old) main list; a hit
// adds into the post-death populations (annihilate_main_list(_initiator)), a
// miss becomes a new-determinant candidate with flagged spaces cleared.
// Lower bound of key k in the sorted main list positions [lo, hi).
__device__ __forceinline__ int64_t main_lower_bound(const uint64_t* __restrict__ det, int64_t cap, int W,
                                                    const uint64_t* k, int64_t lo, int64_t hi) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    int c = 0;
#pragma unroll
    for (int w = FQ_MAX_WORDS - 1; w >= 0; --w)
      if (w < W && c == 0) {
        const uint64_t x = det[w * cap + mid];
        c = (x < k[w]) ? -1 : (x > k[w] ? 1 : 0);
      }
    if (c < 0)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Lower bounds in the main list of the first and last key of each merge
// block's runs (one thread per block of kBlock runs).
__global__ void k_merge_bounds(const int64_t* __restrict__ pnr, int64_t n, int64_t cap, int W,
                               const uint64_t* __restrict__ det, const uint64_t* __restrict__ rdet, int64_t scap,
                               int64_t* __restrict__ mwin, const int64_t* __restrict__ pn) {
  const int64_t nruns = *pnr;
  n = list_len(pn, n);
  const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t r0 = b * kBlock;
  if (r0 >= nruns) return;
  const int64_t r1 = (r0 + kBlock < nruns ? r0 + kBlock : nruns) - 1;
  uint64_t k[FQ_MAX_WORDS] = {0, 0};
  FQ_FOR_WORDS(w, W) k[w] = rdet[w * scap + r0];
  const int64_t lo = main_lower_bound(det, cap, W, k, 0, n);
  FQ_FOR_WORDS(w, W) k[w] = rdet[w * scap + r1];
  mwin[2 * b] = lo;
  mwin[2 * b + 1] = main_lower_bound(det, cap, W, k, lo, n);
}

// Runs are sorted, so a block's runs land between the lower bounds of its
// first and last keys: two searches of the whole list per block, then each
// thread searches only that window (merge-path partition of the two lists).
__global__ __launch_bounds__(kBlock) void k_merge_runs(Sys s, const int64_t* __restrict__ pnr, int64_t n, int64_t cap,
                                                       const uint64_t* __restrict__ det,
                                                       int64_t* __restrict__ pop_new, const uint64_t* __restrict__ rdet,
                                                       int64_t* __restrict__ rpop, const int64_t* __restrict__ rflag,
                                                       int64_t scap, const int64_t* __restrict__ mwin,
                                                       int64_t* __restrict__ rpos, int64_t* __restrict__ rstate,
                                                       const int64_t* __restrict__ pn) {
  const int64_t nruns = *pnr;
  n = list_len(pn, n);
  __shared__ uint64_t wdet[FQ_MAX_WORDS][kMergeWin];  // the window's determinants when it fits
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * blockDim.x;
  if (r0 >= nruns) return;  // whole block past the runs (the grid covers a bound); no window
  // every key of the block lies in [key(r0), key(r1)], so its lower bound
  // lies in [wlo, whi): whi = min(bound(r1) + 1, n)
  const int64_t b0 = mwin[2 * blockIdx.x], b1 = mwin[2 * blockIdx.x + 1];
  const int64_t wlo = b0, whi = b1 < n ? b1 + 1 : n;
  const bool staged = whi - wlo <= kMergeWin;
  if (staged)
    for (int64_t q = threadIdx.x; q < whi - wlo; q += kBlock) FQ_FOR_WORDS(w, s.W) wdet[w][q] = det[w * cap + wlo + q];
  __syncthreads();
  const int64_t r = r0 + threadIdx.x;
  if (r >= nruns) return;
  if (rstate[r] == 0) {
    rpos[r] = 0;
    return;
  }
  uint64_t k[FQ_MAX_WORDS] = {0, 0};
  FQ_FOR_WORDS(w, s.W) k[w] = rdet[w * scap + r];
  int64_t lo;
  bool hit = false;
  if (staged) {
    int a = 0, b = static_cast<int>(whi - wlo);
    while (a < b) {
      const int mid = (a + b) >> 1;
      int c = 0;
#pragma unroll
      for (int w = FQ_MAX_WORDS - 1; w >= 0; --w)
        if (w < s.W && c == 0) {
          const uint64_t x = wdet[w][mid];
          c = (x < k[w]) ? -1 : (x > k[w] ? 1 : 0);
        }
      if (c < 0)
        a = mid + 1;
      else
        b = mid;
    }
    lo = wlo + a;
    if (a < whi - wlo) {
      hit = true;
      FQ_FOR_WORDS(w, s.W) hit = hit && wdet[w][a] == k[w];
    }
  } else {
    lo = main_lower_bound(det, cap, s.W, k, wlo, whi);
    if (lo < n) {
      hit = true;
      FQ_FOR_WORDS(w, s.W) hit = hit && det[w * cap + lo] == k[w];
    }
  }
  rpos[r] = lo;
  const int64_t fl = rflag[r];
  if (hit) {
    for (int sp = 0; sp < s.S; ++sp) {
      int64_t& m = pop_new[sp * cap + lo];
      const int64_t v = rpop[sp * scap + r];
      if (!s.initiator)
        m += v;
      else if (m != 0)
        m += v;
      else if (!((fl >> sp) & 1))
        m = v;
    }
    rstate[r] = 3;  // annihilated into the main list
  } else {
    if (s.initiator) {
      bool discard = true;
      for (int sp = 0; sp < s.S; ++sp) {
        if ((fl >> sp) & 1)
          rpop[sp * scap + r] = 0;
        else
          discard = false;
      }
      if (discard) {
        rstate[r] = 0;
        return;
      }
    }
    rstate[r] = 2;  // new determinant candidate
  }
}

// Number of rounding draws an entry needs: spaces with 0 < |pop| < rf.
__device__ __forceinline__ int round_draws(const int64_t* p, int64_t stride, int64_t i, int S, int64_t rf) {
  int c = 0;
  for (int sp = 0; sp < S; ++sp) {
    const int64_t v = p[sp * stride + i];
    const int64_t av = v < 0 ? -v : v;
    if (av < rf && v != 0) ++c;
  }
  return c;
}

// A main-list entry of the semi-stochastic core (kept whatever its
// population, never rounded: remove_unoccupied_dets, src/annihilation.f90:537-600).
__device__ __forceinline__ bool is_core_entry(const Sys& s, const uint64_t* __restrict__ det, int64_t cap,
                                              int64_t i) {
  if (s.ndeterm == 0) return false;
  uint64_t f[FQ_MAX_WORDS];
#pragma unroll
  for (int w = 0; w < FQ_MAX_WORDS; ++w) f[w] = w < s.W ? det[w * cap + i] : 0;
  return fqd::core_index(s, f) >= 0;
}

__global__ void k_count_round_main(Sys s, int64_t n, int64_t cap, const int64_t* __restrict__ pop_new,
                                   const uint64_t* __restrict__ det, int64_t* __restrict__ mdraw) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  mdraw[i] = (s.real_amps && !is_core_entry(s, det, cap, i)) ? round_draws(pop_new, cap, i, s.S, s.rf) : 0;
}
__global__ void k_count_round_new(Sys s, int64_t nruns, const int64_t* __restrict__ rpop,
                                  const int64_t* __restrict__ rstate, int64_t scap, int64_t* __restrict__ ndraw) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= nruns) return;
  ndraw[r] = (s.real_amps && rstate[r] == 2) ? round_draws(rpop, scap, r, s.S, s.rf) : 0;
}

// Survivor flags are kept as bits, one ballot word per 64 consecutive
// entries, with each word's count; the rank of entry p among survivors is
// off[p >> 6] + the word's bits below p (off: exclusive prefix of the counts).
__device__ __forceinline__ bool bit_get(const uint64_t* __restrict__ mask, int64_t p) {
  return (mask[p >> 6] >> (p & 63)) & 1;
}
__device__ __forceinline__ int64_t bit_rank(const int64_t* __restrict__ off, const uint64_t* __restrict__ mask,
                                            int64_t p) {
  const int b = static_cast<int>(p & 63);
  return off[p >> 6] + (b ? __popcll(mask[p >> 6] & ((1ull << b) - 1)) : 0);
}
// All lanes of a wavefront call this; its 64 entries start at a multiple of 64.
__device__ __forceinline__ void store_ballot(bool alive, int64_t i, int64_t n, uint64_t* __restrict__ mask,
                                             int64_t* __restrict__ cnt) {
  const unsigned long long bal = __ballot(alive);
  if ((threadIdx.x & 63) == 0 && i < n) {
    mask[i >> 6] = bal;
    cnt[i >> 6] = __popcll(bal);
  }
}

// remove_unoccupied_dets: round small real populations, mark survivors.
__global__ __launch_bounds__(kBlock) void k_round_main(Sys s, int64_t n, int64_t cap, int64_t* __restrict__ pop_new,
                                                       int replay, const double* __restrict__ rbuf, int64_t rbase,
                                                       const int64_t* __restrict__ mdraw_off, uint64_t seed,
                                                       uint32_t cyc, uint64_t* __restrict__ mmask,
                                                       int64_t* __restrict__ mcnt, const uint64_t* __restrict__ det,
                                                       const int64_t* __restrict__ pn, const uint32_t* __restrict__ pcyc) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t nl = list_len(pn, n);  // entries past it (up to the bound n) are marked dead
  cyc = cycle_of(pcyc, cyc);
  bool alive = false;
  if (i < nl && is_core_entry(s, det, cap, i)) {
    alive = true;
  } else if (i < nl) {
    if (s.real_amps) {
      if (replay) {
        fqd::ReplayRng rng{rbuf, rbase + mdraw_off[i], rbase + mdraw_off[i + 1], false};
        for (int sp = 0; sp < s.S; ++sp)
          pop_new[sp * cap + i] = fqd::stoch_round_pop(rng, pop_new[sp * cap + i], s.rf);
      } else {
        fqd::Philox rng(seed, static_cast<uint32_t>(i), static_cast<uint32_t>(i >> 32),
                        cyc ^ (fqd::kDrawRoundMain << 24));
        for (int sp = 0; sp < s.S; ++sp)
          pop_new[sp * cap + i] = fqd::stoch_round_pop(rng, pop_new[sp * cap + i], s.rf);
      }
    }
    for (int sp = 0; sp < s.S; ++sp) alive = alive || pop_new[sp * cap + i] != 0;
  }
  store_ballot(alive, i, n, mmask, mcnt);
}

// round_low_population_spawns for new-determinant candidates.
__global__ __launch_bounds__(kBlock) void k_round_new(Sys s, const int64_t* __restrict__ pnr, int64_t rb,
                                                      int64_t* __restrict__ rpop,
                                                      const int64_t* __restrict__ rstate, int64_t scap, int replay,
                                                      const double* __restrict__ rbuf, int64_t rbase,
                                                      const int64_t* __restrict__ ndraw_off, uint64_t seed,
                                                      uint32_t cyc, uint64_t* __restrict__ nmask,
                                                      int64_t* __restrict__ ncnt, const uint32_t* __restrict__ pcyc) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  cyc = cycle_of(pcyc, cyc);
  const int64_t nruns = *pnr;
  bool alive = false;
  if (r < nruns && rstate[r] == 2) {
    if (s.real_amps) {
      if (replay) {
        fqd::ReplayRng rng{rbuf, rbase + ndraw_off[r], rbase + ndraw_off[r + 1], false};
        for (int sp = 0; sp < s.S; ++sp) rpop[sp * scap + r] = fqd::stoch_round_pop(rng, rpop[sp * scap + r], s.rf);
      } else {
        fqd::Philox rng(seed, static_cast<uint32_t>(r), static_cast<uint32_t>(r >> 32),
                        cyc ^ (fqd::kDrawRoundSpawn << 24));
        for (int sp = 0; sp < s.S; ++sp) rpop[sp * scap + r] = fqd::stoch_round_pop(rng, rpop[sp * scap + r], s.rf);
      }
    }
    for (int sp = 0; sp < s.S; ++sp) alive = alive || rpop[sp * scap + r] != 0;
  }
  store_ballot(alive, r, rb, nmask, ncnt);  // words up to the bound rb >= nruns (zero past nruns)
}

__global__ void k_zero(int64_t* __restrict__ p, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Decide the new list length; insertion is skipped when the main list would
// overflow (insert_new_walkers' psip_list%error path).
// pnr: the number of spawned runs (nullptr: none); also recorded in the stats.
__global__ void k_finalize_counts(int64_t n, const int64_t* __restrict__ pnr, const int64_t* __restrict__ moff,
                                  const uint64_t* __restrict__ mmask, const int64_t* __restrict__ noff,
                                  const uint64_t* __restrict__ nmask, int64_t cap, Stats* __restrict__ st,
                                  const int64_t* __restrict__ pn) {
  n = list_len(pn, n);
  const int64_t nruns = pnr ? *pnr : 0;
  st->nruns = nruns;
  const int64_t nmain = bit_rank(moff, mmask, n), nnew = bit_rank(noff, nmask, nruns);
  if (nmain + nnew > cap) {
    st->main_overflow = 1;
    st->new_n = nmain;
  } else {
    st->new_n = nmain + nnew;
  }
}

// Surviving main entries move up by their rank among survivors plus the number
// of surviving new runs inserted at or before them: nrpos lists those runs'
// insertion positions in rank order, non-decreasing because the runs are
// sorted, so the count is an upper-bound search.  Each block takes the counts
// bounding its 256 entries from bstart (k_block_starts) and stages that window
// of nrpos in LDS when it fits (dense insertion), so each thread searches only
// there.
constexpr int kInsWin = 1024;
// bstart[b] = number of surviving new runs inserted before main entry 256 b
// (nrpos < 256 b), b in [0, nblk]: thread k (k <= nnew) fills the blocks whose
// first entry lies in (nrpos[k-1], nrpos[k]].
__global__ void k_block_starts(const int64_t* __restrict__ nrpos, const Stats* __restrict__ st,
                               const int64_t* __restrict__ moff, const uint64_t* __restrict__ mmask, int64_t n,
                               int64_t bound, int64_t nblk, int64_t* __restrict__ bstart,
                               const int64_t* __restrict__ pn) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k > bound) return;
  if (pn) {  // the list's length and blocks on the device
    n = list_len(pn, n);
    nblk = (n + kBlock - 1) / kBlock;
  }
  const int64_t nnew = st->main_overflow ? 0 : st->new_n - bit_rank(moff, mmask, n);
  if (k > nnew) return;
  const int64_t prev = k == 0 ? -1 : nrpos[k - 1];
  const int64_t b_lo = prev < 0 ? 0 : prev / kBlock + 1;
  const int64_t b_hi = k == nnew ? nblk : (nrpos[k] / kBlock < nblk ? nrpos[k] / kBlock : nblk);
  for (int64_t b = b_lo; b <= b_hi; ++b) bstart[b] = k;
}
__device__ __forceinline__ int64_t upper_bound_pos(const int64_t* __restrict__ v, int64_t lo, int64_t hi, int64_t x) {
  while (lo < hi) {  // first k in [lo, hi) with v[k] > x
    const int64_t mid = (lo + hi) >> 1;
    if (v[mid] <= x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}
__global__ __launch_bounds__(kBlock) void k_scatter_main(Sys s, int64_t n, int64_t cap,
                                                         const uint64_t* __restrict__ det,
                                                         const int64_t* __restrict__ pop_new,
                                                         const double* __restrict__ kii,
                                                         const uint64_t* __restrict__ mmask,
                                                         const int64_t* __restrict__ moff,
                                                         const int64_t* __restrict__ nrpos,
                                                         const int64_t* __restrict__ bstart, const Stats* __restrict__ st,
                                                         uint64_t* __restrict__ det_out, int64_t* __restrict__ pop_out,
                                                         double* __restrict__ kii_out, const int64_t* __restrict__ pn) {
  __shared__ int64_t win[kInsWin];
  n = list_len(pn, n);
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * blockDim.x;
  if (i0 >= n) return;  // whole block
  const bool ovf = st->main_overflow != 0;
  // c0 = runs inserted before i0 exactly; c1 >= runs inserted at or before i1
  const int64_t c0 = ovf ? 0 : bstart[blockIdx.x], c1 = ovf ? 0 : bstart[blockIdx.x + 1];
  const bool staged = c1 - c0 <= kInsWin;
  if (staged)
    for (int64_t q = threadIdx.x; q < c1 - c0; q += kBlock) win[q] = nrpos[c0 + q];
  __syncthreads();
  const int64_t i = i0 + threadIdx.x;
  if (i >= n || !bit_get(mmask, i)) return;
  int64_t shift;
  if (staged) {
    int a = 0, b = static_cast<int>(c1 - c0);
    while (a < b) {
      const int mid = (a + b) >> 1;
      if (win[mid] <= i)
        a = mid + 1;
      else
        b = mid;
    }
    shift = c0 + a;
  } else {
    shift = upper_bound_pos(nrpos, c0, c1, i);
  }
  const int64_t o = bit_rank(moff, mmask, i) + shift;
  for (int w = 0; w < s.W; ++w) det_out[w * cap + o] = det[w * cap + i];
  for (int sp = 0; sp < s.S; ++sp) pop_out[sp * cap + o] = pop_new[sp * cap + i];
  kii_out[o] = kii[i];
}

// Surviving new runs in order: list[rank among survivors] = run index.
// The surviving new runs in rank order, with their insertion positions.
__global__ void k_new_list(const int64_t* __restrict__ pnr, const uint64_t* __restrict__ nmask,
                           const int64_t* __restrict__ noff, const int64_t* __restrict__ rpos,
                           int64_t* __restrict__ list, int64_t* __restrict__ nrpos) {
  const int64_t nruns = *pnr;
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r < nruns && bit_get(nmask, r)) {
    const int64_t k = bit_rank(noff, nmask, r);
    list[k] = r;
    nrpos[k] = rpos[r];
  }
}

// Insert the surviving new runs, one thread each (list order = rank order),
// with K_ii from the diagonal rule; a molecular system's dense pair tables are
// staged in LDS when jk_lds (2 * norb^2 doubles) is nonzero.
__global__ __launch_bounds__(kBlock) void k_scatter_new(Sys s, const int64_t* __restrict__ pnr, int64_t cap,
                                                        int64_t scap,
                                                        const uint64_t* __restrict__ rdet,
                                                        const int64_t* __restrict__ rpop,
                                                        const int64_t* __restrict__ rpos,
                                                        const uint64_t* __restrict__ nmask,
                                                        const int64_t* __restrict__ noff,
                                                        const int64_t* __restrict__ list,
                                                        const uint64_t* __restrict__ mmask,
                                                        const int64_t* __restrict__ moff,
                                                        const Stats* __restrict__ st,
                                                        uint64_t* __restrict__ det_out, int64_t* __restrict__ pop_out,
                                                        double* __restrict__ kii_out, const Sys* __restrict__ sys_arr,
                                                        int jk_lds) {
  extern __shared__ double jk[];
  const int64_t nruns = *pnr;
  const int64_t nnew = bit_rank(noff, nmask, nruns);
  const int64_t k0 = static_cast<int64_t>(blockIdx.x) * blockDim.x;
  if (k0 >= nnew || st->main_overflow) return;  // whole block: no surviving run left
  const int nn = s.norb * s.norb;
  if (jk_lds) {
    for (int q = threadIdx.x; q < nn; q += kBlock) {
      jk[q] = s.jmat[q];
      jk[nn + q] = s.kmat[q];
    }
    __syncthreads();
  }
  const int64_t k = k0 + threadIdx.x;
  if (k >= nnew) return;
  const int64_t r = list[k];
  const int64_t o = k + bit_rank(moff, mmask, rpos[r]);
  uint64_t f[FQ_MAX_WORDS] = {0, 0};
  FQ_FOR_WORDS(w, s.W) {
    f[w] = rdet[w * scap + r];
    det_out[w * cap + o] = f[w];
  }
  for (int sp = 0; sp < s.S; ++sp) pop_out[sp * cap + o] = rpop[sp * scap + r];
  if (sys_arr) {  // batched: the instance's diagonal
    const Sys& si = sys_arr[f[0] >> kInstShift];
    f[0] &= kOrbMask;
    kii_out[o] = fqd::diagonal(si, f) - si.H00;
  } else if (jk_lds) {
    kii_out[o] = fqd::diagonal_mol(s, f, jk, jk + nn) - s.H00;
  } else {
    kii_out[o] = fqd::diagonal(s, f) - s.H00;
  }
}

// Sums of |pop| over the new list (its length read from the stats the
// insertion wrote); grid-stride.
__global__ __launch_bounds__(kBlock) void k_abs_pop(const Stats* __restrict__ stn, int64_t cap, int S,
                                                    const int64_t* __restrict__ pop, int64_t* __restrict__ part_i) {
  __shared__ unsigned long long sh[FQ_MAX_SPACES][kBlock];
  const int64_t n = stn->new_n;
  // exact integer sums: any grid and order give the same totals
  for (int sp = 0; sp < FQ_MAX_SPACES; ++sp) {
    unsigned long long v = 0;
    if (sp < S)
      for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
           i += static_cast<int64_t>(gridDim.x) * kBlock) {
        const int64_t p = pop[sp * cap + i];
        v += static_cast<unsigned long long>(p < 0 ? -p : p);
      }
    sh[sp][threadIdx.x] = v;
  }
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k)
      for (int sp = 0; sp < FQ_MAX_SPACES; ++sp) sh[sp][threadIdx.x] += sh[sp][threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int sp = 0; sp < FQ_MAX_SPACES; ++sp) part_i[sp * gridDim.x + blockIdx.x] = static_cast<int64_t>(sh[sp][0]);
}

__global__ __launch_bounds__(kBlock) void k_sum_abs(const int64_t* __restrict__ part_i, int nb, Stats* __restrict__ st) {
  __shared__ unsigned long long sh[FQ_MAX_SPACES][kBlock];
  unsigned long long acc[FQ_MAX_SPACES] = {0, 0};
  for (int b = threadIdx.x; b < nb; b += kBlock)
    for (int sp = 0; sp < FQ_MAX_SPACES; ++sp) acc[sp] += static_cast<unsigned long long>(part_i[sp * nb + b]);
  for (int sp = 0; sp < FQ_MAX_SPACES; ++sp) sh[sp][threadIdx.x] = acc[sp];
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k)
      for (int sp = 0; sp < FQ_MAX_SPACES; ++sp) sh[sp][threadIdx.x] += sh[sp][threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int sp = 0; sp < FQ_MAX_SPACES; ++sp) st->abs_pop[sp] = sh[sp][0];
}

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------
// Make room for `extra` more draws after the live part [rng_pos, rng_end),
// moving the live part to the front of a (possibly larger) buffer.
int reserve_rng(fq_engine* e, int64_t extra) {
  if (e->rng_end + extra <= e->rng_cap) return FQ_OK;
  const int64_t live = e->rng_end - e->rng_pos;
  const int64_t cap = std::max<int64_t>(live + extra, 2 * e->rng_cap);
  double* nb = nullptr;
  if (dalloc(&nb, cap) != FQ_OK) return FQ_ERR_NOMEM;
  if (live > 0)
    FQ_HIP(hipMemcpyAsync(nb, e->rng_buf + e->rng_pos, sizeof(double) * live, hipMemcpyDeviceToDevice, e->stream));
  FQ_HIP(hipStreamSynchronize(e->stream));
  dfree(e->rng_buf);
  e->rng_buf = nb;
  e->rng_cap = cap;
  e->rng_end = live;
  e->rng_pos = 0;
  return FQ_OK;
}

// Profiling event, recorded when the cycle is not being captured into a graph.
void mark(fq_engine* e, hipEvent_t ev) {
  if (!e->capturing) (void)hipEventRecord(ev, e->stream);
}

int sync_stats(fq_engine* e) {
  FQ_HIP(hipMemcpyAsync(&e->h_stats, e->d_stats, sizeof(Stats), hipMemcpyDeviceToHost, e->stream));
  FQ_HIP(hipStreamSynchronize(e->stream));
  return FQ_OK;
}

KeyPlan key_plan(const fq_engine* e) {
  KeyPlan kp{};
  kp.kw = e->sort_kw;
  kp.flag_bits = e->flag_bits;
  kp.rank = e->rank_words ? 1 : 0;
  kp.nel = e->sys.nel;
  kp.nbasis = e->sys.nbasis;
  kp.binom = e->d_binom;
  return kp;
}

// Bytes of the rank table, and whether kernels may stage it in LDS.
size_t binom_bytes(const fq_engine* e) {
  return e->rank_words ? sizeof(uint64_t) * 2 * e->sys.nbasis * e->sys.nel : 0;
}
constexpr size_t kBinomLdsMax = 48 * 1024;

// Digits of the sort: the fewest passes of 8- or 9-bit digits covering the
// key's value field (bits flag_bits .. flag_bits + value_bits - 1).
DigitPlan digit_plan(const fq_engine* e) {
  DigitPlan dp{};
  const int bits = e->value_bits;
  const int passes = e->digit8 ? std::max(1, (bits + 7) / 8) : std::max(1, (bits + 8) / 9);
  const int n9 = e->digit8 ? 0 : std::max(0, std::min(passes, bits - 8 * passes));
  int shift = e->flag_bits;
  for (int p = 0; p < passes; ++p) {
    dp.shift[p] = shift;
    dp.bits[p] = p < n9 ? 9 : 8;
    shift += dp.bits[p];
  }
  dp.np = passes;
  return dp;
}

// Key layout and sort buffers for the engine's value width.
int setup_sort(fq_engine* e) {
  const int vb = e->rank_words ? e->rank_bits : e->key_bits;
  const int fb = e->S;
  if (vb + fb > 128) return FQ_ERR_ARG;
  const int kw = vb + fb > 64 ? 2 : 1;
  if (e->batched && kw != 1) return FQ_ERR_ARG;  // the instance bounds read one key word
  const int64_t scap = e->scap;
  e->value_bits = vb;
  e->flag_bits = fb;
  int rc = FQ_OK;
  // The spawned list as determinant words and records (sdet, srec) is only
  // read before the sort (replay walk, caller-supplied lists, rank exchange);
  // it shares the memory of the sort's second buffers.
  if (kw != e->sort_kw || !e->kbuf[0]) {
    for (int b = 0; b < 2; ++b) {
      dfree(e->kbuf[b]);
      rc |= dalloc(&e->kbuf[b], (b == 0 ? kw : std::max(kw, e->W)) * scap);
    }
  }
  e->sort_kw = kw;
  for (int b = 0; b < 2; ++b)
    if (!e->pbuf[b]) rc |= dalloc(&e->pbuf[b], (b == 0 ? e->S : std::max(e->S, rec_stride(e->S))) * scap);
  e->sdet = e->kbuf[1];
  e->srec = e->pbuf[1];
  if (!e->perm && e->par.rng_mode == FQ_RNG_REPLAY) rc |= dalloc(&e->perm, scap);
  if (!e->pass_cnt) rc |= dalloc(&e->pass_cnt, 512 * kSortBlocks);
  if (!e->pass_off) rc |= dalloc(&e->pass_off, 512 * kSortBlocks + 1);
  if (!e->run_tcnt) rc |= dalloc(&e->run_tcnt, ceil_div(scap, kRunTile) + 2);
  if (!e->run_toff) rc |= dalloc(&e->run_toff, ceil_div(scap, kRunTile) + 2);
  if (!e->run_tfirst) rc |= dalloc(&e->run_tfirst, ceil_div(scap, kRunTile) + 2);
  if (!e->run_tpre) rc |= dalloc(&e->run_tpre, 3 * e->S * (ceil_div(scap, kRunTile) + 2));
  return rc == FQ_OK ? FQ_OK : FQ_ERR_NOMEM;
}

// Sort keys and payload made from the spawned list's determinants and records
// (when the spawning pass did not make them): perm (may be null) gives the order.
void launch_make_keys(fq_engine* e, int64_t ns, const uint32_t* perm) {
  hipLaunchKernelGGL(k_make_keys, dim3(grid_for(ns)), dim3(kBlock), 0, e->stream, key_plan(e), e->S, e->sdet, e->srec,
                     e->W, ns, e->scap, perm, e->kbuf[0], e->pbuf[0]);
}

// One radix pass from buffer `src` to the other.
// ns: the list length, or a bound of it when pns (the device count) is given.
template <int KW, int S, int DB>
int sort_pass(fq_engine* e, int src, int64_t ns, const int64_t* pns, int shift) {
  constexpr int T = kBlock * sort_items(KW, S);
  const int64_t nb = std::min<int64_t>(ceil_div(ns, T), kSortBlocks);
  hipLaunchKernelGGL((k_pass_count<KW, S, DB>), dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, e->stream,
                     e->kbuf[src], pns, ns, e->scap, shift, e->pass_cnt);
  hipLaunchKernelGGL(k_pass_rowscan, dim3(1u << DB), dim3(kSortBlocks), 0, e->stream, static_cast<int>(nb),
                     e->pass_cnt, e->pass_off);
  hipLaunchKernelGGL((k_pass_scatter<KW, S, DB>), dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, e->stream,
                     e->kbuf[src], e->pbuf[src], e->kbuf[1 - src], e->pbuf[1 - src], pns, ns, e->scap, shift,
                     e->pass_cnt, e->pass_off);
  return FQ_OK;
}
template <int KW, int S>
int sort_pass_db(fq_engine* e, int src, int64_t ns, const int64_t* pns, int shift, int db) {
  return db == 9 ? sort_pass<KW, S, 9>(e, src, ns, pns, shift) : sort_pass<KW, S, 8>(e, src, ns, pns, shift);
}

// Sort the spawned list and reduce runs; leaves the device address of the run
// count in p_nruns and a host bound of it in run_bound.  The list holds ns
// entries, or, when pns is given (production, no host wait for the length),
// *pns entries with ns a bound of them.
int sort_and_reduce(fq_engine* e, int64_t ns, const int64_t* pns, bool replay) {
  const bool keys_ready = e->keys_ready;  // made by the spawning pass for this spawned list
  e->keys_ready = false;
  const int64_t scap = e->scap;
  const int kw = e->sort_kw, S = e->S;
  hipStream_t st = e->stream;
  int src = 0;
  e->radix_passes = 0;
  mark(e, e->pev[4]);
  if (ns > 0) {
    if (replay) {
      hipLaunchKernelGGL(k_replay_qsort, dim3(1), dim3(1), 0, st, e->sdet, ns, e->W, scap, e->perm);
      launch_make_keys(e, ns, e->perm);
    } else {
      if (!keys_ready) launch_make_keys(e, ns, nullptr);
      const DigitPlan dp = digit_plan(e);
      for (int p = 0; p < dp.np; ++p) {
        int rc;
        if (kw == 1)
          rc = S == 1 ? sort_pass_db<1, 1>(e, src, ns, pns, dp.shift[p], dp.bits[p])
                      : sort_pass_db<1, 2>(e, src, ns, pns, dp.shift[p], dp.bits[p]);
        else
          rc = S == 1 ? sort_pass_db<2, 1>(e, src, ns, pns, dp.shift[p], dp.bits[p])
                      : sort_pass_db<2, 2>(e, src, ns, pns, dp.shift[p], dp.bits[p]);
        if (rc != FQ_OK) return rc;
        src = 1 - src;
        ++e->radix_passes;
      }
    }
  }
  if (pns && !e->capturing) {  // the passes are queued: now the exact length (copied before them)
    FQ_HIP(hipEventSynchronize(e->ev_len));
    // the attempts cannot exceed the bound the spawning launches covered (every
    // entry makes at most floor(|N|) + 1 of them, |N| for integer populations)
    if (e->h_len[2] > e->spawn_bound) return FQ_ERR_DEVICE;
    ns = std::min<int64_t>(e->h_len[0], e->scap);
    pns = nullptr;
  }
  mark(e, e->pev[5]);
  const int64_t nt = ceil_div(ns, kRunTile);
  int64_t* nruns = e->run_toff + nt;
  if (ns > 0) {
    auto cnt = kw == 1 ? (S == 1 ? k_run_count<1, 1> : k_run_count<1, 2>)
                       : (S == 1 ? k_run_count<2, 1> : k_run_count<2, 2>);
    hipLaunchKernelGGL(cnt, dim3(static_cast<unsigned>(nt)), dim3(kBlock), 0, st, e->kbuf[src], e->pbuf[src],
                       pns, ns, scap, e->flag_bits, e->run_tcnt, e->run_tfirst, e->run_tpre);
    const int rc = scan_exclusive(e, e->run_tcnt, e->run_toff, nt);
    if (rc != FQ_OK) return rc;
    KeyPlan kp = key_plan(e);
    const size_t tab = binom_bytes(e);
    kp.binom_lds = tab > 0 && tab <= kBinomLdsMax;
    // one-word keys: 64-bit rank arithmetic
    auto kern = kw == 1 ? (S == 1 ? k_runs<1, 1, false> : k_runs<1, 2, false>)
                        : (S == 1 ? k_runs<2, 1, true> : k_runs<2, 2, true>);
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(nt)), dim3(kBlock), kp.binom_lds ? tab : 0, st, e->sys, kp,
                       e->kbuf[src], e->pbuf[src], ns, scap, e->run_toff, e->run_tfirst, e->run_tpre, nt, e->rdet,
                       e->rpop, e->rflag, e->rstate, pns);
  }
  // The number of runs stays on the device; kernels size their grids by a
  // host bound: ns in production, the exact count in replay (read there
  // anyway, for the rounding draws).
  int64_t bound = 0;
  e->p_nruns = nullptr;
  if (ns > 0) {
    e->p_nruns = nruns;
    bound = ns;
    if (replay) {
      FQ_HIP(hipMemcpyAsync(&bound, nruns, sizeof(int64_t), hipMemcpyDeviceToHost, st));
      FQ_HIP(hipStreamSynchronize(st));
    }
  }
  mark(e, e->pev[6]);
  e->run_bound = bound;
  return FQ_OK;
}

// Annihilation proper: merge runs, round, compact, insert.  ns spawned
// entries are in the spawned list; pop_new holds the post-death populations.
int annihilate(fq_engine* e, int64_t ns, const int64_t* pns, bool replay, uint32_t cyc, int64_t& rng_pos) {
  int rc = sort_and_reduce(e, ns, pns, replay);
  if (rc != FQ_OK) return rc;
  const int64_t nruns = e->run_bound;  // a bound in production, the count in replay
  const int64_t* pnr = e->p_nruns;
  // a graph-captured cycle: n is a bound of the list's length, read on the device from dn
  const int64_t n = e->n, cap = e->cap, scap = e->scap;
  const int64_t* dn = e->capturing ? e->dn : nullptr;
  const uint32_t* dcyc = e->capturing ? e->dcyc : nullptr;
  const int c = e->cur;
  hipStream_t st = e->stream;
  mark(e, e->pev[7]);
  if (nruns > 0) {
    hipLaunchKernelGGL(k_merge_bounds, dim3(grid_for(grid_for(nruns))), dim3(kBlock), 0, st, pnr, n, cap, e->W,
                       e->det[c], e->rdet, scap, e->mwin, dn);
    hipLaunchKernelGGL(k_merge_runs, dim3(grid_for(nruns)), dim3(kBlock), 0, st, e->sys, pnr, n, cap, e->det[c],
                       e->pop_new, e->rdet, e->rpop, e->rflag, scap, e->mwin, e->rpos, e->rstate, dn);
  }
  // Rounding draws (replay: main list first, then new candidates, in order).
  int64_t main_draws = 0, new_draws = 0;
  if (replay && e->sys.real_amps) {
    hipLaunchKernelGGL(k_count_round_main, dim3(grid_for(n)), dim3(kBlock), 0, st, e->sys, n, cap, e->pop_new,
                       e->det[c], e->mdraw);
    rc = scan_exclusive(e, e->mdraw, e->mdraw_off, n);
    if (rc != FQ_OK) return rc;
    if (nruns > 0) {
      hipLaunchKernelGGL(k_count_round_new, dim3(grid_for(nruns)), dim3(kBlock), 0, st, e->sys, nruns, e->rpop,
                         e->rstate, scap, e->ndraw);
      rc = scan_exclusive(e, e->ndraw, e->ndraw_off, nruns);
      if (rc != FQ_OK) return rc;
    }
    FQ_HIP(hipMemcpyAsync(&main_draws, e->mdraw_off + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    if (nruns > 0) FQ_HIP(hipMemcpyAsync(&new_draws, e->ndraw_off + nruns, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    FQ_HIP(hipStreamSynchronize(st));
    if (rng_pos + main_draws + new_draws > e->rng_end) return FQ_ERR_RNG_EXHAUSTED;
  }
  const uint64_t seed = e->par.philox_seed;
  hipLaunchKernelGGL(k_round_main, dim3(grid_for(n)), dim3(kBlock), 0, st, e->sys, n, cap, e->pop_new,
                     replay ? 1 : 0, e->rng_buf, rng_pos, e->mdraw_off, seed, cyc, e->mmask, e->mcnt, e->det[c], dn,
                     dcyc);
  if (nruns > 0)
    hipLaunchKernelGGL(k_round_new, dim3(grid_for(nruns)), dim3(kBlock), 0, st, e->sys, pnr, nruns, e->rpop, e->rstate,
                       scap, replay ? 1 : 0, e->rng_buf, rng_pos + main_draws, e->ndraw_off, seed, cyc, e->nmask,
                       e->ncnt, dcyc);
  rng_pos += main_draws + new_draws;
  rc = scan_exclusive(e, e->mcnt, e->moff, ceil_div(n, 64));
  if (rc != FQ_OK) return rc;
  if (nruns > 0) {
    rc = scan_exclusive(e, e->ncnt, e->noff, ceil_div(nruns, 64));
    if (rc != FQ_OK) return rc;
  } else {
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(kBlock), 0, st, e->noff, 1);
  }
  if (nruns > 0)
    hipLaunchKernelGGL(k_new_list, dim3(grid_for(nruns)), dim3(kBlock), 0, st, pnr, e->nmask, e->noff, e->rpos,
                       e->nlist, e->nrpos);
  hipLaunchKernelGGL(k_finalize_counts, dim3(1), dim3(1), 0, st, n, nruns > 0 ? pnr : nullptr, e->moff, e->mmask,
                     e->noff, e->nmask, cap, e->d_stats, dn);
  const int o = 1 - c;
  const int64_t nblk = ceil_div(n, kBlock);
  hipLaunchKernelGGL(k_block_starts, dim3(grid_for(nruns + 1)), dim3(kBlock), 0, st, e->nrpos, e->d_stats, e->moff,
                     e->mmask, n, nruns, nblk, e->bstart, dn);
  hipLaunchKernelGGL(k_scatter_main, dim3(grid_for(n)), dim3(kBlock), 0, st, e->sys, n, cap, e->det[c], e->pop_new,
                     e->kii[c], e->mmask, e->moff, e->nrpos, e->bstart, e->d_stats, e->det[o], e->pop[o], e->kii[o],
                     dn);
  if (nruns > 0) {
    // dense pair tables in LDS for the diagonal when they fit in 32 KB
    const size_t jk = e->sys.system == FQ_SYS_READ_IN && !e->batched
                          ? 2 * sizeof(double) * static_cast<size_t>(e->sys.norb) * e->sys.norb
                          : 0;
    const bool use_lds = jk > 0 && jk <= 32768;
    hipLaunchKernelGGL(k_scatter_new, dim3(grid_for(nruns)), dim3(kBlock), use_lds ? jk : 0, st, e->sys, pnr, cap,
                       scap, e->rdet, e->rpop, e->rpos, e->nmask, e->noff, e->nlist, e->mmask, e->moff, e->d_stats,
                       e->det[o], e->pop[o], e->kii[o], e->batched ? e->d_sys_arr : nullptr, use_lds ? 1 : 0);
  }
  mark(e, e->pev[8]);
  // the new list holds at most n + nruns entries
  const unsigned nba = std::min<unsigned>(grid_for(std::min(n + nruns, cap)), 2048);
  hipLaunchKernelGGL(k_abs_pop, dim3(nba), dim3(kBlock), 0, st, e->d_stats, cap, e->S, e->pop[o], e->part_i);
  hipLaunchKernelGGL(k_sum_abs, dim3(1), dim3(kBlock), 0, st, e->part_i, static_cast<int>(nba), e->d_stats);
  if (e->capturing) return FQ_OK;  // the caller reads the stats when the graph has run
  rc = sync_stats(e);
  if (rc != FQ_OK) return rc;
  e->cur = o;
  e->n = e->h_stats.new_n;
  e->att_bound = -1;
  return FQ_OK;
}

void fill_out(fq_engine* e, fq_cycle_out* out, int64_t rng_used) {
  std::memset(out, 0, sizeof(*out));
  const Stats& s = e->h_stats;
  const double rf = static_cast<double>(e->sys.rf);
  for (int sp = 0; sp < e->S; ++sp) {
    out->proj_energy[sp] = s.proj[sp];
    out->D0_population[sp] = s.d0[sp];
    out->nparticles[sp] = static_cast<double>(s.abs_pop[sp]) / rf;
  }
  out->nstates = e->n;
  out->nspawn_events = std::min<int64_t>(s.nspawn, e->scap);
  out->ndeath = s.ndeath;
  out->nattempts = s.total_attempts;
  out->rng_consumed = rng_used;
  out->spawn_overflow = s.spawn_overflow ? 1 : 0;
  out->main_overflow = s.main_overflow ? 1 : 0;
}

// ---- heat-bath tables: init_excit_mol_heat_bath (src/excit_gen_heat_bath_mol.F90:14-248)
// The reference's loops (i, j, a, b in that nesting) run in serial order per
// table entry: every sum below adds its terms in the reference's order.

// One thread per (i, j, a): the b loop; weights(b,a,j,i), weights_tot(a,j,i)
// and hb_ija%weights(a,j,i) (the same sum: invalid b contribute +0.0).
__global__ void k_hb_ijab(Sys s, double* __restrict__ ijab_w, double* __restrict__ ijab_tot,
                          double* __restrict__ ija_w) {
  const int nb = s.nbasis;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<int64_t>(nb) * nb * nb) return;
  const int i = static_cast<int>(t / (static_cast<int64_t>(nb) * nb)) + 1;
  const int j = static_cast<int>((t / nb) % nb) + 1;
  const int a = static_cast<int>(t % nb) + 1;
  double* w = ijab_w + t * nb;
  double tot = 0.0;
  if (i != j && a != i && a != j) {
    const int it = i < j ? i : j, jt = i < j ? j : i;
    const int isymb = s.orb_sym[it - 1] ^ s.orb_sym[jt - 1] ^ s.orb_sym[a - 1];
    for (int b = 1; b <= nb; ++b) {
      double wb = 0.0;
      const bool spin_ok = (fqd::ms_of(it) == fqd::ms_of(a) && fqd::ms_of(jt) == fqd::ms_of(b)) ||
                           (fqd::ms_of(it) == fqd::ms_of(b) && fqd::ms_of(jt) == fqd::ms_of(a));
      if (spin_ok && s.orb_sym[b - 1] == isymb && b != a && b != i && b != j)
        wb = wb + fabs(fqd::sc2_excit(s, it, jt, a < b ? a : b, a < b ? b : a, false));
      w[b - 1] = wb;
      tot = tot + wb;
    }
  } else {
    for (int b = 0; b < nb; ++b) w[b] = 0.0;
  }
  ijab_tot[t] = tot;
  ija_w[t] = tot;
}

// One thread per (i, j): ij_weights(j,i) as the flat (a, b) sum and
// hb_ija%weights_tot(j,i) as the sum over a of the per-a totals.
__global__ void k_hb_ij(Sys s, const double* __restrict__ ijab_w, const double* __restrict__ ija_w,
                        double* __restrict__ ij_w, double* __restrict__ ija_tot) {
  const int nb = s.nbasis;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<int64_t>(nb) * nb) return;
  double flat = 0.0, tot = 0.0;
  if (t / nb != t % nb) {
    const double* w = ijab_w + t * nb * nb;
    for (int64_t ab = 0; ab < static_cast<int64_t>(nb) * nb; ++ab) flat = flat + w[ab];
    for (int a = 0; a < nb; ++a) tot = tot + ija_w[t * nb + a];
  }
  ij_w[t] = flat;
  ija_tot[t] = tot;
}

// One thread per i: i_weights(i) = sum over j of ij_weights(j,i), and the
// reference's bias test (src/excit_gen_heat_bath_mol.F90:185-202) per i.
__global__ void k_hb_i(Sys s, const double* __restrict__ ij_w, const double* __restrict__ ija_w,
                       double* __restrict__ i_w, int* __restrict__ bias) {
  const int nb = s.nbasis;
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  if (i > nb) return;
  double iw = 0.0;
  for (int j = 1; j <= nb; ++j) iw = iw + ij_w[static_cast<int64_t>(i - 1) * nb + (j - 1)];
  i_w[i - 1] = iw;
  int biased = 0;
  for (int a = 1; a <= nb; ++a) {
    int nz = 0;
    for (int j = 1; j <= nb; ++j)
      if (ija_w[(static_cast<int64_t>(i - 1) * nb + (j - 1)) * nb + (a - 1)] > fqd::kDepsilon) ++nz;
    if (nz < nb - s.nel && i != a && s.orb_sym[a - 1] == s.orb_sym[i - 1] && fqd::ms_of(a) == fqd::ms_of(i))
      biased = 1;
  }
  bias[i - 1] = biased;
}

// generate_alias_tables (lib/local/alias.f90:101-184) for ntab tables of n
// entries, one thread per table; table t is built when gate[t / gate_div] and
// tot[t] are nonzero (the reference builds no table otherwise).
__global__ void k_hb_alias(int64_t ntab, int n, const double* __restrict__ w, const double* __restrict__ tot,
                           const double* __restrict__ gate, int64_t gate_div, double* __restrict__ u,
                           int* __restrict__ k) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= ntab) return;
  if (!(fabs(gate[t / gate_div]) > 0.0) || !(fabs(tot[t]) > 0.0)) return;
  fqd::build_alias(n, w + t * n, tot[t], u + t * n, k + t * n);
}

int build_heat_bath(fq_engine* e) {
  Sys& s = e->sys;
  const int64_t nb = s.nbasis, n2 = nb * nb, n3 = n2 * nb, n4 = n3 * nb;
  int rc = FQ_OK;
  rc |= dalloc(&e->d_hb_i, nb);
  rc |= dalloc(&e->d_hb_ij, n2);
  rc |= dalloc(&e->d_hb_ija_tot, n2);
  rc |= dalloc(&e->d_hb_ija, n3);
  rc |= dalloc(&e->d_hb_ija_u, n3);
  rc |= dalloc(&e->d_hb_ija_k, n3);
  rc |= dalloc(&e->d_hb_ijab_tot, n3);
  rc |= dalloc(&e->d_hb_ijab, n4);
  rc |= dalloc(&e->d_hb_ijab_u, n4);
  rc |= dalloc(&e->d_hb_ijab_k, n4);
  rc |= dalloc(&e->d_hb_bias, nb);
  if (rc != FQ_OK) return FQ_ERR_NOMEM;
  hipStream_t st = e->stream;
  FQ_HIP(hipMemsetAsync(e->d_hb_ija_u, 0, sizeof(double) * n3, st));
  FQ_HIP(hipMemsetAsync(e->d_hb_ija_k, 0, sizeof(int) * n3, st));
  FQ_HIP(hipMemsetAsync(e->d_hb_ijab_u, 0, sizeof(double) * n4, st));
  FQ_HIP(hipMemsetAsync(e->d_hb_ijab_k, 0, sizeof(int) * n4, st));
  hipLaunchKernelGGL(k_hb_ijab, dim3(static_cast<unsigned>(ceil_div(n3, kBlock))), dim3(kBlock), 0, st, s,
                     e->d_hb_ijab, e->d_hb_ijab_tot, e->d_hb_ija);
  hipLaunchKernelGGL(k_hb_ij, dim3(static_cast<unsigned>(ceil_div(n2, kBlock))), dim3(kBlock), 0, st, s,
                     e->d_hb_ijab, e->d_hb_ija, e->d_hb_ij, e->d_hb_ija_tot);
  hipLaunchKernelGGL(k_hb_i, dim3(static_cast<unsigned>(ceil_div(nb, kBlock))), dim3(kBlock), 0, st, s, e->d_hb_ij,
                     e->d_hb_ija, e->d_hb_i, e->d_hb_bias);
  hipLaunchKernelGGL(k_hb_alias, dim3(static_cast<unsigned>(ceil_div(n2, kBlock))), dim3(kBlock), 0, st, n2,
                     static_cast<int>(nb), e->d_hb_ija, e->d_hb_ija_tot, e->d_hb_ija_tot, int64_t{1}, e->d_hb_ija_u,
                     e->d_hb_ija_k);
  hipLaunchKernelGGL(k_hb_alias, dim3(static_cast<unsigned>(ceil_div(n3, kBlock))), dim3(kBlock), 0, st, n3,
                     static_cast<int>(nb), e->d_hb_ijab, e->d_hb_ijab_tot, e->d_hb_ija_tot, nb, e->d_hb_ijab_u,
                     e->d_hb_ijab_k);
  FQ_HIP(hipGetLastError());
  std::vector<int> bias(static_cast<size_t>(nb));
  FQ_HIP(hipMemcpyAsync(bias.data(), e->d_hb_bias, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
  FQ_HIP(hipStreamSynchronize(st));
  for (int v : bias)
    if (v) return FQ_ERR_BIASED;
  s.hb_i = e->d_hb_i;
  s.hb_ij = e->d_hb_ij;
  s.hb_ija = e->d_hb_ija;
  s.hb_ija_tot = e->d_hb_ija_tot;
  s.hb_ija_u = e->d_hb_ija_u;
  s.hb_ija_k = e->d_hb_ija_k;
  s.hb_ijab = e->d_hb_ijab;
  s.hb_ijab_tot = e->d_hb_ijab_tot;
  s.hb_ijab_u = e->d_hb_ijab_u;
  s.hb_ijab_k = e->d_hb_ijab_k;
  return FQ_OK;
}

template <class T>
int upload(T** dst, const T* src, int64_t count, hipStream_t st);

// Binomial table and value width of the sort keys (combinatorial ranks).
int setup_rank_keys(fq_engine* e) {
  const int nb = e->sys.nbasis, nel = e->sys.nel;
  if (nel < 1 || nel > nb) return FQ_OK;
  // Pascal's triangle up to row nb in 128-bit integers (C(128, 64) < 2^128)
  std::vector<unsigned __int128> row(static_cast<size_t>(nel) + 1, 0), prev;
  std::vector<uint64_t> tab(static_cast<size_t>(2) * nb * nel, 0);
  row[0] = 1;
  for (int p = 0; p <= nb; ++p) {
    if (p > 0) {
      prev = row;
      for (int i = 1; i <= nel; ++i) row[i] = prev[i] + prev[i - 1];
    }
    if (p < nb)
      for (int i = 1; i <= nel; ++i) {
        tab[2 * (static_cast<size_t>(p) * nel + (i - 1))] = static_cast<uint64_t>(row[i]);
        tab[2 * (static_cast<size_t>(p) * nel + (i - 1)) + 1] = static_cast<uint64_t>(row[i] >> 64);
      }
  }
  unsigned __int128 top = row[nel] - 1;  // largest rank: C(nb, nel) - 1
  int bits = 1;
  while (bits < 128 && (top >> bits) != 0) ++bits;
  e->rank_bits = bits;
  e->rank_words = bits > 64 ? 2 : 1;
  return upload(&e->d_binom, tab.data(), static_cast<int64_t>(tab.size()), e->stream);
}

constexpr int64_t kTrainBatchMax = 16384;  // heat-bath samples per training step

int setup_learned(fq_engine* e) {
  Sys& s = e->sys;
  int rc = fqm::mlp_alloc(e->mlp, s.nbasis, s.W, kTrainBatchMax);
  if (rc != FQ_OK) return rc;
  rc = fqm::mlp_set_classes(e->mlp, s.nbss, s.ssbf, s.nsym_pg, s.max_sym_spin);
  if (rc != FQ_OK) return rc;
  rc = fqm::mlp_init(e->mlp, e->par.philox_seed, e->stream);
  if (rc != FQ_OK) return rc;
  e->mlp_tab_rows = std::min<int64_t>(e->cap, int64_t{1} << 16);
  if (dalloc(&e->mlp_tab, e->mlp_tab_rows * e->mlp.d_out) != FQ_OK || dalloc(&e->mlp_on, e->cap) != FQ_OK ||
      dalloc(&e->mlp_row, e->cap + 1) != FQ_OK || dalloc(&e->mlp_idx, e->cap) != FQ_OK)
    return FQ_ERR_NOMEM;
  s.mlp_tab = e->mlp_tab;
  s.mlp_aff = e->mlp.aff;
  s.mlp_rs = e->mlp.rs;
  s.mlp_crs = e->mlp.crs;
  s.mlp_cmask = e->mlp.cmask;
  s.mlp_au = e->mlp.au;
  s.mlp_ak = e->mlp.ak;
  s.mlp_cu = e->mlp.cu;
  s.mlp_ck = e->mlp.ck;
  {
    const std::vector<float> ones(2 * static_cast<size_t>(s.nbasis), 1.0f);
    if (dalloc(&e->mlp_flat, 2 * static_cast<int64_t>(s.nbasis)) != FQ_OK) return FQ_ERR_NOMEM;
    FQ_HIP(hipMemcpy(e->mlp_flat, ones.data(), sizeof(float) * ones.size(), hipMemcpyHostToDevice));
  }
  s.mlp_flat = e->mlp_flat;
  s.mlp_ld = e->mlp.d_out;
  s.mlp_on = e->mlp_on;
  s.mlp_row = e->mlp_row;
  s.mix = e->par.proposal_mix > 0.0 ? e->par.proposal_mix : 0.1;
  return FQ_OK;
}

// Determinants with at least min_att spawning attempts this cycle get a
// network row; mlp_gather lists them in order.
__global__ void k_mlp_mark(int64_t n, int S, const int64_t* __restrict__ nat, int64_t min_att,
                           int64_t* __restrict__ on) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t a = 0;
  for (int sp = 0; sp < S; ++sp) a += nat[i * S + sp];
  on[i] = a >= min_att ? 1 : 0;
}
__global__ void k_mlp_gather(int64_t n, const int64_t* __restrict__ on, const int64_t* __restrict__ row,
                             int64_t* __restrict__ idx) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && on[i]) idx[row[i]] = i;
}

int scan_cb(void* ctx, const int64_t* in, int64_t* out, int64_t n) {
  return scan_exclusive(static_cast<fq_engine*>(ctx), in, out, n);
}

// One generator call per thread from determinant f (the excitation probe).
__global__ void k_probe(Sys s, uint64_t f0, uint64_t f1, const float* __restrict__ wrow, int64_t n, uint64_t seed,
                        int32_t* __restrict__ exc, double* __restrict__ pgen, double* __restrict__ hmat) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint64_t f[FQ_MAX_WORDS] = {f0, f1};
  fqd::Philox rng(seed, static_cast<uint32_t>(t), static_cast<uint32_t>(t >> 32), fqd::kDrawProbe << 24);
  fqd::Excit conn;
  double h = 0.0, pg = 1.0;
  const bool ok = fqd::generate_any(s, rng, f, wrow, conn, h, pg);
  exc[6 * t] = ok ? 1 : 0;
  exc[6 * t + 1] = ok ? conn.nexcit : 0;
  exc[6 * t + 2] = conn.from[0];
  exc[6 * t + 3] = conn.from[1];
  exc[6 * t + 4] = conn.to[0];
  exc[6 * t + 5] = conn.to[1];
  pgen[t] = pg;
  hmat[t] = h;
}


// Upload the system descriptor to its device copy when it has changed.
int sync_sysp(fq_engine* e) {
  if (!e->sys_dirty) return FQ_OK;
  FQ_HIP(hipMemcpyAsync(e->d_sysp, &e->sys, sizeof(Sys), hipMemcpyHostToDevice, e->stream));
  e->sys_dirty = false;
  return FQ_OK;
}

// Dynamic LDS of the heat-bath spawning kernel: nel weights per thread.
size_t hb_lds_bytes(int nel) { return static_cast<size_t>(nel) * kBlock * sizeof(double); }
// Whether that LDS, with the kernel's static LDS, fits a workgroup here.
bool hb_lds_fits(int device, int nel) {
  hipFuncAttributes fa;
  int maxlds = 0;
  if (hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_spawn<fqd::kGenHeatBath, false>)) != hipSuccess ||
      hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
    return false;
  return fa.sharedSizeBytes + hb_lds_bytes(nel) <= static_cast<size_t>(maxlds);
}

// Dynamic LDS of the learned spawning kernel: the tile's parked child
// determinants, the affinity table and the flat weight row.
size_t learned_lds_bytes(int W, int nbasis) {
  const size_t nb = static_cast<size_t>(nbasis);
  return sizeof(uint64_t) * W * kSpawnTile + sizeof(float) * (nb * nb + 2 * nb);
}
bool learned_lds_fits(int device, int W, int nbasis) {
  hipFuncAttributes fa;
  int maxlds = 0;
  if (hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_spawn<fqd::kGenLearned, false>)) != hipSuccess ||
      hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
    return false;
  return fa.sharedSizeBytes + learned_lds_bytes(W, nbasis) <= static_cast<size_t>(maxlds);
}

// Spawned entries of a batch of nbt attempts starting at t_begin: scan the
// per-tile success counts, pack the tiles in order, advance the spawned count.
// ptot: the attempt total on the device (recorded in the stats).
int pack_spawned(fq_engine* e, int64_t nbt, int64_t t_begin, const int64_t* ptot, int64_t tlim = -1) {
  const int64_t nt = ceil_div(nbt, kSpawnTile);
  int rc = scan_exclusive(e, e->tile_cnt, e->tile_off, nt);
  if (rc != FQ_OK) return rc;
  (void)t_begin;
  KeyPlan kp = key_plan(e);
  const size_t tab = binom_bytes(e);
  kp.binom_lds = tab > 0 && tab <= kBinomLdsMax;
  hipLaunchKernelGGL(k_pack_tiles, dim3(static_cast<unsigned>(std::min<int64_t>(nt, 2048))), dim3(kBlock),
                     kp.binom_lds ? tab : 0, e->stream, e->W, e->S, nt, e->tile_cnt, e->tile_off, e->st_det,
                     e->st_pop, e->st_ent, e->bcap, e->scap, kp, e->kbuf[0], e->pbuf[0], e->d_stats);
  hipLaunchKernelGGL(k_add_spawned, dim3(1), dim3(1), 0, e->stream, e->tile_off, nt, ptot, e->d_stats, tlim);
  return FQ_OK;
}

int alloc_engine(fq_engine* e) {
  const int64_t cap = e->cap, scap = e->scap;
  const int W = e->W, S = e->S;
  int rc = FQ_OK;
  for (int b = 0; b < 2; ++b) {
    rc |= dalloc(&e->det[b], W * cap);
    rc |= dalloc(&e->pop[b], S * cap);
    rc |= dalloc(&e->kii[b], cap);
  }
  rc |= dalloc(&e->pop_new, S * cap);
  rc |= dalloc(&e->nat, S * cap + 1);  // det-major (det, space) entries
  rc |= dalloc(&e->att_off, S * cap + 1);
  rc |= dalloc(&e->flags, cap);
  rc |= dalloc(&e->mmask, cap / 64 + 2);
  rc |= dalloc(&e->mcnt, cap / 64 + 2);
  rc |= dalloc(&e->moff, cap / 64 + 2);
  const bool replay = e->par.rng_mode == FQ_RNG_REPLAY;
  if (replay) {  // rounding-draw counts of the replay stream
    rc |= dalloc(&e->mdraw, cap + 1);
    rc |= dalloc(&e->mdraw_off, cap + 1);
  }
  rc |= dalloc(&e->part_d, 4 * (cap / kBlock + 2));
  rc |= dalloc(&e->part_i, 2 * (cap / kBlock + 2));
  // spawning attempts are staged in batches of at most bcap (one slot each)
  e->bcap = std::min<int64_t>(scap, int64_t{1} << 26);
  rc |= dalloc(&e->st_det, W * e->bcap);
  rc |= dalloc(&e->st_pop, e->bcap);
  rc |= dalloc(&e->st_ent, e->bcap);
  rc |= dalloc(&e->tile_cnt, e->bcap / kSpawnTile + 2);
  rc |= dalloc(&e->tile_off, e->bcap / kSpawnTile + 2);
  rc |= dalloc(&e->tile_win, 2 * (e->bcap / kSpawnTile + 2));
  rc |= dalloc(&e->rdet, W * scap);
  rc |= dalloc(&e->rpop, S * scap);
  rc |= dalloc(&e->rflag, scap);
  rc |= dalloc(&e->rpos, scap);
  rc |= dalloc(&e->rstate, scap);
  rc |= dalloc(&e->nlist, scap);
  rc |= dalloc(&e->nrpos, scap);
  rc |= dalloc(&e->bstart, cap / kBlock + 2);
  rc |= dalloc(&e->d_sysp, 1);
  rc |= dalloc(&e->mwin, 2 * (scap / kBlock + 2));
  rc |= dalloc(&e->nmask, scap / 64 + 2);
  rc |= dalloc(&e->ncnt, scap / 64 + 2);
  rc |= dalloc(&e->noff, scap / 64 + 2);
  if (replay) {
    rc |= dalloc(&e->ndraw, scap + 1);
    rc |= dalloc(&e->ndraw_off, scap + 1);
  }
  rc |= dalloc(&e->d_stats, 1);
  return rc == FQ_OK ? FQ_OK : FQ_ERR_NOMEM;
}

template <class T>
int upload(T** dst, const T* src, int64_t count, hipStream_t st) {
  if (!src || count <= 0) {
    *dst = nullptr;
    return FQ_OK;
  }
  if (dalloc(dst, count) != FQ_OK) return FQ_ERR_NOMEM;
  FQ_HIP(hipMemcpyAsync(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice, st));
  return FQ_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int fq_create(const fq_system* sys, const fq_params* params, fq_engine** out) {
  if (!sys || !params || !out) return FQ_ERR_ARG;
  if (sys->nwords < 1 || sys->nwords > FQ_MAX_WORDS || params->nspaces < 1 || params->nspaces > FQ_MAX_SPACES)
    return FQ_ERR_ARG;
  if (sys->system == FQ_SYS_READ_IN && sys->nsym_pg > 8) return FQ_ERR_ARG;
  if (params->excit_gen < FQ_EXCIT_RENORM || params->excit_gen > FQ_EXCIT_LEARNED) return FQ_ERR_ARG;
  if ((params->excit_gen == FQ_EXCIT_HEAT_BATH || params->excit_gen == FQ_EXCIT_LEARNED) &&
      (sys->system != FQ_SYS_READ_IN || sys->nel > fqd::kHbMaxNel))
    return FQ_ERR_ARG;
  if (params->excit_gen == FQ_EXCIT_LEARNED && (params->rng_mode != FQ_RNG_PHILOX || params->proposal_mix < 0.0 ||
                                                 params->proposal_mix > 1.0 || params->proposal_min_attempts < 0))
    return FQ_ERR_ARG;
  if (sys->system == FQ_SYS_READ_IN && params->excit_gen == FQ_EXCIT_NO_RENORM) return FQ_ERR_ARG;
  if (params->main_capacity < 1 || params->spawn_capacity < 1) return FQ_ERR_ARG;
  // (determinant, space) entries are addressed relative to a window in 32 bits
  if (params->main_capacity * params->nspaces >= (int64_t{1} << 31)) return FQ_ERR_ARG;
  if (params->spawn_capacity > 0xffffffffLL) return FQ_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= params->device) return FQ_ERR_NO_DEVICE;
  auto* e = new fq_engine();
  e->device = params->device;
  if (hipSetDevice(e->device) != hipSuccess) {
    delete e;
    return FQ_ERR_NO_DEVICE;
  }
  if (params->excit_gen == FQ_EXCIT_HEAT_BATH && !hb_lds_fits(e->device, sys->nel)) {
    delete e;
    return FQ_ERR_ARG;
  }
  if (params->excit_gen == FQ_EXCIT_LEARNED && !learned_lds_fits(e->device, sys->nwords, sys->nbasis)) {
    delete e;
    return FQ_ERR_ARG;
  }
  if (hipStreamCreate(&e->stream) != hipSuccess) {
    delete e;
    return FQ_ERR_DEVICE;
  }
  e->fixed_scratch = fixed_scratch_address(e->device);
  e->par = *params;
  e->W = sys->nwords;
  e->S = params->nspaces;
  e->key_bits = sys->nbasis;
  e->cap = params->main_capacity;
  e->scap = params->spawn_capacity;
  Sys& s = e->sys;
  s.system = sys->system;
  s.nbasis = sys->nbasis;
  s.nel = sys->nel;
  s.nalpha = sys->nalpha;
  s.nbeta = sys->nbeta;
  s.nvirt = sys->nvirt;
  s.nvirt_alpha = sys->nvirt_alpha;
  s.nvirt_beta = sys->nvirt_beta;
  s.W = sys->nwords;
  s.S = params->nspaces;
  for (int w = 0; w < FQ_MAX_WORDS; ++w) {
    s.f0[w] = w < sys->nwords ? sys->f0[w] : 0;
    const int bits = std::max(0, std::min(64, sys->nbasis - 64 * w));
    s.valid[w] = bits == 64 ? ~uint64_t{0} : ((uint64_t{1} << bits) - 1);
  }
  s.H00 = sys->H00;
  s.coulomb_k = sys->coulomb_k;
  s.nsym = sys->nsym;
  s.norb = sys->norb;
  s.nsym_pg = sys->nsym_pg;
  s.max_sym_spin = sys->max_sym_spin;
  s.ecore = sys->ecore;
  s.psingle = sys->pattempt_single;
  s.pdouble = sys->pattempt_double;
  s.initiator = params->initiator;
  s.quadrature = params->quadrature_initiator;
  s.real_amps = params->real_amplitudes;
  s.excit_gen = params->excit_gen;
  s.initiator_pop = params->initiator_pop;
  s.rf = int64_t{1} << params->pop_bit_shift;
  s.spawn_cutoff = params->spawn_cutoff;
  int rc = FQ_OK;
  rc |= upload(&e->d_sp_eigv, sys->sp_eigv, sys->nbasis, e->stream);
  if (sys->system == FQ_SYS_HUB_K) {
    rc |= upload(&e->d_sym_table, sys->sym_table, static_cast<int64_t>(sys->nsym) * sys->nsym, e->stream);
    rc |= upload(&e->d_inv_sym, sys->inv_sym, sys->nsym, e->stream);
  } else {
    rc |= upload(&e->d_orb_sym, sys->orb_sym, sys->nbasis, e->stream);
    rc |= upload(&e->d_nbss, sys->nbasis_sym_spin, 2 * sys->nsym_pg, e->stream);
    rc |= upload(&e->d_ssbf, sys->sym_spin_basis_fns, 2 * sys->nsym_pg * sys->max_sym_spin, e->stream);
    rc |= upload(&e->d_h1, sys->h1, static_cast<int64_t>(sys->norb) * sys->norb, e->stream);
    rc |= upload(&e->d_eri, sys->eri, sys->n_eri, e->stream);
    // Dense Coulomb/exchange pair tables for the diagonal rule: entries copied
    // from the packed store (same values, no re-computation).
    const int no = sys->norb;
    std::vector<double> J(static_cast<size_t>(no) * no), K(static_cast<size_t>(no) * no);
    auto tri = [](int64_t i, int64_t j) { return (i * (i - 1)) / 2 + j; };
    auto idx = [&](int p, int q, int r, int t) {  // (pq|rt) packed, 1-based spatial
      const int64_t a = p >= q ? tri(p, q) : tri(q, p), b = r >= t ? tri(r, t) : tri(t, r);
      return (a >= b ? tri(a, b) : tri(b, a)) - 1;
    };
    for (int pp = 1; pp <= no; ++pp)
      for (int qq = 1; qq <= no; ++qq) {
        J[(pp - 1) * no + (qq - 1)] = sys->eri[idx(pp, pp, qq, qq)];  // <ij|ij> = (ii|jj)
        K[(pp - 1) * no + (qq - 1)] = sys->eri[idx(pp, qq, qq, pp)];  // <ij|ji> = (ij|ji)
      }
    // Single-excitation tables: (pa|oo) and (po|oa) for every spatial p, a, o.
    std::vector<double> S1J(static_cast<size_t>(no) * no * no), S1K(S1J.size());
    for (int pp = 1; pp <= no; ++pp)
      for (int aa = 1; aa <= no; ++aa)
        for (int oo = 1; oo <= no; ++oo) {
          const size_t t = (static_cast<size_t>(pp - 1) * no + (aa - 1)) * no + (oo - 1);
          S1J[t] = sys->eri[idx(pp, aa, oo, oo)];
          S1K[t] = sys->eri[idx(pp, oo, oo, aa)];
        }
    if (rc == FQ_OK) rc = upload(&e->d_jmat, J.data(), static_cast<int64_t>(J.size()), e->stream);
    if (rc == FQ_OK) rc = upload(&e->d_kmat, K.data(), static_cast<int64_t>(K.size()), e->stream);
    if (rc == FQ_OK) rc = upload(&e->d_s1j, S1J.data(), static_cast<int64_t>(S1J.size()), e->stream);
    if (rc == FQ_OK) rc = upload(&e->d_s1k, S1K.data(), static_cast<int64_t>(S1K.size()), e->stream);
    // Host-evaluated divisions and square roots of the generators' integer
    // arguments (IEEE division and sqrt are correctly rounded on both sides).
    std::vector<double> R(kRecipN, 0.0);
    for (int k = 1; k < kRecipN; ++k) R[k] = 1.0 / static_cast<double>(k);
    const int npairs = sys->nel * (sys->nel - 1) / 2;
    std::vector<int> JP(static_cast<size_t>(npairs) + 2, 0);
    for (int ind = 1; ind <= npairs + 1; ++ind) JP[ind] = static_cast<int>(1.50 + std::sqrt(2 * ind - 1.750));
    if (rc == FQ_OK) rc = upload(&e->d_recip, R.data(), kRecipN, e->stream);
    if (rc == FQ_OK) rc = upload(&e->d_jpair, JP.data(), static_cast<int64_t>(JP.size()), e->stream);
    if (rc == FQ_OK) FQ_HIP(hipStreamSynchronize(e->stream));
  }
  s.sp_eigv = e->d_sp_eigv;
  s.sym_table = e->d_sym_table;
  s.inv_sym = e->d_inv_sym;
  s.orb_sym = e->d_orb_sym;
  s.nbss = e->d_nbss;
  s.ssbf = e->d_ssbf;
  s.h1 = e->d_h1;
  s.eri = e->d_eri;
  s.jmat = e->d_jmat;
  s.kmat = e->d_kmat;
  s.s1j = e->d_s1j;
  s.s1k = e->d_s1k;
  s.recip = e->d_recip;
  s.nrecip = e->d_recip ? kRecipN : 0;
  s.jpair = e->d_jpair;
  if (rc == FQ_OK) rc = setup_rank_keys(e);
  if (rc == FQ_OK && (params->excit_gen == FQ_EXCIT_HEAT_BATH || params->excit_gen == FQ_EXCIT_LEARNED))
    rc = build_heat_bath(e);
  if (rc == FQ_OK && params->excit_gen == FQ_EXCIT_LEARNED) rc = setup_learned(e);
  if (rc == FQ_OK) rc = alloc_engine(e);
  if (rc == FQ_OK) rc = setup_sort(e);
  for (int k = 0; k < 4 && rc == FQ_OK; ++k)
    if (hipEventCreate(&e->ev[k]) != hipSuccess) rc = FQ_ERR_DEVICE;
  for (int k = 0; k < 9 && rc == FQ_OK; ++k)
    if (hipEventCreate(&e->pev[k]) != hipSuccess) rc = FQ_ERR_DEVICE;
  for (int k = 0; k < 2 * fq_engine::kSpawnEvents && rc == FQ_OK; ++k)
    if (hipEventCreate(&e->sev[k]) != hipSuccess) rc = FQ_ERR_DEVICE;
  if (rc == FQ_OK && hipEventCreateWithFlags(&e->ev_len, hipEventDisableTiming) != hipSuccess) rc = FQ_ERR_DEVICE;
  if (rc == FQ_OK && hipHostMalloc(reinterpret_cast<void**>(&e->h_len), 4 * sizeof(int64_t)) != hipSuccess)
    rc = FQ_ERR_NOMEM;
  if (rc != FQ_OK) {
    fq_destroy(e);
    return rc;
  }
  if (hipStreamSynchronize(e->stream) != hipSuccess) {
    fq_destroy(e);
    return FQ_ERR_DEVICE;
  }
  *out = e;
  return FQ_OK;
}

DetermDev determ_dev(const fq_engine* e) {
  return DetermDev{e->d_drow, e->d_dcol, e->d_dmat, e->d_dvec, e->d_dpos, e->ndeterm, e->determ_separate};
}

void release_determ(fq_engine* e) {
  dfree(e->d_determ);
  dfree(e->d_drow);
  dfree(e->d_dcol);
  dfree(e->d_dmat);
  dfree(e->d_dvec);
  dfree(e->d_dpos);
  if (e->sys.ndeterm > 0) e->sys_dirty = true;
  e->ndeterm = 0;
  e->determ_nnz = 0;
  e->sys.determ = nullptr;
  e->sys.ndeterm = 0;
}

int fq_load(fq_engine* e, const uint64_t* states, const int64_t* pops, const double* dat, int64_t n) {
  if (!e || n < 0 || (n > 0 && (!states || !pops || !dat))) return FQ_ERR_ARG;
  if (n > e->cap) return FQ_ERR_CAPACITY;
  FQ_HIP(hipSetDevice(e->device));
  const int W = e->W, S = e->S;
  std::vector<uint64_t> dw(static_cast<size_t>(W) * n);
  std::vector<int64_t> dp(static_cast<size_t>(S) * n);
  for (int64_t i = 0; i < n; ++i) {
    for (int w = 0; w < W; ++w) dw[w * n + i] = states[i * W + w];
    for (int s = 0; s < S; ++s) dp[s * n + i] = pops[i * S + s];
  }
  e->cur = 0;
  for (int w = 0; w < W && n > 0; ++w)
    FQ_HIP(hipMemcpyAsync(e->det[0] + w * e->cap, dw.data() + w * n, 8 * n, hipMemcpyHostToDevice, e->stream));
  for (int s = 0; s < S && n > 0; ++s)
    FQ_HIP(hipMemcpyAsync(e->pop[0] + s * e->cap, dp.data() + s * n, 8 * n, hipMemcpyHostToDevice, e->stream));
  if (n > 0) FQ_HIP(hipMemcpyAsync(e->kii[0], dat, 8 * n, hipMemcpyHostToDevice, e->stream));
  FQ_HIP(hipStreamSynchronize(e->stream));
  e->n = n;
  e->att_bound = -1;
  release_determ(e);  // a new list: no semi-stochastic core until fq_semi_stoch_init
  return FQ_OK;
}

// init_semi_stoch_t with space "high" (src/semi_stoch.F90:134-377,
// create_high_pop_space :1198-1316, find_most_populated_dets :1318-1385):
// the target_size most populated entries of the current list (sum over
// spaces of |N|).  The reference keeps them with a replace-the-minimum scan
// (a later entry replaces the first smallest kept one when strictly larger);
// a min-heap ordered by (population, slot) replaces the same entries.  The
// core is kept sorted as the main list (every core determinant is in it, so
// add_determ_dets_to_psip_list inserts nothing) and its Hamiltonian is built
// on the device (create_determ_hamil :533-685).
int fq_semi_stoch_init(fq_engine* e, int64_t target_size, int separate) {
  if (!e || target_size < 0 || e->batched || e->pending) return FQ_ERR_ARG;
  FQ_HIP(hipSetDevice(e->device));
  FQ_HIP(hipStreamSynchronize(e->stream));
  release_determ(e);
  const int64_t n = e->n, nd = std::min(target_size, n);
  if (nd == 0) return FQ_OK;
  if (nd > INT32_MAX) return FQ_ERR_ARG;  // column indices are 32-bit
  const int W = e->W, S = e->S, c = e->cur;
  hipStream_t st = e->stream;
  std::vector<int64_t> pw(static_cast<size_t>(S) * n);
  std::vector<uint64_t> dw(static_cast<size_t>(W) * n);
  for (int sp = 0; sp < S; ++sp)
    FQ_HIP(hipMemcpyAsync(pw.data() + sp * n, e->pop[c] + sp * e->cap, 8 * n, hipMemcpyDeviceToHost, st));
  for (int w = 0; w < W; ++w)
    FQ_HIP(hipMemcpyAsync(dw.data() + w * n, e->det[c] + w * e->cap, 8 * n, hipMemcpyDeviceToHost, st));
  FQ_HIP(hipStreamSynchronize(st));
  auto absum = [&](int64_t i) {
    int64_t t = 0;
    for (int sp = 0; sp < S; ++sp) t += std::llabs(pw[sp * n + i]);
    return t;
  };
  using Slot = std::pair<int64_t, int64_t>;  // (population, slot)
  std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap;
  std::vector<int64_t> idx(nd);
  for (int64_t i = 0; i < nd; ++i) {
    idx[i] = i;
    heap.push({absum(i), i});
  }
  for (int64_t i = nd; i < n; ++i) {
    const int64_t a = absum(i);
    if (a > heap.top().first) {
      const int64_t slot = heap.top().second;
      heap.pop();
      idx[slot] = i;
      heap.push({a, slot});
    }
  }
  std::sort(idx.begin(), idx.end());  // list order
  std::vector<uint64_t> core(static_cast<size_t>(W) * nd);
  for (int w = 0; w < W; ++w)
    for (int64_t k = 0; k < nd; ++k) core[w * nd + k] = dw[w * n + idx[k]];
  int rc = dalloc(&e->d_determ, W * nd);
  rc |= dalloc(&e->d_drow, nd + 1);
  rc |= dalloc(&e->d_dvec, 2 * nd * S);  // populations, then (separate) the projected amplitudes
  rc |= dalloc(&e->d_dpos, nd);
  int64_t* cnt = nullptr;
  rc |= dalloc(&cnt, nd);
  if (rc != FQ_OK) {
    dfree(cnt);
    release_determ(e);
    return FQ_ERR_NOMEM;
  }
  FQ_HIP(hipMemcpyAsync(e->d_determ, core.data(), 8 * W * nd, hipMemcpyHostToDevice, st));
  e->sys.determ = e->d_determ;
  e->sys.ndeterm = nd;
  e->sys_dirty = true;
  const int transp = separate ? 1 : 0;
  hipLaunchKernelGGL(k_determ_count, dim3(static_cast<unsigned>(nd)), dim3(kBlock), 0, st, e->sys, transp, cnt);
  rc = scan_exclusive(e, cnt, e->d_drow, nd);
  int64_t nnz = 0;
  if (rc == FQ_OK) {
    FQ_HIP(hipMemcpyAsync(&nnz, e->d_drow + nd, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    FQ_HIP(hipStreamSynchronize(st));
    rc = dalloc(&e->d_dcol, nnz) | dalloc(&e->d_dmat, nnz);
  }
  if (rc == FQ_OK) {
    hipLaunchKernelGGL(k_determ_fill, dim3(static_cast<unsigned>(nd)), dim3(kBlock), 0, st, e->sys, transp,
                       e->d_drow, e->d_dcol, e->d_dmat);
    if (hipStreamSynchronize(st) != hipSuccess) rc = FQ_ERR_DEVICE;
  }
  dfree(cnt);
  if (rc != FQ_OK) {
    release_determ(e);
    return rc == FQ_ERR_DEVICE ? rc : FQ_ERR_NOMEM;
  }
  e->ndeterm = nd;
  e->determ_nnz = nnz;
  e->determ_separate = transp;
  return FQ_OK;
}

int fq_semi_stoch_info(fq_engine* e, int64_t* ndeterm, int64_t* nnz, uint64_t* states, int64_t capacity) {
  if (!e || !ndeterm) return FQ_ERR_ARG;
  *ndeterm = e->ndeterm;
  if (nnz) *nnz = e->determ_nnz;
  if (!states || e->ndeterm == 0) return FQ_OK;
  if (capacity < e->ndeterm) return FQ_ERR_CAPACITY;
  FQ_HIP(hipSetDevice(e->device));
  std::vector<uint64_t> dw(static_cast<size_t>(e->W) * e->ndeterm);
  FQ_HIP(hipMemcpyAsync(dw.data(), e->d_determ, 8 * dw.size(), hipMemcpyDeviceToHost, e->stream));
  FQ_HIP(hipStreamSynchronize(e->stream));
  for (int64_t k = 0; k < e->ndeterm; ++k)
    for (int w = 0; w < e->W; ++w) states[k * e->W + w] = dw[w * e->ndeterm + k];
  return FQ_OK;
}

int fq_push_rng(fq_engine* e, const double* draws, int64_t n) {
  if (!e || n < 0 || (n > 0 && !draws)) return FQ_ERR_ARG;
  if (n == 0) return FQ_OK;
  FQ_HIP(hipSetDevice(e->device));
  int rc = reserve_rng(e, n);
  if (rc != FQ_OK) return rc;
  FQ_HIP(hipMemcpyAsync(e->rng_buf + e->rng_end, draws, sizeof(double) * n, hipMemcpyHostToDevice, e->stream));
  FQ_HIP(hipStreamSynchronize(e->stream));
  e->rng_end += n;
  return FQ_OK;
}

int fq_rng_available(fq_engine* e, int64_t* n) {
  if (!e || !n) return FQ_ERR_ARG;
  *n = e->rng_end - e->rng_pos;
  return FQ_OK;
}

static int cycle_body(fq_engine* e, const double* shift, const double* eproj_old, double tau, fq_cycle_out* out) {
  if (!e || !shift || !eproj_old || !out || e->batched) return FQ_ERR_ARG;
  FQ_HIP(hipSetDevice(e->device));
  const bool replay = e->par.rng_mode == FQ_RNG_REPLAY;
  const int64_t n = e->n, cap = e->cap;
  const int c = e->cur;
  hipStream_t st = e->stream;
  const uint32_t cyc = static_cast<uint32_t>(e->cycle_index++);
  const double sh1 = e->S > 1 ? shift[1] : 0.0, e1 = e->S > 1 ? eproj_old[1] : 0.0;
  // per-cycle counters cleared by a kernel (no memset node in the stream)
  static_assert(sizeof(Stats) % sizeof(int64_t) == 0, "Stats is cleared as int64 words");
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(kBlock), 0, st, reinterpret_cast<int64_t*>(e->d_stats),
                     static_cast<int64_t>(sizeof(Stats) / sizeof(int64_t)));
  FQ_HIP(hipEventRecord(e->ev[0], st));
  FQ_HIP(hipEventRecord(e->pev[0], st));
  int64_t rng_pos = e->rng_pos;
  int64_t ns = 0;
  if (replay) {
    hipLaunchKernelGGL(k_replay_walk, dim3(1), dim3(1), 0, st, e->sys, n, cap, e->det[c], e->pop[c], e->kii[c],
                       e->pop_new, tau, shift[0], sh1, eproj_old[0], e1, e->rng_buf, e->rng_pos, e->rng_end, e->scap,
                       e->sdet, e->srec, e->d_stats, determ_dev(e));
    FQ_HIP(hipEventRecord(e->pev[1], st));
    FQ_HIP(hipEventRecord(e->pev[2], st));
    FQ_HIP(hipEventRecord(e->pev[3], st));
    int rc = sync_stats(e);
    if (rc != FQ_OK) return rc;
    if (e->h_stats.rng_failed) return FQ_ERR_RNG_EXHAUSTED;
    rng_pos = e->h_stats.rng_pos;
    ns = e->h_stats.nspawn;
  } else {
    if (e->online_every > 0 && n > 0 && cyc % static_cast<uint32_t>(e->online_every) == 0) {
      // online training from the current list: this cycle already uses the update
      fqm::Mlp& m = e->mlp;
      for (int32_t k = 0; k < e->online_steps; ++k) {
        int trc = fqm::mlp_sample(m, e->sys, e->det[c], e->pop[c], cap, n, e->online_batch,
                                  static_cast<int64_t>(e->par.proposal_min_attempts),
                                  e->par.philox_seed ^ 0x6a09e667f3bcc908ull, static_cast<uint32_t>(m.adam_t), st,
                                  scan_cb, e);
        if (trc == FQ_OK) trc = fqm::mlp_train_step(m, e->sys, e->online_batch, e->online_lr, true, st);
        if (trc != FQ_OK) return trc;
      }
      ++e->online_runs;
    }
    const unsigned nb = grid_for(n);
    if (n > 0)
      hipLaunchKernelGGL(k_det_pass, dim3(nb), dim3(kBlock), 0, st, e->sys, n, cap, e->det[c], e->pop[c], e->kii[c],
                         e->pop_new, e->nat, e->flags, tau, shift[0], sh1, eproj_old[0], e1, e->par.philox_seed, cyc,
                         e->part_d, e->part_i, e->d_dvec, e->d_dpos);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, st, e->part_d, e->part_i,
                       n > 0 ? static_cast<int>(nb) : 0, e->d_stats);
    FQ_HIP(hipEventRecord(e->pev[1], st));
    if (e->par.excit_gen == FQ_EXCIT_LEARNED && n > 0) {
      hipLaunchKernelGGL(k_mlp_mark, dim3(grid_for(n)), dim3(kBlock), 0, st, n, e->S, e->nat,
                         static_cast<int64_t>(e->par.proposal_min_attempts), e->mlp_on);
      int mrc = scan_exclusive(e, e->mlp_on, e->mlp_row, n);
      if (mrc != FQ_OK) return mrc;
      hipLaunchKernelGGL(k_mlp_gather, dim3(grid_for(n)), dim3(kBlock), 0, st, n, e->mlp_on, e->mlp_row,
                         e->mlp_idx);
      int64_t ninf = 0;
      FQ_HIP(hipMemcpyAsync(&ninf, e->mlp_row + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
      FQ_HIP(hipStreamSynchronize(st));
      e->prof.proposal_rows = ninf;
      if (ninf > e->mlp_tab_rows) {  // grow the weight table (rare: rows follow the inferred determinants)
        const int64_t rows = std::min<int64_t>(e->cap, std::max<int64_t>(ninf, 2 * e->mlp_tab_rows));
        dfree(e->mlp_tab);
        e->mlp_tab_rows = 0;
        if (dalloc(&e->mlp_tab, rows * e->mlp.d_out) != FQ_OK) return FQ_ERR_NOMEM;
        e->mlp_tab_rows = rows;
        e->sys.mlp_tab = e->mlp_tab;
        e->sys_dirty = true;
      }
      mrc = fqm::mlp_infer(e->mlp, e->sys, e->det[c], cap, ninf, e->mlp_tab, st, e->mlp_idx);
      if (mrc != FQ_OK) return mrc;
    }
    // Attempt offsets over the n*S (det, space) entries.
    const int64_t nent = e->S * n;
    int rc = scan_exclusive(e, e->nat, e->att_off, nent);
    if (rc != FQ_OK) return rc;
    // The spawning launches cover attempts [0, bound): after a production
    // cycle the host holds a bound of this cycle's total (every entry makes at
    // most floor(|N|) + 1 attempts, |N| for integer populations), so it need
    // not wait for the prefix sum; otherwise it reads the total.  The kernels
    // clamp to the total on the device.
    int64_t bound = e->att_bound;
    if (bound <= 0) {  // unknown (or an empty list): the exact total
      FQ_HIP(hipMemcpyAsync(&bound, e->att_off + nent, sizeof(int64_t), hipMemcpyDeviceToHost, st));
      FQ_HIP(hipStreamSynchronize(st));
    }
    FQ_HIP(hipEventRecord(e->pev[2], st));
    e->spawn_launches = 0;
    if (int rs = sync_sysp(e); rs != FQ_OK) return rs;
    auto spawn_range = [&](int64_t from, int64_t to) -> int {
      for (int64_t tb = from; tb < to;) {
        const int64_t te = std::min(to, tb + e->bcap);
        const int64_t nbt = te - tb;
        const unsigned g = static_cast<unsigned>(std::min<int64_t>(ceil_div(nbt, kSpawnTile), 1 << 20));
        const int li = e->spawn_launches++;
        hipLaunchKernelGGL(k_tile_bounds, dim3(grid_for(ceil_div(nbt, kSpawnTile))), dim3(kBlock), 0, st, nent,
                           e->att_off, tb, te, e->tile_win, e->tile_cnt);
        if (li < fq_engine::kSpawnEvents) FQ_HIP(hipEventRecord(e->sev[2 * li], st));
        const int gen = fqd::gen_of(e->sys.system, e->sys.excit_gen);
        auto kern = gen == fqd::kGenHubK       ? k_spawn<fqd::kGenHubK, false>
                    : gen == fqd::kGenHeatBath ? k_spawn<fqd::kGenHeatBath, false>
                    : gen == fqd::kGenLearned  ? k_spawn<fqd::kGenLearned, false>
                                               : k_spawn<fqd::kGenMolRenorm, false>;
        const size_t lds = gen == fqd::kGenHeatBath    ? hb_lds_bytes(e->sys.nel)
                           : gen == fqd::kGenMolRenorm ? sizeof(uint64_t) * e->W * kSpawnTile
                           : gen == fqd::kGenLearned   ? learned_lds_bytes(e->W, e->sys.nbasis)
                                                       : 0;
        hipLaunchKernelGGL(kern, dim3(g), dim3(kBlock), lds, st, e->d_sysp, nent, cap, tb, te, e->att_off,
                           e->det[c], e->pop[c], tau, e->par.philox_seed, cyc, e->bcap, e->st_det, e->st_pop,
                           e->st_ent, e->tile_cnt, e->tile_win, e->flags, nullptr, nullptr, nullptr, nullptr);
        if (li < fq_engine::kSpawnEvents) FQ_HIP(hipEventRecord(e->sev[2 * li + 1], st));
        const int prc = pack_spawned(e, nbt, tb, e->att_off + nent);
        if (prc != FQ_OK) return prc;
        tb = te;
      }
      return FQ_OK;
    };
    rc = spawn_range(0, bound);
    if (rc != FQ_OK) return rc;
    int64_t ndet_entries = 0;  // the deterministic projection's spawned entries (combined annihilation)
    if (e->ndeterm > 0) {
      const int64_t m = e->ndeterm * e->S;
      hipLaunchKernelGGL(k_determ_proj, dim3(grid_for(m)), dim3(kBlock), 0, st, e->sys, key_plan(e), determ_dev(e),
                         tau, shift[0], sh1, eproj_old[0], e1, e->par.philox_seed, cyc, cap, e->pop_new,
                         e->kbuf[0], e->pbuf[0], e->scap, e->d_stats);
      if (!e->determ_separate) {
        hipLaunchKernelGGL(k_add_nspawn, dim3(1), dim3(1), 0, st, e->d_stats, m);
        ndet_entries = m;
      }
    }
    FQ_HIP(hipEventRecord(e->pev[3], st));
    e->keys_ready = true;  // the packing kernels wrote the sort keys
    // The spawned-list length and the attempt total are copied to pinned host
    // memory behind the spawning kernels; the radix passes read the length on
    // the device (their grids sized by a bound: successes <= attempts <=
    // bound), so the host waits for the copy only once they are queued, and
    // the device does not idle meanwhile.
    FQ_HIP(hipMemcpyAsync(e->h_len, &e->d_stats->nspawn, 3 * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    FQ_HIP(hipEventRecord(e->ev_len, st));
    e->spawn_bound = bound;
    ns = std::min<int64_t>(bound + ndet_entries, e->scap);
  }
  FQ_HIP(hipEventRecord(e->ev[1], st));
  int rc = annihilate(e, ns, replay ? nullptr : &e->d_stats->nspawn, replay, cyc, rng_pos);
  if (rc != FQ_OK) return rc;
  FQ_HIP(hipEventRecord(e->ev[2], st));
  FQ_HIP(hipEventSynchronize(e->ev[2]));
  ns = std::min<int64_t>(e->h_stats.nspawn, e->scap);
  if (!replay) {  // next cycle's attempts <= sum floor(|N|) (+1 per entry with real amplitudes)
    unsigned long long sum = 0;
    for (int sp = 0; sp < e->S; ++sp) sum += e->h_stats.abs_pop[sp];
    e->att_bound = static_cast<int64_t>(sum / static_cast<unsigned long long>(e->sys.rf)) +
                   (e->sys.real_amps ? e->n * e->S : 0);
  }
  const int64_t used = rng_pos - e->rng_pos;
  e->rng_pos = rng_pos;
  (void)hipEventElapsedTime(&e->t_spawn, e->ev[0], e->ev[1]);
  (void)hipEventElapsedTime(&e->t_annih, e->ev[1], e->ev[2]);
  (void)hipEventElapsedTime(&e->t_total, e->ev[0], e->ev[2]);
  {
    float t[6] = {0, 0, 0, 0, 0, 0};
    (void)hipEventElapsedTime(&t[0], e->pev[0], e->pev[1]);
    (void)hipEventElapsedTime(&t[1], e->pev[2], e->pev[3]);
    (void)hipEventElapsedTime(&t[2], e->pev[4], e->pev[5]);
    (void)hipEventElapsedTime(&t[3], e->pev[5], e->pev[6]);
    (void)hipEventElapsedTime(&t[4], e->pev[7], e->pev[8]);
    (void)hipEventElapsedTime(&t[5], e->pev[0], e->pev[8]);
    fq_profile_out& pr = e->prof;
    pr.det_pass_ms = t[0];
    pr.spawn_ms = t[1];
    pr.sort_ms = t[2];
    pr.reduce_ms = t[3];
    pr.merge_ms = t[4];
    pr.total_ms = t[5];
    float tp = 0.0f;
    if (!replay) (void)hipEventElapsedTime(&tp, e->pev[1], e->pev[2]);
    pr.proposal_ms = e->par.excit_gen == FQ_EXCIT_LEARNED ? tp : 0.0;
    if (e->par.excit_gen != FQ_EXCIT_LEARNED || n == 0) pr.proposal_rows = 0;
    pr.nstates_in = n;
    pr.nattempts = e->h_stats.total_attempts;
    pr.nspawned = ns;
    pr.nruns = e->h_stats.nruns;
    pr.nstates_out = e->n;
    pr.radix_passes = e->radix_passes;
    pr.spawn_launches = replay ? 0 : std::min(e->spawn_launches, fq_engine::kSpawnEvents);
    pr.spawn_kernel_ms = 0.0;
    for (int li = 0; li < pr.spawn_launches; ++li) {
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e->sev[2 * li], e->sev[2 * li + 1]);
      pr.spawn_kernel_ms += ms;
    }
  }
  fill_out(e, out, used);
  return FQ_OK;
}

int fq_cycle(fq_engine* e, const double* shift, const double* eproj_old, double tau, fq_cycle_out* out) {
  if (e && e->pending) return FQ_ERR_ARG;
  return cycle_body(e, shift, eproj_old, tau, out);
}

int fq_cycle_spawn(fq_engine* e, const double* shift, const double* eproj_old, double tau, fq_cycle_out* out) {
  if (!e || !shift || !eproj_old || !out || e->batched || e->pending) return FQ_ERR_ARG;
  if (e->par.rng_mode != FQ_RNG_REPLAY) return FQ_ERR_ARG;
  FQ_HIP(hipSetDevice(e->device));
  const int c = e->cur;
  hipStream_t st = e->stream;
  const uint32_t cyc = static_cast<uint32_t>(e->cycle_index++);
  const double sh1 = e->S > 1 ? shift[1] : 0.0, e1 = e->S > 1 ? eproj_old[1] : 0.0;
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(kBlock), 0, st, reinterpret_cast<int64_t*>(e->d_stats),
                     static_cast<int64_t>(sizeof(Stats) / sizeof(int64_t)));
  hipLaunchKernelGGL(k_replay_walk, dim3(1), dim3(1), 0, st, e->sys, e->n, e->cap, e->det[c], e->pop[c], e->kii[c],
                     e->pop_new, tau, shift[0], sh1, eproj_old[0], e1, e->rng_buf, e->rng_pos, e->rng_end, e->scap,
                     e->sdet, e->srec, e->d_stats, determ_dev(e));
  int rc = sync_stats(e);
  if (rc != FQ_OK) return rc;
  if (e->h_stats.rng_failed) return FQ_ERR_RNG_EXHAUSTED;
  e->pending = true;
  e->pend_cyc = cyc;
  e->pend_rng_pos = e->h_stats.rng_pos;
  e->pend_ns = std::min<int64_t>(e->h_stats.nspawn, e->scap);
  e->pend_stats = e->h_stats;
  std::memset(out, 0, sizeof(*out));
  for (int sp = 0; sp < e->S; ++sp) {
    out->proj_energy[sp] = e->h_stats.proj[sp];
    out->D0_population[sp] = e->h_stats.d0[sp];
  }
  out->nspawn_events = e->pend_ns;
  out->ndeath = e->h_stats.ndeath;
  out->nattempts = e->h_stats.total_attempts;
  out->rng_consumed = e->pend_rng_pos - e->rng_pos;
  out->spawn_overflow = e->h_stats.spawn_overflow ? 1 : 0;
  return FQ_OK;
}

int fq_spawned_list(fq_engine* e, uint64_t* states, int64_t* pops, int64_t* flags, int64_t capacity, int64_t* n) {
  if (!e || !n || !e->pending) return FQ_ERR_ARG;
  const int64_t ns = e->pend_ns;
  *n = ns;
  if (capacity < ns) return FQ_ERR_CAPACITY;
  if (ns == 0) return FQ_OK;
  FQ_HIP(hipSetDevice(e->device));
  const int W = e->W, S = e->S, RS = rec_stride(S);
  std::vector<uint64_t> dw(static_cast<size_t>(W) * ns);
  std::vector<int64_t> rec(static_cast<size_t>(RS) * ns);
  for (int w = 0; w < W; ++w)
    FQ_HIP(hipMemcpyAsync(dw.data() + w * ns, e->sdet + w * e->scap, 8 * ns, hipMemcpyDeviceToHost, e->stream));
  FQ_HIP(hipMemcpyAsync(rec.data(), e->srec, 8 * RS * ns, hipMemcpyDeviceToHost, e->stream));
  FQ_HIP(hipStreamSynchronize(e->stream));
  for (int64_t i = 0; i < ns; ++i) {
    if (states)
      for (int w = 0; w < W; ++w) states[i * W + w] = dw[w * ns + i];
    if (pops)
      for (int sp = 0; sp < S; ++sp) pops[i * S + sp] = rec[i * RS + sp];
    if (flags) flags[i] = rec[i * RS + S];
  }
  return FQ_OK;
}

int fq_cycle_finish(fq_engine* e, const uint64_t* sstates, const int64_t* spops, const int64_t* sflags, int64_t ns,
                    fq_cycle_out* out) {
  if (!e || !out || !e->pending || ns < 0 || (ns > 0 && (!sstates || !spops))) return FQ_ERR_ARG;
  if (ns > e->scap) return FQ_ERR_CAPACITY;
  FQ_HIP(hipSetDevice(e->device));
  const int W = e->W, S = e->S, RS = rec_stride(S);
  std::vector<uint64_t> dw(static_cast<size_t>(W) * std::max<int64_t>(ns, 1));
  std::vector<int64_t> rec(static_cast<size_t>(RS) * std::max<int64_t>(ns, 1), 0);
  for (int64_t i = 0; i < ns; ++i) {
    for (int w = 0; w < W; ++w) dw[w * ns + i] = sstates[i * W + w];
    for (int sp = 0; sp < S; ++sp) rec[i * RS + sp] = spops[i * S + sp];
    rec[i * RS + S] = (sflags && e->par.initiator) ? sflags[i] : 0;
  }
  hipStream_t st = e->stream;
  for (int w = 0; w < W && ns > 0; ++w)
    FQ_HIP(hipMemcpyAsync(e->sdet + w * e->scap, dw.data() + w * ns, 8 * ns, hipMemcpyHostToDevice, st));
  if (ns > 0) FQ_HIP(hipMemcpyAsync(e->srec, rec.data(), 8 * RS * ns, hipMemcpyHostToDevice, st));
  e->pending = false;
  int64_t rng_pos = e->pend_rng_pos;
  int rc = annihilate(e, ns, nullptr, true, e->pend_cyc, rng_pos);
  if (rc != FQ_OK) return rc;
  const int64_t used = rng_pos - e->rng_pos;
  e->rng_pos = rng_pos;
  const Stats& k = e->pend_stats;
  for (int sp = 0; sp < S; ++sp) {
    e->h_stats.proj[sp] = k.proj[sp];
    e->h_stats.d0[sp] = k.d0[sp];
  }
  e->h_stats.nspawn = e->pend_ns;
  e->h_stats.ndeath = k.ndeath;
  e->h_stats.total_attempts = k.total_attempts;
  e->h_stats.spawn_overflow = k.spawn_overflow;
  fill_out(e, out, used);
  return FQ_OK;
}

int fq_cycle_noalloc(fq_engine* e, const double* shift, const double* eproj_old, double tau, fq_cycle_out* out) {
  if (!e || !shift || !eproj_old || !out || e->batched) return FQ_ERR_ARG;
  // replay streams and the learned proposal's weight table grow on demand
  if (e->par.rng_mode != FQ_RNG_PHILOX || e->par.excit_gen == FQ_EXCIT_LEARNED) return FQ_ERR_ARG;
  if (!e->fixed_scratch) return FQ_ERR_DEVICE;
  e->noalloc = true;
  e->scratch_short = false;
  const int rc = cycle_body(e, shift, eproj_old, tau, out);
  e->noalloc = false;
  if (e->scratch_short) {
    fprintf(stderr, "fqmc: fq_cycle_noalloc: this cycle needs more scratch than the fixed %lld bytes hold; "
            "the walker list is unchanged\n", static_cast<long long>(sizeof(int64_t) * kScratchElems));
    (void)hipStreamSynchronize(e->stream);
    return FQ_ERR_SCRATCH;
  }
  return rc;
}

int fq_annihilate(fq_engine* e, const uint64_t* sstates, const int64_t* spops, const int64_t* sflags, int64_t ns,
                  fq_cycle_out* out) {
  if (!e || !out || ns < 0 || (ns > 0 && (!sstates || !spops)) || e->pending) return FQ_ERR_ARG;
  if (ns > e->scap) return FQ_ERR_CAPACITY;
  FQ_HIP(hipSetDevice(e->device));
  const bool replay = e->par.rng_mode == FQ_RNG_REPLAY;
  const int W = e->W, S = e->S;
  const int64_t n = e->n, cap = e->cap;
  std::vector<uint64_t> dw(static_cast<size_t>(W) * std::max<int64_t>(ns, 1));
  const int RS = rec_stride(S);
  std::vector<int64_t> rec(static_cast<size_t>(RS) * std::max<int64_t>(ns, 1), 0);
  for (int64_t i = 0; i < ns; ++i) {
    for (int w = 0; w < W; ++w) dw[w * ns + i] = sstates[i * W + w];
    for (int s = 0; s < S; ++s) rec[i * RS + s] = spops[i * S + s];
    rec[i * RS + S] = (sflags && e->par.initiator) ? sflags[i] : 0;
  }
  hipStream_t st = e->stream;
  for (int w = 0; w < W && ns > 0; ++w)
    FQ_HIP(hipMemcpyAsync(e->sdet + w * e->scap, dw.data() + w * ns, 8 * ns, hipMemcpyHostToDevice, st));
  if (ns > 0) FQ_HIP(hipMemcpyAsync(e->srec, rec.data(), 8 * RS * ns, hipMemcpyHostToDevice, st));
  // post-death populations = current populations
  for (int s = 0; s < S && n > 0; ++s)
    FQ_HIP(hipMemcpyAsync(e->pop_new + s * cap, e->pop[e->cur] + s * cap, 8 * n, hipMemcpyDeviceToDevice, st));
  FQ_HIP(hipMemsetAsync(e->d_stats, 0, sizeof(Stats), st));
  FQ_HIP(hipStreamSynchronize(st));
  const uint32_t cyc = static_cast<uint32_t>(e->cycle_index++);
  int64_t rng_pos = e->rng_pos;
  int rc = annihilate(e, ns, nullptr, replay, cyc, rng_pos);
  if (rc != FQ_OK) return rc;
  const int64_t used = rng_pos - e->rng_pos;
  e->rng_pos = rng_pos;
  e->h_stats.nspawn = ns;
  fill_out(e, out, used);
  return FQ_OK;
}

int fq_nstates(fq_engine* e, int64_t* n) {
  if (!e || !n) return FQ_ERR_ARG;
  *n = e->n;
  return FQ_OK;
}

int fq_download(fq_engine* e, uint64_t* states, int64_t* pops, double* dat, int64_t capacity, int64_t* nout) {
  if (!e || !nout) return FQ_ERR_ARG;
  const int64_t n = e->n;
  *nout = n;
  if (capacity < n) return FQ_ERR_CAPACITY;
  if (n == 0) return FQ_OK;
  FQ_HIP(hipSetDevice(e->device));
  const int W = e->W, S = e->S;
  std::vector<uint64_t> dw(static_cast<size_t>(W) * n);
  std::vector<int64_t> dp(static_cast<size_t>(S) * n);
  const int c = e->cur;
  for (int w = 0; w < W; ++w)
    FQ_HIP(hipMemcpyAsync(dw.data() + w * n, e->det[c] + w * e->cap, 8 * n, hipMemcpyDeviceToHost, e->stream));
  for (int s = 0; s < S; ++s)
    FQ_HIP(hipMemcpyAsync(dp.data() + s * n, e->pop[c] + s * e->cap, 8 * n, hipMemcpyDeviceToHost, e->stream));
  if (dat) FQ_HIP(hipMemcpyAsync(dat, e->kii[c], 8 * n, hipMemcpyDeviceToHost, e->stream));
  FQ_HIP(hipStreamSynchronize(e->stream));
  for (int64_t i = 0; i < n; ++i) {
    if (states)
      for (int w = 0; w < W; ++w) states[i * W + w] = dw[w * n + i];
    if (pops)
      for (int s = 0; s < S; ++s) pops[i * S + s] = dp[s * n + i];
  }
  return FQ_OK;
}

int fq_last_profile(fq_engine* e, fq_profile_out* out) {
  if (!e || !out) return FQ_ERR_ARG;
  *out = e->prof;
  return FQ_OK;
}

int fq_last_timing(fq_engine* e, double* spawn_ms, double* annihilate_ms, double* total_ms) {
  if (!e) return FQ_ERR_ARG;
  if (spawn_ms) *spawn_ms = e->t_spawn;
  if (annihilate_ms) *annihilate_ms = e->t_annih;
  if (total_ms) *total_ms = e->t_total;
  return FQ_OK;
}

// ---- batched independent instances ---------------------------------------------
int fq_create_batch(const fq_system* systems, int32_t ninst, const fq_params* params, fq_engine** out) {
  if (!systems || !params || !out || ninst < 1 || ninst > kMaxInst) return FQ_ERR_ARG;
  if (params->rng_mode != FQ_RNG_PHILOX || params->real_amplitudes) return FQ_ERR_ARG;
  if (params->excit_gen == FQ_EXCIT_HEAT_BATH || params->excit_gen == FQ_EXCIT_LEARNED) return FQ_ERR_ARG;
  for (int k = 0; k < ninst; ++k)
    if (systems[k].nwords != 1 || systems[k].nbasis > kInstShift || systems[k].system != systems[0].system)
      return FQ_ERR_ARG;
  fq_engine* e = nullptr;
  int rc = fq_create(&systems[0], params, &e);
  if (rc != FQ_OK) return rc;
  e->batched = true;
  e->ninst = ninst;
  e->rank_words = 0;  // the instance index is part of the key: sort the encoded words
  int ib = 0;
  while ((1 << ib) < ninst) ++ib;
  e->key_bits = kInstShift + std::max(ib, 1);
  rc = setup_sort(e);
  if (rc != FQ_OK) {
    fq_destroy(e);
    return rc;
  }
  fq_params small = *params;
  small.main_capacity = small.spawn_capacity = 1;
  std::vector<Sys> hs(static_cast<size_t>(ninst));
  for (int k = 0; k < ninst && rc == FQ_OK; ++k) {
    fq_engine* sub = nullptr;
    rc = fq_create(&systems[k], &small, &sub);
    if (rc == FQ_OK) {
      e->subs.push_back(sub);
      hs[k] = sub->sys;
    }
  }
  const int64_t S = e->S;
  if (rc == FQ_OK) rc = upload(&e->d_sys_arr, hs.data(), ninst, e->stream);
  if (rc == FQ_OK) {
    rc |= dalloc(&e->d_ipar, (1 + 2 * S) * ninst);
    rc |= dalloc(&e->d_ibound, 2 * (ninst + 1));
    rc |= dalloc(&e->d_isb, ninst + 1);
    rc |= dalloc(&e->d_iatt, ninst);
    rc |= dalloc(&e->d_cterm, 2 * S * e->cap);
    rc |= dalloc(&e->d_cdeath, e->cap);
    e->ipb = e->cap / kBlock + ninst + 2;
    rc |= dalloc(&e->d_iblk_cnt, ninst + 1);
    rc |= dalloc(&e->d_iblk_off, ninst + 1);
    rc |= dalloc(&e->d_nblk_off, ninst + 1);
    rc |= dalloc(&e->d_ipart_d, 2 * S * e->ipb);
    rc |= dalloc(&e->d_ipart_i, (1 + S) * e->ipb);
    rc |= dalloc(&e->d_iout_d, 2 * S * ninst);
    rc |= dalloc(&e->d_iout_i, (4 + S) * ninst);
    if (rc != FQ_OK) rc = FQ_ERR_NOMEM;
  }
  if (rc == FQ_OK && hipStreamSynchronize(e->stream) != hipSuccess) rc = FQ_ERR_DEVICE;
  if (rc != FQ_OK) {
    fq_destroy(e);
    return rc;
  }
  *out = e;
  return FQ_OK;
}

int fq_load_batch(fq_engine* e, const int32_t* inst, const uint64_t* states, const int64_t* pops, const double* dat,
                  int64_t n) {
  if (!e || !e->batched || n < 0 || (n > 0 && (!inst || !states || !pops || !dat))) return FQ_ERR_ARG;
  const int S = e->S;
  std::vector<int64_t> order(static_cast<size_t>(n));
  std::vector<uint64_t> key(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    if (inst[i] < 0 || inst[i] >= e->ninst || (states[i] >> kInstShift) != 0) return FQ_ERR_ARG;
    key[i] = states[i] | (static_cast<uint64_t>(inst[i]) << kInstShift);
    order[i] = i;
  }
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
  for (int64_t i = 1; i < n; ++i)
    if (key[order[i]] == key[order[i - 1]]) return FQ_ERR_ARG;
  std::vector<uint64_t> st(static_cast<size_t>(n));
  std::vector<int64_t> po(static_cast<size_t>(S) * n);
  std::vector<double> da(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    st[i] = key[order[i]];
    for (int q = 0; q < S; ++q) po[i * S + q] = pops[order[i] * S + q];
    da[i] = dat[order[i]];
  }
  return fq_load(e, st.data(), po.data(), da.data(), n);  // the encoded, sorted list
}

// Per-instance estimators and counts of a batched cycle (after annihilation;
// old-list bounds ob, new-list bounds nbd, nold and nnew bounds of the two
// lengths, pnew the new list's populations).
static int inst_stats(fq_engine* e, const int64_t* ob, const int64_t* nbd, int64_t nold, int64_t nnew,
                      const int64_t* pnew) {
  const int ninst = e->ninst, S = e->S;
  hipStream_t st = e->stream;
  hipLaunchKernelGGL(k_inst_nblk, dim3(grid_for(ninst)), dim3(kBlock), 0, st, ninst, ob, e->d_iblk_cnt);
  int rc = scan_exclusive(e, e->d_iblk_cnt, e->d_iblk_off, ninst);
  if (rc != FQ_OK) return rc;
  hipLaunchKernelGGL(k_inst_nblk, dim3(grid_for(ninst)), dim3(kBlock), 0, st, ninst, nbd, e->d_iblk_cnt);
  rc = scan_exclusive(e, e->d_iblk_cnt, e->d_nblk_off, ninst);
  if (rc != FQ_OK) return rc;
  // blocks of either list: at most (length / 256) + ninst
  const int64_t nblk = std::min<int64_t>(std::max<int64_t>(nold, nnew) / kBlock + ninst + 1, e->ipb);
  hipLaunchKernelGGL(k_inst_part, dim3(static_cast<unsigned>(nblk)), dim3(kBlock), 0, st, ninst, S, e->cap, ob,
                     e->d_iblk_off, e->d_cterm, e->d_cdeath, nbd, e->d_nblk_off, pnew, e->ipb,
                     e->d_ipart_d, e->d_ipart_i);
  hipLaunchKernelGGL(k_inst_stats, dim3(ninst), dim3(kBlock), 0, st, ninst, S, ob, nbd, e->d_isb, e->d_iblk_off,
                     e->d_nblk_off, e->ipb, e->d_ipart_d, e->d_ipart_i, e->att_off, e->d_iout_d, e->d_iout_i);
  return FQ_OK;
}

// Per-instance results of a batched cycle from the k_inst_stats outputs.
static void batch_outputs(const fq_engine* e, bool spawn_ovf, bool main_ovf, const double* od, const int64_t* oi,
                          fq_cycle_out* out) {
  const int S = e->S;
  const double rf = static_cast<double>(e->sys.rf);
  for (int k = 0; k < e->ninst; ++k) {
    fq_cycle_out& r = out[k];
    std::memset(&r, 0, sizeof(r));
    const int64_t* q = oi + k * (4 + S);
    for (int sp = 0; sp < S; ++sp) {
      r.proj_energy[sp] = od[k * 2 * S + 2 * sp];
      r.D0_population[sp] = od[k * 2 * S + 2 * sp + 1];
      r.nparticles[sp] = static_cast<double>(q[4 + sp]) / rf;
    }
    r.ndeath = q[0];
    r.nattempts = q[1];
    r.nspawn_events = q[2];
    r.nstates = q[3];
    r.spawn_overflow = spawn_ovf ? 1 : 0;
    r.main_overflow = main_ovf ? 1 : 0;
  }
}

// ---- graph-captured batched cycles ---------------------------------------------
// Launch bound of a length: the next multiple of a quarter of its octave (at
// most 25% past it), so one graph serves while the length stays in its step.
static int64_t len_step(int64_t x) {
  if (x <= 4096) return 4096;
  const int e = 63 - __builtin_clzll(static_cast<unsigned long long>(x - 1));  // 2^e <= x - 1 < 2^(e+1)
  const int64_t step = int64_t{1} << (e - 2);
  return ceil_div(x, step) * step;
}
// Candidate schedules: (8-bit digits only, spawning grid cap); index 0 is the
// one ordinary cycles use.
constexpr int kNumSched = 6;
constexpr int kSchedDigit8[kNumSched] = {0, 0, 0, 1, 1, 1};
constexpr int kSchedSpawnBlocks[kNumSched] = {0, 2048, 1024, 0, 2048, 1024};
constexpr int kTuneReps = 3;  // timed graph cycles per candidate

static size_t graph_in_bytes(const fq_engine* e) { return 16 + sizeof(double) * (1 + 2 * e->S) * e->ninst; }

// The whole batched cycle for list buffer c, launches sized by nb (main list)
// and tb (attempts), lengths and the cycle index read on the device, inputs
// copied from h_in first and outputs to the pinned buffers last.  Called while
// the stream is captured.
static int enqueue_graph_cycle(fq_engine* e, int64_t nb, int64_t tb) {
  const int ninst = e->ninst, S = e->S;
  const int64_t cap = e->cap;
  const int c = e->cur;
  hipStream_t st = e->stream;
  const int64_t* dn = reinterpret_cast<const int64_t*>(e->d_in);
  const uint32_t* dcyc = reinterpret_cast<const uint32_t*>(e->d_in + 8);
  const double* par = reinterpret_cast<const double*>(e->d_in + 16);
  FQ_HIP(hipMemcpyAsync(e->d_in, e->h_in, graph_in_bytes(e), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(kBlock), 0, st, reinterpret_cast<int64_t*>(e->d_stats),
                     static_cast<int64_t>(sizeof(Stats) / sizeof(int64_t)));
  int64_t* ob = e->d_ibound;
  int64_t* nbd = e->d_ibound + (ninst + 1);
  hipLaunchKernelGGL(k_inst_bounds, dim3(grid_for(ninst + 1)), dim3(kBlock), 0, st, e->det[c], nb, ninst, kInstShift,
                     ob, dn);
  hipLaunchKernelGGL(k_det_pass_batch, dim3(grid_for(nb)), dim3(kBlock), 0, st, e->sys, e->d_sys_arr, ob, par, ninst,
                     nb, cap, e->det[c], e->pop[c], e->kii[c], e->pop_new, e->nat, e->flags, e->par.philox_seed, 0u,
                     e->d_cterm, e->d_cdeath, dn, dcyc);
  // entries past the list have no attempts, so att_off[S nb] is the total
  const int64_t nent = S * nb;
  int rc = scan_exclusive(e, e->nat, e->att_off, nent);
  if (rc != FQ_OK) return rc;
  hipLaunchKernelGGL(k_inst_att, dim3(grid_for(ninst)), dim3(kBlock), 0, st, ninst, S, ob, e->att_off, e->d_iatt);
  const unsigned g = static_cast<unsigned>(std::min<int64_t>(
      ceil_div(tb, kSpawnTile), e->spawn_blocks > 0 ? e->spawn_blocks : int64_t{1} << 20));
  hipLaunchKernelGGL(k_tile_bounds, dim3(grid_for(ceil_div(tb, kSpawnTile))), dim3(kBlock), 0, st, nent, e->att_off,
                     int64_t{0}, tb, e->tile_win, e->tile_cnt);
  auto kern = e->sys.system == FQ_SYS_HUB_K ? k_spawn<fqd::kGenHubK, true> : k_spawn<fqd::kGenMolRenorm, true>;
  hipLaunchKernelGGL(kern, dim3(g), dim3(kBlock), 0, st, e->d_sysp, nent, cap, int64_t{0}, tb, e->att_off, e->det[c],
                     e->pop[c], 0.0, e->par.philox_seed, 0u, e->bcap, e->st_det, e->st_pop, e->st_ent, e->tile_cnt,
                     e->tile_win, e->flags, e->d_sys_arr, e->d_iatt, par, dcyc);
  rc = pack_spawned(e, tb, 0, e->att_off + nent, tb);
  if (rc != FQ_OK) return rc;
  e->keys_ready = true;
  const int64_t nsb = std::min<int64_t>(tb, e->scap);
  const int64_t* pns = &e->d_stats->nspawn;
  hipLaunchKernelGGL(k_inst_bounds, dim3(grid_for(ninst + 1)), dim3(kBlock), 0, st, e->kbuf[0], nsb, ninst,
                     kInstShift + e->flag_bits, e->d_isb, pns);
  int64_t rng_pos = e->rng_pos;
  rc = annihilate(e, nsb, pns, false, 0u, rng_pos);
  if (rc != FQ_OK) return rc;
  const int o = 1 - c;
  hipLaunchKernelGGL(k_inst_bounds, dim3(grid_for(ninst + 1)), dim3(kBlock), 0, st, e->det[o],
                     std::min<int64_t>(nb + nsb, cap), ninst, kInstShift, nbd, &e->d_stats->new_n);
  rc = inst_stats(e, ob, nbd, nb, std::min<int64_t>(nb + nsb, cap), e->pop[o]);
  if (rc != FQ_OK) return rc;
  FQ_HIP(hipMemcpyAsync(e->h_gstats, e->d_stats, sizeof(Stats), hipMemcpyDeviceToHost, st));
  FQ_HIP(hipMemcpyAsync(e->h_god, e->d_iout_d, sizeof(double) * 2 * S * ninst, hipMemcpyDeviceToHost, st));
  FQ_HIP(hipMemcpyAsync(e->h_goi, e->d_iout_i, sizeof(int64_t) * (4 + S) * ninst, hipMemcpyDeviceToHost, st));
  return FQ_OK;
}

// The graph of (list buffer, bounds, schedule), captured on first use.
static int graph_for(fq_engine* e, int64_t nb, int64_t tb, int sched, hipGraphExec_t* out) {
  for (const auto& g : e->graphs)
    if (g.c == e->cur && g.nb == nb && g.tb == tb && g.sched == sched) {
      *out = g.exec;
      return FQ_OK;
    }
  // no allocation may happen inside the capture: size the scan scratch for
  // the longest scan of the cycle first
  const int64_t longest = std::max<int64_t>(e->S * nb, tb);
  if (longest > kSmallScanMax) {
    const int64_t ntile = ceil_div(longest, kScanTile);
    if (!scan_scratch(e, 0, ntile) || !scan_scratch(e, 1, ntile + 1)) return FQ_ERR_NOMEM;
  }
  e->digit8 = kSchedDigit8[sched];
  e->spawn_blocks = kSchedSpawnBlocks[sched];
  const int64_t n_keep = e->n;
  e->n = nb;  // annihilation sizes its launches by e->n
  e->dn = reinterpret_cast<const int64_t*>(e->d_in);
  e->dcyc = reinterpret_cast<const uint32_t*>(e->d_in + 8);
  e->capturing = true;
  hipGraph_t graph = nullptr;
  int rc = FQ_OK;
  if (hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
    rc = FQ_ERR_DEVICE;
  } else {
    rc = enqueue_graph_cycle(e, nb, tb);
    if (hipStreamEndCapture(e->stream, &graph) != hipSuccess && rc == FQ_OK) rc = FQ_ERR_DEVICE;
  }
  e->capturing = false;
  e->dn = nullptr;
  e->dcyc = nullptr;
  e->n = n_keep;
  e->digit8 = 0;
  e->spawn_blocks = 0;
  hipGraphExec_t exec = nullptr;
  if (rc == FQ_OK && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) rc = FQ_ERR_DEVICE;
  if (graph) (void)hipGraphDestroy(graph);
  if (rc != FQ_OK) return rc;
  e->graphs.push_back({e->cur, nb, tb, sched, exec});
  *out = exec;
  return FQ_OK;
}

static float median_of(std::vector<float> v) {
  std::sort(v.begin(), v.end());
  return v.empty() ? 0.0f : v[v.size() / 2];
}

// One batched cycle as a graph launch; *done = false when the cycle must run
// as ordinary launches (attempt bound unknown or past the staging capacity).
static int graph_cycle(fq_engine* e, const double* shift, const double* eproj_old, const double* tau,
                       fq_cycle_out* out, bool* done) {
  *done = false;
  const int ninst = e->ninst, S = e->S;
  if (e->n <= 0 || e->att_bound <= 0) return FQ_OK;
  const int64_t nb = std::min<int64_t>(len_step(e->n), e->cap);
  const int64_t tb = std::min<int64_t>(len_step(e->att_bound), e->bcap);
  if (e->att_bound > tb) return FQ_OK;
  if (int rs = sync_sysp(e); rs != FQ_OK) return rs;
  // the tuner works per octave of the attempt bound (the schedules' relative
  // cost changes with the workload's size): candidates in turn until each has
  // run kTuneReps times, then the fastest by median
  const int oct = 63 - __builtin_clzll(static_cast<unsigned long long>(tb));
  e->last_oct = oct;
  int sched = e->sched[oct];
  if (sched < 0) sched = e->tune_iter[oct] % kNumSched;
  hipGraphExec_t exec = nullptr;
  int rc = graph_for(e, nb, tb, sched, &exec);
  if (rc != FQ_OK) return rc;
  // inputs: the list length, the cycle index, then tau, shift, E_proj per instance
  const uint32_t cyc = static_cast<uint32_t>(e->cycle_index++);
  int64_t* hi = reinterpret_cast<int64_t*>(e->h_in);
  hi[0] = e->n;
  hi[1] = static_cast<int64_t>(cyc);
  double* par = reinterpret_cast<double*>(e->h_in + 16);
  for (int k = 0; k < ninst; ++k) {
    par[k] = tau[k];
    for (int sp = 0; sp < S; ++sp) {
      par[ninst + sp * ninst + k] = shift[k * S + sp];
      par[(1 + S) * ninst + sp * ninst + k] = eproj_old[k * S + sp];
    }
  }
  FQ_HIP(hipEventRecord(e->gev[0], e->stream));
  FQ_HIP(hipGraphLaunch(exec, e->stream));
  FQ_HIP(hipEventRecord(e->gev[1], e->stream));
  FQ_HIP(hipStreamSynchronize(e->stream));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e->gev[0], e->gev[1]);
  ++e->graph_cycles;
  if (e->sched[oct] < 0) {
    e->tune_ms[oct][sched].push_back(ms);
    if (++e->tune_iter[oct] >= kNumSched * kTuneReps) {
      int best = 0;
      for (int k = 1; k < kNumSched; ++k)
        if (median_of(e->tune_ms[oct][k]) < median_of(e->tune_ms[oct][best])) best = k;
      e->sched[oct] = best;
    }
  }
  e->h_stats = *e->h_gstats;
  if (e->h_stats.spawn_overflow && e->h_stats.total_attempts > tb) return FQ_ERR_DEVICE;  // the bound was wrong
  e->cur = 1 - e->cur;
  e->n = e->h_stats.new_n;
  unsigned long long sum = 0;
  for (int sp = 0; sp < S; ++sp) sum += e->h_stats.abs_pop[sp];
  e->att_bound = static_cast<int64_t>(sum / static_cast<unsigned long long>(e->sys.rf));
  e->t_total = ms;
  e->prof = fq_profile_out{};
  e->prof.total_ms = ms;
  e->prof.nattempts = e->h_stats.total_attempts;
  e->prof.nspawned = std::min<int64_t>(e->h_stats.nspawn, e->scap);
  e->prof.nruns = e->h_stats.nruns;
  e->prof.nstates_out = e->n;
  batch_outputs(e, e->h_stats.spawn_overflow != 0, e->h_stats.main_overflow != 0, e->h_god, e->h_goi, out);
  *done = true;
  return FQ_OK;
}

int fq_batch_graph(fq_engine* e, int32_t enable) {
  if (!e || !e->batched) return FQ_ERR_ARG;
  FQ_HIP(hipSetDevice(e->device));
  if (enable && !e->h_in) {
    const size_t nin = graph_in_bytes(e);
    if (hipHostMalloc(reinterpret_cast<void**>(&e->h_in), nin) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&e->h_gstats), sizeof(Stats)) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&e->h_god), sizeof(double) * 2 * e->S * e->ninst) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&e->h_goi), sizeof(int64_t) * (4 + e->S) * e->ninst) != hipSuccess)
      return FQ_ERR_NOMEM;
    if (dalloc(&e->d_in, static_cast<int64_t>(nin)) != FQ_OK) return FQ_ERR_NOMEM;
    for (int k = 0; k < 2; ++k)
      if (hipEventCreate(&e->gev[k]) != hipSuccess) return FQ_ERR_DEVICE;
    for (int o = 0; o < 64; ++o) {
      e->sched[o] = -1;
      e->tune_iter[o] = 0;
    }
  }
  e->graph_on = enable != 0;
  return FQ_OK;
}

int fq_batch_schedule(fq_engine* e, fq_schedule_out* out) {
  if (!e || !e->batched || !out) return FQ_ERR_ARG;
  std::memset(out, 0, sizeof(*out));
  const int oct = e->last_oct;  // the last graph cycle's octave
  const int ch = oct >= 0 ? e->sched[oct] : -1;
  const int k = ch < 0 ? 0 : ch;
  out->graph = e->graph_on ? 1 : 0;
  out->tuned = ch >= 0 ? 1 : 0;
  out->chosen = ch;
  out->octave = oct;
  for (int o = 0; o < 64; ++o) out->tuned_octaves += e->h_in && e->sched[o] >= 0 ? 1 : 0;
  out->digit_bits = kSchedDigit8[k] ? 8 : 9;
  out->spawn_blocks = kSchedSpawnBlocks[k];
  out->ncand = kNumSched;
  out->graph_cycles = e->graph_cycles;
  out->graphs = static_cast<int64_t>(e->graphs.size());
  out->eager_cycles = e->eager_cycles;
  for (int q = 0; q < kNumSched; ++q) {
    out->cand_ms[q] = oct >= 0 ? median_of(e->tune_ms[oct][q]) : 0.0;
    out->cand_digit_bits[q] = kSchedDigit8[q] ? 8 : 9;
    out->cand_spawn_blocks[q] = kSchedSpawnBlocks[q];
  }
  return FQ_OK;
}

int fq_cycle_batch(fq_engine* e, const double* shift, const double* eproj_old, const double* tau,
                   fq_cycle_out* out) {
  if (!e || !e->batched || !shift || !eproj_old || !tau || !out) return FQ_ERR_ARG;
  FQ_HIP(hipSetDevice(e->device));
  if (e->graph_on) {
    bool done = false;
    const int grc = graph_cycle(e, shift, eproj_old, tau, out, &done);
    if (grc != FQ_OK || done) return grc;
  }
  ++e->eager_cycles;
  const int ninst = e->ninst, S = e->S;
  const int64_t n = e->n, cap = e->cap;
  const int c = e->cur;
  hipStream_t st = e->stream;
  const uint32_t cyc = static_cast<uint32_t>(e->cycle_index++);
  std::vector<double> par(static_cast<size_t>((1 + 2 * S) * ninst));
  for (int k = 0; k < ninst; ++k) {
    par[k] = tau[k];
    for (int sp = 0; sp < S; ++sp) {
      par[ninst + sp * ninst + k] = shift[k * S + sp];
      par[(1 + S) * ninst + sp * ninst + k] = eproj_old[k * S + sp];
    }
  }
  FQ_HIP(hipMemcpyAsync(e->d_ipar, par.data(), sizeof(double) * par.size(), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(kBlock), 0, st, reinterpret_cast<int64_t*>(e->d_stats),
                     static_cast<int64_t>(sizeof(Stats) / sizeof(int64_t)));
  FQ_HIP(hipEventRecord(e->ev[0], st));
  FQ_HIP(hipEventRecord(e->pev[0], st));
  int64_t* ob = e->d_ibound;
  int64_t* nbd = e->d_ibound + (ninst + 1);
  hipLaunchKernelGGL(k_inst_bounds, dim3(grid_for(ninst + 1)), dim3(kBlock), 0, st, e->det[c], n, ninst, kInstShift, ob,
                     nullptr);
  if (n > 0)
    hipLaunchKernelGGL(k_det_pass_batch, dim3(grid_for(n)), dim3(kBlock), 0, st, e->sys, e->d_sys_arr, ob, e->d_ipar,
                       ninst, n, cap, e->det[c], e->pop[c], e->kii[c], e->pop_new, e->nat, e->flags,
                       e->par.philox_seed, cyc, e->d_cterm, e->d_cdeath, nullptr, nullptr);
  FQ_HIP(hipEventRecord(e->pev[1], st));
  const int64_t nent = S * n;
  int rc = scan_exclusive(e, e->nat, e->att_off, nent);
  if (rc != FQ_OK) return rc;
  hipLaunchKernelGGL(k_inst_att, dim3(grid_for(ninst)), dim3(kBlock), 0, st, ninst, S, ob, e->att_off, e->d_iatt);
  int64_t total = 0;
  FQ_HIP(hipMemcpyAsync(&total, e->att_off + nent, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  FQ_HIP(hipStreamSynchronize(st));
  FQ_HIP(hipEventRecord(e->pev[2], st));
  e->spawn_launches = 0;
  if (int rs = sync_sysp(e); rs != FQ_OK) return rs;
  for (int64_t tb = 0; tb < total; tb += e->bcap) {
    const int64_t te = std::min(total, tb + e->bcap);
    const int64_t nbt = te - tb;
    const unsigned g = static_cast<unsigned>(std::min<int64_t>(ceil_div(nbt, kSpawnTile), 1 << 20));
    const int li = e->spawn_launches++;
    hipLaunchKernelGGL(k_tile_bounds, dim3(grid_for(ceil_div(nbt, kSpawnTile))), dim3(kBlock), 0, st, nent,
                       e->att_off, tb, te, e->tile_win, e->tile_cnt);
    if (li < fq_engine::kSpawnEvents) FQ_HIP(hipEventRecord(e->sev[2 * li], st));
    auto kern = e->sys.system == FQ_SYS_HUB_K ? k_spawn<fqd::kGenHubK, true> : k_spawn<fqd::kGenMolRenorm, true>;
    hipLaunchKernelGGL(kern, dim3(g), dim3(kBlock), 0, st, e->d_sysp, nent, cap, tb, te, e->att_off, e->det[c],
                       e->pop[c], 0.0, e->par.philox_seed, cyc, e->bcap, e->st_det, e->st_pop, e->st_ent,
                       e->tile_cnt, e->tile_win, e->flags, e->d_sys_arr, e->d_iatt, e->d_ipar, nullptr);
    if (li < fq_engine::kSpawnEvents) FQ_HIP(hipEventRecord(e->sev[2 * li + 1], st));
    rc = pack_spawned(e, nbt, tb, e->att_off + nent);
    if (rc != FQ_OK) return rc;
  }
  FQ_HIP(hipEventRecord(e->pev[3], st));
  e->keys_ready = true;  // the packing kernels wrote the sort keys
  rc = sync_stats(e);
  if (rc != FQ_OK) return rc;
  const int64_t ns = std::min<int64_t>(e->h_stats.nspawn, e->scap);
  // spawned entries are in attempt order, hence grouped by instance
  hipLaunchKernelGGL(k_inst_bounds, dim3(grid_for(ninst + 1)), dim3(kBlock), 0, st, e->kbuf[0], ns, ninst,
                     kInstShift + e->flag_bits, e->d_isb, nullptr);
  const Stats keep = e->h_stats;
  int64_t rng_pos = e->rng_pos;
  rc = annihilate(e, ns, nullptr, false, cyc, rng_pos);
  if (rc != FQ_OK) return rc;
  const int o = e->cur;
  hipLaunchKernelGGL(k_inst_bounds, dim3(grid_for(ninst + 1)), dim3(kBlock), 0, st, e->det[o], e->n, ninst, kInstShift,
                     nbd, nullptr);
  rc = inst_stats(e, ob, nbd, n, e->n, e->pop[o]);
  if (rc != FQ_OK) return rc;
  std::vector<double> od(static_cast<size_t>(2 * S * ninst));
  std::vector<int64_t> oi(static_cast<size_t>((4 + S) * ninst));
  FQ_HIP(hipMemcpyAsync(od.data(), e->d_iout_d, sizeof(double) * od.size(), hipMemcpyDeviceToHost, st));
  FQ_HIP(hipMemcpyAsync(oi.data(), e->d_iout_i, sizeof(int64_t) * oi.size(), hipMemcpyDeviceToHost, st));
  FQ_HIP(hipEventRecord(e->ev[2], st));
  FQ_HIP(hipStreamSynchronize(st));
  (void)hipEventElapsedTime(&e->t_total, e->ev[0], e->ev[2]);
  {
    fq_profile_out& pr = e->prof;
    float t[6] = {0, 0, 0, 0, 0, 0};
    (void)hipEventElapsedTime(&t[0], e->pev[0], e->pev[1]);
    (void)hipEventElapsedTime(&t[1], e->pev[2], e->pev[3]);
    (void)hipEventElapsedTime(&t[2], e->pev[4], e->pev[5]);
    (void)hipEventElapsedTime(&t[3], e->pev[5], e->pev[6]);
    (void)hipEventElapsedTime(&t[4], e->pev[7], e->pev[8]);
    (void)hipEventElapsedTime(&t[5], e->pev[0], e->pev[8]);
    pr.det_pass_ms = t[0];
    pr.spawn_ms = t[1];
    pr.sort_ms = t[2];
    pr.reduce_ms = t[3];
    pr.merge_ms = t[4];
    pr.total_ms = t[5];
    pr.nstates_in = n;
    pr.nattempts = total;
    pr.nspawned = ns;
    pr.nruns = e->h_stats.nruns;
    pr.nstates_out = e->n;
    pr.radix_passes = e->radix_passes;
    pr.spawn_launches = std::min(e->spawn_launches, fq_engine::kSpawnEvents);
    pr.spawn_kernel_ms = 0.0;
    for (int li = 0; li < pr.spawn_launches; ++li) {
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e->sev[2 * li], e->sev[2 * li + 1]);
      pr.spawn_kernel_ms += ms;
    }
  }
  batch_outputs(e, keep.spawn_overflow != 0, e->h_stats.main_overflow != 0, od.data(), oi.data(), out);
  // the next cycle's attempt total: sum |N| over the new list (integer populations)
  unsigned long long sum = 0;
  for (int sp = 0; sp < S; ++sp) sum += e->h_stats.abs_pop[sp];
  e->att_bound = static_cast<int64_t>(sum / static_cast<unsigned long long>(e->sys.rf));
  return FQ_OK;
}

// ---- learned proposal ---------------------------------------------------------
int fq_mlp_nparams(fq_engine* e, int64_t* n) {
  if (!e || !n) return FQ_ERR_ARG;
  if (e->par.excit_gen != FQ_EXCIT_LEARNED) return FQ_ERR_ARG;
  *n = e->mlp.nparams;
  return FQ_OK;
}

int fq_mlp_set_params(fq_engine* e, const float* p, int64_t n) {
  if (!e || !p || e->par.excit_gen != FQ_EXCIT_LEARNED || n != e->mlp.nparams) return FQ_ERR_ARG;
  FQ_HIP(hipSetDevice(e->device));
  return fqm::mlp_set(e->mlp, p, e->stream);
}

int fq_mlp_get_params(fq_engine* e, float* p, int64_t n) {
  if (!e || !p || e->par.excit_gen != FQ_EXCIT_LEARNED || n != e->mlp.nparams) return FQ_ERR_ARG;
  FQ_HIP(hipSetDevice(e->device));
  FQ_HIP(hipMemcpyAsync(p, e->mlp.p, sizeof(float) * n, hipMemcpyDeviceToHost, e->stream));
  FQ_HIP(hipStreamSynchronize(e->stream));
  return FQ_OK;
}

int fq_mlp_affinity(fq_engine* e, float* out, int64_t n) {
  if (!e || !out || e->par.excit_gen != FQ_EXCIT_LEARNED) return FQ_ERR_ARG;
  if (n != static_cast<int64_t>(e->mlp.nb) * e->mlp.nb) return FQ_ERR_ARG;
  FQ_HIP(hipSetDevice(e->device));
  FQ_HIP(hipMemcpyAsync(out, e->mlp.aff, sizeof(float) * n, hipMemcpyDeviceToHost, e->stream));
  FQ_HIP(hipStreamSynchronize(e->stream));
  return FQ_OK;
}

int fq_mlp_forward(fq_engine* e, const uint64_t* dets, int64_t n, float* out) {
  if (!e || !dets || !out || n < 1 || e->par.excit_gen != FQ_EXCIT_LEARNED) return FQ_ERR_ARG;
  FQ_HIP(hipSetDevice(e->device));
  const int W = e->W;
  std::vector<uint64_t> soa(static_cast<size_t>(W) * n);
  for (int64_t i = 0; i < n; ++i)
    for (int w = 0; w < W; ++w) soa[w * n + i] = dets[i * W + w];
  uint64_t* d_det = nullptr;
  float* d_out = nullptr;
  int rc = dalloc(&d_det, W * n);
  rc |= dalloc(&d_out, n * e->mlp.d_out);
  if (rc == FQ_OK) {
    hipStream_t st = e->stream;
    if (hipMemcpyAsync(d_det, soa.data(), sizeof(uint64_t) * W * n, hipMemcpyHostToDevice, st) != hipSuccess)
      rc = FQ_ERR_DEVICE;
    if (rc == FQ_OK) rc = fqm::mlp_infer(e->mlp, e->sys, d_det, n, n, d_out, st);
    if (rc == FQ_OK &&
        hipMemcpyAsync(out, d_out, sizeof(float) * n * e->mlp.d_out, hipMemcpyDeviceToHost, st) != hipSuccess)
      rc = FQ_ERR_DEVICE;
    if (hipStreamSynchronize(st) != hipSuccess) rc = FQ_ERR_DEVICE;
  } else {
    rc = FQ_ERR_NOMEM;
  }
  dfree(d_det);
  dfree(d_out);
  return rc;
}

static int loss_grad(fq_engine* e, const uint64_t* dets, const int32_t* exc, const double* w, const double* habs,
                     const int32_t* flat, int64_t n, double* loss, float* grad) {
  if (!e || !dets || !exc || !w || !loss || n < 1 || e->par.excit_gen != FQ_EXCIT_LEARNED) return FQ_ERR_ARG;
  fqm::Mlp& m = e->mlp;
  if (n > m.bcap) return FQ_ERR_CAPACITY;
  const int W = e->W;
  // every weighted sample: a double i<j (occupied) -> a<b (unoccupied), orbitals in 1..nbasis
  for (int64_t k = 0; k < n; ++k) {
    if (!(w[k] > 0.0)) continue;
    const int32_t* x = exc + 4 * k;
    auto occ = [&](int o) { return (dets[k * W + (o - 1) / 64] >> ((o - 1) % 64)) & 1; };
    for (int q = 0; q < 4; ++q)
      if (x[q] < 1 || x[q] > e->sys.nbasis) return FQ_ERR_ARG;
    if (!(x[0] < x[1] && x[2] < x[3]) || !occ(x[0]) || !occ(x[1]) || occ(x[2]) || occ(x[3])) return FQ_ERR_ARG;
  }
  FQ_HIP(hipSetDevice(e->device));
  hipStream_t st = e->stream;
  std::vector<uint64_t> soa(static_cast<size_t>(W) * n);
  for (int64_t i = 0; i < n; ++i)
    for (int q = 0; q < W; ++q) soa[q * n + i] = dets[i * W + q];
  for (int q = 0; q < W; ++q)
    FQ_HIP(hipMemcpyAsync(m.sdet + q * m.bcap, soa.data() + q * n, sizeof(uint64_t) * n, hipMemcpyHostToDevice, st));
  FQ_HIP(hipMemcpyAsync(m.sexc, exc, sizeof(int32_t) * 4 * n, hipMemcpyHostToDevice, st));
  FQ_HIP(hipMemcpyAsync(m.sw, w, sizeof(double) * n, hipMemcpyHostToDevice, st));
  if (habs) FQ_HIP(hipMemcpyAsync(m.sh, habs, sizeof(double) * n, hipMemcpyHostToDevice, st));
  const std::vector<int32_t> fl = flat ? std::vector<int32_t>(flat, flat + n) : std::vector<int32_t>(n, 0);
  FQ_HIP(hipMemcpyAsync(m.sflat, fl.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
  const int keep = m.objective;
  m.objective = habs ? 1 : 0;
  int rc = fqm::mlp_train_step(m, e->sys, n, 0.0, false, st);
  m.objective = keep;
  if (rc != FQ_OK) return rc;
  double sums[2] = {0.0, 0.0};
  FQ_HIP(hipMemcpyAsync(sums, m.dsum, sizeof(sums), hipMemcpyDeviceToHost, st));
  if (grad) FQ_HIP(hipMemcpyAsync(grad, m.g, sizeof(float) * m.nparams, hipMemcpyDeviceToHost, st));
  FQ_HIP(hipStreamSynchronize(st));
  *loss = sums[1];
  return FQ_OK;
}

int fq_mlp_loss_grad(fq_engine* e, const uint64_t* dets, const int32_t* exc, const double* w, const int32_t* flat,
                     int64_t n, double* loss, float* grad) {
  return loss_grad(e, dets, exc, w, nullptr, flat, n, loss, grad);
}

int fq_mlp_variance_grad(fq_engine* e, const uint64_t* dets, const int32_t* exc, const double* w,
                         const double* habs, const int32_t* flat, int64_t n, double* loss, float* grad) {
  if (!habs) return FQ_ERR_ARG;
  return loss_grad(e, dets, exc, w, habs, flat, n, loss, grad);
}

int fq_mlp_online(fq_engine* e, int32_t every, int32_t steps, int64_t batch, double lr, int32_t objective) {
  if (!e || e->par.excit_gen != FQ_EXCIT_LEARNED || every < 0 || steps < 0 || (objective != 0 && objective != 1))
    return FQ_ERR_ARG;
  if (every > 0 && (batch < 1 || !(lr > 0.0))) return FQ_ERR_ARG;
  if (batch > e->mlp.bcap) return FQ_ERR_CAPACITY;
  e->online_every = every;
  e->online_steps = steps;
  e->online_batch = batch;
  e->online_lr = lr;
  e->mlp.objective = objective;
  return FQ_OK;
}

int fq_mlp_train(fq_engine* e, int64_t batch, int32_t steps, double lr, uint64_t seed, double* losses) {
  if (!e || batch < 1 || steps < 0 || !(lr > 0.0) || e->par.excit_gen != FQ_EXCIT_LEARNED) return FQ_ERR_ARG;
  fqm::Mlp& m = e->mlp;
  if (batch > m.bcap) return FQ_ERR_CAPACITY;
  if (e->n < 1) return FQ_ERR_ARG;
  FQ_HIP(hipSetDevice(e->device));
  hipStream_t st = e->stream;
  const int c = e->cur;
  for (int32_t k = 0; k < steps; ++k) {
    int rc = fqm::mlp_sample(m, e->sys, e->det[c], e->pop[c], e->cap, e->n, batch,
                             static_cast<int64_t>(e->par.proposal_min_attempts), seed,
                             static_cast<uint32_t>(m.adam_t), st, scan_cb, e);
    if (rc != FQ_OK) return rc;
    rc = fqm::mlp_train_step(m, e->sys, batch, lr, true, st);
    if (rc != FQ_OK) return rc;
    if (losses) {
      double sums[2] = {0.0, 0.0};
      FQ_HIP(hipMemcpyAsync(sums, m.dsum, sizeof(sums), hipMemcpyDeviceToHost, st));
      FQ_HIP(hipStreamSynchronize(st));
      losses[k] = sums[1];
    }
  }
  FQ_HIP(hipStreamSynchronize(st));
  return FQ_OK;
}

int fq_probe_excitations(fq_engine* e, const uint64_t* det, int64_t n, uint64_t seed, int32_t* exc, double* pgen,
                         double* hmatel) {
  if (!e || !det || !exc || !pgen || !hmatel || n < 1) return FQ_ERR_ARG;
  if (e->sys.system != FQ_SYS_READ_IN) return FQ_ERR_ARG;
  const int W = e->W;
  {  // the determinant must hold nel electrons in orbitals 1..nbasis
    int nel = 0;
    for (int w = 0; w < W; ++w) {
      if (det[w] & ~e->sys.valid[w]) return FQ_ERR_ARG;
      nel += __builtin_popcountll(det[w]);
    }
    if (nel != e->sys.nel) return FQ_ERR_ARG;
  }
  FQ_HIP(hipSetDevice(e->device));
  hipStream_t st = e->stream;
  int32_t* d_exc = nullptr;
  double *d_pg = nullptr, *d_h = nullptr;
  float* d_row = nullptr;
  uint64_t* d_det = nullptr;
  int rc = dalloc(&d_exc, 6 * n);
  rc |= dalloc(&d_pg, n);
  rc |= dalloc(&d_h, n);
  rc |= dalloc(&d_det, W);
  rc |= dalloc(&d_row, 2 * static_cast<int64_t>(e->sys.nbasis));
  if (rc == FQ_OK && hipMemcpyAsync(d_det, det, sizeof(uint64_t) * W, hipMemcpyHostToDevice, st) != hipSuccess)
    rc = FQ_ERR_DEVICE;
  // learned: the network's row, or (proposal_min_attempts > 1, as for a determinant with one
  // attempt in a cycle) none, so that the flat weights are drawn from
  const bool row = e->par.excit_gen == FQ_EXCIT_LEARNED && e->par.proposal_min_attempts <= 1;
  if (rc == FQ_OK && row) rc = fqm::mlp_infer(e->mlp, e->sys, d_det, 1, 1, d_row, st);
  if (rc == FQ_OK) {
    hipLaunchKernelGGL(k_probe, dim3(grid_for(n)), dim3(kBlock), 0, st, e->sys, det[0], W > 1 ? det[1] : 0,
                       row ? d_row : nullptr, n, seed, d_exc, d_pg, d_h);
    if (hipGetLastError() != hipSuccess) rc = FQ_ERR_DEVICE;
  }
  if (rc == FQ_OK) {
    if (hipMemcpyAsync(exc, d_exc, sizeof(int32_t) * 6 * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(pgen, d_pg, sizeof(double) * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(hmatel, d_h, sizeof(double) * n, hipMemcpyDeviceToHost, st) != hipSuccess)
      rc = FQ_ERR_DEVICE;
  }
  if (hipStreamSynchronize(st) != hipSuccess) rc = FQ_ERR_DEVICE;
  dfree(d_exc);
  dfree(d_pg);
  dfree(d_h);
  dfree(d_det);
  dfree(d_row);
  return rc;
}

void fq_destroy(fq_engine* e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  for (int b = 0; b < 2; ++b) {
    dfree(e->det[b]);
    dfree(e->pop[b]);
    dfree(e->kii[b]);
    dfree(e->kbuf[b]);
  }
  dfree(e->pop_new);
  dfree(e->nat);
  dfree(e->att_off);
  dfree(e->flags);
  release_determ(e);
  dfree(e->mmask);
  dfree(e->mcnt);
  dfree(e->moff);
  dfree(e->mdraw);
  dfree(e->mdraw_off);
  dfree(e->part_d);
  dfree(e->part_i);
  dfree(e->st_det);
  dfree(e->st_pop);
  dfree(e->st_ent);
  dfree(e->tile_cnt);
  dfree(e->tile_off);
  dfree(e->tile_win);
  e->sdet = nullptr;  // aliases of kbuf[1], pbuf[1]
  e->srec = nullptr;
  dfree(e->perm);
  dfree(e->pbuf[0]);
  dfree(e->pbuf[1]);
  dfree(e->pass_cnt);
  dfree(e->pass_off);
  dfree(e->run_tcnt);
  dfree(e->run_toff);
  dfree(e->run_tfirst);
  dfree(e->run_tpre);
  dfree(e->rdet);
  dfree(e->rpop);
  dfree(e->rflag);
  dfree(e->rpos);
  dfree(e->rstate);
  dfree(e->nlist);
  dfree(e->nrpos);
  dfree(e->bstart);
  dfree(e->d_sysp);
  dfree(e->mwin);
  dfree(e->nmask);
  dfree(e->ncnt);
  dfree(e->noff);
  dfree(e->ndraw);
  dfree(e->ndraw_off);
  for (auto*& p : e->scan_tmp) dfree(p);
  dfree(e->rng_buf);
  dfree(e->d_stats);
  dfree(e->d_sp_eigv);
  dfree(e->d_sym_table);
  dfree(e->d_inv_sym);
  dfree(e->d_orb_sym);
  dfree(e->d_nbss);
  dfree(e->d_ssbf);
  dfree(e->d_h1);
  dfree(e->d_eri);
  dfree(e->d_jmat);
  dfree(e->d_kmat);
  dfree(e->d_s1j);
  dfree(e->d_s1k);
  dfree(e->d_recip);
  dfree(e->d_jpair);
  dfree(e->d_hb_i);
  dfree(e->d_hb_ij);
  dfree(e->d_hb_ija);
  dfree(e->d_hb_ija_tot);
  dfree(e->d_hb_ija_u);
  dfree(e->d_hb_ija_k);
  dfree(e->d_hb_ijab);
  dfree(e->d_hb_ijab_tot);
  dfree(e->d_hb_ijab_u);
  dfree(e->d_hb_ijab_k);
  dfree(e->d_hb_bias);
  fqm::mlp_free(e->mlp);
  dfree(e->mlp_tab);
  dfree(e->mlp_flat);
  dfree(e->mlp_on);
  dfree(e->mlp_row);
  dfree(e->mlp_idx);
  dfree(e->d_binom);
  dfree(e->d_sys_arr);
  dfree(e->d_ipar);
  dfree(e->d_ibound);
  dfree(e->d_isb);
  dfree(e->d_iatt);
  dfree(e->d_cterm);
  dfree(e->d_cdeath);
  dfree(e->d_iblk_cnt);
  dfree(e->d_iblk_off);
  dfree(e->d_nblk_off);
  dfree(e->d_ipart_d);
  dfree(e->d_ipart_i);
  dfree(e->d_iout_d);
  dfree(e->d_iout_i);
  for (fq_engine* sub : e->subs) fq_destroy(sub);
  e->subs.clear();
  for (int k = 0; k < 4; ++k)
    if (e->ev[k]) (void)hipEventDestroy(e->ev[k]);
  for (int k = 0; k < 9; ++k)
    if (e->pev[k]) (void)hipEventDestroy(e->pev[k]);
  for (int k = 0; k < 2 * fq_engine::kSpawnEvents; ++k)
    if (e->sev[k]) (void)hipEventDestroy(e->sev[k]);
  if (e->ev_len) (void)hipEventDestroy(e->ev_len);
  if (e->h_len) (void)hipHostFree(e->h_len);
  for (auto& g : e->graphs) (void)hipGraphExecDestroy(g.exec);
  for (int k = 0; k < 2; ++k)
    if (e->gev[k]) (void)hipEventDestroy(e->gev[k]);
  if (e->h_in) (void)hipHostFree(e->h_in);
  if (e->h_gstats) (void)hipHostFree(e->h_gstats);
  if (e->h_god) (void)hipHostFree(e->h_god);
  if (e->h_goi) (void)hipHostFree(e->h_goi);
  dfree(e->d_in);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

}  // extern "C"
