// Learned double-excitation proposal: a 2-hidden-layer MLP (256 wide, GELU)
// from a determinant's occupations to per-orbital hole and particle weights.
//
// Parameters (float32 master copy, one flat vector):
//   W1[d_in][256], b1[256], W2[256][256], b2[256], W3[256][d_out], b3[d_out],
//   theta[nbasis][nbasis]
// with d_in = d_out = 2*nbasis; input x_k = n_k, x_{nb+k} = n_k * eps_k.
// theta is the determinant-independent hole-particle affinity (logits; zero
// at initialisation, where the proposal is the product of hole and particle
// weights); the generator reads aff = exp(clamp(theta, -60, 60)) as float.
// Inference uses bf16 copies of W1..W3 (transposed to [N][K], K padded to a
// multiple of 32) on the bf16 MFMA with fp32 accumulation; the stored
// proposal weights are exp(clamp(z3, -60, 60)).  Training (fp32) minimises
// the weighted negative log-likelihood of heat-bath samples under the
// network's double-excitation distribution, with Adam.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace fqm {

constexpr int kHidden = 256;
constexpr float kClip = 60.0f;
constexpr int kAffChunk = 512;        // samples per block of the affinity-gradient reduction
constexpr double kAffLrScale = 10.0;  // Adam step of theta relative to the network's

struct Mlp {
  int nb = 0, d_in = 0, d_out = 0, k1 = 0, n3 = 0;  // k1: d_in padded to 32; n3: d_out padded to 16
  int64_t nparams = 0, o_w1 = 0, o_b1 = 0, o_w2 = 0, o_b2 = 0, o_w3 = 0, o_b3 = 0, o_th = 0;
  float* p = nullptr;  // master parameters
  float* g = nullptr;  // gradient
  float* m = nullptr;  // Adam moments
  float* v = nullptr;
  int64_t adam_t = 0;
  uint16_t* w1t = nullptr;  // bf16 bits [256][k1]
  uint16_t* w2t = nullptr;  // [256][256]
  uint16_t* w3t = nullptr;  // [n3][256]
  float* aff = nullptr;     // [nb][nb] exp(clamp(theta)), read by the generator
  // affinity row sums for flat-weight draws (refreshed with aff): rs[h] over
  // all orbitals, crs[h][c] over spin/irrep class c, class bit masks cmask[c][w]
  double* rs = nullptr;
  double* crs = nullptr;
  uint64_t* cmask = nullptr;
  // alias tables of the affinity rows for flat-weight draws: whole row
  // (au/ak [h][p]) and per class (cu/ck [h][c][k] over the class's members)
  double *au = nullptr, *cu = nullptr;
  int *ak = nullptr, *ck = nullptr;
  const int* nbss = nullptr;  // the system's class tables (device), set before mlp_init
  const int* ssbf = nullptr;
  int nsym_pg = 1, max_sym_spin = 0;
  // training batch storage (capacity bcap samples)
  int64_t bcap = 0;
  float *x = nullptr, *z1 = nullptr, *a1 = nullptr, *z2 = nullptr, *a2 = nullptr, *z3 = nullptr;
  float *d3 = nullptr, *d2 = nullptr, *d1 = nullptr;
  float* dth = nullptr;       // [bcap][2][nb]: per-sample gradients of the affinity rows of holes i, j
  float* aff_part = nullptr;  // [nchunk][nb][nb]: chunk sums of dth
  int nchunk = 0;
  uint64_t* sdet = nullptr;  // [W][bcap]
  int32_t* sexc = nullptr;   // [bcap][4]: i, j, a, b (i<j, a<b)
  double* sw = nullptr;      // sample weights
  double* sh = nullptr;      // |H_ij| of each sample (the spawning-variance objective's factor)
  int objective = 0;         // 0: weighted cross-entropy to q* ~ |H|; 1: spawning variance sum |H|^2 / q
  int32_t* sflat = nullptr;  // 1: sample scored with flat weights (its determinant gets no network row)
  double* sloss = nullptr;   // per-sample weighted NLL
  double* dsum = nullptr;    // [2]: sum of weights, loss
  int64_t* swts = nullptr;   // determinant sampling weights
  int64_t* scum = nullptr;   // their exclusive prefix (n + 1)
  int64_t scum_cap = 0;
};

int mlp_alloc(Mlp& m, int nb, int W, int64_t bcap);
// The system's spin/irrep classes (device tables), for the flat-draw sums.
int mlp_set_classes(Mlp& m, const int* nbss, const int* ssbf, int nsym_pg, int max_sym_spin);
void mlp_free(Mlp& m);
// Deterministic initial parameters from seed (host RNG), uploaded; Adam reset.
int mlp_init(Mlp& m, uint64_t seed, hipStream_t st);
int mlp_set(Mlp& m, const float* host, hipStream_t st);
// bf16 inference copies from the master parameters.
int mlp_refresh(Mlp& m, hipStream_t st);
// Proposal weights of determinants [0, n) of a structure-of-arrays list (row
// r of tab for determinant r), or of determinants idx[0..n) when idx is given.
int mlp_infer(const Mlp& m, const fqd::Sys& s, const uint64_t* det, int64_t cap, int64_t n, float* tab,
              hipStream_t st, const int64_t* idx = nullptr);
// Forward/backward over the first ns stored samples (sdet/sexc/sw): fills g,
// leaves sum(w) and the weighted loss in dsum.  apply: one Adam step (lr).
int mlp_train_step(Mlp& m, const fqd::Sys& s, int64_t ns, double lr, bool apply, hipStream_t st);
// ns heat-bath samples from the main list (determinant chosen with
// probability proportional to 1 + floor(|pop|/real_factor), summed over
// spaces); a sample whose determinant has 1 + sum floor|N| < min_att is flagged
// flat (scored with the weights production gives determinants without a row).
int mlp_sample(Mlp& m, const fqd::Sys& s, const uint64_t* det, const int64_t* pop, int64_t cap, int64_t n,
               int64_t ns, int64_t min_att, uint64_t seed, uint32_t step, hipStream_t st,
               int (*scan)(void* ctx, const int64_t* in, int64_t* out, int64_t n), void* scan_ctx);

}  // namespace fqm
