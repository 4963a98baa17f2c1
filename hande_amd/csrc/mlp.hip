// Learned double-excitation proposal: inference (bf16 MFMA, gfx950),
// training (fp32 forward/backward, Adam) and heat-bath training samples.
#include "mlp.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace fqm {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;
constexpr int kInferRows = 64;  // determinants per inference block: 4 waves x 16 rows

#define MLP_HIP(call)                                                                                   \
  do {                                                                                                  \
    hipError_t err_ = (call);                                                                           \
    if (err_ != hipSuccess) {                                                                           \
      fprintf(stderr, "fqmc mlp: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(err_), __FILE__, \
              __LINE__);                                                                                \
      return FQ_ERR_DEVICE;                                                                             \
    }                                                                                                   \
  } while (0)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline unsigned blocks(int64_t n) { return static_cast<unsigned>(std::max<int64_t>(1, ceil_div(n, kBlock))); }

// float -> bfloat16 bits, round to nearest even (NaN kept quiet).
__host__ __device__ inline uint16_t f2bf(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__device__ inline float gelu_f(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_g(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

// ---- inference ---------------------------------------------------------------
// Rows [row, row+16) of the block's tile times W (stored transposed [N][K]):
// acc[t] is the 16x16 output tile t (v_mfma_f32_16x16x32_bf16: lane l feeds
// A[row l%16][k 8(l/16)..+7] and B[k 8(l/16)..+7][col l%16] and holds
// D[rows 4(l/16)..+3][col l%16]).  Each 32-wide slab of K of W is staged once
// per block in LDS (slab: the activation buffer this layer does not use),
// so the block's four waves share one read of the weights; every thread of
// the block calls this.
constexpr int kSlabLd = 40;  // slab row stride (bf16): 32 + 8 keeps 16-byte rows and spreads banks
template <int NT>
__device__ inline void mfma_rows16(const uint16_t* A, int lda, int K, const uint16_t* __restrict__ Wt, int ntiles,
                                   f32x4 (&acc)[NT], int row, int kq, int cl, uint16_t* slab) {
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int ncols = ntiles * 16;
  for (int k0 = 0; k0 < K; k0 += 32) {
    for (int e = threadIdx.x; e < ncols * 4; e += blockDim.x) {  // 4 x 8 bf16 per row
      const int n = e >> 2, part = (e & 3) * 8;
      *reinterpret_cast<bf16x8*>(slab + n * kSlabLd + part) =
          *reinterpret_cast<const bf16x8*>(Wt + static_cast<int64_t>(n) * K + k0 + part);
    }
    __syncthreads();
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(A + row * lda + k0 + kq);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (t < ntiles) {
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(slab + (t * 16 + cl) * kSlabLd + kq);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_mlp_infer(int nb, int k1, int n3, int d_out,
                                                      const uint64_t* __restrict__ det, int64_t cap, int64_t n,
                                                      const int64_t* __restrict__ idx,
                                                      const double* __restrict__ eps,
                                                      const uint16_t* __restrict__ w1t, const float* __restrict__ b1,
                                                      const uint16_t* __restrict__ w2t, const float* __restrict__ b2,
                                                      const uint16_t* __restrict__ w3t, const float* __restrict__ b3,
                                                      float* __restrict__ tab, int64_t ld) {
  __shared__ __attribute__((aligned(16))) uint16_t bufA[kInferRows * kHidden];
  __shared__ __attribute__((aligned(16))) uint16_t bufB[kInferRows * kHidden];
  const int64_t row0 = static_cast<int64_t>(blockIdx.x) * kInferRows;
  for (int el = threadIdx.x; el < kInferRows * k1; el += blockDim.x) {
    const int r = el / k1, c = el - r * k1;
    const int64_t d = row0 + r;
    float x = 0.0f;
    if (d < n && c < 2 * nb) {
      const int o = c < nb ? c : c - nb;
      const int64_t src = idx ? idx[d] : d;
      const bool occ = (det[static_cast<int64_t>(o >> 6) * cap + src] >> (o & 63)) & 1;
      x = occ ? (c < nb ? 1.0f : static_cast<float>(eps[o])) : 0.0f;
    }
    bufA[r * k1 + c] = f2bf(x);
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = wave * 16 + (lane & 15), kq = (lane >> 4) * 8, cl = lane & 15;
  const int orow = wave * 16 + (lane >> 4) * 4;
  f32x4 acc[16];
  static_assert(kHidden * kSlabLd <= kInferRows * kHidden, "a weight slab fits the idle activation buffer");
  mfma_rows16<16>(bufA, k1, k1, w1t, 16, acc, row, kq, cl, bufB);
#pragma unroll
  for (int t = 0; t < 16; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int col = t * 16 + cl;
      bufB[(orow + r) * kHidden + col] = f2bf(gelu_f(acc[t][r] + b1[col]));
    }
  __syncthreads();
  mfma_rows16<16>(bufB, kHidden, kHidden, w2t, 16, acc, row, kq, cl, bufA);
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 16; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int col = t * 16 + cl;
      bufA[(orow + r) * kHidden + col] = f2bf(gelu_f(acc[t][r] + b2[col]));
    }
  __syncthreads();
  const int nt3 = n3 / 16;
  mfma_rows16<16>(bufA, kHidden, kHidden, w3t, nt3, acc, row, kq, cl, bufB);
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    if (t >= nt3) break;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int col = t * 16 + cl;
      const int64_t d = row0 + orow + r;
      if (d < n && col < d_out) {
        const float z = acc[t][r] + b3[col];
        tab[d * ld + col] = expf(fminf(fmaxf(z, -kClip), kClip));
      }
    }
  }
}

// bf16 inference copies: w1t[n][k] = W1[k][n] (k < d_in, else 0), w2t[n][k] =
// W2[k][n], w3t[n][k] = W3[k][n] (n < d_out, else 0).
__global__ void k_mlp_pack(int d_in, int k1, int d_out, int n3, const float* __restrict__ p, int64_t o_w1,
                           int64_t o_w2, int64_t o_w3, uint16_t* __restrict__ w1t, uint16_t* __restrict__ w2t,
                           uint16_t* __restrict__ w3t) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n1 = static_cast<int64_t>(kHidden) * k1, n2 = static_cast<int64_t>(kHidden) * kHidden;
  const int64_t n3t = static_cast<int64_t>(n3) * kHidden;
  if (t < n1) {
    const int n = static_cast<int>(t / k1), k = static_cast<int>(t % k1);
    w1t[t] = k < d_in ? f2bf(p[o_w1 + static_cast<int64_t>(k) * kHidden + n]) : 0;
  } else if (t < n1 + n2) {
    const int64_t u = t - n1;
    const int n = static_cast<int>(u / kHidden), k = static_cast<int>(u % kHidden);
    w2t[u] = f2bf(p[o_w2 + static_cast<int64_t>(k) * kHidden + n]);
  } else if (t < n1 + n2 + n3t) {
    const int64_t u = t - n1 - n2;
    const int n = static_cast<int>(u / kHidden), k = static_cast<int>(u % kHidden);
    w3t[u] = n < d_out ? f2bf(p[o_w3 + static_cast<int64_t>(k) * d_out + n]) : 0;
  }
}

// ---- training (fp32) ------------------------------------------------------------
__global__ void k_mlp_feat(int64_t ns, int nb, int d_in, const uint64_t* __restrict__ sdet, int64_t bcap,
                           const double* __restrict__ eps, float* __restrict__ x) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= ns * d_in) return;
  const int64_t s = t / d_in;
  const int c = static_cast<int>(t % d_in);
  const int o = c < nb ? c : c - nb;
  const bool occ = (sdet[static_cast<int64_t>(o >> 6) * bcap + s] >> (o & 63)) & 1;
  x[t] = occ ? (c < nb ? 1.0f : static_cast<float>(eps[o])) : 0.0f;
}

// C[M][N] = op(A)[M][K] op(B)[K][N] (+ bias[N]); op(A)[i][k] = ta ? A[k][i] : A[i][k],
// on fp32 matrix cores: a 64x64 tile of C per block, each
// of its four waves a 32x32 quarter as 2x2 tiles of v_mfma_f32_16x16x4_f32
// (lane l feeds A[row l%16][k l/16] and B[k l/16][col l%16] and holds
// D[rows 4(l/16)..+3][col l%16]); K in slabs of 16 staged through LDS with
// the transposes applied on load.  The accumulation order is fixed by the
// tile shape, so results are deterministic.
constexpr int kGemmTile = 64, kGemmSlab = 16;
__global__ __launch_bounds__(kBlock) void k_gemm_mfma(int M, int N, int K, const float* __restrict__ A, int lda,
                                                      int ta, const float* __restrict__ B, int ldb, int tb,
                                                      const float* __restrict__ bias, float* __restrict__ C, int ldc) {
  __shared__ float As[kGemmTile][kGemmSlab + 1];
  __shared__ float Bs[kGemmSlab][kGemmTile + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i0 = blockIdx.y * kGemmTile, j0 = blockIdx.x * kGemmTile;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += kGemmSlab) {
    // 64 x 16 of op(A) and 16 x 64 of op(B): 1024 elements each, 4 per thread
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = q * kBlock + tid;
      {  // A: consecutive threads along k (row-major A) or along i (transposed A)
        const int r = ta ? (e & 63) : (e >> 4), kk = ta ? (e >> 6) : (e & 15);
        const int gi = i0 + r, gk = k0 + kk;
        As[r][kk] = (gi < M && gk < K)
                        ? (ta ? A[static_cast<int64_t>(gk) * lda + gi] : A[static_cast<int64_t>(gi) * lda + gk])
                        : 0.0f;
      }
      {  // B: along j (row-major B) or along k (transposed B)
        const int cidx = tb ? (e >> 4) : (e & 63), kk = tb ? (e & 15) : (e >> 6);
        const int gj = j0 + cidx, gk = k0 + kk;
        Bs[kk][cidx] = (gj < N && gk < K)
                           ? (tb ? B[static_cast<int64_t>(gj) * ldb + gk] : B[static_cast<int64_t>(gk) * ldb + gj])
                           : 0.0f;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kGemmSlab; kk += 4) {
      const int kq = kk + (lane >> 4);
      float av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = As[wr + a * 16 + (lane & 15)][kq];
#pragma unroll
      for (int b = 0; b < 2; ++b) bv[b] = Bs[kq][wc + b * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = i0 + wr + a * 16 + (lane >> 4) * 4 + r, gj = j0 + wc + b * 16 + (lane & 15);
        if (gi < M && gj < N) C[static_cast<int64_t>(gi) * ldc + gj] = acc[a][b][r] + (bias ? bias[gj] : 0.0f);
      }
}

__global__ void k_gelu_fwd(int64_t n, const float* __restrict__ z, float* __restrict__ a) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < n) a[t] = gelu_f(z[t]);
}
__global__ void k_gelu_bwd(int64_t n, const float* __restrict__ z, float* __restrict__ d) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < n) d[t] = d[t] * gelu_g(z[t]);
}
// out[j] = sum_i D[i][j], rows in order (one thread per column).
__global__ void k_colsum(int M, int N, const float* __restrict__ D, int ldd, float* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  float acc = 0.0f;
  for (int i = 0; i < M; ++i) acc = acc + D[static_cast<int64_t>(i) * ldd + j];
  out[j] = acc;
}
// Fixed-order sum of n doubles on one block.
__global__ void k_dsum(int64_t n, const double* __restrict__ x, double* __restrict__ out) {
  __shared__ double sh[kBlock];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kBlock) acc = acc + x[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int k = kBlock / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// Weighted NLL of each sample's double under the network distribution and its
// gradient with respect to the logits (d3) and to the affinity rows of the
// sample's two holes (dth: row i then row j, nb each), one thread per sample.
// The network component sums four paths (h1, h2, p1, p2): the first particle
// p1 is drawn with h1's affinity over the unoccupied orbitals (normaliser
// Z1(h1)), its partner p2 with h2's affinity over p1's partner class (Z2).
// A sample flagged flat (sflat, may be null) is scored as a determinant the
// network does not run on: hole and particle weights 1 (logits 0), so only the
// affinity receives its gradient.
__global__ void k_mlp_loss_grad(fqd::Sys s, int64_t ns, int nb, int d_out, const uint64_t* __restrict__ sdet,
                                int64_t bcap, const int32_t* __restrict__ sexc, const double* __restrict__ sw,
                                const int32_t* __restrict__ sflat, const double* __restrict__ wsum,
                                const float* __restrict__ z3, const float* __restrict__ th, float* __restrict__ d3,
                                float* __restrict__ dth, double* __restrict__ sloss, const double* __restrict__ sh,
                                int variance) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= ns) return;
  const float* z = z3 + t * d_out;
  float* d = d3 + t * d_out;
  float* gth = dth + t * 2 * nb;
  const double wt = sw[t] / *wsum;
  if (!(sw[t] > 0.0)) {
    for (int c = 0; c < d_out; ++c) d[c] = 0.0f;
    for (int c = 0; c < 2 * nb; ++c) gth[c] = 0.0f;
    sloss[t] = 0.0;
    return;
  }
  uint64_t f[FQ_MAX_WORDS] = {0, 0};
  for (int w = 0; w < s.W; ++w) f[w] = sdet[static_cast<int64_t>(w) * bcap + t];
  const int i = sexc[4 * t], j = sexc[4 * t + 1], a = sexc[4 * t + 2], b = sexc[4 * t + 3];
  const bool flat = sflat && sflat[t];
  auto ew = [&](int o) { return flat ? 1.0 : exp(static_cast<double>(fminf(fmaxf(z[o - 1], -kClip), kClip))); };
  auto fw = [&](int o) { return flat ? 1.0 : exp(static_cast<double>(fminf(fmaxf(z[nb + o - 1], -kClip), kClip))); };
  auto aff = [&](int hole, int o) {
    return exp(static_cast<double>(fminf(fmaxf(th[(hole - 1) * nb + o - 1], -kClip), kClip)));
  };
  const fqd::PartnerClass cb = fqd::partner_class(s, i, j, a), ca = fqd::partner_class(s, i, j, b);
  auto in_class = [&](const fqd::PartnerClass& c, int o, int x) {
    return c.ok && o != x && !fqd::btest(f, o) && fqd::ims_of(o) == c.ims && s.orb_sym[o - 1] == c.sym;
  };
  // Z1(i), Z1(j); Z2 of holes i, j over b's class given a first (zbi, zbj) and a's given b (zai, zaj)
  double zh = 0.0, z1i = 0.0, z1j = 0.0, zbi = 0.0, zbj = 0.0, zai = 0.0, zaj = 0.0;
  for (int o = 1; o <= nb; ++o) {
    if (fqd::btest(f, o)) {
      zh = zh + ew(o);
    } else {
      const double v = fw(o), wi = v * aff(i, o), wj = v * aff(j, o);
      z1i = z1i + wi;
      z1j = z1j + wj;
      if (in_class(cb, o, a)) {
        zbi = zbi + wi;
        zbj = zbj + wj;
      }
      if (in_class(ca, o, b)) {
        zai = zai + wi;
        zaj = zaj + wj;
      }
    }
  }
  const double ei = ew(i), ej = ew(j), fa = fw(a), fb = fw(b);
  const double aia = aff(i, a), aib = aff(i, b), aja = aff(j, a), ajb = aff(j, b);
  const double pij = ei / zh * ej / (zh - ei), pji = ej / zh * ei / (zh - ej);
  const double t1 = pij * (fa * aia / z1i) * (fb * ajb / zbj);  // (i, j, a, b)
  const double t2 = pij * (fb * aib / z1i) * (fa * aja / zaj);  // (i, j, b, a)
  const double t3 = pji * (fa * aja / z1j) * (fb * aib / zbi);  // (j, i, a, b)
  const double t4 = pji * (fb * ajb / z1j) * (fa * aia / zai);  // (j, i, b, a)
  const double q = t1 + t2 + t3 + t4;
  const double r1 = t1 / q, r2 = t2 / q, r3 = t3 / q, r4 = t4 / q;
  // the sample's loss and the factor of its -log q gradient: the NLL weighted
  // by wt, or the spawning-variance term wt |H| / q, whose gradient is the NLL
  // gradient weighted by that term
  const double c = variance ? wt * sh[t] / q : wt;
  sloss[t] = variance ? c : wt * -log(q);
  for (int o = 1; o <= nb; ++o) {
    double gu = 0.0, gi = 0.0, gj = 0.0;
    if (fqd::btest(f, o)) {
      const double e = ew(o);
      double gij = -e / zh - (o != i ? e / (zh - ei) : 0.0);
      double gji = -e / zh - (o != j ? e / (zh - ej) : 0.0);
      if (o == i || o == j) {
        gij += 1.0;
        gji += 1.0;
      }
      gu = (r1 + r2) * gij + (r3 + r4) * gji;
    } else {
      const double v = fw(o), wi = v * aff(i, o), wj = v * aff(j, o);
      const double da = o == a ? 1.0 : 0.0, db = o == b ? 1.0 : 0.0;
      const bool inb = in_class(cb, o, a), ina = in_class(ca, o, b);
      gi = r1 * (da - wi / z1i) + r2 * (db - wi / z1i) + r3 * (db - (inb ? wi / zbi : 0.0)) +
           r4 * (da - (ina ? wi / zai : 0.0));
      gj = r1 * (db - (inb ? wj / zbj : 0.0)) + r2 * (da - (ina ? wj / zaj : 0.0)) + r3 * (da - wj / z1j) +
           r4 * (db - wj / z1j);
    }
    const bool in_h = !flat && fabsf(z[o - 1]) < kClip, in_p = !flat && fabsf(z[nb + o - 1]) < kClip;
    d[o - 1] = in_h ? static_cast<float>(-c * gu) : 0.0f;
    d[nb + o - 1] = in_p ? static_cast<float>(-c * (gi + gj)) : 0.0f;
    gth[o - 1] = fabsf(th[(i - 1) * nb + o - 1]) < kClip ? static_cast<float>(-c * gi) : 0.0f;
    gth[nb + o - 1] = fabsf(th[(j - 1) * nb + o - 1]) < kClip ? static_cast<float>(-c * gj) : 0.0f;
  }
  for (int k = 2 * nb; k < d_out; ++k) d[k] = 0.0f;
}

// Affinity gradient, deterministic: block (h, c) sums, in sample order, the
// row gradients of the samples [c*chunk, (c+1)*chunk) whose hole i or j is
// h + 1 into part[c][h][.]; k_aff_grad_sum adds the chunks in order.
__global__ void k_aff_grad_part(int64_t ns, int nb, int64_t chunk, const int32_t* __restrict__ sexc,
                                const double* __restrict__ sw, const float* __restrict__ dth,
                                float* __restrict__ part) {
  const int h = blockIdx.x + 1, c = blockIdx.y, o = threadIdx.x;
  if (o >= nb) return;
  const int64_t s0 = c * chunk, s1 = s0 + chunk < ns ? s0 + chunk : ns;
  float acc = 0.0f;
  for (int64_t t = s0; t < s1; ++t) {
    if (!(sw[t] > 0.0)) continue;
    if (sexc[4 * t] == h) acc += dth[t * 2 * nb + o];
    if (sexc[4 * t + 1] == h) acc += dth[t * 2 * nb + nb + o];
  }
  part[(static_cast<int64_t>(c) * nb + (h - 1)) * nb + o] = acc;
}

__global__ void k_aff_grad_sum(int nb, int nchunk, const float* __restrict__ part, float* __restrict__ g) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n = static_cast<int64_t>(nb) * nb;
  if (k >= n) return;
  float acc = 0.0f;
  for (int c = 0; c < nchunk; ++c) acc += part[c * n + k];
  g[k] = acc;
}

// Row sums of the stored affinity (one thread per hole h, orbital order) and
// per spin/irrep class (members in the class table's order); thread 0 also
// writes the class bit masks.
__global__ void k_aff_sums(int nb, const float* __restrict__ aff, const int* __restrict__ nbss,
                           const int* __restrict__ ssbf, int nsym_pg, int max_sym_spin, double* __restrict__ rs,
                           double* __restrict__ crs, uint64_t* __restrict__ cmask) {
  const int h = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int ncls = 2 * nsym_pg;
  if (h == 0)
    for (int c = 0; c < ncls; ++c) {
      uint64_t mw[FQ_MAX_WORDS] = {0, 0};
      for (int k = 0; k < nbss[c]; ++k) {
        const int o = ssbf[c * max_sym_spin + k] - 1;
        mw[o >> 6] |= uint64_t{1} << (o & 63);
      }
      for (int w = 0; w < FQ_MAX_WORDS; ++w) cmask[c * FQ_MAX_WORDS + w] = mw[w];
    }
  if (h >= nb) return;
  const float* row = aff + static_cast<int64_t>(h) * nb;
  double z = 0.0;
  for (int p = 0; p < nb; ++p) z = z + static_cast<double>(row[p]);
  rs[h] = z;
  for (int c = 0; c < ncls; ++c) {
    double zc = 0.0;
    for (int k = 0; k < nbss[c]; ++k) zc = zc + static_cast<double>(row[ssbf[c * max_sym_spin + k] - 1]);
    crs[h * ncls + c] = zc;
  }
}

// Alias tables of the affinity rows (one thread per (hole, table): table 0 the
// whole row, table 1 + c class c's members), after k_aff_sums.
__global__ void k_aff_alias(int nb, const float* __restrict__ aff, const int* __restrict__ nbss,
                            const int* __restrict__ ssbf, int nsym_pg, int max_sym_spin,
                            const double* __restrict__ rs, const double* __restrict__ crs, double* __restrict__ au,
                            int* __restrict__ ak, double* __restrict__ cu, int* __restrict__ ck) {
  const int ncls = 2 * nsym_pg;
  const int t = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nb * (1 + ncls)) return;
  const int h = t / (1 + ncls), c = t % (1 + ncls) - 1;
  const float* row = aff + static_cast<int64_t>(h) * nb;
  double w[128];
  if (c < 0) {
    for (int p = 0; p < nb; ++p) w[p] = static_cast<double>(row[p]);
    fqd::build_alias(nb, w, rs[h], au + static_cast<int64_t>(h) * nb, ak + static_cast<int64_t>(h) * nb);
  } else {
    const int n = nbss[c];
    if (n == 0) return;
    for (int q = 0; q < n; ++q) w[q] = static_cast<double>(row[ssbf[c * max_sym_spin + q] - 1]);
    const int64_t o = (static_cast<int64_t>(h) * ncls + c) * max_sym_spin;
    fqd::build_alias(n, w, crs[h * ncls + c], cu + o, ck + o);
  }
}

// The stored affinity table: exp(clamp(theta, -60, 60)) rounded to float.
__global__ void k_aff_table(int64_t n, const float* __restrict__ th, float* __restrict__ aff) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) aff[k] = static_cast<float>(exp(static_cast<double>(fminf(fmaxf(th[k], -kClip), kClip))));
}

__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float c1, float c2) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.0f - b1) * gt;
  const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] = p[t] - lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

// ---- heat-bath training samples -------------------------------------------------
__global__ void k_det_weight(int64_t n, int64_t cap, int S, const int64_t* __restrict__ pop, int shift,
                             int64_t* __restrict__ out) {
  const int64_t d = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (d >= n) return;
  int64_t w = 1;
  for (int q = 0; q < S; ++q) {
    const int64_t p = pop[q * cap + d];
    w += (p < 0 ? -p : p) >> shift;
  }
  out[d] = w;
}

// One heat-bath training sample per thread; its determinant is flagged flat
// when even rounded up it makes fewer than min_att attempts (the network does
// not run on it in production).
__global__ void k_hb_sample(fqd::Sys s, const uint64_t* __restrict__ det, int64_t cap, int64_t n,
                            const int64_t* __restrict__ cum, const int64_t* __restrict__ dw, int64_t min_att,
                            int64_t ns, uint64_t seed, uint32_t step, uint64_t* __restrict__ sdet, int64_t bcap,
                            int32_t* __restrict__ sexc, double* __restrict__ sw, int32_t* __restrict__ sflat,
                            double* __restrict__ sh) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= ns) return;
  fqd::Philox rng(seed, static_cast<uint32_t>(t), static_cast<uint32_t>(t >> 32), step ^ (fqd::kDrawTrain << 24));
  const int64_t total = cum[n];
  const int64_t target = static_cast<int64_t>(rng.next() * static_cast<double>(total));
  int64_t lo = 0, hi = n - 1;  // last d with cum[d] <= target
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (cum[mid] <= target)
      lo = mid;
    else
      hi = mid - 1;
  }
  uint64_t f[FQ_MAX_WORDS] = {0, 0};
  FQ_FOR_WORDS(w, s.W) f[w] = det[static_cast<int64_t>(w) * cap + lo];
  fqd::Excit conn;
  double h = 0.0, pgen = 1.0;
  double wloc[fqd::kHbMaxNel];
  fqd::gen_heat_bath(s, rng, f, conn, h, pgen, fqd::HbBuf{wloc, 1});
  FQ_FOR_WORDS(w, s.W) sdet[static_cast<int64_t>(w) * bcap + t] = f[w];
  const bool use = conn.nexcit == 2 && h != 0.0;
  sexc[4 * t] = conn.from[0];
  sexc[4 * t + 1] = conn.from[1];
  sexc[4 * t + 2] = conn.to[0];
  sexc[4 * t + 3] = conn.to[1];
  sw[t] = use ? fabs(h) / pgen : 0.0;
  sh[t] = use ? fabs(h) : 0.0;
  sflat[t] = dw[lo] < min_att ? 1 : 0;  // dw = 1 + sum of floor|N|
}

template <class T>
int dalloc(T** p, int64_t count) {
  if (count < 1) count = 1;
  return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * static_cast<size_t>(count)) == hipSuccess
             ? FQ_OK
             : FQ_ERR_NOMEM;
}
template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void gemm(int M, int N, int K, const float* A, int lda, int ta, const float* B, int ldb, int tb, const float* bias,
          float* C, int ldc, hipStream_t st) {
  hipLaunchKernelGGL(k_gemm_mfma, dim3(static_cast<unsigned>(ceil_div(N, kGemmTile)),
                                      static_cast<unsigned>(ceil_div(M, kGemmTile))),
                     dim3(kBlock), 0, st, M, N, K, A, lda, ta, B, ldb, tb, bias, C, ldc);
}

// splitmix64 -> standard normal (Box--Muller) for parameter initialisation.
struct HostNormal {
  uint64_t x;
  uint64_t next_u64() {
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (static_cast<double>(next_u64() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double normal() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(6.283185307179586 * uniform()); }
};

}  // namespace

int mlp_alloc(Mlp& m, int nb, int W, int64_t bcap) {
  m.nb = nb;
  m.d_in = m.d_out = 2 * nb;
  m.k1 = static_cast<int>(ceil_div(m.d_in, 32) * 32);
  m.n3 = static_cast<int>(ceil_div(m.d_out, 16) * 16);
  if (m.k1 > kHidden || m.n3 > kHidden) return FQ_ERR_ARG;
  m.o_w1 = 0;
  m.o_b1 = m.o_w1 + static_cast<int64_t>(m.d_in) * kHidden;
  m.o_w2 = m.o_b1 + kHidden;
  m.o_b2 = m.o_w2 + static_cast<int64_t>(kHidden) * kHidden;
  m.o_w3 = m.o_b2 + kHidden;
  m.o_b3 = m.o_w3 + static_cast<int64_t>(kHidden) * m.d_out;
  m.o_th = m.o_b3 + m.d_out;
  m.nparams = m.o_th + static_cast<int64_t>(nb) * nb;
  m.bcap = bcap;
  m.nchunk = static_cast<int>(ceil_div(bcap, kAffChunk));
  int rc = FQ_OK;
  rc |= dalloc(&m.p, m.nparams);
  rc |= dalloc(&m.g, m.nparams);
  rc |= dalloc(&m.m, m.nparams);
  rc |= dalloc(&m.v, m.nparams);
  rc |= dalloc(&m.w1t, static_cast<int64_t>(kHidden) * m.k1);
  rc |= dalloc(&m.w2t, static_cast<int64_t>(kHidden) * kHidden);
  rc |= dalloc(&m.w3t, static_cast<int64_t>(m.n3) * kHidden);
  rc |= dalloc(&m.x, bcap * m.d_in);
  rc |= dalloc(&m.z1, bcap * kHidden);
  rc |= dalloc(&m.a1, bcap * kHidden);
  rc |= dalloc(&m.z2, bcap * kHidden);
  rc |= dalloc(&m.a2, bcap * kHidden);
  rc |= dalloc(&m.z3, bcap * m.d_out);
  rc |= dalloc(&m.d3, bcap * m.d_out);
  rc |= dalloc(&m.d2, bcap * kHidden);
  rc |= dalloc(&m.d1, bcap * kHidden);
  rc |= dalloc(&m.sdet, bcap * W);
  rc |= dalloc(&m.sexc, bcap * 4);
  rc |= dalloc(&m.sw, bcap);
  rc |= dalloc(&m.sh, bcap);
  rc |= dalloc(&m.sflat, bcap);
  rc |= dalloc(&m.sloss, bcap);
  rc |= dalloc(&m.dsum, 2);
  rc |= dalloc(&m.aff, static_cast<int64_t>(nb) * nb);
  rc |= dalloc(&m.dth, bcap * 2 * nb);
  rc |= dalloc(&m.aff_part, static_cast<int64_t>(m.nchunk) * nb * nb);
  return rc == FQ_OK ? FQ_OK : FQ_ERR_NOMEM;
}

void mlp_free(Mlp& m) {
  dfree(m.p);
  dfree(m.g);
  dfree(m.m);
  dfree(m.v);
  dfree(m.w1t);
  dfree(m.w2t);
  dfree(m.w3t);
  dfree(m.x);
  dfree(m.z1);
  dfree(m.a1);
  dfree(m.z2);
  dfree(m.a2);
  dfree(m.z3);
  dfree(m.d3);
  dfree(m.d2);
  dfree(m.d1);
  dfree(m.sdet);
  dfree(m.sexc);
  dfree(m.sw);
  dfree(m.sh);
  dfree(m.sflat);
  dfree(m.sloss);
  dfree(m.dsum);
  dfree(m.aff);
  dfree(m.rs);
  dfree(m.crs);
  dfree(m.cmask);
  dfree(m.au);
  dfree(m.ak);
  dfree(m.cu);
  dfree(m.ck);
  dfree(m.dth);
  dfree(m.aff_part);
  dfree(m.scum);
  dfree(m.swts);
  m.scum_cap = 0;
}

int mlp_refresh(Mlp& m, hipStream_t st) {
  const int64_t total = static_cast<int64_t>(kHidden) * m.k1 + static_cast<int64_t>(kHidden) * kHidden +
                        static_cast<int64_t>(m.n3) * kHidden;
  hipLaunchKernelGGL(k_mlp_pack, dim3(blocks(total)), dim3(kBlock), 0, st, m.d_in, m.k1, m.d_out, m.n3, m.p, m.o_w1,
                     m.o_w2, m.o_w3, m.w1t, m.w2t, m.w3t);
  const int64_t na = static_cast<int64_t>(m.nb) * m.nb;
  hipLaunchKernelGGL(k_aff_table, dim3(blocks(na)), dim3(kBlock), 0, st, na, m.p + m.o_th, m.aff);
  if (m.nbss) {
    hipLaunchKernelGGL(k_aff_sums, dim3(blocks(m.nb)), dim3(kBlock), 0, st, m.nb, m.aff, m.nbss, m.ssbf, m.nsym_pg,
                       m.max_sym_spin, m.rs, m.crs, m.cmask);
    const int ntab = m.nb * (1 + 2 * m.nsym_pg);
    hipLaunchKernelGGL(k_aff_alias, dim3(blocks(ntab)), dim3(kBlock), 0, st, m.nb, m.aff, m.nbss, m.ssbf,
                       m.nsym_pg, m.max_sym_spin, m.rs, m.crs, m.au, m.ak, m.cu, m.ck);
  }
  MLP_HIP(hipGetLastError());
  return FQ_OK;
}

int mlp_set_classes(Mlp& m, const int* nbss, const int* ssbf, int nsym_pg, int max_sym_spin) {
  m.nbss = nbss;
  m.ssbf = ssbf;
  m.nsym_pg = nsym_pg;
  m.max_sym_spin = max_sym_spin;
  int rc = dalloc(&m.rs, m.nb);
  rc |= dalloc(&m.crs, static_cast<int64_t>(m.nb) * 2 * nsym_pg);
  rc |= dalloc(&m.cmask, static_cast<int64_t>(2) * nsym_pg * FQ_MAX_WORDS);
  rc |= dalloc(&m.au, static_cast<int64_t>(m.nb) * m.nb);
  rc |= dalloc(&m.ak, static_cast<int64_t>(m.nb) * m.nb);
  rc |= dalloc(&m.cu, static_cast<int64_t>(m.nb) * 2 * nsym_pg * max_sym_spin);
  rc |= dalloc(&m.ck, static_cast<int64_t>(m.nb) * 2 * nsym_pg * max_sym_spin);
  return rc == FQ_OK ? FQ_OK : FQ_ERR_NOMEM;
}

int mlp_set(Mlp& m, const float* host, hipStream_t st) {
  MLP_HIP(hipMemcpyAsync(m.p, host, sizeof(float) * m.nparams, hipMemcpyHostToDevice, st));
  MLP_HIP(hipMemsetAsync(m.m, 0, sizeof(float) * m.nparams, st));
  MLP_HIP(hipMemsetAsync(m.v, 0, sizeof(float) * m.nparams, st));
  m.adam_t = 0;
  int rc = mlp_refresh(m, st);
  if (rc != FQ_OK) return rc;
  MLP_HIP(hipStreamSynchronize(st));
  return FQ_OK;
}

int mlp_init(Mlp& m, uint64_t seed, hipStream_t st) {
  std::vector<float> h(static_cast<size_t>(m.nparams), 0.0f);
  HostNormal rng{seed ^ 0x5DEECE66DULL};
  auto fill = [&](int64_t off, int64_t count, double sd) {
    for (int64_t k = 0; k < count; ++k) h[off + k] = static_cast<float>(sd * rng.normal());
  };
  fill(m.o_w1, static_cast<int64_t>(m.d_in) * kHidden, 1.0 / std::sqrt(static_cast<double>(m.d_in)));
  fill(m.o_w2, static_cast<int64_t>(kHidden) * kHidden, 1.0 / std::sqrt(static_cast<double>(kHidden)));
  // small output layer: the initial proposal is close to uniform
  fill(m.o_w3, static_cast<int64_t>(kHidden) * m.d_out, 0.1 / std::sqrt(static_cast<double>(kHidden)));
  return mlp_set(m, h.data(), st);
}

int mlp_infer(const Mlp& m, const fqd::Sys& s, const uint64_t* det, int64_t cap, int64_t n, float* tab,
              hipStream_t st, const int64_t* idx) {
  if (n <= 0) return FQ_OK;
  const unsigned g = static_cast<unsigned>(ceil_div(n, kInferRows));
  hipLaunchKernelGGL(k_mlp_infer, dim3(g), dim3(kBlock), 0, st, m.nb, m.k1, m.n3, m.d_out, det, cap, n, idx,
                     s.sp_eigv,
                     m.w1t, m.p + m.o_b1, m.w2t, m.p + m.o_b2, m.w3t, m.p + m.o_b3, tab,
                     static_cast<int64_t>(m.d_out));
  MLP_HIP(hipGetLastError());
  return FQ_OK;
}

int mlp_train_step(Mlp& m, const fqd::Sys& s, int64_t ns, double lr, bool apply, hipStream_t st) {
  if (ns < 1 || ns > m.bcap) return FQ_ERR_ARG;
  const int H = kHidden, di = m.d_in, dout = m.d_out, B = static_cast<int>(ns);
  hipLaunchKernelGGL(k_mlp_feat, dim3(blocks(ns * di)), dim3(kBlock), 0, st, ns, m.nb, di, m.sdet, m.bcap,
                     s.sp_eigv, m.x);
  gemm(B, H, di, m.x, di, 0, m.p + m.o_w1, H, 0, m.p + m.o_b1, m.z1, H, st);
  hipLaunchKernelGGL(k_gelu_fwd, dim3(blocks(ns * H)), dim3(kBlock), 0, st, ns * H, m.z1, m.a1);
  gemm(B, H, H, m.a1, H, 0, m.p + m.o_w2, H, 0, m.p + m.o_b2, m.z2, H, st);
  hipLaunchKernelGGL(k_gelu_fwd, dim3(blocks(ns * H)), dim3(kBlock), 0, st, ns * H, m.z2, m.a2);
  gemm(B, dout, H, m.a2, H, 0, m.p + m.o_w3, dout, 0, m.p + m.o_b3, m.z3, dout, st);
  hipLaunchKernelGGL(k_dsum, dim3(1), dim3(kBlock), 0, st, ns, m.sw, m.dsum);
  hipLaunchKernelGGL(k_mlp_loss_grad, dim3(blocks(ns)), dim3(kBlock), 0, st, s, ns, m.nb, dout, m.sdet, m.bcap,
                     m.sexc, m.sw, m.sflat, m.dsum, m.z3, m.p + m.o_th, m.d3, m.dth, m.sloss, m.sh, m.objective);
  hipLaunchKernelGGL(k_dsum, dim3(1), dim3(kBlock), 0, st, ns, m.sloss, m.dsum + 1);
  const int nchunk = static_cast<int>(ceil_div(ns, kAffChunk));
  hipLaunchKernelGGL(k_aff_grad_part, dim3(static_cast<unsigned>(m.nb), static_cast<unsigned>(nchunk)),
                     dim3(kHidden / 2), 0, st, ns, m.nb, static_cast<int64_t>(kAffChunk), m.sexc, m.sw, m.dth,
                     m.aff_part);
  const int64_t na = static_cast<int64_t>(m.nb) * m.nb;
  hipLaunchKernelGGL(k_aff_grad_sum, dim3(blocks(na)), dim3(kBlock), 0, st, m.nb, nchunk, m.aff_part,
                     m.g + m.o_th);
  // backward
  gemm(H, dout, B, m.a2, H, 1, m.d3, dout, 0, nullptr, m.g + m.o_w3, dout, st);
  hipLaunchKernelGGL(k_colsum, dim3(blocks(dout)), dim3(kBlock), 0, st, B, dout, m.d3, dout, m.g + m.o_b3);
  gemm(B, H, dout, m.d3, dout, 0, m.p + m.o_w3, dout, 1, nullptr, m.d2, H, st);
  hipLaunchKernelGGL(k_gelu_bwd, dim3(blocks(ns * H)), dim3(kBlock), 0, st, ns * H, m.z2, m.d2);
  gemm(H, H, B, m.a1, H, 1, m.d2, H, 0, nullptr, m.g + m.o_w2, H, st);
  hipLaunchKernelGGL(k_colsum, dim3(blocks(H)), dim3(kBlock), 0, st, B, H, m.d2, H, m.g + m.o_b2);
  gemm(B, H, H, m.d2, H, 0, m.p + m.o_w2, H, 1, nullptr, m.d1, H, st);
  hipLaunchKernelGGL(k_gelu_bwd, dim3(blocks(ns * H)), dim3(kBlock), 0, st, ns * H, m.z1, m.d1);
  gemm(di, H, B, m.x, di, 1, m.d1, H, 0, nullptr, m.g + m.o_w1, H, st);
  hipLaunchKernelGGL(k_colsum, dim3(blocks(H)), dim3(kBlock), 0, st, B, H, m.d1, H, m.g + m.o_b1);
  MLP_HIP(hipGetLastError());
  if (apply) {
    m.adam_t += 1;
    const float c1 = static_cast<float>(1.0 - std::pow(0.9, static_cast<double>(m.adam_t)));
    const float c2 = static_cast<float>(1.0 - std::pow(0.999, static_cast<double>(m.adam_t)));
    hipLaunchKernelGGL(k_adam, dim3(blocks(m.o_th)), dim3(kBlock), 0, st, m.o_th, m.p, m.g, m.m, m.v,
                       static_cast<float>(lr), c1, c2);
    const int64_t na = m.nparams - m.o_th;
    hipLaunchKernelGGL(k_adam, dim3(blocks(na)), dim3(kBlock), 0, st, na, m.p + m.o_th, m.g + m.o_th,
                       m.m + m.o_th, m.v + m.o_th, static_cast<float>(lr * kAffLrScale), c1, c2);
    MLP_HIP(hipGetLastError());
    return mlp_refresh(m, st);
  }
  return FQ_OK;
}

int mlp_sample(Mlp& m, const fqd::Sys& s, const uint64_t* det, const int64_t* pop, int64_t cap, int64_t n,
               int64_t ns, int64_t min_att, uint64_t seed, uint32_t step, hipStream_t st,
               int (*scan)(void* ctx, const int64_t* in, int64_t* out, int64_t n), void* scan_ctx) {
  if (n < 1 || ns < 1 || ns > m.bcap) return FQ_ERR_ARG;
  if (m.scum_cap < n + 1) {
    dfree(m.scum);
    dfree(m.swts);
    m.scum_cap = 0;
    if (dalloc(&m.scum, n + 1) != FQ_OK || dalloc(&m.swts, n + 1) != FQ_OK) return FQ_ERR_NOMEM;
    m.scum_cap = n + 1;
  }
  int shift = 0;
  while ((int64_t{1} << shift) < s.rf) ++shift;
  // per-determinant weights, their exclusive prefix (total at scum[n]), then the samples
  hipLaunchKernelGGL(k_det_weight, dim3(blocks(n)), dim3(kBlock), 0, st, n, cap, s.S, pop, shift, m.swts);
  int rc = scan(scan_ctx, m.swts, m.scum, n);
  if (rc != FQ_OK) return rc;
  hipLaunchKernelGGL(k_hb_sample, dim3(blocks(ns)), dim3(kBlock), 0, st, s, det, cap, n, m.scum, m.swts, min_att,
                     ns, seed, step, m.sdet, m.bcap, m.sexc, m.sw, m.sflat, m.sh);
  MLP_HIP(hipGetLastError());
  return FQ_OK;
}

}  // namespace fqm
