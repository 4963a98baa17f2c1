// dSFMT-19937 double-precision stream, close-open [0,1), as consumed by the
// reference's buffered RNG wrapper.
//
// Reference behaviour being reproduced:
//   lib/dSFMT_F03_interface/dSFMT_interface.F90:192-250  dSFMT_init(seed, 50000)
//   lib/dSFMT_F03_interface/dSFMT_interface.F90:387-409  get_rand_close_open:
//       refill a 50000-double store with dsfmt_fill_array_close_open, hand out
//       one double per call.
//   lib/dSFMT-src-2.2.3/dSFMT.c (dsfmt_chk_init_gen_rand, gen_rand_array_c0o1,
//       period_certification), dSFMT-params19937.h, dSFMT-common.h (do_recursion)
//
// The generator is the published dSFMT recursion (Saito & Matsumoto, 2009) with
// the MEXP=19937 parameter set.  fill_array over any even size produces the
// same 128-bit sequence as block-wise generation, so this implementation keeps
// the state as a ring of N 128-bit words and emits doubles two at a time; the
// close-open value is the [1,2) double minus 1.0.  Pinned against golden
// vectors written by the reference's own C source (see tests/golden).
#pragma once
#include <cstdint>
#include <cstring>
#include <vector>

namespace fqmc {

class Dsfmt19937 {
 public:
  static constexpr int kMexp = 19937;
  static constexpr int kN = (kMexp - 128) / 104 + 1;  // 191 128-bit words
  static constexpr int kPos1 = 117;
  static constexpr int kSl1 = 19;
  static constexpr int kSr = 12;
  static constexpr uint64_t kMsk1 = 0x000ffafffffffb3fULL;
  static constexpr uint64_t kMsk2 = 0x000ffdfffc90fffdULL;
  static constexpr uint64_t kFix1 = 0x90014964b32f4329ULL;
  static constexpr uint64_t kFix2 = 0x3b8d12ac548a7c7aULL;
  static constexpr uint64_t kPcv1 = 0x3d84e1ac0dc82880ULL;
  static constexpr uint64_t kPcv2 = 0x0000000000000001ULL;
  static constexpr uint64_t kLowMask = 0x000FFFFFFFFFFFFFULL;
  static constexpr uint64_t kHighConst = 0x3FF0000000000000ULL;

  explicit Dsfmt19937(uint32_t seed) { seed_state(seed); }

  // Next close-open double of the stream.
  double next() {
    if (pos_ >= 2 * kN) refill();
    uint64_t bits = state_[pos_ >> 1][pos_ & 1];
    ++pos_;
    double d;
    std::memcpy(&d, &bits, sizeof d);
    return d - 1.0;
  }

  void fill(double* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = next();
  }

  uint64_t consumed() const { return consumed_blocks_ * 2 * kN + pos_ - 2 * kN; }

 private:
  uint64_t state_[kN][2];
  uint64_t lung_[2];
  int pos_ = 2 * kN;  // index of next double within the current block
  uint64_t consumed_blocks_ = 0;

  void seed_state(uint32_t seed) {
    // 32-bit linear-congruential fill of the (N+1) 128-bit words, word 0 = seed.
    uint32_t w[(kN + 1) * 4];
    w[0] = seed;
    for (int i = 1; i < (kN + 1) * 4; ++i)
      w[i] = 1812433253u * (w[i - 1] ^ (w[i - 1] >> 30)) + static_cast<uint32_t>(i);
    uint64_t all[kN + 1][2];
    for (int i = 0; i < kN + 1; ++i)
      for (int h = 0; h < 2; ++h)
        all[i][h] = static_cast<uint64_t>(w[4 * i + 2 * h]) |
                    (static_cast<uint64_t>(w[4 * i + 2 * h + 1]) << 32);
    // Mask the state words into [1,2) doubles (the lung word is not masked).
    for (int i = 0; i < kN; ++i)
      for (int h = 0; h < 2; ++h) all[i][h] = (all[i][h] & kLowMask) | kHighConst;
    // Period certification on the lung word.
    uint64_t t0 = all[kN][0] ^ kFix1, t1 = all[kN][1] ^ kFix2;
    uint64_t inner = (t0 & kPcv1) ^ (t1 & kPcv2);
    for (int s = 32; s > 0; s >>= 1) inner ^= inner >> s;
    if ((inner & 1) == 0) all[kN][1] ^= 1;  // PCV2 is odd: flip its lowest bit
    for (int i = 0; i < kN; ++i) {
      state_[i][0] = all[i][0];
      state_[i][1] = all[i][1];
    }
    lung_[0] = all[kN][0];
    lung_[1] = all[kN][1];
    pos_ = 2 * kN;
    consumed_blocks_ = 0;
  }

  static inline void recursion(uint64_t r[2], const uint64_t a[2], const uint64_t b[2],
                               uint64_t lung[2]) {
    const uint64_t t0 = a[0], t1 = a[1], l0 = lung[0], l1 = lung[1];
    lung[0] = (t0 << kSl1) ^ (l1 >> 32) ^ (l1 << 32) ^ b[0];
    lung[1] = (t1 << kSl1) ^ (l0 >> 32) ^ (l0 << 32) ^ b[1];
    r[0] = (lung[0] >> kSr) ^ (lung[0] & kMsk1) ^ t0;
    r[1] = (lung[1] >> kSr) ^ (lung[1] & kMsk2) ^ t1;
  }

  // Advance the whole ring by N words (in place; word i depends on the new
  // value of word i+POS1-N once i+POS1 wraps).
  void refill() {
    for (int i = 0; i < kN; ++i) {
      uint64_t out[2];
      recursion(out, state_[i], state_[(i + kPos1) % kN], lung_);
      state_[i][0] = out[0];
      state_[i][1] = out[1];
    }
    pos_ = 0;
    ++consumed_blocks_;
  }
};

}  // namespace fqmc
