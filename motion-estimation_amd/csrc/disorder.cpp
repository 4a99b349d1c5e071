// DISORDER sample ordering (simulation.py:192-245), host code.
//
// The reference draws each tile's within-tile order with Python's
// random.Random(seed).shuffle (simulation.py:223, 233), so bit-exact parity
// needs the same generator and the same draw procedure:
//   * MT19937 (Matsumoto & Nishimura 1998) seeded by init_by_array with the
//     seed's 32-bit little-endian words (CPython's integer seeding);
//   * shuffle: Fisher-Yates from the last index down, j = randbelow(i + 1);
//   * randbelow(n): k = bit_length(n); draw r = genrand() >> (32 - k) until
//     r < n (rejection sampling, CPython's _randbelow_with_getrandbits).
#include <stdint.h>

#include <string>
#include <vector>

#include "me_internal.h"

namespace me {
namespace {

class MT19937 {
 public:
  explicit MT19937(unsigned long long seed) {
    std::vector<uint32_t> key;
    unsigned long long s = seed;
    do {
      key.push_back((uint32_t)(s & 0xffffffffu));
      s >>= 32;
    } while (s);
    init_by_array(key);
  }
  uint32_t next() {
    if (idx_ >= kN) twist();
    uint32_t y = mt_[idx_++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
  uint32_t randbelow(uint32_t n) {
    int k = 0;
    for (uint32_t v = n; v; v >>= 1) ++k;
    uint32_t r = next() >> (32 - k);
    while (r >= n) r = next() >> (32 - k);
    return r;
  }

 private:
  static constexpr int kN = 624, kM = 397;
  uint32_t mt_[kN];
  int idx_ = kN + 1;

  void init_genrand(uint32_t s) {
    mt_[0] = s;
    for (int i = 1; i < kN; ++i) mt_[i] = 1812433253u * (mt_[i - 1] ^ (mt_[i - 1] >> 30)) + (uint32_t)i;
    idx_ = kN;
  }
  void init_by_array(const std::vector<uint32_t>& key) {
    init_genrand(19650218u);
    int i = 1, j = 0;
    const int klen = (int)key.size();
    for (int k = (kN > klen ? kN : klen); k; --k) {
      mt_[i] = (mt_[i] ^ ((mt_[i - 1] ^ (mt_[i - 1] >> 30)) * 1664525u)) + key[j] + (uint32_t)j;
      ++i;
      ++j;
      if (i >= kN) {
        mt_[0] = mt_[kN - 1];
        i = 1;
      }
      if (j >= klen) j = 0;
    }
    for (int k = kN - 1; k; --k) {
      mt_[i] = (mt_[i] ^ ((mt_[i - 1] ^ (mt_[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
      ++i;
      if (i >= kN) {
        mt_[0] = mt_[kN - 1];
        i = 1;
      }
    }
    mt_[0] = 0x80000000u;
    idx_ = kN;
  }
  void twist() {
    for (int i = 0; i < kN; ++i) {
      const uint32_t y = (mt_[i] & 0x80000000u) | (mt_[(i + 1) % kN] & 0x7fffffffu);
      mt_[i] = mt_[(i + kM) % kN] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    idx_ = 0;
  }
};

}  // namespace
}  // namespace me

extern "C" {

// Fills out_line/out_part with the DISORDER order (simulation.py:201-245) and
// returns its length in *out_count, the padded line/partition counts in
// *out_L / *out_P.  Fails if capacity is too small (out_count still set).
int me_disorder_order(int max_line_number, int max_partition_number, int pat_lines_to_measure, int block_lin,
                      int block_par, unsigned long long seed, int* out_line, int* out_part, long long capacity,
                      long long* out_count, int* out_L, int* out_P) {
  ME_REQUIRE(block_lin >= 1 && block_par >= 1, "disorder: block sizes must be positive");
  ME_REQUIRE(max_line_number >= 0 && max_partition_number >= 0 && pat_lines_to_measure >= 0,
             "disorder: negative sizes");
  // pad to block multiples (simulation.py:207-210)
  const long long L = ((max_line_number + 1 + block_lin - 1) / block_lin) * (long long)block_lin;
  const long long P = ((max_partition_number + 1 + block_par - 1) / block_par) * (long long)block_par;
  const long long mln = L - 1;
  const long long n_reorder = (long long)pat_lines_to_measure * P;  // simulation.py:213
  const long long per_group_max = (long long)block_lin * block_par;
  ME_REQUIRE(per_group_max < (1LL << 31), "disorder: tile too large");
  const long long n_lin_g = L / block_lin, n_par_g = P / block_par;
  const long long n_groups = n_lin_g * n_par_g;
  const long long per_group = n_reorder / n_groups;  // integer division (simulation.py:220)
  const long long total = per_group * n_groups;
  if (out_count) *out_count = total;
  if (out_L) *out_L = (int)L;
  if (out_P) *out_P = (int)P;
  ME_REQUIRE(per_group <= per_group_max, "disorder: more draws per tile than tile entries");
  ME_REQUIRE(out_line && out_part, "disorder: null output");
  ME_REQUIRE(capacity >= total, "disorder: output capacity too small");
  me::MT19937 rng(seed);
  std::vector<int> idx((size_t)per_group_max);
  for (long long linG = 0; linG < n_lin_g; ++linG) {
    for (long long parG = 0; parG < n_par_g; ++parG) {
      for (long long i = 0; i < per_group_max; ++i) idx[(size_t)i] = (int)i;
      for (long long i = per_group_max - 1; i >= 1; --i) {  // random.shuffle
        const uint32_t j = rng.randbelow((uint32_t)(i + 1));
        const int tmp = idx[(size_t)i];
        idx[(size_t)i] = idx[j];
        idx[j] = tmp;
      }
      for (long long j = 0; j < per_group; ++j) {
        const long long local = idx[(size_t)j];
        const long long ll = local % block_lin, lp = local / block_lin;
        // simulation.py:239-242
        const long long line = ll + (linG % block_par) * (mln + 1) / block_par + (linG / block_par) * block_lin;
        const long long part = lp + parG * block_par;
        const long long pos = j * n_groups + linG * n_par_g + parG;
        out_line[pos] = (int)line;
        out_part[pos] = (int)part;
      }
    }
  }
  return 0;
}

}  // extern "C"
