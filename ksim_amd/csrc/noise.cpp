// noise.cpp — host construction of the device noise tables.
#include "noise.hpp"

#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>

namespace ksim {

namespace {

constexpr uint64_t U64_MAX = std::numeric_limits<uint64_t>::max();

// ceil(x 2^64) - 1 clamped to [0, 2^64 - 1], for x in [0, 1]
uint64_t threshold_le(long double x) {
  if (!(x > 0.0L)) return 0;  // never (v <= 0 has probability 2^-64)
  if (x >= 1.0L) return U64_MAX;
  const long double y = std::ceil(std::ldexp(x, 64));
  if (y >= 18446744073709551616.0L) return U64_MAX;
  return (uint64_t)y - 1;
}

// floor(x 2^64) clamped, for x in [0, 1]
uint64_t threshold_lt(long double x) {
  if (!(x > 0.0L)) return 0;
  if (x >= 1.0L) return U64_MAX;
  const long double y = std::floor(std::ldexp(x, 64));
  if (y >= 18446744073709551616.0L) return U64_MAX;
  return (uint64_t)y;
}

}  // namespace

NoiseHost build_noise(const double *probs, const int32_t *n_outcomes, const uint8_t *patterns,
                      int32_t n_channels, int32_t num_f) {
  if (n_channels < 0 || num_f < 0) throw std::invalid_argument("negative noise sizes");
  if (n_channels > 0 && (!probs || !n_outcomes)) throw std::invalid_argument("NULL noise arrays");
  NoiseHost h;
  h.num_f = num_f;
  h.wf = (num_f + 31) / 32;
  size_t po = 0, pr = 0;  // offsets into probs and pattern rows
  struct Kept { int src; double p; size_t po, pr; int m; };
  std::vector<Kept> kept;
  for (int c = 0; c < n_channels; ++c) {
    const int n = n_outcomes[c];
    if (n < 1 || (n & (n - 1))) throw std::invalid_argument("channel outcome count is not a power of two");
    double sum = 0;
    for (int o = 0; o < n; ++o) {
      const double q = probs[po + o];
      if (!(q >= -1e-6 && q <= 1.0 + 1e-6))
        throw std::invalid_argument("channel " + std::to_string(c) + ": probability outside [0, 1]");
      sum += q;
    }
    if (std::fabs(sum - 1.0) > 1e-6)
      throw std::invalid_argument("channel " + std::to_string(c) + ": probabilities do not sum to 1");
    const double p_fire = 1.0 - probs[po];
    bool flips = false;
    if (n > 1 && num_f > 0 && patterns)
      for (size_t i = 0; i < (size_t)(n - 1) * num_f && !flips; ++i) flips = patterns[pr * num_f + i] & 1u;
    if (p_fire > 1e-15 && n > 1 && flips) kept.push_back(Kept{c, p_fire, po, pr, n - 1});
    po += n;
    pr += n - 1;
  }
  h.C = (int32_t)kept.size();
  for (const Kept &k : kept) h.p_max = std::max(h.p_max, std::min(k.p, 1.0));
  h.skip.assign((size_t)h.C + 1, U64_MAX);
  const long double keep = 1.0L - (long double)h.p_max;
  long double pw = 1.0L;
  for (int k = 1; k <= h.C; ++k) {
    pw *= keep;
    h.skip[k] = threshold_lt(pw);
  }
  for (const Kept &k : kept) {
    NoiseChan ch{};
    ch.acc = threshold_le((long double)std::min(k.p, 1.0) / (long double)h.p_max);
    ch.n_out = (uint32_t)k.m;
    ch.off_cdf = (uint32_t)h.cdf.size();
    ch.off_pat = (uint32_t)h.pat.size();
    // conditional CDF over the non-identity outcomes, normalised to end at 1
    // (channels.py:609-610)
    long double tot = 0;
    for (int o = 1; o <= k.m; ++o) tot += (long double)probs[k.po + o];
    long double run = 0;
    for (int o = 1; o <= k.m; ++o) {
      run += (long double)probs[k.po + o];
      h.cdf.push_back(o == k.m ? U64_MAX : threshold_le(tot > 0 ? run / tot : 0.0L));
      for (int w = 0; w < h.wf; ++w) {
        uint32_t word = 0;
        for (int b = 0; b < 32 && 32 * w + b < num_f; ++b)
          word |= (uint32_t)(patterns[(k.pr + o - 1) * num_f + 32 * w + b] & 1u) << b;
        h.pat.push_back(word);
      }
    }
    h.chan.push_back(ch);
    h.source.push_back(k.src);
  }
  return h;
}

}  // namespace ksim
