// noise.hpp — device-side noise: the f-bit channel sampler's tables.
//
// The reference draws the f bits on the host before sampling
// (src/python/ksim/compile/channels.py:503-658, ChannelSampler: per channel a
// fire probability, a conditional outcome CDF and an XOR pattern over the f
// bits; geometric skips along the shot axis with numpy's PCG64).  Here each
// shot draws its own noise on the device from a counter-keyed SplitMix64
// stream (Vigna's published generator), walking the channels with
// geometric skips at the largest fire probability p_max and thinning each
// candidate to its own probability, all in 64-bit integer comparisons:
//
//   skip[k]  = floor((1 - p_max)^k 2^64)      P(next k channels all skip)
//   acc[c]   = ceil(p_c / p_max 2^64) - 1     candidate c fires iff v <= acc[c]
//   cdf[c,o] = ceil(Q_o 2^64) - 1             outcome = first o with v <= cdf[c,o]
//
// so a channel fires with probability p_c and picks outcome o with
// probability probs[o] / p_c, as in the reference, up to 2^-64 rounding.
#pragma once
#include <cstdint>
#include <vector>

namespace ksim {

struct NoiseChan {
  uint64_t acc;      // thinning threshold
  uint32_t n_out;    // non-identity outcomes
  uint32_t off_cdf;  // into cdf (u64)
  uint32_t off_pat;  // into pat (u32 words), n_out rows of wf words
  uint32_t pad;
};

struct NoiseHost {
  int32_t C = 0, num_f = 0, wf = 0;
  double p_max = 0;
  std::vector<uint64_t> skip;  // C + 1 entries, skip[0] = 2^64 - 1 (always)
  std::vector<NoiseChan> chan;
  std::vector<uint64_t> cdf;
  std::vector<uint32_t> pat;
  std::vector<int32_t> source;  // index of each kept channel in the input
};

// probs: concatenated outcome probabilities (sum of n_outcomes), outcome 0 =
// identity; patterns: for every channel, n_outcomes - 1 rows of num_f bytes
// (bit 0 used): the f bits outcome o flips.  Channels that cannot fire
// (p_fire <= 1e-15, as channels.py:606-607) or flip nothing are dropped.
NoiseHost build_noise(const double *probs, const int32_t *n_outcomes, const uint8_t *patterns,
                      int32_t n_channels, int32_t num_f);

}  // namespace ksim
