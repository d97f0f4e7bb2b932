// squiggle.cpp -- expected-current reference squiggle from a sequence and a
// k-mer pore model, both strands (dtwax_squiggle_build).
//
// Same arithmetic as load_reference::load_ref_coeffs
// (src/include/load_reference.hpp:33-108): per strand, over the first
// ref_len k-mers, float sequential sums of the model level and of its
// square; mean = sum/ref_len, sd = sqrt(sumsq/ref_len - mean^2) in float;
// coeff = (level - mean)/sd in float.  The reference finds each k-mer by a
// linear scan of a std::map (O(ref_len * 4^k)); here a 2-bit rolling code
// indexes the level table directly, which visits the same levels in the same
// order.  K-mers with a non-ACGT base add nothing to the sums, still count in
// ref_len, and leave their output slot at 0 (the reference never writes it).
// The reverse strand is the reverse complement (load_reference.hpp:111-128,
// 196-199: only A, C, G, T are complemented).

#include <cmath>
#include <string>
#include <vector>

#include "dtwax.h"

static int base_code(char ch) {
  switch (ch) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    default: return -1;
  }
}

static char complement_base(char ch) {
  switch (ch) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    default: return ch;
  }
}

// codes[i] = 2-bit code of the k-mer starting at i, or -1 (invalid / short).
static void kmer_codes(const std::string& s, int k, int64_t n, std::vector<int32_t>& codes) {
  codes.assign((size_t)n, -1);
  const int64_t L = (int64_t)s.size();
  const uint32_t mask = (k >= 16) ? 0xffffffffu : ((1u << (2 * k)) - 1u);
  uint32_t code = 0;
  int valid = 0;  // consecutive valid bases ending at position j
  for (int64_t j = 0; j < L && j < n + k - 1; j++) {
    int b = base_code(s[(size_t)j]);
    if (b < 0) {
      valid = 0;
      code = 0;
      continue;
    }
    code = ((code << 2) | (uint32_t)b) & mask;
    valid++;
    const int64_t start = j - k + 1;
    if (valid >= k && start >= 0 && start < n) codes[(size_t)start] = (int32_t)code;
  }
}


extern "C" int dtwax_squiggle_build(const char* seq, int64_t seq_len, int32_t k, const float* model_means,
                                    int64_t ref_len, float* out, int64_t* n_skipped) {
  if (!seq || !model_means || !out || k < 1 || k > 12 || ref_len < 1 || seq_len < 0) return DTWAX_EINVAL;
  std::string fwd(seq, (size_t)seq_len);
  std::string rev(fwd.rbegin(), fwd.rend());
  for (auto& ch : rev) ch = complement_base(ch);
  const std::string* target[2] = {&fwd, &rev};
  int64_t skipped = 0;
  std::vector<int32_t> codes;
  for (int j = 0; j < 2; j++) {
    kmer_codes(*target[j], k, ref_len, codes);
    float mean1 = 0.0f, stdev1 = 0.0f;
    for (int64_t i = 0; i < ref_len; i++) {
      const int32_t c = codes[(size_t)i];
      if (c < 0) continue;
      const float lv = model_means[c];
      mean1 += lv;
      stdev1 = stdev1 + lv * lv;
    }
    mean1 = mean1 / (float)ref_len;
    stdev1 = stdev1 / (float)ref_len;
    stdev1 = std::sqrt(stdev1 - (mean1 * mean1));
    float* o = out + (int64_t)j * ref_len;
    for (int64_t i = 0; i < ref_len; i++) {
      const int32_t c = codes[(size_t)i];
      if (c < 0) {
        o[i] = 0.0f;
        skipped++;
        continue;
      }
      o[i] = (model_means[c] - mean1) / stdev1;
    }
  }
  if (n_skipped) *n_skipped = skipped;
  return DTWAX_OK;
}
