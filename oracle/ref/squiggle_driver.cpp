// Driver for the reference's own squiggle build (test infrastructure).
//
// Compiled against the reference checkout's src/include/load_reference.hpp
// in its binary32 mode (no -DFP16: common.hpp then needs no CUDA header),
// where load_ref_coeffs stores the z-scored float levels of both strands at
// offsets 0 and REF_LEN/2 -- the values its FP16 build rounds to half
// (load_reference.hpp:33-108).  Usage:
//   ref_squiggle <kmer_model> <ref.fasta> <out.f32>
// writes REF_LEN floats (forward then reverse complement, REF_LEN/2 each).
#include <cstdint>  // the header uses int64_t/int8_t and std::istringstream
#include <cstdio>   // without including them (its CUDA build gets them elsewhere)
#include <sstream>

#include "load_reference.hpp"

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s <kmer_model> <ref.fasta> <out.f32>\n", argv[0]);
    return 2;
  }
  load_reference loader;
  loader.ref_loader(argv[2]);
  loader.read_kmer_model(argv[1]);
  reference_coefficients* ref = new reference_coefficients[REF_LEN];
  for (int i = 0; i < REF_LEN; i++) ref[i].coeff1 = 0.0f;
  loader.load_ref_coeffs(ref);
  std::FILE* f = std::fopen(argv[3], "wb");
  if (!f) return 1;
  for (int i = 0; i < REF_LEN; i++) std::fwrite(&ref[i].coeff1, sizeof(float), 1, f);
  std::fclose(f);
  delete[] ref;
  return 0;
}
