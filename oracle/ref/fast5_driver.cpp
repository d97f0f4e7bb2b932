// Driver for the reference's own FAST5 reader (test infrastructure):
// fast5/nanopolish_fast5_io.cpp, used as src/include/generate_load_squiggle.hpp:
// 172-233 uses it.  Usage: ref_fast5 <file.fast5>...
// For every read group (HDF5 name order) prints one line:
//   <group> <n_samples> <samples [1952, 4000) if n >= 4000>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "nanopolish_fast5_io.h"

int main(int argc, char** argv) {
  for (int a = 1; a < argc; a++) {
    fast5_file f = fast5_open(argv[a]);
    if (!fast5_is_open(f)) continue;
    std::vector<std::string> groups = fast5_get_multi_read_groups(f);
    for (const std::string& g : groups) {
      const std::string name = g.substr(5);
      fast5_raw_scaling sc = fast5_get_channel_params(f, name);
      raw_table rt = fast5_get_raw_samples(f, name, sc);
      std::printf("%s %lld", g.c_str(), (long long)rt.n);
      if (rt.n >= 4000)
        for (int i = 1952; i < 4000; i++) std::printf(" %g", (double)rt.raw[i]);
      std::printf("\n");
      free(rt.raw);
    }
    fast5_close(f);
  }
  return 0;
}
