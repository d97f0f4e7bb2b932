// fast5_reader.h -- multi-read FAST5 input (HDF5 C API) for the drop-in CLI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace dtwax_io {

struct ChannelParams {
  double digitisation = 0, offset = 0, range = 0, sample_rate = 0;
  bool ok = false;
};

struct Fast5Read {
  std::string group;          // "read_<uuid>"
  std::vector<float> samples; // raw ADC values as float (no pA scaling)
  ChannelParams channel;
};

// "*.fast5" entries of `dir` (and "*.pod5" ones when with_pod5) in readdir
// order, each as dir + name (the folder argument is expected to end in '/',
// as the reference concatenates it; src/include/generate_load_squiggle.hpp:132-157).
std::vector<std::string> list_fast5(const std::string& dir, bool* opened, bool with_pod5 = false);

// Read groups of one multi-read file in HDF5 name order
// (fast5/nanopolish_fast5_io.cpp:72-112).  Returns false if the file cannot
// be opened.  err receives a message for unreadable reads or files.
bool read_fast5_file(const std::string& path, std::vector<Fast5Read>& out, std::string& err);

}  // namespace dtwax_io
