// MT(19,9) parameter rows for the RNG streams.
//
// Built-in: the 1024 twist vectors of mt_params_table.inc (own primitive
// search, see mt_params_gen.cpp) with the MT19937 tempering masks.
// Optional: a file of 16-byte records {matrix_a, mask_b, mask_c, seed} in
// the layout of the reference's mt_struct_stripped (include/MersenneTwister.h:
// 23-28), e.g. its data/MersenneTwister.dat.
#include <fstream>

#include "host_internal.h"

namespace dsmc_host {

static const uint32_t kBuiltinA[] = {
#include "mt_params_table.inc"
};

const uint32_t* builtin_mt_table(uint32_t* rows) {
  *rows = (uint32_t)(sizeof(kBuiltinA) / sizeof(kBuiltinA[0]));
  return kBuiltinA;
}

bool load_mt_rows(const std::string& path, std::vector<uint32_t>& a, std::vector<uint32_t>& b,
                  std::vector<uint32_t>& c, std::string& err) {
  std::ifstream in(path, std::ios::binary);
  if (!in) {
    err = "cannot open MT parameter file " + path;
    return false;
  }
  std::vector<uint32_t> buf;
  uint32_t rec[4];
  while (in.read((char*)rec, sizeof(rec))) {
    a.push_back(rec[0]);
    b.push_back(rec[1]);
    c.push_back(rec[2]);
  }
  if (a.empty()) {
    err = "empty MT parameter file " + path;
    return false;
  }
  for (uint32_t x : a)
    if (!(x & 0x80000000u)) {
      err = "MT parameter row with a singular twist (top bit clear) in " + path;
      return false;
    }
  return true;
}

}  // namespace dsmc_host
