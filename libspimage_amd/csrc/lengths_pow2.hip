// Kernel instantiations for one family of transform lengths.
#include "passes.h"

namespace spx {
// host-only table (a function-local static keeps it out of device code)
const LengthEntry* table_lengths_pow2(int* n) {
  static const LengthEntry t[] = {
      length_entry<2>(),
      length_entry<4>(),
      length_entry<8>(),
      length_entry<16>(),
      length_entry<32>(),
      length_entry<64>(),
      length_entry<128>(),
      length_entry<256>(),
      length_entry<512>(),
      length_entry<1024>(),
      length_entry<2048>(),
      length_entry<4096>()};
  *n = (int)(sizeof(t) / sizeof(t[0]));
  return t;
}
}  // namespace spx
