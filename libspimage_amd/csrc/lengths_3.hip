// Kernel instantiations for one family of transform lengths.
#include "passes.h"

namespace spx {
// host-only table (a function-local static keeps it out of device code)
const LengthEntry* table_lengths_3(int* n) {
  static const LengthEntry t[] = {
      length_entry<3>(),
      length_entry<6>(),
      length_entry<9>(),
      length_entry<12>(),
      length_entry<18>(),
      length_entry<24>(),
      length_entry<27>(),
      length_entry<36>(),
      length_entry<48>(),
      length_entry<54>(),
      length_entry<72>(),
      length_entry<96>(),
      length_entry<108>(),
      length_entry<144>(),
      length_entry<192>(),
      length_entry<216>(),
      length_entry<288>(),
      length_entry<384>(),
      length_entry<432>(),
      length_entry<576>(),
      length_entry<768>(),
      length_entry<864>(),
      length_entry<1152>(),
      length_entry<1536>(),
      length_entry<1728>(),
      length_entry<2304>(),
      length_entry<3072>()};
  *n = (int)(sizeof(t) / sizeof(t[0]));
  return t;
}
}  // namespace spx
