// Kernel instantiations for one family of transform lengths.
#include "passes.h"

namespace spx {
// host-only table (a function-local static keeps it out of device code)
const LengthEntry* table_lengths_15(int* n) {
  static const LengthEntry t[] = {
      length_entry<15>(),
      length_entry<30>(),
      length_entry<45>(),
      length_entry<60>(),
      length_entry<75>(),
      length_entry<90>(),
      length_entry<120>(),
      length_entry<135>(),
      length_entry<150>(),
      length_entry<180>(),
      length_entry<225>(),
      length_entry<240>(),
      length_entry<270>(),
      length_entry<300>(),
      length_entry<360>(),
      length_entry<375>(),
      length_entry<405>(),
      length_entry<450>(),
      length_entry<480>(),
      length_entry<540>(),
      length_entry<600>(),
      length_entry<720>(),
      length_entry<750>(),
      length_entry<810>(),
      length_entry<900>(),
      length_entry<960>(),
      length_entry<1080>(),
      length_entry<1200>(),
      length_entry<1440>(),
      length_entry<1500>(),
      length_entry<1620>(),
      length_entry<1800>(),
      length_entry<1920>(),
      length_entry<2160>(),
      length_entry<2400>(),
      length_entry<2880>(),
      length_entry<3000>(),
      length_entry<3240>(),
      length_entry<3600>(),
      length_entry<3840>()};
  *n = (int)(sizeof(t) / sizeof(t[0]));
  return t;
}
}  // namespace spx
