// Kernel instantiations for one family of transform lengths.
#include "passes.h"

namespace spx {
// host-only table (a function-local static keeps it out of device code)
const LengthEntry* table_lengths_5(int* n) {
  static const LengthEntry t[] = {
      length_entry<5>(),
      length_entry<10>(),
      length_entry<20>(),
      length_entry<25>(),
      length_entry<40>(),
      length_entry<50>(),
      length_entry<80>(),
      length_entry<100>(),
      length_entry<125>(),
      length_entry<160>(),
      length_entry<200>(),
      length_entry<250>(),
      length_entry<320>(),
      length_entry<400>(),
      length_entry<500>(),
      length_entry<640>(),
      length_entry<800>(),
      length_entry<1000>(),
      length_entry<1280>(),
      length_entry<1600>(),
      length_entry<2000>(),
      length_entry<2560>(),
      length_entry<3200>()};
  *n = (int)(sizeof(t) / sizeof(t[0]));
  return t;
}
}  // namespace spx
