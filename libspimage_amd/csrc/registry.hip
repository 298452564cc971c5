// Registry of the supported transform lengths (products of 2, 3 and 5).
#include "passes.h"

namespace spx {
const LengthEntry* table_lengths_15(int* n);
const LengthEntry* table_lengths_3(int* n);
const LengthEntry* table_lengths_5(int* n);
const LengthEntry* table_lengths_pow2(int* n);

const LengthEntry* find_length(int L) {
  typedef const LengthEntry* (*TableFn)(int*);
  static const TableFn fns[] = {table_lengths_15, table_lengths_3, table_lengths_5, table_lengths_pow2};
  for (TableFn f : fns) {
    int n = 0;
    const LengthEntry* t = f(&n);
    for (int i = 0; i < n; ++i)
      if (t[i].L == L) return &t[i];
  }
  return nullptr;
}

}  // namespace spx
