// Leaf panels for large stacks (at least four matrices per CU, m <= 256 rows):
// one or two waves per matrix, working precision double.
#include "lu_leaf.hpp"

namespace npph {

int leaf_launch_thin(const LeafArgs<double>& la, int m, dim3 grid, size_t shm, hipStream_t s) {
  if (m <= 64) launch_leaf<double, 1, 64, 64>(grid, shm, s, la);
  else if (m <= 128) launch_leaf<double, 1, 64, 128>(grid, shm, s, la);
  // FP64 rows take twice the registers: 16-column inner blocks between 128 and 256 rows (measured)
  else if (m <= 256) launch_leaf<double, 2, 16, 128>(grid, shm, s, la);
  else { set_error("leaf: thin panel of %d rows", m); return NPPH_ERR_INVALID; }
  return NPPH_OK;
}

}  // namespace npph
