// Leaf panels for large stacks (at least four matrices per CU, m <= 256 rows):
// one or two waves per matrix, working precision float.
#include "lu_leaf.hpp"

namespace npph {

int leaf_launch_thin(const LeafArgs<float>& la, int m, dim3 grid, size_t shm, hipStream_t s) {
  if (m <= 64) launch_leaf<float, 1, 64, 64>(grid, shm, s, la);
  else if (m <= 128) launch_leaf<float, 1, 64, 128>(grid, shm, s, la);
  // two rows per lane up to 256 rows
  else if (m <= 256) launch_leaf<float, 2, 64, 128>(grid, shm, s, la);
  else { set_error("leaf: thin panel of %d rows", m); return NPPH_ERR_INVALID; }
  return NPPH_OK;
}

}  // namespace npph
