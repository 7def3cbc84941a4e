// Leaf panels with 256 or 512 threads per matrix (any height up to 8 x 512
// rows), working precision double. Inner blocks are as wide as the registers
// allow (fewer block-level global round trips per leaf); FP64 rows take twice the registers of FP32 ones.
#include "lu_leaf.hpp"

namespace npph {

int leaf_launch_wide(const LeafArgs<double>& la, int m, dim3 grid, size_t shm, hipStream_t s) {
  // short panels: 256 threads (half of a 512-thread workgroup would hold no
  // rows, and the per-column barrier spans fewer waves); taller panels
  // measured faster with 512 threads and fewer rows per thread
  // Panels above 1024 rows (four rows per thread) take 16-column inner
  // blocks: each block re-reads the L columns left of it from HBM with one CU
  // per matrix, and 16 instead of 8 columns per block halve that re-read
  // (measured: 64 x 2048^2 FP64 panels 6.85 -> 5.91 ms per factorisation)
  if (m <= PT / 2) launch_leaf<double, 1, 64, PT / 2>(grid, shm, s, la);
  else if (m <= PT) launch_leaf<double, 1, 64>(grid, shm, s, la);
  else if (m <= 2 * PT) launch_leaf<double, 2, 16>(grid, shm, s, la);
  else if (m <= 4 * PT) launch_leaf<double, 4, 16>(grid, shm, s, la);
  else if (m <= 8 * PT) launch_leaf<double, 8, 4>(grid, shm, s, la);
  else { set_error("leaf: panel of %d rows", m); return NPPH_ERR_INVALID; }
  return NPPH_OK;
}

}  // namespace npph
