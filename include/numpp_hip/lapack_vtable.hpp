// numpp::LapackVTable as provided by the MI355X HIP backend.
//
// Layout: the reference struct (include/numpp/backend/lapack_vtable.hpp:14-27 in
// NumPP) member for member. This provider fills gesv and inv; potrf, gesdd,
// geev and heevd are null (NumPP's portable kernels run for them). When built
// inside NumPP, the real numpp/core/dtype.hpp provides DTypeId (see
// gpu_vtable.hpp in this directory for the standalone declaration).
#pragma once

#include "numpp_hip/gpu_vtable.hpp"

namespace numpp {

struct LapackVTable {
  // Solve A X = B in place (A: n x n -> packed LU, B: n x nrhs -> X), in dt.
  bool (*gesv)(int n, int nrhs, DTypeId dt, void* A, void* B);
  // Inverse of A (n x n) into Ainv, in dt.
  bool (*inv)(int n, DTypeId dt, const void* A, void* Ainv);
  bool (*potrf)(int n, DTypeId dt, const void* A, void* L);
  bool (*gesdd)(int m, int n, DTypeId dt, const void* A, void* U, void* S, void* Vh, bool full);
  bool (*geev)(int n, DTypeId dt, const void* A, void* W, void* V);
  bool (*heevd)(int n, DTypeId dt, const void* A, void* W, void* V);
};

// Null unless a usable gfx950 device is present (lapack_vtable.hpp:30).
const LapackVTable* lapack_vtable();

}  // namespace numpp
