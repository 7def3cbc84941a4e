// Drop-in check of numpp::gpu_vtable() exactly as NumPP's dispatcher would use
// it (backend.cpp:144-152 calls vt->gemm on host buffers). Built with g++
// against libnumpp_hip.so by tests/test_abi.py.
//   no device: the accessor must be null (gpu_vtable.hpp:56 contract) -> exit 0
//   device:    gemm, gesv_batched, det, slogdet and inv slots must reproduce
//              small known answers.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "numpp_hip/gpu_vtable.hpp"

int main(int argc, char** argv) {
  const bool expect_device = argc > 1 && std::strcmp(argv[1], "--device") == 0;
  const numpp::GpuVTable* vt = numpp::gpu_vtable();
  if (!vt) {
    std::printf("vtable: null (no device)\n");
    return expect_device ? 1 : 0;
  }
  if (!expect_device) { std::printf("vtable present on a host without device?\n"); return 1; }
  std::printf("vtable: %s\n", vt->name);
  // [[1,2,3],[4,5,6]] @ [[1,2],[3,4],[5,6]] = [[22,28],[49,64]] (test_backend.cpp:46-55)
  double a[6] = {1, 2, 3, 4, 5, 6}, b[6] = {1, 2, 3, 4, 5, 6}, c[4] = {0, 0, 0, 0};
  if (!vt->gemm || !vt->gemm(numpp::DTypeId::Float64, 2, 2, 3, a, b, c)) return 2;
  if (c[0] != 22 || c[1] != 28 || c[2] != 49 || c[3] != 64) { std::printf("gemm wrong\n"); return 3; }
  float af[6] = {1, 2, 3, 4, 5, 6}, bf[6] = {1, 2, 3, 4, 5, 6}, cf[4];
  if (!vt->gemm(numpp::DTypeId::Float32, 2, 2, 3, af, bf, cf) || cf[3] != 64.0f) return 4;
  if (vt->gemm(numpp::DTypeId::Int64, 2, 2, 3, a, b, c)) return 5;  // unsupported dtype -> false
  // stack3 = arange(12).reshape(3,2,2), b = arange(1,7).reshape(3,2) (test_batched_linalg.cpp:35-42)
  double A[12], B[6];
  for (int i = 0; i < 12; ++i) A[i] = i;
  for (int i = 0; i < 6; ++i) B[i] = i + 1;
  int32_t info[3] = {9, 9, 9};
  if (!vt->gesv_batched || !vt->gesv_batched(numpp::DTypeId::Float64, 3, 2, 1, A, B, nullptr, info)) return 6;
  // each det = -2; x = A^-1 b checked by reconstruction
  for (int m = 0; m < 3; ++m) {
    if (info[m] != 0) return 7;
    const double a00 = 4 * m, a01 = 4 * m + 1, a10 = 4 * m + 2, a11 = 4 * m + 3;
    const double r0 = a00 * B[2 * m] + a01 * B[2 * m + 1] - (2 * m + 1);
    const double r1 = a10 * B[2 * m] + a11 * B[2 * m + 1] - (2 * m + 2);
    if (std::fabs(r0) > 1e-12 || std::fabs(r1) > 1e-12) { std::printf("gesv wrong m=%d\n", m); return 8; }
  }
  // det / slogdet / inv of the same stack (test_batched_linalg.cpp:21-32, 66-74):
  // every det is -2, slogdet (-1, log 2), and A inv(A) = I
  for (int i = 0; i < 12; ++i) A[i] = i;
  double det[3], sg[3], la[3], Ai[12];
  if (!vt->det_batched || !vt->det_batched(numpp::DTypeId::Float64, 3, 2, A, det, info)) return 9;
  if (!vt->slogdet_batched(numpp::DTypeId::Float64, 3, 2, A, sg, la, info)) return 10;
  if (!vt->inv_batched(numpp::DTypeId::Float64, 3, 2, A, Ai, info)) return 11;
  for (int m = 0; m < 3; ++m) {
    if (std::fabs(det[m] + 2.0) > 1e-13 || sg[m] != -1.0 || std::fabs(la[m] - std::log(2.0)) > 1e-14) {
      std::printf("det/slogdet wrong m=%d: %g %g %g\n", m, det[m], sg[m], la[m]);
      return 12;
    }
    const double* a = A + 4 * m;
    const double* x = Ai + 4 * m;
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        if (std::fabs(a[2 * i] * x[j] + a[2 * i + 1] * x[2 + j] - (i == j)) > 1e-12) return 13;
  }
  // singular: zeros (test_linalg.cpp:56-59) -> det 0, inv reports info > 0
  double Z[4] = {0, 0, 0, 0}, dz = 1.0, Zi[4];
  int32_t iz = 0;
  if (!vt->det_batched(numpp::DTypeId::Float64, 1, 2, Z, &dz, &iz) || dz != 0.0 || iz != 1) return 14;
  if (!vt->inv_batched(numpp::DTypeId::Float64, 1, 2, Z, Zi, &iz) || iz != 1) return 15;
  std::printf("vtable slots ok\n");
  return 0;
}
