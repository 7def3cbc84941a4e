// Drop-in check of numpp::gpu_vtable() exactly as NumPP's dispatcher would use
// it (backend.cpp:144-152 calls vt->gemm on host buffers). Built with g++
// against libnumpp_hip.so by tests/test_abi.py.
//   no device: the accessor must be null (gpu_vtable.hpp:56 contract) -> exit 0
//   device:    gemm and gesv_batched slots must reproduce small known answers.
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
  std::printf("vtable slots ok\n");
  return 0;
}
