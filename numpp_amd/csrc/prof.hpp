// Optional per-kernel-class timing with HIP events (off by default). Each
// launch wrapper opens a ProfScope with the algorithmic flops and bytes of that
// launch; npph_prof_get sums elapsed time per class after synchronising.
#pragma once

#include "common.hpp"

namespace npph {

enum ProfKind {
  kProfGemmF64 = 0,
  kProfGemmF32,
  kProfGemmBf16,
  kProfPanel,
  kProfMoveTrsm,
  kProfGetrs,
  kProfResidual,
  kProfSmallLu,
  kProfVbatch,
  kProfKinds
};

bool prof_enabled();
void prof_begin(int kind, double flops, double bytes, hipStream_t s, void** token);
void prof_end(void* token, hipStream_t s);

struct ProfScope {
  void* tok = nullptr;
  hipStream_t s;
  ProfScope(int kind, double flops, double bytes, hipStream_t st) : s(st) {
    if (kind >= 0 && kind < kProfKinds && prof_enabled()) prof_begin(kind, flops, bytes, st, &tok);
  }
  ~ProfScope() {
    if (tok) prof_end(tok, s);
  }
};

}  // namespace npph
