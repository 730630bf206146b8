// Pipelines for float64 data (forward and the optimal rule's backward).
#include "pd_kernels.inl"

namespace pd {

int run_f64(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, hipStream_t st,
          StageTimer& tm) {
  return run_chunks<double, double, double>(P, pl, ws, count, st, tm);
}

int backward_f64(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, const void* gout,
          const float* w4, void* gin, double* glam, hipStream_t st, StageTimer& tm) {
  return run_backward<double, double, double>(P, pl, ws, count, gout, w4, gin, glam, st, tm);
}

}  // namespace pd
