// Pipelines for float32 data with a float64 eigen stage (forward and the
// optimal rule's backward).
#include "pd_kernels.inl"

namespace pd {

int run_f32_f64(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, hipStream_t st,
          StageTimer& tm) {
  return run_chunks<float, double, float>(P, pl, ws, count, st, tm);
}

int backward_f32_f64(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, const void* gout,
          const float* w4, void* gin, double* glam, hipStream_t st, StageTimer& tm) {
  return run_backward<float, double, float>(P, pl, ws, count, gout, w4, gin, glam, st, tm);
}

}  // namespace pd
