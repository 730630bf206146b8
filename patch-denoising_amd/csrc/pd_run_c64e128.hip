// Forward pipeline for complex64 data with a complex128 eigen stage.
#include "pd_kernels.inl"

namespace pd {

int run_c64_c128(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, hipStream_t st,
          StageTimer& tm) {
  return run_chunks<cpx<float>, cpx<double>, float>(P, pl, ws, count, st, tm);
}

}  // namespace pd
