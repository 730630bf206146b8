// Forward pipeline for complex64 data with a complex64 eigen stage.
#include "pd_kernels.inl"

namespace pd {

int run_c64_c64(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, hipStream_t st,
          StageTimer& tm) {
  return run_chunks<cpx<float>, cpx<float>, float>(P, pl, ws, count, st, tm);
}

}  // namespace pd
