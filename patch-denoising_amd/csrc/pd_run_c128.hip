// Forward pipeline for complex128 data.
#include "pd_kernels.inl"

namespace pd {

int run_c128(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, hipStream_t st,
          StageTimer& tm) {
  return run_chunks<cpx<double>, cpx<double>, double>(P, pl, ws, count, st, tm);
}

}  // namespace pd
