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

int tridiag_f64(double* H, double* DE, double* TAU, int batch, int n, hipStream_t st) {
  Plan pl;
  pl.n = n;
  pl.tile = n <= 32 ? 2 : (n <= 64 ? 4 : (n <= 128 ? 8 : 0));
  GramArgs ga{};
  ga.np = batch;
  ga.n = n;
  ga.H = H;
  ga.DE = DE;
  ga.TAU = TAU;
  ga.np_dev = nullptr;
  return launch_tridiag_tile<double>(pl, ga, st);
}

}  // namespace pd
