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

int refine_f32_f64(pdamd_problem* P, const Plan& pl, char* ws, hipStream_t st) {
  const int* cnt = reinterpret_cast<const int*>(ws + pl.off_RCNT);
  GramArgs ga = gram_args(P, pl, ws, reinterpret_cast<const int32_t*>(ws + pl.off_RIDX), pl.rcap);
  ga.np_dev = cnt;
  ga.H = ws + pl.off_RH;
  ga.DE = ws + pl.off_RDE;
  ga.TAU = ws + pl.off_RTAU;
  ga.MEAN = ws + pl.off_RMEAN;
  ga.MEANW = nullptr;
  int rc = pl.tside ? launch_gram_tile<float, double, true>(pl, P->input_dev, ga, st)
                    : launch_gram_tile<float, double, false>(pl, P->input_dev, ga, st);
  if (rc) return rc;
  rc = launch_tridiag_tile<double>(pl, ga, st);
  if (rc) return rc;
  EigArgs ea = eig_args(P, pl, ws, ga);
  ea.np_dev = cnt;
  ea.Y = ws + pl.off_RY;
  ea.Z = ws + pl.off_RZ;
  ea.SCR = ws + pl.off_RSCR;
  ea.WC = reinterpret_cast<double*>(ws + pl.off_RWC);
  ea.NCOMP = reinterpret_cast<int*>(ws + pl.off_RNC);
  ea.THETA = reinterpret_cast<double*>(ws + pl.off_RTH);
  ea.RANK = reinterpret_cast<int32_t*>(ws + pl.off_RRK);
  ea.NOISE = reinterpret_cast<float*>(ws + pl.off_RNZ);
  ea.EIG = reinterpret_cast<float*>(ws + pl.off_REG);
  return launch_eig<double, float>(pl, ea, st);
}

}  // namespace pd
