// Shared host-side types of the denoising engine: the plan of one problem
// (grid, Gram side, element types, workspace layout), per-stage timers, the
// error channel, and the per-element-type pipelines, each built in a
// translation unit of its own (pd_run_*.hip) so that they compile in parallel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "patch_denoise_amd.h"
#include "pd_common.h"

namespace pd {

// Message of the last failure on this thread (pdamd_last_error).
extern thread_local std::string g_err;
// Records the formatted message and returns `code`.
int fail(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess) return fail(PDAMD_EHIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

// Optional per-stage timing with HIP events on the work stream.
struct StageTimer {
  bool on = false;
  hipStream_t st = nullptr;
  struct Rec {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  int cur = -1;
  hipEvent_t cur_a = nullptr;
  void begin(int stage) {
    if (!on) return;
    cur = stage;
    hipEventCreate(&cur_a);
    hipEventRecord(cur_a, st);
  }
  void end() {
    if (!on || cur < 0) return;
    hipEvent_t b;
    hipEventCreate(&b);
    hipEventRecord(b, st);
    recs.push_back({cur, cur_a, b});
    cur = -1;
  }
  int collect(float* ms, int64_t* launches) {
    if (!on) return 0;
    hipError_t e = hipStreamSynchronize(st);
    for (auto& r : recs) {
      float t = 0.f;
      if (e == hipSuccess) hipEventElapsedTime(&t, r.a, r.b);
      if (ms) ms[r.stage] += t;
      if (launches) launches[r.stage] += 1;
      hipEventDestroy(r.a);
      hipEventDestroy(r.b);
    }
    recs.clear();
    return e == hipSuccess ? 0 : -1;
  }
};

struct Plan {
  Geom g;
  int64_t npatch, nvox;
  int n, kmax, kb, tile;
  int per_cu[4] = {0, 0, 0, 0};  // launch overrides: gram, tridiag, eig, synth (0 = default)
  bool tside;
  int etype;  // 0 float, 1 double, 2 cpx<float>, 3 cpx<double>
  int rule_f64;
  bool bf16x3 = false;  // split-BF16 Gram (FP32 data, M-side MFMA path)
  bool gather = false;  // gather recombination (pd_gather.h) instead of the atomic scatter
  int64_t chunk;
  size_t ebytes, rbytes;
  // workspace offsets
  size_t off_MEAN = 0;
  size_t off_cnt, off_flags, off_btot, off_idx, off_count, off_H, off_DE, off_TAU, off_Y, off_Z,
      off_WC, off_NC, off_TH, off_w3, off_SCR, total;
  // gather recombination only (per call, indexed by selected-patch slot)
  size_t off_GMAP = 0, off_GNC = 0, off_GTH = 0, off_GMEAN = 0, off_GY = 0, off_GV = 0, off_GRF = 0,
         off_GW3 = 0;
  // FP64 re-check of near-threshold decisions (float32 data, FP32 eigen
  // stage, MP-PCA / Hybrid): per-patch flags, the list of at most rcap
  // flagged patches, and that list's FP64 Gram / Householder / eigen buffers
  bool refine = false;
  int rcap = 0;
  size_t off_RFLAG = 0, off_RCNT = 0, off_RSLOT = 0, off_RIDX = 0, off_RH = 0, off_RDE = 0,
         off_RTAU = 0, off_RSCR = 0, off_RMEAN = 0, off_RY = 0, off_RZ = 0, off_RWC = 0,
         off_RNC = 0, off_RTH = 0, off_RRK = 0, off_RNZ = 0, off_REG = 0;
  // backward only
  size_t off_P1 = 0, off_P2 = 0, off_PE = 0, off_GS = 0, off_S1 = 0, off_UV = 0, off_THB = 0,
         off_GH = 0;
};


// Forward pipeline over the selected patches (chunks of Gram, Householder,
// eigen-rule and synthesis launches, then the recombination), one per
// (input, eigen-stage) element type: float32 / float32, float32 / float64,
// float64, complex64 / complex64, complex64 / complex128, complex128.
using RunFn = int (*)(pdamd_problem*, const Plan&, char*, int64_t, hipStream_t, StageTimer&);
int run_f32_f32(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, hipStream_t st,
                StageTimer& tm);
int run_f32_f64(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, hipStream_t st,
                StageTimer& tm);
int run_f64(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, hipStream_t st,
            StageTimer& tm);
int run_c64_c64(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, hipStream_t st,
                StageTimer& tm);
int run_c64_c128(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, hipStream_t st,
                 StageTimer& tm);
int run_c128(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, hipStream_t st,
             StageTimer& tm);

// FP64 re-check (pd_run_f32e64.hip) of the listed near-threshold patches of
// a float32 problem: the linear patch indices at ws + off_RIDX, their count
// (<= rcap) at ws + off_RCNT; Gram, tridiagonalisation and eigen rule in
// FP64 into the off_R* buffers (eigenvectors n x kmax per list entry).
int refine_f32_f64(pdamd_problem* P, const Plan& pl, char* ws, hipStream_t st);

// Householder tridiagonalisation (pd_run_f64.hip) of `batch` symmetric
// float64 n x n matrices at H (both triangles; overwritten by reflectors):
// d | e | e^2 rows of 3n into DE, tau into TAU.
int tridiag_f64(double* H, double* DE, double* TAU, int batch, int n, hipStream_t st);

// Backward of the optimal rule (real data): float32 / float32, float32 /
// float64, float64.
int backward_f32_f32(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, const void* gout,
                     const float* w4, void* gin, double* glam, hipStream_t st, StageTimer& tm);
int backward_f32_f64(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, const void* gout,
                     const float* w4, void* gin, double* glam, hipStream_t st, StageTimer& tm);
int backward_f64(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, const void* gout,
                 const float* w4, void* gin, double* glam, hipStream_t st, StageTimer& tm);

}  // namespace pd
