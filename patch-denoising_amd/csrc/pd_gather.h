// Gather-based recombination (pd_gather.hip): replaces the atomic
// scatter-accumulate of the synthesis for float32 patches with at most 10
// components (MP-PCA, Hybrid), one temporal patch, weighted / average
// recombination.
//
// Per selected patch (slot j, patches in increasing linear index), M-side
// (Gram Xc Xc^T, eigenvectors U):
//   Lf_j = U_j (M x KM rows of the Gram's kept eigenvectors, zero past the
//          kept count) -- the eigen stage's per-call eigenvector buffer;
//   Rf_j = (KM + 1) x T: row 0 = theta_j m_j (per-frame mean), row 1 + k =
//          theta_j wc_k (U_j^T Xc_j)[k] (zero past the kept count);
// T-side (Gram Xc^T Xc, eigenvectors V, T x KM in a per-call buffer):
//   Lf_j = Xc_j V_j diag(wc) (M x KM), Rf_j row 0 = theta_j m_j, row 1 + k
//          = theta_j V_j[:, k]^T;
// so that theta_j * (denoised patch j) = Lf_j Rf_j[1:] + 1 Rf_j[0].
// Every output element is then the sum of its covering patches' rows in
// increasing patch index (the reference's C-order loop), divided by the
// float32 weight sum, without atomics: results repeat bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pd_common.h"

namespace pd {

constexpr int kGatherKM = 10;  // components per patch in the factor layout

struct GatherArgs {
  Geom g;
  int64_t shape[3];      // spatial sizes
  int n;                 // rows of a left factor (patch voxels M)
  int np;                // patches of this launch (factor kernel: the chunk)
  int slot0;             // factor kernel: slot of the chunk's first patch
  int recomb;            // PDAMD_WEIGHTED / PDAMD_AVERAGE
  int theta_weak;        // 1: weights add theta in float32 (Hybrid's Python-float theta)
  const int32_t* idx;    // selected linear patch indices (all slots)
  const float* vol;      // input volume
  const float* MEAN;     // slots x T per-frame means
  const float* Y;        // slots x M x KM left factors Lf (M-side: the eigenvectors)
  float* YT;             // T-side: the left factors, written by the factor kernel
  const float* V;        // T-side: slots x T x KM eigenvectors (zero past the kept count)
  int tside;             // 1: T-side factors
  const double* WC;      // chunk x KM component weights (factor kernel)
  const int* NCOMP;      // slots: kept components
  const double* THETA;   // slots: 1 / (2 + rank)
  float* RF;             // slots x (KM + 1) x T
  int32_t* MAP;          // linear patch index -> slot * 16 + ncomp, or -1
  float* W3;             // spatial float32 weight sums
  const uint8_t* mask;   // mask bytes (kind below) or null
  int mask_kind;
  float* out;            // output volume
  float* wout;           // float32 weights, volume shaped
  // factor kernel over a list instead of a chunk: blocks q < *list_cnt
  // take slot list[q] (component weights 1, as for the eigen rules); the
  // launch has list_cap blocks
  const int32_t* list = nullptr;
  const int* list_cnt = nullptr;
  int list_cap = 0;
};

// Each enqueues on `st` and returns 0, or -1 if the launch failed.
int launch_gather_factors(const GatherArgs& a, hipStream_t st);
int launch_gather_map(const GatherArgs& a, int64_t npatch, int64_t nsel, hipStream_t st);
int launch_gather_weights(const GatherArgs& a, hipStream_t st);
int launch_gather_recombine(const GatherArgs& a, hipStream_t st);

}  // namespace pd
