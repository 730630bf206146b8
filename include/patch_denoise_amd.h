/*
 * patch_denoise_amd.h -- C ABI of the MI355X (gfx950) space-time low-rank
 * patch denoiser.
 *
 * Plain pointers and sizes only.  Every pointer named *_dev is device
 * memory on the current HIP device; `stream` is a hipStream_t (NULL = the
 * null stream).  All work is enqueued on that stream.  Functions return 0 on
 * success and a negative PDAMD_E* code on error; pdamd_last_error() returns
 * a message for the last error on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the reference's
 * src/patch_denoise):
 *   pdamd_denoise            BaseSpaceTimeDenoiser.denoise, space_time/base.py:177-270,
 *                            with the per-method _patch_processing of
 *                            space_time/lowrank.py:42-63 (MP-PCA), :104-119 (Hybrid),
 *                            :170-184 (Raw SVT / NORDIC), :379-407 (Optimal)
 *   pdamd_select_patches     patch selection loop, space_time/base.py:201-227
 *   pdamd_workspace_size     (no counterpart: the caller allocates device scratch)
 *   pdamd_last_error, pdamd_version  (no counterpart: error reporting)
 *
 * A "patch batch" of independent Casorati matrices (the plugin method
 * _patch_processing, space_time/base.py:352-357) is expressed through
 * pdamd_denoise with a volume of shape (B, M, 1, T), patch (1, M, 1, T),
 * step (1, 1, 1, 1), average recombination and no mask: the output is then the
 * denoised batch itself.
 */
#ifndef PATCH_DENOISE_AMD_H
#define PATCH_DENOISE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types of the 4-D volume */
enum { PDAMD_F32 = 0, PDAMD_F64 = 1, PDAMD_C64 = 2, PDAMD_C128 = 3 };

/* threshold rules */
enum {
  PDAMD_MPPCA = 0,   /* lowrank.py:42-63   Marchenko-Pastur PCA on the top-10 eigenvalues */
  PDAMD_HYBRID = 1,  /* lowrank.py:104-119 Hybrid-PCA against a prior variance map       */
  PDAMD_HARD = 2,    /* lowrank.py:170-184 hard singular-value threshold (raw SVT, NORDIC) */
  PDAMD_OPTIMAL = 3, /* lowrank.py:379-407 optimal shrinkage (Gavish-Donoho)             */
  PDAMD_ADAPTIVE = 4 /* lowrank.py:410-636 adaptive trace-norm shrinkage (SURE/GSURE/QUT) */
};

/* adaptive parameter selection (lowrank.py:410-540) */
enum { PDAMD_ADAPT_SURE = 0, PDAMD_ADAPT_GSURE = 1, PDAMD_ADAPT_QUT = 2 };

/* recombination (base.py:243-256) */
enum { PDAMD_WEIGHTED = 0, PDAMD_AVERAGE = 1, PDAMD_CENTER = 2 };

/* optimal-shrinkage losses (lowrank.py:297-303) */
enum { PDAMD_LOSS_FRO = 0, PDAMD_LOSS_NUC = 1, PDAMD_LOSS_OPE = 2 };

/* mask kinds */
enum { PDAMD_MASK_NONE = 0, PDAMD_MASK_SPATIAL = 1, PDAMD_MASK_FULL = 2 };

/* eigen-stage precision */
/* PDAMD_PREC_BF16X3: FP32 eigen stage with the Gram from split-BF16
   three-product MFMA with FP32 accumulation (complex: four real products)
   for FP32 real data with Gram size 65..128 (either side) and for FP32
   real / complex64 data with Gram size 129..512; other shapes use the FP32
   Gram.                                                                     */
enum { PDAMD_PREC_AUTO = 0, PDAMD_PREC_FP32 = 1, PDAMD_PREC_FP64 = 2, PDAMD_PREC_BF16X3 = 3 };

/* error codes */
enum {
  PDAMD_OK = 0,
  PDAMD_EINVAL = -1,      /* bad argument */
  PDAMD_EUNSUPPORTED = -2,/* shape / size outside what this build handles */
  PDAMD_EHIP = -3,        /* a HIP runtime call failed */
  PDAMD_EWORKSPACE = -4   /* workspace too small */
};

typedef struct pdamd_problem {
  /* volume: C order (d0, d1, d2, T); fewer spatial dims are padded with 1s in front */
  int dtype;
  int64_t shape[4];
  int64_t patch[4];   /* patch shape, already clipped to the volume */
  int64_t step[4];    /* patch_shape - overlap per dim (0 allowed) */

  int method;         /* PDAMD_MPPCA ... */
  int recombination;  /* PDAMD_WEIGHTED ... */
  int precision;      /* PDAMD_PREC_* */

  /* threshold operands.  *_f64 flags: 1 when the reference compares in
     float64 (a strong float64 operand), 0 when it compares in the data's
     own precision (a Python float operand).                               */
  double threshold_scale2; /* MP-PCA: threshold_scale**2 */
  int    scale_f64;
  double tau;              /* HARD: threshold value * scale */
  int    tau_f64;
  int    loss;             /* OPTIMAL */
  double mp_median;        /* OPTIMAL: Marchenko-Pastur median for beta */

  /* per-voxel prior variance map (d0*d1*d2 float64), HYBRID and OPTIMAL
     with a noise level; NULL otherwise */
  const double* var_map_dev;

  /* selection */
  const uint8_t* mask_dev; /* d0*d1*d2 (SPATIAL) or d0*d1*d2*T (FULL) bytes, 0/1 */
  int    mask_kind;
  double mask_threshold;

  /* data */
  const void* input_dev;   /* volume, dtype */
  void*  output_dev;       /* volume, dtype (written) */
  float* weights_dev;      /* volume, float32 (written) */

  /* optional per-selected-patch outputs (NULL to skip), length >= count */
  int32_t* patch_index_dev; /* linear patch index, increasing */
  int32_t* rank_dev;        /* retained rank (the value that sets the weight) */
  float*   noise_dev;       /* noise variance estimate (NaN when the rule has none) */
  float*   eig_dev;         /* 10 per patch: top eigenvalues, ascending (MP-PCA/Hybrid) */

  /* out: number of selected patches */
  int64_t selected;

  /* OPTIMAL: per-patch threshold scale lambda (float64, indexed by linear
     patch index; NULL = 1).  The noise scale c of every patch becomes
     lambda_p * c, so lambda = 1 is the reference's rule.  No counterpart in
     the reference (added for the differentiable path).                     */
  const double* lam_dev;

  /* optional stage timing (host arrays of PDAMD_NSTAGES, NULL to skip): when
     set, HIP events are recorded on `stream` around every kernel launch and,
     after the call completes, stage_ms[s] += summed elapsed milliseconds and
     stage_launches[s] += number of launches of stage s.  Adds one stream
     synchronisation at the end of the call.                                 */
  float*   stage_ms;
  int64_t* stage_launches;

  /* ADAPTIVE (lowrank.py:410-636; the prior variance map var_map_dev is
     required: the patch noise level is sqrt(mean variance over its voxels)).
     SURE / GSURE: tau minimises the SURE / GSURE cost in log(tau) from
     tau0_log (NaN = log median singular value) for each exponent in
     gamma_dev[0 .. n_gamma) (n_gamma >= 1); the lowest cost wins.
     QUT: tau from qut_tau_dev (float64, by linear patch index, computed by
     the caller from its Monte-Carlo draws); gamma = gamma_dev[0] when
     n_gamma == 1, else minimises the SURE cost in log(gamma - 1) from 0.   */
  int    adaptive_method;  /* PDAMD_ADAPT_* */
  int    n_gamma;
  const double* gamma_dev;
  double tau0_log;
  const double* qut_tau_dev;

  /* launch shape: workgroups per CU for the gram, tridiagonalisation, eigen
     and synthesis launches (0 = the library's default; capped by what the
     kernel's LDS allows; 1 .. 32).  Set by an autotuner (the Python
     package's patch_denoise.autotune); results do not depend on it beyond
     the order of the output's float atomic additions.  No counterpart in
     the reference.                                                          */
  int launch_per_cu[4];

  /* recombination strategy (results agree to float32 rounding; no
     counterpart in the reference):
     PDAMD_RECOMBINE_AUTO    gather when eligible, else the atomic scatter;
     PDAMD_RECOMBINE_SCATTER every patch adds into the output with float
                             atomics (order-dependent rounding);
     PDAMD_RECOMBINE_GATHER  every output element sums its covering patches
                             in increasing patch index, no atomics, results
                             repeat bit for bit.  Eligible: float32 data, one
                             temporal patch, weighted / average
                             recombination, MP-PCA or Hybrid, a Gram of at
                             most 128 (voxel side) or 256 (frame side, at
                             most 512 frames), an uncapped workspace;
                             otherwise PDAMD_EUNSUPPORTED.                  */
  int recombine_mode;

  /* FP64 re-check of near-threshold decisions (no counterpart in the
     reference).  Float32 data with an FP32 eigen stage, MP-PCA or Hybrid:
     a patch where a comparison its rule makes -- an eigenvalue against
     sigma^2 * threshold_scale^2, the Marchenko-Pastur iteration's scaled
     mean against the eigenvalue spread, Hybrid's mean against the prior --
     has its two sides within refine_rel of each other (relative to the
     larger magnitude) is recomputed with an FP64 Gram, tridiagonalisation
     and eigen stage (eigenvalues rounded to float32 for the rule, as the
     reference's float32 spectrum), and that result is kept.  At most a fixed
     number of patches per call (gather recombination) or per chunk is
     re-checked (256 MiB of FP64 buffers, at most 1024 patches): the first in
     patch order; the others keep the FP32 decision.  0 = off.             */
  double refine_rel;
  int32_t* refined_dev;  /* optional, per selected patch: 1 where re-checked */

  /* 1 when the reference's prior variance map is float32 (a float32
     noise-level array: lowrank.py:94-97, :351-354): each patch's prior
     (Hybrid: mean variance; Optimal, Adaptive: mean standard deviation) is
     rounded to float32, the precision the reference averages it in.      */
  int var_f32;
} pdamd_problem;

enum { PDAMD_RECOMBINE_AUTO = 0, PDAMD_RECOMBINE_SCATTER = 1, PDAMD_RECOMBINE_GATHER = 2 };

/* stages for pdamd_problem.stage_ms */
enum {
  PDAMD_STAGE_SELECT = 0,   /* mask window + select + scan + compact          */
  PDAMD_STAGE_GRAM = 1,     /* gram_kernel: gather, centre, Gram               */
  PDAMD_STAGE_EIG = 2,      /* trieig_kernel: eigenvalues, rule, eigenvectors  */
  PDAMD_STAGE_SYNTH = 3,    /* synth_kernel: projection + scatter-accumulate
                               (gather recombination: the per-patch factors) */
  PDAMD_STAGE_FINALIZE = 4, /* finalize_kernel: normalisation and mask
                               (gather recombination: patch map and weight
                               sums)                                       */
  PDAMD_STAGE_TRIDIAG = 5,  /* tridiag_kernel: Householder tridiagonalisation  */
  PDAMD_STAGE_ADJ_CROSS = 6,    /* backward: Xc^T Gbar per patch                */
  PDAMD_STAGE_ADJ_SPECTRAL = 7, /* backward: singular-vector adjoint            */
  PDAMD_STAGE_ADJ_SYNTH = 8,    /* backward: dL/dX scatter-accumulate           */
  PDAMD_STAGE_RECOMBINE = 9,    /* gather recombination: every output element
                                   summed from its covering patches' factors */
  PDAMD_STAGE_REFINE = 10,      /* FP64 re-check of near-threshold patches     */
  PDAMD_NSTAGES = 11
};

/* Bytes of device workspace pdamd_denoise needs for this problem. */
int pdamd_workspace_size(const pdamd_problem* prob, size_t* bytes);

/* Selection -> per-patch low-rank denoising -> recombination -> normalisation.
   `workspace_dev` must hold pdamd_workspace_size() bytes.  Synchronises the
   stream once to read the number of selected patches.                     */
int pdamd_denoise(pdamd_problem* prob, void* workspace_dev, size_t workspace_bytes, void* stream);

/* Selection only: writes the increasing linear indices of selected patches to
   index_dev (capacity: number of patches) and their count to *count (host).
   Uses prob's shape/patch/step/mask fields.                               */
int pdamd_select_patches(const pdamd_problem* prob, int32_t* index_dev, int64_t* count,
                         void* workspace_dev, size_t workspace_bytes, void* stream);

/* Batches of independent volumes (replaces a Python loop of
   BaseSpaceTimeDenoiser.denoise calls, space_time/base.py:177-270, one per
   volume).  Every volume is zeroed and selected on the caller's stream, all
   patch counts are read with ONE host synchronisation, then each volume's
   Gram / eigen / synthesis chain runs on its own stream forked from the
   caller's stream (up to 8 concurrently) and joined back to it.  With
   use_graph != 0 that post-selection work is captured into a hipGraph the
   first time a given set of problems (fields, pointers and patch counts) is
   seen and replayed afterwards (a small per-process cache).  Each problem's
   `selected` is set; stage timing is not collected in batches.  The
   workspace is the sum of the per-volume workspaces (each rounded up to
   256 bytes).                                                              */
int pdamd_batch_workspace_size(const pdamd_problem* probs, int count, size_t* bytes);
int pdamd_denoise_batch(pdamd_problem* probs, int count, void* workspace_dev,
                        size_t workspace_bytes, int use_graph, void* stream);

/* Backward of OPTIMAL shrinkage (weighted / average / center recombination,
   real data): given dL/d(output) (grad_out_dev, input dtype) and the forward's
   weights (weights_dev), accumulate dL/d(input) into grad_in_dev (input dtype,
   zeroed by this call) and write dL/d(lambda) per linear patch index into
   grad_lam_dev (float64, zeroed by this call; NULL to skip).  U and S are
   recomputed; the singular-vector adjoint uses guarded
   (g_i -/+ g_j) / (s_i -/+ s_j) terms.  No counterpart in the reference (its
   path is not differentiable).                                             */
int pdamd_backward_workspace_size(const pdamd_problem* prob, size_t* bytes);
int pdamd_optimal_backward(pdamd_problem* prob, const void* grad_out_dev,
                           const float* weights_dev, void* grad_in_dev, double* grad_lam_dev,
                           void* workspace_dev, size_t workspace_bytes, void* stream);

/* Sliding-window noise map, replaces space_time/utils.py:207-242
   (estimate_noise): out[c] = population standard deviation of the
   block_size^d x T elements of the stride-1 window centred at voxel c, for
   every window inside the volume; other voxels of out are left untouched
   (the reference leaves them uninitialised).  in_dev: shape4 = (X, Y, Z, T)
   with leading spatial sizes 1 when nspatial < 3, dtype PDAMD_F32 / F64 /
   C64 / C128; out_dev: float64 X x Y x Z.  Every one of the nspatial spatial
   sizes must exceed block_size (the reference raises otherwise).          */
int pdamd_noise_workspace_size(const int64_t* shape4, size_t* bytes);
int pdamd_estimate_noise(const void* in_dev, int dtype, const int64_t* shape4, int nspatial,
                         int block_size, double* out_dev, void* workspace_dev,
                         size_t workspace_bytes, void* stream);

/* Largest singular value of each of `batch` float64 matrices (rows x cols,
   C order, consecutive in a_dev) into sv_dev (float64, batch): the square
   root of the top eigenvalue of the min-side Gram, formed in FP64 (split
   contraction summed in a fixed order: results repeat bit for bit),
   tridiagonalised and bisected to full precision.  Replaces the
   max(svd(draw, compute_uv=False)) of the Monte-Carlo thresholds in
   space_time/lowrank.py:221-232 (NORDIC) and :446-456 (QUT); the draws stay
   with the caller.  min(rows, cols) must be in 2 .. 2048, batch <= 65535. */
int pdamd_top_sv_workspace_size(int batch, int64_t rows, int64_t cols, size_t* bytes);
int pdamd_top_singular_values(const double* a_dev, int batch, int64_t rows, int64_t cols,
                              double* sv_dev, void* workspace_dev, size_t workspace_bytes,
                              void* stream);

/* pdamd_denoise without any device allocation, for callers that cannot call
   hipMalloc (code under HIP graph capture, tight serving loops, harnesses
   that forbid allocation).  No counterpart in the reference.  The workspace
   is a fixed device array of pdamd_noalloc_scratch_bytes() bytes (256 MiB)
   held by every GPU the library runs on until the process ends; patches are
   processed in chunks sized to fit it.  Problems whose per-volume arrays
   plus one patch exceed it fail with PDAMD_EUNSUPPORTED.  The caller still
   passes the input, output and weight buffers (and the optional per-patch
   arrays) in `prob`; stage timing is not available here.  Calls on one GPU
   share its array: calls that overlap on one GPU race.  Results equal
   pdamd_denoise's up to the summation order of overlapping patches.      */
int pdamd_noalloc_scratch_bytes(size_t* bytes);
int pdamd_denoise_noalloc(pdamd_problem* prob, void* stream);

const char* pdamd_last_error(void);
const char* pdamd_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PATCH_DENOISE_AMD_H */
