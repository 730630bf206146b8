// C ABI of the space-time denoiser (include/patch_denoise_amd.h): problem
// planning and workspace layout, patch selection, dispatch to the
// per-element-type pipelines (pd_run_*.hip), batches of volumes on forked
// streams with cached hipGraphs, and the backward entry point.
#include "pd_kernels.inl"

namespace pd {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

}  // namespace pd

namespace {

// bwd: layout for pdamd_optimal_backward (all eigenpairs, adjoint buffers).
// chunk_cap > 0 limits the patches per chunk (workspace of a fixed size).
int make_plan(const pdamd_problem* P, Plan* pl, bool bwd = false, int64_t chunk_cap = 0) {
  if (!P) return fail(PDAMD_EINVAL, "null problem");
  if (P->dtype < 0 || P->dtype > 3) return fail(PDAMD_EINVAL, "bad dtype %d", P->dtype);
  for (int i = 0; i < 4; ++i) {
    if (P->shape[i] < 1 || P->patch[i] < 1 || P->patch[i] > P->shape[i] || P->step[i] < 0)
      return fail(PDAMD_EINVAL, "bad shape/patch/step in dim %d", i);
  }
  if (P->method < 0 || P->method > PDAMD_ADAPTIVE) return fail(PDAMD_EINVAL, "bad method %d", P->method);
  if (P->method == PDAMD_ADAPTIVE) {
    if (P->adaptive_method < 0 || P->adaptive_method > PDAMD_ADAPT_QUT)
      return fail(PDAMD_EINVAL, "bad adaptive method %d", P->adaptive_method);
    if (P->adaptive_method != PDAMD_ADAPT_QUT && (P->n_gamma < 1 || !P->gamma_dev))
      return fail(PDAMD_EINVAL, "SURE / GSURE need at least one gamma candidate");
    if (P->n_gamma > 0 && !P->gamma_dev) return fail(PDAMD_EINVAL, "gamma_dev missing");
  }
  if (P->recombination < 0 || P->recombination > 2)
    return fail(PDAMD_EINVAL, "recombination must be one of 'weighted', 'average', 'center'");
  Geom& g = pl->g;
  const int64_t T = P->shape[3];
  g.T = T;
  g.vs2 = T;
  g.vs1 = P->shape[2] * T;
  g.vs0 = P->shape[1] * P->shape[2] * T;
  int* gg[4] = {&g.g0, &g.g1, &g.g2, &g.gt};
  int* ss[4] = {&g.s0, &g.s1, &g.s2, &g.st};
  int* pp[4] = {&g.p0, &g.p1, &g.p2, &g.pt};
  pl->npatch = 1;
  for (int i = 0; i < 4; ++i) {
    const int64_t S = P->shape[i], p = P->patch[i], st = P->step[i];
    const int64_t gd = (p < S) ? (st > 0 ? (S - p) / st + 1 : 1) : 1;
    *gg[i] = (int)gd;
    *ss[i] = (int)((p < S) ? st : 0);
    *pp[i] = (int)p;
    pl->npatch *= gd;
  }
  if (pl->npatch >= (int64_t)1 << 31) return fail(PDAMD_EUNSUPPORTED, "too many patches");
  g.M = g.p0 * g.p1 * g.p2;
  pl->nvox = P->shape[0] * P->shape[1] * P->shape[2];
  const int M = g.M, pt = g.pt;
  pl->tside = pt < M;
  pl->n = pl->tside ? pt : M;
  const int n = pl->n;
  if (n > kNBig)
    return fail(PDAMD_EUNSUPPORTED,
                "min(patch voxels, patch frames) = %d exceeds %d (supported maximum)", n, kNBig);
  if (n < 2) return fail(PDAMD_EUNSUPPORTED, "patch too small (min side %d)", n);
  if ((P->method == PDAMD_MPPCA || P->method == PDAMD_HYBRID) && pt < 10)
    return fail(PDAMD_EINVAL, "the eigen rules need at least 10 frames per patch");
  pl->tile = n <= 32 ? 2 : (n <= 64 ? 4 : (n <= 128 ? 8 : 0));  // 0: large-Gram kernels
  for (int i = 0; i < 4; ++i) {
    if (P->launch_per_cu[i] < 0 || P->launch_per_cu[i] > 32)
      return fail(PDAMD_EINVAL, "launch_per_cu[%d] = %d outside 0..32", i, P->launch_per_cu[i]);
    pl->per_cu[i] = P->launch_per_cu[i];
  }
  const bool cplx = P->dtype == PDAMD_C64 || P->dtype == PDAMD_C128;
  const bool in64 = P->dtype == PDAMD_F64 || P->dtype == PDAMD_C128;
  pl->rule_f64 = in64 ? 1 : 0;
  bool e64;
  if (P->precision < 0 || P->precision > PDAMD_PREC_BF16X3)
    return fail(PDAMD_EINVAL, "bad precision %d", P->precision);
  if (P->precision == PDAMD_PREC_FP32 || P->precision == PDAMD_PREC_BF16X3)
    e64 = in64;  // never below the input precision
  else if (P->precision == PDAMD_PREC_FP64) e64 = true;
  else e64 = in64 || (P->method == PDAMD_HARD || P->method == PDAMD_OPTIMAL ||
                      P->method == PDAMD_ADAPTIVE);
  if (cplx && n > 64 && e64 && !in64) {
    // complex double Gram of n > 64 does not fit LDS with this kernel
  }
  pl->etype = (cplx ? 2 : 0) + (e64 ? 1 : 0);
  // split-BF16: FP32 real Grams of 65..128 (gram_kernel) and every
  // FP32 real / complex Gram above 128 (gram_big_kernel)
  pl->bf16x3 = P->precision == PDAMD_PREC_BF16X3 && !e64 && (!cplx || n > 128);
  pl->rbytes = e64 ? 8 : 4;
  pl->ebytes = pl->rbytes * (cplx ? 2 : 1);
  pl->kmax = (P->method == PDAMD_MPPCA || P->method == PDAMD_HYBRID) ? std::min(10, n) : n;
  if (bwd) {
    if (P->method != PDAMD_OPTIMAL)
      return fail(PDAMD_EUNSUPPORTED, "the backward pass is built for the optimal rule only");
    if (cplx) return fail(PDAMD_EUNSUPPORTED, "the backward pass is built for real data only");
    if (!pl->tside)
      return fail(PDAMD_EUNSUPPORTED,
                  "the backward pass needs fewer frames than patch voxels (T-side Gram)");
    if (n > 128)
      return fail(PDAMD_EUNSUPPORTED, "the backward pass supports min(patch voxels, frames) <= 128");
    pl->kmax = n;
  }
  pl->kb = std::min(pl->tile ? kKB : (n > kNMid ? 1 : kKBBig), pl->kmax);
  // recombination by gather: float32 patches (M-side Grams up to 128,
  // T-side up to 256 over at most 512 frames), one temporal patch, weighted /
  // average, the eigen rules (10 components), not in a workspace of capped
  // size (the per-call factor buffers would not fit)
  if (P->recombine_mode < 0 || P->recombine_mode > 2)
    return fail(PDAMD_EINVAL, "bad recombine_mode %d", P->recombine_mode);
  {
    const bool elig = !bwd && P->dtype == PDAMD_F32 && pl->etype == 0 && g.gt == 1 &&
                      (P->recombination == PDAMD_WEIGHTED || P->recombination == PDAMD_AVERAGE) &&
                      (P->method == PDAMD_MPPCA || P->method == PDAMD_HYBRID) &&
                      pl->kmax == kGatherKM &&
                      (pl->tside ? (n <= 256 && g.pt <= 512) : n <= 128) &&
                      pl->nvox * g.T < ((int64_t)1 << 31);
    if (P->recombine_mode == PDAMD_RECOMBINE_GATHER && (!elig || chunk_cap > 0))
      return fail(PDAMD_EUNSUPPORTED,
                  "gather recombination needs float32 data, one temporal patch, weighted or "
                  "average recombination, the MP-PCA or Hybrid rule, a Gram of at most 128 "
                  "(voxel side) or 256 (frame side, at most 512 frames) and an uncapped "
                  "workspace");
    pl->gather = elig && P->recombine_mode != PDAMD_RECOMBINE_SCATTER && chunk_cap == 0;
  }
  // FP64 re-check of near-threshold decisions: float32 data whose eigen
  // rule runs on FP32 eigenvalues
  pl->refine = !bwd && P->refine_rel > 0 && P->dtype == PDAMD_F32 && pl->etype == 0 &&
               (P->method == PDAMD_MPPCA || P->method == PDAMD_HYBRID);
  // chunk: keep the per-chunk buffers under ~4 GiB (24 GiB for the backward)
  size_t per = (size_t)n * n * pl->ebytes + (size_t)n * pl->kmax * (pl->ebytes + pl->rbytes) +
               3 * n * pl->ebytes + pl->kmax * 8 + 16;
  if (bwd) per += (3 * (size_t)M * n + (size_t)n * n + 2 * (size_t)n) * pl->ebytes + 8;
  int64_t chunk = (int64_t)((bwd ? (24ull << 30) : (12ull << 30)) / per);
  chunk = std::max<int64_t>(256, std::min<int64_t>(chunk, 65536));
  pl->chunk = std::min<int64_t>(chunk, std::max<int64_t>(pl->npatch, 1));
  if (chunk_cap > 0) pl->chunk = std::min<int64_t>(pl->chunk, chunk_cap);
  // workspace layout
  size_t o = 0;
  auto take = [&](size_t b) {
    size_t r = o;
    o += ((b + 255) / 256) * 256;
    return r;
  };
  const int64_t nsel_blocks = (pl->npatch + kSelPerBlock - 1) / kSelPerBlock;
  pl->off_cnt = take((size_t)pl->nvox * g.gt * sizeof(int));
  pl->off_flags = take((size_t)pl->npatch);
  pl->off_btot = take((size_t)nsel_blocks * sizeof(int));
  pl->off_idx = take((size_t)pl->npatch * sizeof(int32_t));
  pl->off_count = take(sizeof(int64_t));
  const size_t C = (size_t)pl->chunk;
  pl->off_H = take(C * n * n * pl->ebytes);
  pl->off_DE = take(C * 3 * n * pl->rbytes);
  pl->off_SCR = take(C * 3 * n * (bwd ? n : pl->kb) * pl->rbytes);
  pl->off_TAU = take(C * n * pl->ebytes);
  if (pl->tside) {  // per-frame means of T-side patches, from the Gram to the synthesis
    const size_t inb = P->dtype == PDAMD_F32 ? 4 : (P->dtype == PDAMD_C128 ? 16 : 8);
    pl->off_MEAN = take(C * (size_t)pt * inb);
  }
  pl->off_Y = take(C * n * pl->kmax * pl->ebytes);
  pl->off_Z = take(C * n * pl->kmax * pl->rbytes);
  pl->off_WC = take(C * pl->kmax * sizeof(double));
  pl->off_NC = take(C * sizeof(int));
  pl->off_TH = take(C * sizeof(double));
  pl->off_w3 = take((size_t)pl->nvox * (g.gt == 1 ? 1 : g.T) * sizeof(double));
  if (pl->gather) {
    const size_t NPs = (size_t)pl->npatch;
    pl->off_GMAP = take(NPs * sizeof(int32_t));
    pl->off_GNC = take(NPs * sizeof(int));
    pl->off_GTH = take(NPs * sizeof(double));
    pl->off_GMEAN = take(NPs * pt * sizeof(float));
    pl->off_GY = take(NPs * M * kGatherKM * sizeof(float));
    if (pl->tside) pl->off_GV = take(NPs * n * kGatherKM * sizeof(float));
    pl->off_GRF = take(NPs * (kGatherKM + 1) * pt * sizeof(float));
    pl->off_GW3 = take((size_t)pl->nvox * sizeof(float));
  }
  if (pl->refine) {
    // per listed patch: FP64 Gram / reflectors, tridiagonal, tau,
    // inverse-iteration scratch, eigenvectors (Gram and tridiagonal), rule
    // outputs; at most 1024 patches and 256 MiB
    const size_t per_r = (size_t)n * n * 8 + 3 * (size_t)n * 8 + (size_t)n * 8 +
                         3 * (size_t)n * pl->kb * 8 + (size_t)pt * 4 +
                         2 * (size_t)n * pl->kmax * 8 + (size_t)pl->kmax * 8 + 64 + 40 + 1024;
    const int64_t cap = std::max<int64_t>(1, (int64_t)((256ull << 20) / per_r));
    pl->rcap = (int)std::min<int64_t>({cap, (int64_t)1024, pl->chunk});
    const size_t R = (size_t)pl->rcap;
    pl->off_RFLAG = take((size_t)pl->npatch);
    pl->off_RCNT = take(2 * sizeof(int));
    pl->off_RSLOT = take(R * sizeof(int32_t));
    pl->off_RIDX = take(R * sizeof(int32_t));
    pl->off_RH = take(R * n * n * 8);
    pl->off_RDE = take(R * 3 * n * 8);
    pl->off_RTAU = take(R * n * 8);
    pl->off_RSCR = take(R * 3 * n * pl->kb * 8);
    pl->off_RMEAN = take(R * pt * 4);
    pl->off_RY = take(R * n * pl->kmax * 8);
    pl->off_RZ = take(R * n * pl->kmax * 8);
    pl->off_RWC = take(R * pl->kmax * 8);
    pl->off_RNC = take(R * sizeof(int));
    pl->off_RTH = take(R * sizeof(double));
    pl->off_RRK = take(R * sizeof(int32_t));
    pl->off_RNZ = take(R * sizeof(float));
    pl->off_REG = take(R * 10 * sizeof(float));
  }
  if (bwd) {
    pl->off_GH = take((size_t)pl->nvox * g.T * pl->ebytes);
    pl->off_P1 = take(C * M * n * pl->ebytes);
    pl->off_P2 = take(C * M * n * pl->ebytes);
    pl->off_PE = take(C * M * n * pl->ebytes);
    pl->off_GS = take(C * n * pl->ebytes);
    pl->off_S1 = take(C * n * n * pl->ebytes);
    pl->off_UV = take(C * n * pl->ebytes);
    pl->off_THB = take(C * sizeof(double));
  }
  pl->total = o;
  return PDAMD_OK;
}

// Enqueue the selection kernels; the patch count lands in ws + off_count.
int run_select_async(const pdamd_problem* P, const Plan& pl, char* ws, int32_t* idx_out,
                     hipStream_t st, StageTimer* tm = nullptr) {
  int* cnt = reinterpret_cast<int*>(ws + pl.off_cnt);
  uint8_t* flags = reinterpret_cast<uint8_t*>(ws + pl.off_flags);
  int* btot = reinterpret_cast<int*>(ws + pl.off_btot);
  int64_t* dcount = reinterpret_cast<int64_t*>(ws + pl.off_count);
  const int kind = P->mask_kind;
  if (kind != PDAMD_MASK_NONE && !P->mask_dev) return fail(PDAMD_EINVAL, "mask pointer missing");
  if (tm) tm->begin(PDAMD_STAGE_SELECT);
  if (kind != PDAMD_MASK_NONE) {
    const int64_t tot = pl.nvox * pl.g.gt;
    const int nb = (int)((tot + 255) / 256);
    mask_window_kernel<<<nb, 256, 0, st>>>(P->mask_dev, kind, pl.nvox, pl.g.T, pl.g.gt, pl.g.st,
                                           pl.g.pt, cnt);
  }
  const int nblk = (int)((pl.npatch + kSelPerBlock - 1) / kSelPerBlock);
  const double psize = (double)pl.g.M * (double)pl.g.pt;
  select_kernel<<<nblk, kNT, 0, st>>>(cnt, pl.g, pl.npatch, kind, P->mask_threshold, psize,
                                      flags, btot);
  scan_kernel<<<1, 1024, 0, st>>>(btot, nblk, dcount);
  compact_kernel<<<nblk, kNT, 0, st>>>(flags, pl.npatch, btot, idx_out);
  if (tm) tm->end();
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
}

int run_select(const pdamd_problem* P, const Plan& pl, char* ws, int32_t* idx_out,
               int64_t* count_host, hipStream_t st, StageTimer* tm = nullptr) {
  const int rc = run_select_async(P, pl, ws, idx_out, st, tm);
  if (rc) return rc;
  const int64_t* dcount = reinterpret_cast<const int64_t*>(ws + pl.off_count);
  HIP_TRY(hipMemcpyAsync(count_host, dcount, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return PDAMD_OK;
}

// Zero fill by a kernel (capturable; used where hipMemsetAsync could end up in
// a graph).  bytes must be a multiple of 4.
int dispatch_chunks(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, hipStream_t st,
                    StageTimer& tm) {
  switch (P->dtype * 4 + pl.etype) {
    case PDAMD_F32 * 4 + 0: return run_f32_f32(P, pl, ws, count, st, tm);
    case PDAMD_F32 * 4 + 1: return run_f32_f64(P, pl, ws, count, st, tm);
    case PDAMD_F64 * 4 + 1: return run_f64(P, pl, ws, count, st, tm);
    case PDAMD_C64 * 4 + 2: return run_c64_c64(P, pl, ws, count, st, tm);
    case PDAMD_C64 * 4 + 3: return run_c64_c128(P, pl, ws, count, st, tm);
    case PDAMD_C128 * 4 + 3: return run_c128(P, pl, ws, count, st, tm);
    default:
      return fail(PDAMD_EUNSUPPORTED, "dtype/precision combination not built");
  }
}

size_t elem_bytes(int dtype) {
  return dtype == PDAMD_F32 ? 4 : (dtype == PDAMD_C128 ? 16 : 8);
}

// ---------------------------------------------------------------------------
// Batches of volumes: one stream per volume (forked from and joined back to
// the caller's stream), optional capture of the post-selection work into a
// cached hipGraph.
// ---------------------------------------------------------------------------
constexpr int kBatchStreams = 8;

struct BatchStreams {
  int dev = -1;
  hipStream_t s[kBatchStreams] = {};
  hipEvent_t ev[kBatchStreams] = {};
  hipEvent_t fork = nullptr;
  hipStream_t cap = nullptr;  // graph capture / replay stream (the caller's may be the null stream)
  hipEvent_t pre = nullptr, done = nullptr;
};

int batch_streams(BatchStreams** out) {
  static BatchStreams bs[16];
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 16) return fail(PDAMD_EUNSUPPORTED, "device index %d", dev);
  BatchStreams& b = bs[dev];
  if (b.dev < 0) {
    for (int i = 0; i < kBatchStreams; ++i) {
      HIP_TRY(hipStreamCreateWithFlags(&b.s[i], hipStreamNonBlocking));
      HIP_TRY(hipEventCreateWithFlags(&b.ev[i], hipEventDisableTiming));
    }
    HIP_TRY(hipEventCreateWithFlags(&b.fork, hipEventDisableTiming));
    HIP_TRY(hipStreamCreateWithFlags(&b.cap, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&b.pre, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&b.done, hipEventDisableTiming));
    b.dev = dev;
  }
  *out = &b;
  return PDAMD_OK;
}

struct GraphEntry {
  int dev;
  std::vector<char> key;
  hipGraphExec_t exec;
};

std::vector<GraphEntry>& graph_cache() {
  static std::vector<GraphEntry> g;
  return g;
}

}  // namespace

namespace pdamd_detail {
void set_error(const char* msg) { g_err = msg; }

// The body of pdamd_denoise for an already-made plan.
int denoise_planned(pdamd_problem* P, const void* plan, void* workspace_dev,
                    size_t workspace_bytes, void* stream) {
  const Plan& pl = *reinterpret_cast<const Plan*>(plan);
  if (workspace_bytes < pl.total)
    return fail(PDAMD_EWORKSPACE, "workspace %zu < %zu bytes", workspace_bytes, pl.total);
  if (!P->input_dev || !P->output_dev || !P->weights_dev)
    return fail(PDAMD_EINVAL, "null data pointer");
  if ((P->method == PDAMD_HYBRID || P->method == PDAMD_ADAPTIVE) && !P->var_map_dev)
    return fail(PDAMD_EINVAL, "hybrid / adaptive rules need a variance map");
  if (P->method == PDAMD_ADAPTIVE && P->adaptive_method == PDAMD_ADAPT_QUT && !P->qut_tau_dev)
    return fail(PDAMD_EINVAL, "QUT needs the per-patch threshold array");
  if (P->lam_dev && P->method != PDAMD_OPTIMAL)
    return fail(PDAMD_EINVAL, "a threshold scale map is only used by the optimal rule");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  char* ws = reinterpret_cast<char*>(workspace_dev);
  const size_t esz = (P->dtype == PDAMD_F32) ? 4 : (P->dtype == PDAMD_C128 ? 16 : 8);
  if (!pl.gather) {  // the gather recombination writes every output element and weight
    HIP_TRY(hipMemsetAsync(P->output_dev, 0, (size_t)pl.nvox * pl.g.T * esz, st));
    if (P->recombination == PDAMD_CENTER)  // otherwise the normalisation writes every weight
      HIP_TRY(hipMemsetAsync(P->weights_dev, 0, (size_t)pl.nvox * pl.g.T * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(ws + pl.off_w3, 0,
                           (size_t)pl.nvox * (pl.g.gt == 1 ? 1 : pl.g.T) * sizeof(double), st));
  }
  int32_t* idx = reinterpret_cast<int32_t*>(ws + pl.off_idx);
  int64_t count = 0;
  StageTimer tm;
  tm.on = P->stage_ms != nullptr || P->stage_launches != nullptr;
  tm.st = st;
  int rc = run_select(P, pl, ws, idx, &count, st, &tm);
  if (rc) return rc;
  P->selected = count;
  if (P->patch_index_dev && count > 0)
    HIP_TRY(hipMemcpyAsync(P->patch_index_dev, idx, count * sizeof(int32_t),
                           hipMemcpyDeviceToDevice, st));
  rc = dispatch_chunks(P, pl, ws, count, st, tm);
  if (rc) return rc;
  if (tm.collect(P->stage_ms, P->stage_launches))
    return fail(PDAMD_EHIP, "stream failed while collecting stage timings");
  return PDAMD_OK;
}

// pdamd_denoise in a workspace of exactly `cap` bytes at `ws`: the largest
// chunk of patches whose layout fits.  Fails when even one patch per chunk
// (plus the per-volume arrays) exceeds the cap.
int denoise_capped(pdamd_problem* P, void* ws, size_t cap, void* stream) {
  Plan pl;
  int rc = make_plan(P, &pl);
  if (rc) return rc;
  if (pl.total > cap) {
    int64_t lo = 0, hi = pl.chunk;  // lo fits (0 = nothing yet), hi does not
    while (hi - lo > 1) {
      const int64_t mid = lo + (hi - lo) / 2;
      Plan t;
      rc = make_plan(P, &t, false, mid);
      if (rc) return rc;
      if (t.total <= cap) lo = mid;
      else hi = mid;
    }
    if (lo == 0)
      return fail(PDAMD_EUNSUPPORTED,
                  "problem needs more than the %zu-byte no-allocation scratch even one patch "
                  "at a time (volume-sized arrays are part of it)", cap);
    rc = make_plan(P, &pl, false, lo);
    if (rc) return rc;
  }
  return denoise_planned(P, &pl, ws, cap, stream);
}
}  // namespace pdamd_detail

extern "C" {

const char* pdamd_last_error(void) { return g_err.c_str(); }

const char* pdamd_version(void) { return "patch_denoise_amd 0.1 (gfx950)"; }

int pdamd_workspace_size(const pdamd_problem* prob, size_t* bytes) {
  Plan pl;
  const int rc = make_plan(prob, &pl);
  if (rc) return rc;
  *bytes = pl.total;
  return PDAMD_OK;
}

int pdamd_select_patches(const pdamd_problem* prob, int32_t* index_dev, int64_t* count,
                         void* workspace_dev, size_t workspace_bytes, void* stream) {
  Plan pl;
  int rc = make_plan(prob, &pl);
  if (rc) return rc;
  if (workspace_bytes < pl.total) return fail(PDAMD_EWORKSPACE, "workspace too small");
  return run_select(prob, pl, reinterpret_cast<char*>(workspace_dev), index_dev, count,
                    reinterpret_cast<hipStream_t>(stream));
}

int pdamd_denoise(pdamd_problem* P, void* workspace_dev, size_t workspace_bytes, void* stream) {
  Plan pl;
  int rc = make_plan(P, &pl);
  if (rc) return rc;
  return pdamd_detail::denoise_planned(P, &pl, workspace_dev, workspace_bytes, stream);
}

int pdamd_batch_workspace_size(const pdamd_problem* probs, int count, size_t* bytes) {
  if (!probs || count < 1) return fail(PDAMD_EINVAL, "empty batch");
  size_t tot = 0;
  for (int i = 0; i < count; ++i) {
    Plan pl;
    const int rc = make_plan(&probs[i], &pl);
    if (rc) return fail(rc, "volume %d: %s", i, g_err.c_str());
    tot += ((pl.total + 255) / 256) * 256;
  }
  *bytes = tot;
  return PDAMD_OK;
}

int pdamd_denoise_batch(pdamd_problem* probs, int count, void* workspace_dev,
                        size_t workspace_bytes, int use_graph, void* stream) {
  if (!probs || count < 1) return fail(PDAMD_EINVAL, "empty batch");
  std::vector<Plan> pls(count);
  std::vector<size_t> off(count);
  size_t tot = 0;
  for (int i = 0; i < count; ++i) {
    pdamd_problem* P = &probs[i];
    int rc = make_plan(P, &pls[i]);
    if (rc) return fail(rc, "volume %d: %s", i, g_err.c_str());
    if (!P->input_dev || !P->output_dev || !P->weights_dev)
      return fail(PDAMD_EINVAL, "volume %d: null data pointer", i);
    if ((P->method == PDAMD_HYBRID || P->method == PDAMD_ADAPTIVE) && !P->var_map_dev)
      return fail(PDAMD_EINVAL, "volume %d: hybrid / adaptive rules need a variance map", i);
    if (P->method == PDAMD_ADAPTIVE && P->adaptive_method == PDAMD_ADAPT_QUT && !P->qut_tau_dev)
      return fail(PDAMD_EINVAL, "volume %d: QUT needs the per-patch threshold array", i);
    if (P->lam_dev && P->method != PDAMD_OPTIMAL)
      return fail(PDAMD_EINVAL, "volume %d: a threshold scale map is only used by the optimal rule", i);
    off[i] = tot;
    tot += ((pls[i].total + 255) / 256) * 256;
  }
  if (workspace_bytes < tot)
    return fail(PDAMD_EWORKSPACE, "workspace %zu < %zu bytes", workspace_bytes, tot);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  char* ws = reinterpret_cast<char*>(workspace_dev);
  // 1. zero outputs and select, every volume on the caller's stream
  for (int i = 0; i < count; ++i) {
    pdamd_problem* P = &probs[i];
    const Plan& pl = pls[i];
    char* w = ws + off[i];
    int rc = zero_async(P->output_dev, (size_t)pl.nvox * pl.g.T * elem_bytes(P->dtype), st);
    if (!rc && P->recombination == PDAMD_CENTER)
      rc = zero_async(P->weights_dev, (size_t)pl.nvox * pl.g.T * sizeof(float), st);
    if (!rc)
      rc = zero_async(w + pl.off_w3, (size_t)pl.nvox * (pl.g.gt == 1 ? 1 : pl.g.T) * sizeof(double), st);
    if (!rc) rc = run_select_async(P, pl, w, reinterpret_cast<int32_t*>(w + pl.off_idx), st);
    if (rc) return rc;
  }
  // 2. one synchronisation for every patch count
  std::vector<int64_t> counts(count);
  for (int i = 0; i < count; ++i)
    HIP_TRY(hipMemcpyAsync(&counts[i], ws + off[i] + pls[i].off_count, sizeof(int64_t),
                           hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  for (int i = 0; i < count; ++i) probs[i].selected = counts[i];
  // 3. per-volume kernel chains on forked streams
  BatchStreams* bs = nullptr;
  int rc = batch_streams(&bs);
  if (rc) return rc;
  StageTimer off_tm;  // no per-stage timing in batches
  auto enqueue = [&](hipStream_t st) -> int {
    HIP_TRY(hipEventRecord(bs->fork, st));
    const int ns = std::min(count, kBatchStreams);
    for (int k = 0; k < ns; ++k) HIP_TRY(hipStreamWaitEvent(bs->s[k], bs->fork, 0));
    for (int i = 0; i < count; ++i) {
      hipStream_t si = bs->s[i % kBatchStreams];
      pdamd_problem* P = &probs[i];
      char* w = ws + off[i];
      if (P->patch_index_dev && counts[i] > 0)
        HIP_TRY(hipMemcpyAsync(P->patch_index_dev, w + pls[i].off_idx, counts[i] * sizeof(int32_t),
                               hipMemcpyDeviceToDevice, si));
      const int r = dispatch_chunks(P, pls[i], w, counts[i], si, off_tm);
      if (r) return r;
    }
    for (int k = 0; k < ns; ++k) {
      HIP_TRY(hipEventRecord(bs->ev[k], bs->s[k]));
      HIP_TRY(hipStreamWaitEvent(st, bs->ev[k], 0));
    }
    return PDAMD_OK;
  };
  if (!use_graph) return enqueue(st);
  // graph key: every problem field that shapes the launches, and the counts
  std::vector<char> key;
  auto put = [&](const void* p, size_t b) {
    const char* c = reinterpret_cast<const char*>(p);
    key.insert(key.end(), c, c + b);
  };
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  for (int i = 0; i < count; ++i) {
    pdamd_problem q = probs[i];
    q.stage_ms = nullptr;
    q.stage_launches = nullptr;
    q.selected = 0;
    put(&q, sizeof q);
    put(&counts[i], sizeof(int64_t));
  }
  put(&workspace_dev, sizeof workspace_dev);
  auto& cache = graph_cache();
  hipGraphExec_t exec = nullptr;
  for (auto& e : cache)
    if (e.dev == dev && e.key == key) exec = e.exec;
  hipStream_t cs = bs->cap;
  if (!exec) {
    HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
    const int r = enqueue(cs);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &graph);
    if (r) {
      if (graph) hipGraphDestroy(graph);
      return r;
    }
    if (ec != hipSuccess) return fail(PDAMD_EHIP, "graph capture: %s", hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(PDAMD_EHIP, "graph instantiate: %s", hipGetErrorString(ei));
    if (cache.size() >= 8) {
      hipGraphExecDestroy(cache.front().exec);
      cache.erase(cache.begin());
    }
    cache.push_back({dev, key, exec});
  }
  // replay on the capture stream, ordered after the selection on the caller's
  // stream and before whatever the caller enqueues next
  HIP_TRY(hipEventRecord(bs->pre, st));
  HIP_TRY(hipStreamWaitEvent(cs, bs->pre, 0));
  HIP_TRY(hipGraphLaunch(exec, cs));
  HIP_TRY(hipEventRecord(bs->done, cs));
  HIP_TRY(hipStreamWaitEvent(st, bs->done, 0));
  return PDAMD_OK;
}

int pdamd_backward_workspace_size(const pdamd_problem* prob, size_t* bytes) {
  Plan pl;
  const int rc = make_plan(prob, &pl, true);
  if (rc) return rc;
  *bytes = pl.total;
  return PDAMD_OK;
}

int pdamd_optimal_backward(pdamd_problem* P, const void* grad_out_dev, const float* weights_dev,
                           void* grad_in_dev, double* grad_lam_dev, void* workspace_dev,
                           size_t workspace_bytes, void* stream) {
  Plan pl;
  int rc = make_plan(P, &pl, true);
  if (rc) return rc;
  if (workspace_bytes < pl.total)
    return fail(PDAMD_EWORKSPACE, "workspace %zu < %zu bytes", workspace_bytes, pl.total);
  if (!P->input_dev || !grad_out_dev || !grad_in_dev)
    return fail(PDAMD_EINVAL, "null data pointer");
  if (!weights_dev && P->recombination != PDAMD_CENTER)
    return fail(PDAMD_EINVAL, "the forward weights are needed for this recombination");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  char* ws = reinterpret_cast<char*>(workspace_dev);
  const size_t esz = (P->dtype == PDAMD_F32) ? 4 : 8;
  HIP_TRY(hipMemsetAsync(grad_in_dev, 0, (size_t)pl.nvox * pl.g.T * esz, st));
  if (grad_lam_dev) HIP_TRY(hipMemsetAsync(grad_lam_dev, 0, (size_t)pl.npatch * sizeof(double), st));
  int32_t* idx = reinterpret_cast<int32_t*>(ws + pl.off_idx);
  int64_t count = 0;
  StageTimer tm;
  tm.on = P->stage_ms != nullptr || P->stage_launches != nullptr;
  tm.st = st;
  rc = run_select(P, pl, ws, idx, &count, st, &tm);
  if (rc) return rc;
  P->selected = count;
  switch (P->dtype * 4 + pl.etype) {
    case PDAMD_F32 * 4 + 0:
      rc = backward_f32_f32(P, pl, ws, count, grad_out_dev, weights_dev,
                                             grad_in_dev, grad_lam_dev, st, tm);
      break;
    case PDAMD_F32 * 4 + 1:
      rc = backward_f32_f64(P, pl, ws, count, grad_out_dev, weights_dev,
                                              grad_in_dev, grad_lam_dev, st, tm);
      break;
    case PDAMD_F64 * 4 + 1:
      rc = backward_f64(P, pl, ws, count, grad_out_dev, weights_dev,
                                                grad_in_dev, grad_lam_dev, st, tm);
      break;
    default:
      return fail(PDAMD_EUNSUPPORTED, "dtype/precision combination not built for the backward");
  }
  if (rc) return rc;
  if (tm.collect(P->stage_ms, P->stage_launches))
    return fail(PDAMD_EHIP, "stream failed while collecting stage timings");
  return PDAMD_OK;
}

}  // extern "C"
