// Largest singular value of each float64 matrix of a batch.
//
// The Monte-Carlo thresholds of NORDIC (reference: space_time/lowrank.py:221-232,
// the mean over draws of max(svd(randn(N, T)))) and of QUT
// (space_time/lowrank.py:446-456, the quantile over draws of
// max(svd(randn(M, T) * sigma)))) need the top singular value of Gaussian
// draws.  Here it is the square root of the top eigenvalue of the min-side
// Gram:
//   1. syrk_part_kernel   split-K FP64 Gram, 64 x 64 lower-triangle tiles per
//                         workgroup over one slice of the contraction index
//                         (VALU FP64: the MI355X's vector and matrix FP64
//                         rates are the same);
//   2. syrk_reduce_kernel the slices summed in slice order (no atomics: the
//                         result repeats bit for bit), both triangles written;
//   3. Householder tridiagonalisation by the denoiser's FP64 kernels;
//   4. top_eig_kernel     multisection on Sturm counts, one wavefront per
//                         matrix, to the last bit of the tridiagonal's top
//                         eigenvalue.
#include <hip/hip_runtime.h>

#include <float.h>

#include <algorithm>

#include "patch_denoise_amd.h"
#include "pd_engine.h"

using namespace pd;

namespace {

constexpr int kST = 64;    // Gram output tile
constexpr int kSK = 16;    // contraction indices per staged chunk
constexpr int kSNT = 256;  // threads per tile workgroup (4 x 4 outputs each)

struct SyrkArgs {
  const double* A;  // batch x rows x cols, C order
  int64_t rows, cols;
  int n;            // Gram size min(rows, cols)
  int64_t K;        // contraction length max(rows, cols)
  int tall;         // 1: G = A^T A (contract over rows), 0: G = A A^T
  int ntile;        // lower-triangle tiles per Gram
  int splits;       // contraction slices
  int64_t kper;     // indices per slice
  double* part;     // batch x splits x ntile x 64 x 64
};

// Element (k, i) of the contraction view: the k-th term of Gram index i.
__device__ __forceinline__ double view(const SyrkArgs& a, const double* A, int64_t k, int i) {
  return a.tall ? A[k * a.cols + i] : A[(int64_t)i * a.cols + k];
}

__global__ __launch_bounds__(kSNT) void syrk_part_kernel(SyrkArgs a) {
  const int b = blockIdx.y;
  const int tile = blockIdx.x % a.ntile, sp = blockIdx.x / a.ntile;
  int bi = 0;
  while ((bi + 1) * (bi + 2) / 2 <= tile) ++bi;
  const int bj = tile - bi * (bi + 1) / 2;
  const int64_t k0 = (int64_t)sp * a.kper, k1 = min(a.K, k0 + a.kper);
  const double* A = a.A + (size_t)b * a.rows * a.cols;
  __shared__ double As[kSK][kST + 1];
  __shared__ double Bs[kSK][kST + 1];
  const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;
  for (int64_t kc = k0; kc < k1; kc += kSK) {
    // stage kSK x 64 of both operands: tall reads 4 consecutive columns of
    // one row per thread, wide 4 consecutive columns of one Gram row
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      int kk, ii;
      if (a.tall) {
        kk = tid / 16;
        ii = (tid % 16) * 4 + u;
      } else {
        ii = tid / 4;
        kk = (tid % 4) * 4 + u;
      }
      const int64_t k = kc + kk;
      const int gi = bi * kST + ii, gj = bj * kST + ii;
      As[kk][ii] = (k < k1 && gi < a.n) ? view(a, A, k, gi) : 0.0;
      Bs[kk][ii] = (k < k1 && gj < a.n) ? view(a, A, k, gj) : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kSK; ++kk) {
      double x[4], y[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) x[r] = As[kk][ty * 4 + r];
#pragma unroll
      for (int c = 0; c < 4; ++c) y[c] = Bs[kk][tx * 4 + c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = fma(x[r], y[c], acc[r][c]);
    }
    __syncthreads();
  }
  double* P = a.part + (((size_t)b * a.splits + sp) * a.ntile + tile) * kST * kST;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) P[(ty * 4 + r) * kST + tx * 4 + c] = acc[r][c];
}

// G[b] (n x n, both triangles) = sum over slices, in slice order.
__global__ void syrk_reduce_kernel(SyrkArgs a, int batch, double* __restrict__ G) {
  const int64_t nn = (int64_t)a.n * a.n;
  const int64_t total = (int64_t)batch * nn;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / nn);
    const int i = (int)((e % nn) / a.n), j = (int)(e % a.n);
    const int hi = max(i, j), lo = min(i, j);
    const int bi = hi / kST, bj = lo / kST;
    const int tile = bi * (bi + 1) / 2 + bj;
    const int li = hi - bi * kST, lj = lo - bj * kST;
    const double* P = a.part + ((size_t)b * a.splits * a.ntile + tile) * kST * kST + li * kST + lj;
    double s = 0.0;
    for (int sp = 0; sp < a.splits; ++sp) s += P[(size_t)sp * a.ntile * kST * kST];
    G[e] = s;
  }
}

// Number of eigenvalues of the tridiagonal (d, e^2) below x (Sturm sequence).
__device__ int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
  int c = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  c += q < 0.0;
  for (int i = 1; i < n; ++i) {
    q = d[i] - x - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0;
  }
  return c;
}

// One wavefront per matrix: the top eigenvalue of its tridiagonal (d | e | e^2
// rows of 3n) by multisection, 64 Sturm counts per step; sv = sqrt(max(l, 0)).
__global__ __launch_bounds__(kWave) void top_eig_kernel(const double* __restrict__ DE, int n,
                                                       double* __restrict__ sv) {
  const int b = blockIdx.x, lane = threadIdx.x;
  const double* d = DE + (size_t)b * 3 * n;
  const double* e = d + n;
  const double* e2 = d + 2 * n;
  double glo = INFINITY, ghi = -INFINITY, emax2 = 0.0;
  for (int i = lane; i < n; i += kWave) {
    const double off = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
    glo = fmin(glo, d[i] - off);
    ghi = fmax(ghi, d[i] + off);
    if (i < n - 1) emax2 = fmax(emax2, e2[i]);
  }
  for (int m = kWave / 2; m >= 1; m >>= 1) {
    glo = fmin(glo, __shfl_xor(glo, m));
    ghi = fmax(ghi, __shfl_xor(ghi, m));
    emax2 = fmax(emax2, __shfl_xor(emax2, m));
  }
  const double tnorm = fmax(fabs(glo), fabs(ghi));
  const double pivmin = DBL_MIN * fmax(1.0, emax2);
  double lo = glo - 4 * DBL_EPSILON * tnorm - 4 * pivmin;
  double hi = ghi + 4 * DBL_EPSILON * tnorm + 4 * pivmin;
  // invariant: count(lo) <= n - 1 (the top eigenvalue is above lo), count(hi) == n
  for (int it = 0; it < 64; ++it) {
    const double w = hi - lo;
    if (w <= 2 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + pivmin) break;
    const double x = lo + w * (double)(lane + 1) / (double)(kWave + 1);
    const bool below = sturm_count(d, e2, n, x, pivmin) < n;  // top eigenvalue >= x
    const unsigned long long m = __ballot(below);
    // points are increasing in lane: `below` holds for a prefix of lanes
    const int nb = __popcll(m);
    const double nlo = nb > 0 ? lo + w * (double)nb / (double)(kWave + 1) : lo;
    const double nhi = nb < kWave ? lo + w * (double)(nb + 1) / (double)(kWave + 1) : hi;
    if (nlo == lo && nhi == hi) break;  // no representable progress
    lo = nlo;
    hi = nhi;
  }
  if (lane == 0) sv[b] = sqrt(fmax(0.5 * (lo + hi), 0.0));
}

struct TopSvPlan {
  SyrkArgs s;
  size_t off_part, off_H, off_DE, off_TAU, total;
};

int top_sv_plan(int batch, int64_t rows, int64_t cols, TopSvPlan* pl) {
  if (batch < 1 || rows < 1 || cols < 1) return fail(PDAMD_EINVAL, "empty batch or matrix");
  if (batch > 65535) return fail(PDAMD_EUNSUPPORTED, "batch of %d matrices exceeds 65535", batch);
  const int64_t n = std::min(rows, cols);
  if (n < 2) return fail(PDAMD_EUNSUPPORTED, "matrices need at least 2 rows and 2 columns");
  if (n > 2048) return fail(PDAMD_EUNSUPPORTED, "min(rows, cols) = %lld exceeds 2048", (long long)n);
  SyrkArgs& s = pl->s;
  s.rows = rows;
  s.cols = cols;
  s.n = (int)n;
  s.tall = rows >= cols ? 1 : 0;
  s.K = std::max(rows, cols);
  const int nbt = (s.n + kST - 1) / kST;
  s.ntile = nbt * (nbt + 1) / 2;
  // slices: enough workgroups to fill the chip, at least 64 indices each
  const int64_t want = std::max<int64_t>(1, 2048 / ((int64_t)batch * s.ntile));
  const int64_t maxs = std::max<int64_t>(1, s.K / 64);
  s.splits = (int)std::min<int64_t>({want, maxs, 4096});
  s.kper = (s.K + s.splits - 1) / s.splits;
  s.kper = ((s.kper + kSK - 1) / kSK) * kSK;
  s.splits = (int)((s.K + s.kper - 1) / s.kper);
  size_t o = 0;
  auto take = [&](size_t b) {
    const size_t r = o;
    o += ((b + 255) / 256) * 256;
    return r;
  };
  pl->off_part = take((size_t)batch * s.splits * s.ntile * kST * kST * sizeof(double));
  pl->off_H = take((size_t)batch * n * n * sizeof(double));
  pl->off_DE = take((size_t)batch * 3 * n * sizeof(double));
  pl->off_TAU = take((size_t)batch * n * sizeof(double));
  pl->total = o;
  return PDAMD_OK;
}

}  // namespace

extern "C" {

int pdamd_top_sv_workspace_size(int batch, int64_t rows, int64_t cols, size_t* bytes) {
  TopSvPlan pl;
  const int rc = top_sv_plan(batch, rows, cols, &pl);
  if (rc) return rc;
  *bytes = pl.total;
  return PDAMD_OK;
}

int pdamd_top_singular_values(const double* a_dev, int batch, int64_t rows, int64_t cols,
                              double* sv_dev, void* workspace_dev, size_t workspace_bytes,
                              void* stream) {
  TopSvPlan pl;
  int rc = top_sv_plan(batch, rows, cols, &pl);
  if (rc) return rc;
  if (!a_dev || !sv_dev) return fail(PDAMD_EINVAL, "null data pointer");
  if (workspace_bytes < pl.total)
    return fail(PDAMD_EWORKSPACE, "workspace %zu < %zu bytes", workspace_bytes, pl.total);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  char* ws = reinterpret_cast<char*>(workspace_dev);
  SyrkArgs s = pl.s;
  s.A = a_dev;
  s.part = reinterpret_cast<double*>(ws + pl.off_part);
  double* H = reinterpret_cast<double*>(ws + pl.off_H);
  double* DE = reinterpret_cast<double*>(ws + pl.off_DE);
  double* TAU = reinterpret_cast<double*>(ws + pl.off_TAU);
  syrk_part_kernel<<<dim3(s.ntile * s.splits, batch), kSNT, 0, st>>>(s);
  HIP_TRY(hipGetLastError());
  const int64_t tot = (int64_t)batch * s.n * s.n;
  syrk_reduce_kernel<<<(int)std::min<int64_t>((tot + 255) / 256, 65536), 256, 0, st>>>(s, batch, H);
  HIP_TRY(hipGetLastError());
  rc = tridiag_f64(H, DE, TAU, batch, s.n, st);
  if (rc) return rc;
  top_eig_kernel<<<batch, kWave, 0, st>>>(DE, s.n, sv_dev);
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
}

}  // extern "C"
