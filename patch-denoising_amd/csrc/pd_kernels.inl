// Space-time low-rank patch denoising on MI355X (gfx950).
//
// Forward pipeline for one volume (all launches on the caller's stream):
//   1. mask_window_kernel / select_kernel / scan / compact   patch selection
//   2. gram_kernel      per patch: gather the Casorati matrix, centre it
//                       (per-frame spatial mean, same summation order as
//                       NumPy), form the min-side Gram (FP32 on MFMA 32x32x2)
//   3. tridiag_kernel   per patch: register-resident Householder
//                       tridiagonalisation of the Gram
//   4. trieig_kernel    per patch, one wavefront: multisection on the
//                       tridiagonal (Sturm counts), the method's threshold
//                       rule, inverse iteration for the kept components,
//                       back-transformation to Gram eigenvectors
//   5. synth_kernel     per patch: re-gather, project on the kept
//                       components, scatter-accumulate theta * patch into
//                       the output and theta into float64 weight sums
//   6. finalize_kernel  float32 weights, out /= weights, masked voxels -> 0
// Backward of the optimal rule (pd_adjoint.inl): steps 1-3 again, then
// trieig_all_kernel + backtransform_kernel (every eigenpair), ghat_kernel,
// adj_project_kernel, adj_spectral_kernel, adj_synth_kernel.
#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <type_traits>
#include <vector>

#include "patch_denoise_amd.h"
#include "pd_common.h"
#include "pd_engine.h"
#include "pd_gather.h"
#include "pd_tridiag.h"

using namespace pd;

namespace {

constexpr int kNT = 256;  // threads per block for the per-patch block kernels

// ===========================================================================
// 1. Selection (base.py:201-227)
// ===========================================================================

// cnt[v * gt + it] = number of set mask entries of voxel v inside temporal
// window it.  A spatial mask counts pt per set voxel.
__global__ void mask_window_kernel(const uint8_t* __restrict__ mask, int kind, int64_t nvox,
                                   int64_t T, int gt, int st, int pt, int* __restrict__ cnt) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nvox * gt) return;
  const int64_t v = q / gt;
  const int it = (int)(q % gt);
  int c = 0;
  if (kind == PDAMD_MASK_SPATIAL) {
    c = mask[v] ? pt : 0;
  } else {
    const uint8_t* m = mask + v * T + (int64_t)it * st;
    for (int t = 0; t < pt; ++t) c += m[t] ? 1 : 0;
  }
  cnt[q] = c;
}

// flags[L] = 100 * count / size > thr ; per-block totals for the compaction.
// Each block covers kSelPerBlock consecutive patches.
constexpr int kSelPerThread = 4;
constexpr int kSelPerBlock = kNT * kSelPerThread;

__global__ __launch_bounds__(kNT) void select_kernel(const int* __restrict__ cnt, Geom g,
                                                     int64_t npatch, int kind, double thr,
                                                     double psize, uint8_t* __restrict__ flags,
                                                     int* __restrict__ block_tot) {
  __shared__ int red[kNT / kWave];
  int local = 0;
  const int64_t base = (int64_t)blockIdx.x * kSelPerBlock + threadIdx.x * kSelPerThread;
  for (int k = 0; k < kSelPerThread; ++k) {
    const int64_t L = base + k;
    if (L >= npatch) break;
    bool sel;
    if (kind == PDAMD_MASK_NONE) {
      sel = 100.0 > thr;
    } else {
      int64_t sp;
      int t0;
      patch_origin(g, L, sp, t0);
      const int it = (int)(L % g.gt);
      int64_t c = 0;
      for (int r = 0; r < g.M; ++r) {
        const int64_t v = (sp + row_offset(g, r)) / g.T;
        c += cnt[v * g.gt + it];
      }
      sel = (100.0 * (double)c) / psize > thr;
    }
    flags[L] = sel ? 1 : 0;
    local += sel ? 1 : 0;
  }
  const int tot = block_sum<kNT>(local, red);
  if (threadIdx.x == 0) block_tot[blockIdx.x] = tot;
}

// Exclusive scan of the per-block totals (single block); total -> *count.
__global__ __launch_bounds__(1024) void scan_kernel(int* __restrict__ tot, int nb,
                                                    int64_t* __restrict__ count) {
  __shared__ int buf[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int b0 = 0; b0 < nb; b0 += 1024) {
    const int i = b0 + threadIdx.x;
    const int v = i < nb ? tot[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int a = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += a;
      __syncthreads();
    }
    if (i < nb) tot[i] = carry + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = carry;
}

// Write the increasing list of selected patch indices.
__global__ __launch_bounds__(kNT) void compact_kernel(const uint8_t* __restrict__ flags,
                                                      int64_t npatch,
                                                      const int* __restrict__ block_off,
                                                      int32_t* __restrict__ out) {
  __shared__ int wsum[kNT / kWave];
  const int64_t base = (int64_t)blockIdx.x * kSelPerBlock + threadIdx.x * kSelPerThread;
  int f[kSelPerThread];
  int local = 0;
  for (int k = 0; k < kSelPerThread; ++k) {
    const int64_t L = base + k;
    f[k] = (L < npatch) ? flags[L] : 0;
    local += f[k];
  }
  // inclusive scan of `local` across the block
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  int incl = local;
  for (int off = 1; off < kWave; off <<= 1) {
    const int a = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += a;
  }
  if (lane == kWave - 1) wsum[wid] = incl;
  __syncthreads();
  int wbase = 0;
  for (int w = 0; w < wid; ++w) wbase += wsum[w];
  int pos = block_off[blockIdx.x] + wbase + incl - local;
  for (int k = 0; k < kSelPerThread; ++k) {
    if (f[k]) out[pos++] = (int32_t)(base + k);
  }
}

// ===========================================================================
// 2. Gram + Householder tridiagonalisation (utils.py:102-137 / :58-78)
// ===========================================================================

// Offset of packed Householder row k (columns k+1 .. n-1) in an n x n reduction.
__host__ __device__ __forceinline__ int hh_row(int k, int n) { return k * (n - 1) - (k * (k - 1)) / 2; }
__host__ __device__ __forceinline__ size_t hh_size(int n) { return (size_t)n * (n - 1) / 2; }

// a - x*y - z*t as explicit fused multiply-adds.
__device__ __forceinline__ float rank2_update(float a, float x, float y, float z, float t) {
  return fmaf(-z, t, fmaf(-x, y, a));
}
__device__ __forceinline__ double rank2_update(double a, double x, double y, double z, double t) {
  return fma(-z, t, fma(-x, y, a));
}
template <class R>
__device__ __forceinline__ cpx<R> rank2_update(cpx<R> a, cpx<R> x, cpx<R> y, cpx<R> z, cpx<R> t) {
  R re = fma(-x.x, y.x, a.x);
  re = fma(x.y, y.y, re);
  re = fma(-z.x, t.x, re);
  re = fma(z.y, t.y, re);
  R im = fma(-x.x, y.y, a.y);
  im = fma(-x.y, y.x, im);
  im = fma(-z.x, t.y, im);
  im = fma(-z.y, t.x, im);
  return {re, im};
}

// Householder tridiagonalisation of an n x n Hermitian matrix (n <= 128)
// held in registers.  Thread (w = tid / 64, l = tid % 64) owns the entries
// A[r][c] for rows r = w + 4 jj (jj < RJ) and columns c in {l, l + 64}.  The
// reduction follows LAPACK ?hetd2 (lower): T = Q^H A Q with
// Q = H_0 H_1 ... H_{n-2}, H_k = I - tau_k v_k v_k^H.
//
// Per step k, with two workgroup barriers:
//   publish  the wave owning row k wrote conj(row k) = column k into X (and
//            the diagonal into dd) at the end of step k-1's update;
//   (b)      every wave forms the reflector from X (DPP sums give the same
//            bits in every wave) and keeps v in registers for its own columns
//            and, gathered by row slot, in its private LDS copy Vr;
//   (c)      partial products p_c = sum over own rows of conj(A[r][c]) v_r -> P;
//   (d)      every wave: p = tau sum_w P, K = -tau/2 p^H v, w = p + K v (own
//            columns in registers, own rows in Wr);
//   (e)      A -= v w^H + w v^H on own entries; the owner of row k+1 publishes
//            it for the next step.
// v and w are zero at indices <= k and >= n, so no entry needs masking; whole
// quarters of inactive rows are skipped with wave-uniform branches.
// Outputs: dd / ee (real tridiagonal), tt (tau_k), and v_k in packed row k of
// Hrow (LDS; row k starts at hh_row(k, n) and holds columns k+1 .. n-1,
// v_k[0] = 1).
template <class E, int RJ>
__device__ void tridiag_reg(E (&a)[RJ][2], int n, typename real_of<E>::type* dd,
                            typename real_of<E>::type* ee, E* tt, E* X, E* Vall, E* Wall, E* P,
                            E* Hrow) {
  using R = typename real_of<E>::type;
  constexpr bool TWO = RJ * 4 > 64;  // second column set exists
  constexpr int QJ = RJ / 4;
  const int tid = threadIdx.x, l = tid & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(tid / kWave);  // wave-uniform (SGPR)
  const int i0 = l, i1 = l + kWave;  // absolute columns held by this lane
  const E zero = E(R(0));
  E* Vr = Vall + w * 32;  // v at this wave's rows, by row slot jj
  E* Wr = Wall + w * 32;
  // publish row k of the current matrix (conjugated) as column k
  auto publish = [&](int jj, int k) {
    const E c0 = a[jj][0], c1 = a[jj][1];
    if (i0 > k && i0 < n) X[i0] = conj_(c0);
    if (TWO && i1 > k && i1 < n) X[i1] = conj_(c1);
    if (i0 == k) dd[k] = re_(c0);
    if (TWO && i1 == k) dd[k] = re_(c1);
  };
  if (w == 0) publish(0, 0);
  for (int k = 0; k < n - 1; ++k) {
    lds_barrier();
    // (b) every wave: reflector from x = X[k+1 .. n-1]
    const E x0 = (i0 > k && i0 < n) ? X[i0] : zero;
    const E x1 = (i1 > k && i1 < n) ? X[i1] : zero;
    const R xn2 = dpp_sum((i0 > k + 1 ? abs2_(x0) : R(0)) + (i1 > k + 1 ? abs2_(x1) : R(0)));
    const E alpha = X[k + 1];
    const R ar = re_(alpha), ai = im_(alpha);
    const bool refl = (xn2 > R(0)) || (ai != R(0));
    E v0 = zero, v1 = zero, tk = zero;
    R beta = ar;
    if (refl) {
      beta = -copysign(sqrt(ar * ar + ai * ai + xn2), ar);
      const R ibeta = R(1) / beta;
      tk = mk<E>((beta - ar) * ibeta, -ai * ibeta);
      const E scal = recip_(alpha - E(beta));
      v0 = (i0 > k && i0 < n) ? ((i0 == k + 1) ? E(R(1)) : x0 * scal) : zero;
      v1 = (i1 > k && i1 < n) ? ((i1 == k + 1) ? E(R(1)) : x1 * scal) : zero;
    }
    if ((i0 & 3) == w) Vr[i0 >> 2] = v0;
    if ((i1 & 3) == w) Vr[i1 >> 2] = v1;
    if (w == 0) {
      const int pk = hh_row(k, n) - (k + 1);  // packed row k, absolute column index
      if (i0 > k && i0 < n) Hrow[pk + i0] = v0;
      if (i1 > k && i1 < n) Hrow[pk + i1] = v1;
      if (l == 0) {
        tt[k] = tk;
        ee[k] = beta;
      }
    }
    if (!refl) {
      // H_k = I (uniform: every wave computed the same xn2, alpha).  Row k+1
      // is unchanged; publish it after every wave has read X.
      lds_barrier();
      if (w == ((k + 1) & 3)) {
#pragma unroll
        for (int jj = 0; jj < RJ; ++jj)
          if (w + 4 * jj == k + 1) publish(jj, k + 1);
      }
      continue;
    }
    wave_sync();
    // (c) partial p_c = sum over my rows r of conj(A[r][c]) v_r
    E p0a = zero, p0b = zero, p1a = zero, p1b = zero;
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      if (w + 4 * (qq * QJ + QJ - 1) > k) {
#pragma unroll
        for (int u = 0; u < QJ; u += 2) {
          const int jj = qq * QJ + u;
          const E va = Vr[jj], vb = Vr[jj + 1];
          p0a += conj_(a[jj][0]) * va;
          p0b += conj_(a[jj + 1][0]) * vb;
          if (TWO) {
            p1a += conj_(a[jj][1]) * va;
            p1b += conj_(a[jj + 1][1]) * vb;
          }
        }
      }
    }
    P[w * 128 + i0] = p0a + p0b;
    P[w * 128 + i1] = p1a + p1b;
    lds_barrier();
    // (d) every wave: p = tau * sum_w P, K = -tau/2 p^H v, w = p + K v
    E q0 = P[i0] + P[128 + i0] + P[256 + i0] + P[384 + i0];
    E q1 = P[i1] + P[128 + i1] + P[256 + i1] + P[384 + i1];
    q0 = (i0 > k && i0 < n) ? tk * q0 : zero;
    q1 = (i1 > k && i1 < n) ? tk * q1 : zero;
    const E pv = dpp_sum(conj_(q0) * v0 + conj_(q1) * v1);
    const E kc = (R(-0.5) * tk) * pv;
    const E w0 = q0 + kc * v0, w1 = q1 + kc * v1;
    if ((i0 & 3) == w) Wr[i0 >> 2] = w0;
    if ((i1 & 3) == w) Wr[i1 >> 2] = w1;
    wave_sync();
    // (e) A -= v w^H + w v^H on my entries; publish row k+1
    const bool set0 = k < kWave - 1;  // some column l > k exists
    const E cw0 = conj_(w0), cv0 = conj_(v0), cw1 = conj_(w1), cv1 = conj_(v1);
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      if (w + 4 * (qq * QJ + QJ - 1) > k) {
#pragma unroll
        for (int u = 0; u < QJ; ++u) {
          const int jj = qq * QJ + u;
          const E vr = Vr[jj], wr = Wr[jj];
          if (set0) a[jj][0] = rank2_update(a[jj][0], vr, cw0, wr, cv0);
          if (TWO) a[jj][1] = rank2_update(a[jj][1], vr, cw1, wr, cv1);
          if (w + 4 * jj == k + 1) publish(jj, k + 1);
        }
      }
    }
  }
  lds_barrier();
}

// s + a[0] + a[stride] + ... + a[(count-1) stride], added left to right
// (NumPy's order for an axis-0 reduction); loads are issued 8 at a time.
template <class T>
__device__ __forceinline__ T seq_add(T s, const T* a, int count, int stride) {
  int r = 0;
  for (; r + 8 <= count; r += 8) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = a[(r + u) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; r < count; ++r) s += a[r * stride];
  return s;
}

// s + a[off[0]] + a[off[1]] + ... in that order, 8 loads in flight.
template <class T>
__device__ __forceinline__ T seq_add(T s, const T* a, const int64_t* off, int count) {
  int r = 0;
  for (; r + 8 <= count; r += 8) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = a[off[r + u]];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; r < count; ++r) s += a[off[r]];
  return s;
}

typedef float floatx16 __attribute__((ext_vector_type(16)));

struct GramArgs {
  Geom g;
  const int32_t* idx;  // selected patch indices (this chunk)
  int np;              // patches in this chunk
  int n;               // Gram size = min(M, pt)
  int scale;           // 1: divide the Gram by (M - 1) (eigen rules)
  void* H;             // np x n x n   Householder rows
  void* DE;            // np x 3n      d | e | e^2 (real type)
  void* TAU;           // np x n
  void* MEAN;          // np x pt per-frame patch means (input type), reused by the synthesis
  void* MEANW;         // M-side: np x pt per-frame means out (gather recombination), or null
  const int* np_dev;   // device patch count (<= np, read by the kernel), or null: np
};

// Staging geometry: M-side stages kTC frames x (all rows); T-side stages kRC
// rows x (all frames).
template <class E> struct Stage { static constexpr int TC = 32, RC = 32; };
template <> struct Stage<double> { static constexpr int TC = 8, RC = 8; };
template <> struct Stage<cpx<float>> { static constexpr int TC = 8, RC = 8; };
template <> struct Stage<cpx<double>> { static constexpr int TC = 8, RC = 8; };

template <class T> __device__ __forceinline__ T* lds_carve(char*& p, size_t count) {
  T* r = reinterpret_cast<T*>(p);
  p += ((count * sizeof(T) + 15) / 16) * 16;
  return r;
}

// Divide by a positive integer count in the element's precision (NumPy
// true_divide; for complex values each component is divided).
__device__ __forceinline__ float div_n(float a, int n) { return a / (float)n; }
__device__ __forceinline__ double div_n(double a, int n) { return a / (double)n; }
template <class R> __device__ __forceinline__ cpx<R> div_n(cpx<R> a, int n) {
  return {a.x / (R)n, a.y / (R)n};
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int kBfW = 32 + 8;  // bf16 staging row: SC frames + pad (80 B, 16-B aligned)

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Lower-triangle 16 x 16 tiles (I >= J) of a padded 128 x 128 Gram, nine per
// wave (36 in all): wave 0 the leading 4 x 4 block triangle without (3, 3),
// wave 1 the trailing one without (7, 7), waves 2 / 3 block rows 4..7 x block
// columns {0, 1} / {2, 3} plus (3, 3) / (7, 7).
template <int W> struct Tri16;
template <> struct Tri16<0> {
  static constexpr int I[9] = {0, 1, 1, 2, 2, 2, 3, 3, 3}, J[9] = {0, 0, 1, 0, 1, 2, 0, 1, 2};
};
template <> struct Tri16<1> {
  static constexpr int I[9] = {4, 5, 5, 6, 6, 6, 7, 7, 7}, J[9] = {4, 4, 5, 4, 5, 6, 4, 5, 6};
};
template <> struct Tri16<2> {
  static constexpr int I[9] = {4, 4, 5, 5, 6, 6, 7, 7, 3}, J[9] = {0, 1, 0, 1, 0, 1, 0, 1, 3};
};
template <> struct Tri16<3> {
  static constexpr int I[9] = {4, 4, 5, 5, 6, 6, 7, 7, 7}, J[9] = {2, 3, 2, 3, 2, 3, 2, 3, 7};
};
template <int W> constexpr bool tri16_uses(int b) {
  for (int t = 0; t < 9; ++t)
    if (Tri16<W>::I[t] == b || Tri16<W>::J[t] == b) return true;
  return false;
}

// acc[t] += Xc[16 I_t .., k] Xc[16 J_t .., k]^T over reduction indices
// k < kcount on mfma_f32_16x16x4f32, centring while the operands are read:
// S holds the raw staged values [k][i] (row length ldw), Xc = S - mean with
// the mean by reduction index (M-side: per frame) or by Gram index (T-side);
// indices k >= kcount contribute zero.  Each 16-index block the wave needs is
// read once per step of 4.
template <int W, bool MSIDE>
__device__ __forceinline__ void tri16_accum(const float* __restrict__ S, int ldw, int kcount,
                                            int lane, const float* __restrict__ mean,
                                            floatx4 (&acc)[9]) {
  float mi[8];
#pragma unroll
  for (int b = 0; b < 8; ++b) mi[b] = (!MSIDE && tri16_uses<W>(b)) ? mean[16 * b + (lane & 15)] : 0.f;
  auto load = [&](int k, float (&v)[8]) {
    const int kk = k + (lane >> 4);
    const bool kv = kk < kcount;
    const float* row = S + kk * ldw + (lane & 15);
    const float mk = MSIDE ? mean[kk] : 0.f;
#pragma unroll
    for (int b = 0; b < 8; ++b)
      v[b] = (tri16_uses<W>(b) && kv) ? row[16 * b] - (MSIDE ? mk : mi[b]) : 0.f;
  };
  for (int k = 0; k < kcount; k += 4) {
    float cur[8];
    load(k, cur);
#pragma unroll
    for (int t = 0; t < 9; ++t)
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur[Tri16<W>::I[t]], cur[Tri16<W>::J[t]],
                                                    acc[t], 0, 0, 0);
  }
}

// FP32 16 x 16 tiles: the staged chunk Gram-index-major, S[i * LD + p8(k)]
// for Gram index i and chunk reduction index k (M-side: frames, LD = 66;
// T-side: voxel rows, LD = 34), where p8 stores k and k + 4 of each group of
// 8 side by side: one 64-bit read gives a lane its operands for two
// consecutive MFMA steps (the same operands in the same order as a read per
// step).  With these LD a 16-lane read is bank-conflict free.
constexpr int kPLD = 66, kPLDT = 34;
__device__ __forceinline__ int p8(int t) { return (t & ~7) | ((t & 3) << 1) | ((t >> 2) & 1); }

// MSIDE: mean by reduction index (mean[p8(k)]); else by Gram index (mean[i])
template <int W, bool MSIDE>
__device__ __forceinline__ void tri16_accum_p8(const float* __restrict__ S, int kcount, int lane,
                                               const float* __restrict__ mean,
                                               floatx4 (&acc)[9]) {
  constexpr int LD = MSIDE ? kPLD : kPLDT;
  const int j = lane & 15, g = lane >> 4;
  float mi[8];
#pragma unroll
  for (int b = 0; b < 8; ++b) mi[b] = (!MSIDE && tri16_uses<W>(b)) ? mean[16 * b + j] : 0.f;
  for (int k8 = 0; k8 < kcount; k8 += 8) {
    const float2 m2 = MSIDE ? *reinterpret_cast<const float2*>(mean + k8 + 2 * g) : float2{0.f, 0.f};
    float2 v2[8];
#pragma unroll
    for (int b = 0; b < 8; ++b)
      v2[b] = tri16_uses<W>(b) ? *reinterpret_cast<const float2*>(S + (16 * b + j) * LD + k8 + 2 * g)
                               : float2{0.f, 0.f};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const bool kv = k8 + 4 * h + g < kcount;  // this lane's reduction index at this step
      float cur[8];
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const float mk = MSIDE ? (h ? m2.y : m2.x) : mi[b];
        cur[b] = (tri16_uses<W>(b) && kv) ? (h ? v2[b].y : v2[b].x) - mk : 0.f;
      }
#pragma unroll
      for (int t = 0; t < 9; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(cur[Tri16<W>::I[t]], cur[Tri16<W>::J[t]],
                                                      acc[t], 0, 0, 0);
    }
  }
}

template <class IN, class E, bool TSIDE, int TILE, bool BF = false>
__global__ __launch_bounds__(kNT, 2) void gram_kernel(const IN* __restrict__ vol,
                                                           GramArgs a) {
  using R = typename real_of<E>::type;
  constexpr int NB = 16 * TILE;  // padded Gram size
  constexpr bool SAME = std::is_same<IN, E>::value;
  // M-side FP32 MFMA path (no split): 64 frames per staged chunk, so 64
  // threads compute the per-frame means (sequential over rows) at a time
  constexpr int SC = TSIDE ? Stage<E>::RC
                           : ((std::is_same<E, float>::value && TILE == 8 && !BF) ? 64 : Stage<E>::TC);
  constexpr int LDS_W = NB + 1;  // staging row length (padded)
  const Geom& g = a.g;
  const int n = a.n;
  const int tid = threadIdx.x;
  const int ty = tid / 16, tx = tid % 16;
  // FP32 real Grams of size 65..128 use MFMA 32x32x2: wave w owns the 64x64
  // block (w >> 1, w & 1), i.e. 2x2 tiles of 32x32 (16 accumulators per lane each).
  constexpr bool MF = std::is_same<E, float>::value && TILE == 8;
  // Split-BF16 (both sides): each centred value x = hi + lo with hi = bf16(x),
  // lo = bf16(x - hi); G += hi hi^T + hi lo^T + lo hi^T on MFMA 32x32x16 bf16
  // with FP32 accumulation (the lo lo^T term, ~2^-16 relative, is dropped).
  constexpr bool BFM = BF && MF && SC == 32;  // M-side: planes [row][frame]; T-side: [frame][row]
  // FP32 (no split): lower-triangle 16 x 16 tiles on MFMA 16x16x4 (Tri16)
  constexpr bool M16 = MF && !BF;
  const int wv = tid / kWave, lane = tid % kWave;
  const int bi = wv >> 1, bj = wv & 1, lhi = lane >> 5, llo = lane & 31;

  extern __shared__ __align__(16) char smem[];
  char* cur = smem;
  // FP32 16 x 16 tiles (P8): the Gram-index-major permuted staging
  constexpr bool P8 = M16 && (TSIDE ? SC == 32 : SC == 64);
  constexpr int PLD = TSIDE ? kPLDT : kPLD;
  IN* Sraw = lds_carve<IN>(cur, P8 ? (size_t)NB * PLD : (size_t)SC * LDS_W);
  E* Sc = SAME ? reinterpret_cast<E*>(Sraw) : lds_carve<E>(cur, (size_t)SC * LDS_W);
  IN* mean = lds_carve<IN>(cur, NB);
  int64_t* roff = lds_carve<int64_t>(cur, TSIDE ? g.M : NB);  // row offsets of the patch
  __bf16* Sh = BFM ? lds_carve<__bf16>(cur, (size_t)NB * kBfW) : nullptr;  // Sh[i][k]
  __bf16* Sl = BFM ? lds_carve<__bf16>(cur, (size_t)NB * kBfW) : nullptr;

  for (int p = xcd_rank(), np_ = a.np_dev ? *a.np_dev : a.np; p < np_; p += gridDim.x) {
    int64_t sp;
    int t0;
    patch_origin(g, a.idx[p], sp, t0);
    const IN* X = vol + sp + t0;
    const int M = g.M, T = g.pt;

    E acc[MF ? 1 : TILE][MF ? 1 : TILE];
    floatx16 mc[2][2];
    floatx4 m16[M16 ? 9 : 1];
    if constexpr (M16) {
#pragma unroll
      for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) m16[t][q] = 0.0f;
    } else if constexpr (MF) {
#pragma unroll
      for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
        for (int b2 = 0; b2 < 2; ++b2)
#pragma unroll
          for (int q = 0; q < 16; ++q) mc[a2][b2][q] = 0.0f;
    } else {
#pragma unroll
      for (int i = 0; i < TILE; ++i)
#pragma unroll
        for (int j = 0; j < TILE; ++j) acc[i][j] = E(R(0));
    }
    // one reduction index k of the staged S[k][i] (i = Gram row/column index)
    auto accum = [&](const E* Sk, int kcount) {
      if constexpr (BFM) {
        // lane: rows 64 b + 32 h + llo, frames k + 8 lhi .. + 7 (one 16-B load each)
        auto ld = [&](const __bf16* S, int row, int k) {
          return *reinterpret_cast<const bf16x8*>(S + row * kBfW + k + 8 * lhi);
        };
#pragma unroll
        for (int k = 0; k < SC; k += 16) {  // frames >= kcount are zero in Sh / Sl
          const bf16x8 ha[2] = {ld(Sh, 64 * bi + llo, k), ld(Sh, 64 * bi + 32 + llo, k)};
          const bf16x8 la[2] = {ld(Sl, 64 * bi + llo, k), ld(Sl, 64 * bi + 32 + llo, k)};
          const bf16x8 hb[2] = {ld(Sh, 64 * bj + llo, k), ld(Sh, 64 * bj + 32 + llo, k)};
          const bf16x8 lb[2] = {ld(Sl, 64 * bj + llo, k), ld(Sl, 64 * bj + 32 + llo, k)};
#pragma unroll
          for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
            for (int b2 = 0; b2 < 2; ++b2) {
              mc[a2][b2] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha[a2], hb[b2], mc[a2][b2], 0, 0, 0);
              mc[a2][b2] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha[a2], lb[b2], mc[a2][b2], 0, 0, 0);
              mc[a2][b2] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(la[a2], hb[b2], mc[a2][b2], 0, 0, 0);
            }
        }
      } else if constexpr (M16) {  // Sk: the raw staging, centred on the fly
        const float* S = reinterpret_cast<const float*>(Sk);
        const float* mf = reinterpret_cast<const float*>(mean);
        if constexpr (P8) {
          switch (wv) {
            case 0: tri16_accum_p8<0, !TSIDE>(S, kcount, lane, mf, m16); break;
            case 1: tri16_accum_p8<1, !TSIDE>(S, kcount, lane, mf, m16); break;
            case 2: tri16_accum_p8<2, !TSIDE>(S, kcount, lane, mf, m16); break;
            default: tri16_accum_p8<3, !TSIDE>(S, kcount, lane, mf, m16); break;
          }
          return;
        }
        switch (wv) {
          case 0: tri16_accum<0, !TSIDE>(S, LDS_W, kcount, lane, mf, m16); break;
          case 1: tri16_accum<1, !TSIDE>(S, LDS_W, kcount, lane, mf, m16); break;
          case 2: tri16_accum<2, !TSIDE>(S, LDS_W, kcount, lane, mf, m16); break;
          default: tri16_accum<3, !TSIDE>(S, LDS_W, kcount, lane, mf, m16); break;
        }
      }
    };

    constexpr int NL = (SC * NB + kNT - 1) / kNT;  // staged elements per thread
    IN buf[NL];
    if (!TSIDE) {
      // ---- M-side: G = Xc Xc^H (M x M), frames streamed in chunks of SC.
      // Global loads of chunk c+1 are issued before chunk c is reduced.
      for (int r = tid; r < NB; r += kNT) roff[r] = r < M ? row_offset(g, r) : 0;
      lds_barrier();
      auto load = [&](int c0) {
        const int tc = min(SC, T - c0);
#pragma unroll
        for (int u = 0; u < NL; ++u) {
          const int q = tid + u * kNT;
          const int t = q % SC, r = q / SC;
          buf[u] = (q < SC * NB && t < tc && r < M) ? X[roff[r] + c0 + t] : IN(0);
        }
      };
      load(0);
      for (int c0 = 0; c0 < T; c0 += SC) {
        const int tc = min(SC, T - c0);
#pragma unroll
        for (int u = 0; u < NL; ++u) {
          const int q = tid + u * kNT;
          if (q < SC * NB) {
            if constexpr (P8) Sraw[(q / SC) * kPLD + p8(q % SC)] = buf[u];
            else Sraw[(q % SC) * LDS_W + q / SC] = buf[u];
          }
        }
        lds_barrier();
        if (c0 + SC < T) load(c0 + SC);
        if (tid < tc) {  // sequential over rows, as numpy's axis-0 reduction
          IN mv;
          if constexpr (P8) {
            mv = div_n(seq_add(Sraw[p8(tid)], Sraw + kPLD + p8(tid), M - 1, kPLD), M);
            mean[p8(tid)] = mv;
          } else {
            mv = div_n(seq_add(Sraw[tid * LDS_W], Sraw + tid * LDS_W + 1, M - 1, 1), M);
            mean[tid] = mv;
          }
          if (a.MEANW) reinterpret_cast<IN*>(a.MEANW)[(size_t)p * T + c0 + tid] = mv;
        }
        lds_barrier();
        if constexpr (M16) {  // centred while the MFMA operands are read
          accum(Sc, tc);
          lds_barrier();
          continue;
        }
        for (int q = tid; q < SC * NB; q += kNT) {
          const int t = q / NB, r = q % NB;
          const IN v = Sraw[t * LDS_W + r];
          const E x = (t < tc && r < M) ? cvt<E>(v - mean[t]) : E(R(0));
          if constexpr (BFM) {
            const __bf16 hi = (__bf16)x;
            Sh[r * kBfW + t] = hi;
            Sl[r * kBfW + t] = (__bf16)(x - (float)hi);
          } else {
            Sc[t * LDS_W + r] = x;
          }
        }
        lds_barrier();
        if constexpr (MF) {
          accum(Sc, tc);  // rows t >= tc of the staging are zero
        } else {
          for (int t = 0; t < tc; ++t) {
            E xa[TILE], xb[TILE];
#pragma unroll
            for (int i = 0; i < TILE; ++i) xa[i] = Sc[t * LDS_W + ty + 16 * i];
#pragma unroll
            for (int j = 0; j < TILE; ++j) xb[j] = conj_(Sc[t * LDS_W + tx + 16 * j]);
#pragma unroll
            for (int i = 0; i < TILE; ++i)
#pragma unroll
              for (int j = 0; j < TILE; ++j) acc[i][j] += xa[i] * xb[j];
          }
        }
        lds_barrier();
      }
    } else {
      // ---- T-side: G = Xc^H Xc (T x T); pass 1 = per-frame mean over all rows
      for (int r = tid; r < M; r += kNT) roff[r] = row_offset(g, r);
      lds_barrier();
      auto load = [&](int r0) {
        const int rc = min(SC, M - r0);
#pragma unroll
        for (int u = 0; u < NL; ++u) {
          const int q = tid + u * kNT;
          const int r = q / NB, t = q % NB;
          buf[u] = (q < SC * NB && r < rc && t < T) ? X[roff[r0 + r] + t] : IN(0);
        }
      };
      auto store = [&]() {
#pragma unroll
        for (int u = 0; u < NL; ++u) {
          const int q = tid + u * kNT;
          if (q < SC * NB) {
            if constexpr (P8) Sraw[(q % NB) * PLD + p8(q / NB)] = buf[u];
            else Sraw[(q / NB) * LDS_W + q % NB] = buf[u];
          }
        }
      };
      IN s = IN(0);
      load(0);
      for (int r0 = 0; r0 < M; r0 += SC) {
        const int rc = min(SC, M - r0);
        store();
        lds_barrier();
        if (r0 + SC < M) load(r0 + SC);
        if (tid < T) {
          if constexpr (P8) {  // rows in order, as seq_add
            const IN* col = Sraw + tid * PLD;
            int r = 0;
            if (r0 == 0) s = col[p8(r++)];
            for (; r < rc; ++r) s += col[p8(r)];
          } else {
            if (r0 == 0) s = seq_add(Sraw[tid], Sraw + LDS_W + tid, rc - 1, LDS_W);
            else s = seq_add(s, Sraw + tid, rc, LDS_W);
          }
        }
        lds_barrier();
      }
      if (tid < NB) mean[tid] = tid < T ? div_n(s, M) : IN(0);
      if (tid < T) reinterpret_cast<IN*>(a.MEAN)[(size_t)p * T + tid] = mean[tid];
      load(0);
      for (int r0 = 0; r0 < M; r0 += SC) {
        const int rc = min(SC, M - r0);
        store();
        lds_barrier();
        if (r0 + SC < M) load(r0 + SC);
        if constexpr (M16) {  // centred while the MFMA operands are read
          accum(Sc, rc);
          lds_barrier();
          continue;
        }
        for (int q = tid; q < SC * NB; q += kNT) {
          const int r = q / NB, t = q % NB;
          E c = E(R(0));
          if (r < rc && t < T) c = cvt<E>(Sraw[r * LDS_W + t] - mean[t]);
          if constexpr (BFM) {  // Gram index t, reduction index r
            const __bf16 hi = (__bf16)c;
            Sh[t * kBfW + r] = hi;
            Sl[t * kBfW + r] = (__bf16)(c - (float)hi);
          } else {
            Sc[r * LDS_W + t] = c;
          }
        }
        lds_barrier();
        if constexpr (MF) {
          accum(Sc, rc);  // rows r >= rc of the staging are zero
        } else {
          for (int r = 0; r < rc; ++r) {
            E xa[TILE], xb[TILE];
#pragma unroll
            for (int i = 0; i < TILE; ++i) xa[i] = conj_(Sc[r * LDS_W + ty + 16 * i]);
#pragma unroll
            for (int j = 0; j < TILE; ++j) xb[j] = Sc[r * LDS_W + tx + 16 * j];
#pragma unroll
            for (int i = 0; i < TILE; ++i)
#pragma unroll
              for (int j = 0; j < TILE; ++j) acc[i][j] += xa[i] * xb[j];
          }
        }
        lds_barrier();
      }
    }

    // ---- Gram -> this patch's H block (row-major, optionally / (M-1) as the
    //      reference's eigen path)
    E* G = reinterpret_cast<E*>(a.H) + (size_t)p * n * n;
    if constexpr (M16) {
      // 16x16 MFMA output: accumulator q of lane l is C[4(l/16) + q][l%16].
      // Only the lower-triangle tiles are written (diagonal tiles whole): the
      // one-wavefront Householder that takes every FP32 Gram of this size
      // reads the blocks on and below the diagonal only, and the mirrored
      // stores (4-byte writes a row apart) cost more than the rest of the
      // Gram's output
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        int I, J;
        switch (wv) {
          case 0: I = Tri16<0>::I[t]; J = Tri16<0>::J[t]; break;
          case 1: I = Tri16<1>::I[t]; J = Tri16<1>::J[t]; break;
          case 2: I = Tri16<2>::I[t]; J = Tri16<2>::J[t]; break;
          default: I = Tri16<3>::I[t]; J = Tri16<3>::J[t]; break;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = 16 * I + 4 * (lane >> 4) + q, c = 16 * J + (lane & 15);
          if (r < n && c < n) {
            float v = m16[t][q];
            if (a.scale) v = div_n(v, M - 1);
            float* Gf = reinterpret_cast<float*>(G);
            Gf[(size_t)r * n + c] = v;
          }
        }
      }
    } else if constexpr (MF) {
      // 32x32 MFMA output: accumulator q of lane l is C[8(q/4) + 4(l/32) + q%4][l%32]
#pragma unroll
      for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
        for (int b2 = 0; b2 < 2; ++b2)
#pragma unroll
          for (int q = 0; q < 16; ++q) {
            const int r = 64 * bi + 32 * a2 + 8 * (q >> 2) + 4 * lhi + (q & 3);
            const int c = 64 * bj + 32 * b2 + llo;
            if (r < n && c < n) {
              float v = mc[a2][b2][q];
              if (a.scale) v = div_n(v, M - 1);
              reinterpret_cast<float*>(G)[(size_t)r * n + c] = v;
            }
          }
    } else {
#pragma unroll
    for (int i = 0; i < TILE; ++i) {
      const int r = ty + 16 * i;
#pragma unroll
      for (int j = 0; j < TILE; ++j) {
        const int c = tx + 16 * j;
        if (r < n && c < n) {
          E v = acc[i][j];
          if (a.scale) v = div_n(v, M - 1);
          G[(size_t)r * n + c] = v;
        }
      }
    }
    }
  }
}

template <class IN, class E, bool TSIDE, int TILE, bool BF = false>
size_t gram_lds_bytes(int n, int M) {
  constexpr int NB = 16 * TILE;
  constexpr int SC = TSIDE ? Stage<E>::RC
                           : ((std::is_same<E, float>::value && TILE == 8 && !BF) ? 64 : Stage<E>::TC);
  constexpr int LDS_W = NB + 1;
  constexpr bool SAME = std::is_same<IN, E>::value;
  auto al = [](size_t b) { return ((b + 15) / 16) * 16; };
  constexpr bool P8 = std::is_same<E, float>::value && TILE == 8 && !BF && (TSIDE ? SC == 32 : SC == 64);
  size_t s = al((P8 ? (size_t)NB * (TSIDE ? kPLDT : kPLD) : (size_t)SC * LDS_W) * sizeof(IN)) +
             (SAME ? 0 : al((size_t)SC * LDS_W * sizeof(E)));
  s += al(NB * sizeof(IN));
  s += al((TSIDE ? (size_t)M : (size_t)NB) * sizeof(int64_t));
  if (BF && std::is_same<E, float>::value && TILE == 8 && SC == 32)
    s += 2 * al((size_t)NB * kBfW * 2);
  return s;
}

// Householder tridiagonalisation of each patch's Gram (read from its H block),
// leaving d | e | e^2, tau and the packed reflectors in that block.
template <class E, int RJ>
__global__ __launch_bounds__(kNT) void tridiag_kernel(GramArgs a) {
  using R = typename real_of<E>::type;
  const int n = a.n;
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid / kWave), l = tid % kWave;
  extern __shared__ __align__(16) char smem[];
  char* cur = smem;
  E* Hl = lds_carve<E>(cur, hh_size(n));
  E* Xb = lds_carve<E>(cur, 128);
  E* V = lds_carve<E>(cur, 4 * 32);
  E* W = lds_carve<E>(cur, 4 * 32);
  E* P = lds_carve<E>(cur, 4 * 128);
  R* dd = lds_carve<R>(cur, 128);
  R* ee = lds_carve<R>(cur, 128);
  E* tt = lds_carve<E>(cur, 128);
  for (int p = blockIdx.x, np_ = a.np_dev ? *a.np_dev : a.np; p < np_; p += gridDim.x) {
    E* Hp = reinterpret_cast<E*>(a.H) + (size_t)p * n * n;
    E areg[RJ][2];
#pragma unroll
    for (int jj = 0; jj < RJ; ++jj) {
      const int r = w + 4 * jj;
      areg[jj][0] = (r < n && l < n) ? Hp[(size_t)r * n + l] : E(R(0));
      areg[jj][1] = (RJ * 4 > 64 && r < n && l + kWave < n) ? Hp[(size_t)r * n + l + kWave]
                                                             : E(R(0));
    }
    tridiag_reg<E, RJ>(areg, n, dd, ee, tt, Xb, V, W, P, Hl);
    __syncthreads();  // every read of the Gram precedes the overwrite below
    for (int q = tid; q < (int)hh_size(n); q += kNT) Hp[q] = Hl[q];
    R* DEp = reinterpret_cast<R*>(a.DE) + (size_t)p * 3 * n;
    E* TAUp = reinterpret_cast<E*>(a.TAU) + (size_t)p * n;
    for (int q = tid; q < n; q += kNT) {
      const R ev = q < n - 1 ? ee[q] : R(0);
      DEp[q] = dd[q];
      DEp[n + q] = ev;
      DEp[2 * n + q] = ev * ev;
      TAUp[q] = q < n - 1 ? tt[q] : E(R(0));
    }
    lds_barrier();
  }
}

template <class E>
size_t tridiag_lds_bytes(int n) {
  using R = typename real_of<E>::type;
  auto al = [](size_t b) { return ((b + 15) / 16) * 16; };
  return al(hh_size(n) * sizeof(E)) + al(128 * sizeof(E)) + 2 * al(4 * 32 * sizeof(E)) +
         al(4 * 128 * sizeof(E)) + 2 * al(128 * sizeof(R)) + al(128 * sizeof(E));
}

// ---------------------------------------------------------------------------
// Gram sizes 129 .. kNBig (e.g. 11^3-voxel patches of 200 frames, or 7^3-voxel
// patches of 300 frames): the Gram no longer fits one workgroup's registers or
// LDS, so it lives in the patch's H block in global memory (L2 / MALL
// resident while the workgroup works on it).
//   gram_big_kernel     per patch: per-frame means (rows added in order), then
//                       the Gram in 64 x 64 output tiles (lower tiles, each
//                       mirrored), reduction staged through LDS 32 at a time;
//                       FP32 real / complex on MFMA 32x32x2 (a wave per 32 x 32
//                       quarter; complex = 4 real products), FP64 on the VALU
//   tridiag_big_kernel  per patch: unblocked Householder (LAPACK ?hetd2,
//                       lower) on the full Hermitian matrix in global memory,
//                       one thread per column for the product and the rank-2
//                       update (coalesced row sweeps); reflector k is stored
//                       in row k, columns k+1 .. n-1, a part of the matrix
//                       that no later step reads
// ---------------------------------------------------------------------------
constexpr int kNBig = 2048;  // largest supported Gram size
constexpr int kNMid = 512;   // largest Gram size of the 8-rows-per-lane eigenvector kernel
constexpr int kGBT = 64;    // Gram output tile
constexpr int kGBK = 32;    // reduction indices per staged chunk

// Split-BF16 staging row (kGBK indices + pad: 80 bytes, 16-byte aligned rows).
constexpr int kGBW = kGBK + 8;
typedef short short8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ bf16x8 neg_bf16x8(bf16x8 v) {
  return __builtin_bit_cast(bf16x8, __builtin_bit_cast(short8, v) ^ (short8)(short)0x8000);
}

// BF: split-BF16 three-product Gram (each centred value x = hi + lo, hi =
// bf16(x), lo = bf16(x - hi); sum hi hi + hi lo + lo hi on MFMA 32x32x16 bf16
// with FP32 accumulation; complex as four real products), FP32 data only.
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <class IN, class E, bool TSIDE, bool BF = false>
__global__ __launch_bounds__(kNT) void gram_big_kernel(const IN* __restrict__ vol, GramArgs a) {
  using R = typename real_of<E>::type;
  const Geom& g = a.g;
  const int n = a.n, M = g.M, T = g.pt;
  const int K = TSIDE ? M : T;  // reduction length
  const int tid = threadIdx.x, ty = tid / 16, tx = tid % 16;
  constexpr int LW = kGBT + 1;
  constexpr bool MF = std::is_same<R, float>::value;  // FP32 real / complex on MFMA
  constexpr bool MD = std::is_same<R, double>::value;  // FP64 real / complex on MFMA 16x16x4
  constexpr bool CPX = is_cpx<E>::value;
  constexpr bool BFM = BF && MF;
  constexpr int NPL = CPX ? 2 : 1;  // real planes per operand (re, im)
  const int wv = tid / kWave, lane = tid % kWave;
  const int wi = wv >> 1, wj = wv & 1, lhi = lane >> 5, llo = lane & 31;
  extern __shared__ __align__(16) char smem[];
  char* cur = smem;
  IN* mean = lds_carve<IN>(cur, T);
  int64_t* roff = lds_carve<int64_t>(cur, M);
  E* As = BFM ? nullptr : lds_carve<E>(cur, (size_t)kGBK * LW);  // As[k][i]
  E* Bs = BFM ? nullptr : lds_carve<E>(cur, (size_t)kGBK * LW);
  // split-BF16 planes [operand A/B][re/im][hi/lo], each kGBT rows of kGBW
  __bf16* Pl = BFM ? lds_carve<__bf16>(cur, (size_t)2 * NPL * 2 * kGBT * kGBW) : nullptr;
  auto plane = [&](int op, int ri, int hl) {
    return Pl + ((size_t)((op * NPL + ri) * 2 + hl)) * kGBT * kGBW;
  };
  const int nb = (n + kGBT - 1) / kGBT;
  const int ntile = nb * (nb + 1) / 2;
  for (int p = xcd_rank(), np_ = a.np_dev ? *a.np_dev : a.np; p < np_; p += gridDim.x) {
    int64_t sp;
    int t0;
    patch_origin(g, a.idx[p], sp, t0);
    const IN* X = vol + sp + t0;
    for (int r = tid; r < M; r += kNT) roff[r] = row_offset(g, r);
    lds_barrier();
    // per-frame mean over the rows, added in row order (numpy's axis-0 reduction)
    for (int t = tid; t < T; t += kNT) {
      mean[t] = div_n(seq_add(X[roff[0] + t], X + t, roff + 1, M - 1), M);
      if (TSIDE) reinterpret_cast<IN*>(a.MEAN)[(size_t)p * T + t] = mean[t];
    }
    lds_barrier();
    E* G = reinterpret_cast<E*>(a.H) + (size_t)p * n * n;
    // element (i, k) of the n x K operand, centred: T-side Xc[k][i], M-side Xc[i][k]
    auto stage = [&](E* S, int b0, int k0, int kc) {
      for (int q = tid; q < kGBK * kGBT; q += kNT) {
        int i, k;
        if (TSIDE) {  // consecutive threads: consecutive frames of one row
          k = q / kGBT;
          i = q % kGBT;
        } else {  // consecutive threads: consecutive frames of one row
          i = q / kGBK;
          k = q % kGBK;
        }
        E v = E(R(0));
        if (k < kc && b0 + i < n) {
          const int rr = TSIDE ? k0 + k : b0 + i, tt = TSIDE ? b0 + i : k0 + k;
          v = cvt<E>(X[roff[rr] + tt] - mean[tt]);
        }
        S[k * LW + i] = v;
      }
    };
    // split-BF16 staging of operand op (0: rows b0.., 1: columns b0..)
    auto stage_bf = [&](int op, int b0, int k0, int kc) {
      for (int q = tid; q < kGBK * kGBT; q += kNT) {
        int i, k;
        if (TSIDE) {
          k = q / kGBT;
          i = q % kGBT;
        } else {
          i = q / kGBK;
          k = q % kGBK;
        }
        E v = E(R(0));
        if (k < kc && b0 + i < n) {
          const int rr = TSIDE ? k0 + k : b0 + i, tt = TSIDE ? b0 + i : k0 + k;
          v = cvt<E>(X[roff[rr] + tt] - mean[tt]);
        }
        const float comp[2] = {(float)re_(v), (float)im_(v)};
#pragma unroll
        for (int ri = 0; ri < NPL; ++ri) {
          const __bf16 hi = (__bf16)comp[ri];
          plane(op, ri, 0)[i * kGBW + k] = hi;
          plane(op, ri, 1)[i * kGBW + k] = (__bf16)(comp[ri] - (float)hi);
        }
      }
    };
    for (int tp = 0; tp < ntile; ++tp) {
      int bi = 0;
      while ((bi + 1) * (bi + 2) / 2 <= tp) ++bi;
      const int bj = tp - bi * (bi + 1) / 2;
      const int i0 = bi * kGBT, j0 = bj * kGBT;
      E acc[MF || MD ? 1 : 4][MF || MD ? 1 : 4];
      floatx16 cre, cim;
      // FP64: wave (wi, wj) computes the 32 x 32 quadrant of the 64 x 64 tile
      // as 2 x 2 MFMA tiles of 16 x 16 (k in steps of 4; lane l supplies
      // A[16 ta + l % 16][k + l / 16] and B[k + l / 16][16 tb + l % 16]);
      // quadrants past n or above the diagonal are not computed
      doublex4 dre[2][2], dim[2][2];
      const bool qlive = i0 + 32 * wi < n && j0 + 32 * wj < n && !(bi == bj && wj > wi);
      if constexpr (MD) {
#pragma unroll
        for (int ta = 0; ta < 2; ++ta)
#pragma unroll
          for (int tb = 0; tb < 2; ++tb) dre[ta][tb] = dim[ta][tb] = doublex4{0.0, 0.0, 0.0, 0.0};
      } else if constexpr (MF) {
#pragma unroll
        for (int q = 0; q < 16; ++q) cre[q] = cim[q] = 0.0f;
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[u][v] = E(R(0));
      }
      if constexpr (MD) {
        // the next chunk's centred values are loaded into registers while the
        // current chunk is multiplied
        constexpr int NE = kGBK * kGBT / kNT;
        IN ra[NE], rb[NE];
        auto fetch = [&](IN* rr, int b0, int k0, int kc) {
#pragma unroll
          for (int u = 0; u < NE; ++u) {
            const int q = tid + u * kNT;
            const int i = TSIDE ? q % kGBT : q / kGBK, k = TSIDE ? q / kGBT : q % kGBK;
            const int row = TSIDE ? k0 + k : b0 + i, tt = TSIDE ? b0 + i : k0 + k;
            rr[u] = (k < kc && b0 + i < n) ? IN(X[roff[row] + tt] - mean[tt]) : IN(0);
          }
        };
        auto stash = [&](E* S, const IN* rr) {
#pragma unroll
          for (int u = 0; u < NE; ++u) {
            const int q = tid + u * kNT;
            const int i = TSIDE ? q % kGBT : q / kGBK, k = TSIDE ? q / kGBT : q % kGBK;
            S[k * LW + i] = cvt<E>(rr[u]);
          }
        };
        fetch(ra, i0, 0, min(kGBK, K));
        fetch(rb, j0, 0, min(kGBK, K));
        const int l16 = lane & 15, l4 = lane >> 4;
        for (int k0 = 0; k0 < K; k0 += kGBK) {
          const int kc = min(kGBK, K - k0);
          stash(As, ra);
          stash(Bs, rb);
          lds_barrier();
          if (k0 + kGBK < K) {
            fetch(ra, i0, k0 + kGBK, min(kGBK, K - k0 - kGBK));
            fetch(rb, j0, k0 + kGBK, min(kGBK, K - k0 - kGBK));
          }
          // rows k >= kc of the staging are zero
          if (qlive) {
            for (int k = 0; k < kc; k += 4) {
              E fa[2], fb[2];
#pragma unroll
              for (int t = 0; t < 2; ++t) {
                fa[t] = As[(k + l4) * LW + 32 * wi + 16 * t + l16];
                fb[t] = Bs[(k + l4) * LW + 32 * wj + 16 * t + l16];
              }
#pragma unroll
              for (int ta = 0; ta < 2; ++ta)
#pragma unroll
                for (int tb = 0; tb < 2; ++tb) {
                  const double ar = re_(fa[ta]), br = re_(fb[tb]);
                  dre[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, dre[ta][tb], 0, 0, 0);
                  if constexpr (CPX) {
                    const double ai = im_(fa[ta]), bim = im_(fb[tb]);
                    dre[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bim, dre[ta][tb], 0, 0, 0);
                    if (TSIDE) {  // conj(a) b: im = ar bi - ai br
                      dim[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bim, dim[ta][tb], 0, 0, 0);
                      dim[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, br, dim[ta][tb], 0, 0, 0);
                    } else {  // a conj(b): im = ai br - ar bi
                      dim[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, dim[ta][tb], 0, 0, 0);
                      dim[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar, bim, dim[ta][tb], 0, 0, 0);
                    }
                  }
                }
            }
          }
          lds_barrier();
        }
      }
      for (int k0 = 0; k0 < (MD ? 0 : K); k0 += kGBK) {
        const int kc = min(kGBK, K - k0);
        if constexpr (BFM) {
          stage_bf(0, i0, k0, kc);
          stage_bf(1, j0, k0, kc);
          lds_barrier();
          auto ld = [&](int op, int ri, int hl, int row, int k) {
            return *reinterpret_cast<const bf16x8*>(plane(op, ri, hl) + row * kGBW + k + 8 * lhi);
          };
          // x y on 3 products: xh yh + xh yl + xl yh
          auto prod3 = [&](floatx16 c, bf16x8 xh, bf16x8 xl, bf16x8 yh, bf16x8 yl) {
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yh, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yl, c, 0, 0, 0);
            return __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl, yh, c, 0, 0, 0);
          };
          const int ra = 32 * wi + llo, rb = 32 * wj + llo;
#pragma unroll
          for (int k = 0; k < kGBK; k += 16) {  // indices >= kc are zero in the planes
            const bf16x8 arh = ld(0, 0, 0, ra, k), arl = ld(0, 0, 1, ra, k);
            const bf16x8 brh = ld(1, 0, 0, rb, k), brl = ld(1, 0, 1, rb, k);
            cre = prod3(cre, arh, arl, brh, brl);
            if constexpr (CPX) {
              const bf16x8 aih = ld(0, 1, 0, ra, k), ail = ld(0, 1, 1, ra, k);
              const bf16x8 bih = ld(1, 1, 0, rb, k), bil = ld(1, 1, 1, rb, k);
              cre = prod3(cre, aih, ail, bih, bil);
              if (TSIDE) {  // conj(a) b: im = ar bi - ai br
                cim = prod3(cim, arh, arl, bih, bil);
                cim = prod3(cim, neg_bf16x8(aih), neg_bf16x8(ail), brh, brl);
              } else {  // a conj(b): im = ai br - ar bi
                cim = prod3(cim, aih, ail, brh, brl);
                cim = prod3(cim, neg_bf16x8(arh), neg_bf16x8(arl), bih, bil);
              }
            }
          }
          lds_barrier();
          continue;
        } else {
          stage(As, i0, k0, kc);
          stage(Bs, j0, k0, kc);
          lds_barrier();
        }
        if constexpr (MF) {
          // rows k >= kc of the staging are zero
          for (int k = 0; k < kc; k += 2) {
            const E xa = As[(k + lhi) * LW + 32 * wi + llo];
            const E xb = Bs[(k + lhi) * LW + 32 * wj + llo];
            if constexpr (CPX) {
              const float ar = re_(xa), ai = im_(xa), br = re_(xb), bim = im_(xb);
              cre = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, br, cre, 0, 0, 0);
              cre = __builtin_amdgcn_mfma_f32_32x32x2f32(ai, bim, cre, 0, 0, 0);
              if (TSIDE) {  // conj(a) b: im = ar bi - ai br
                cim = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, bim, cim, 0, 0, 0);
                cim = __builtin_amdgcn_mfma_f32_32x32x2f32(-ai, br, cim, 0, 0, 0);
              } else {  // a conj(b): im = ai br - ar bi
                cim = __builtin_amdgcn_mfma_f32_32x32x2f32(ai, br, cim, 0, 0, 0);
                cim = __builtin_amdgcn_mfma_f32_32x32x2f32(-ar, bim, cim, 0, 0, 0);
              }
            } else {
              cre = __builtin_amdgcn_mfma_f32_32x32x2f32(re_(xa), re_(xb), cre, 0, 0, 0);
            }
          }
          lds_barrier();
          continue;
        }
        for (int k = 0; k < kc; ++k) {
          E xa[4], xb[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            xa[u] = As[k * LW + ty + 16 * u];
            xb[u] = Bs[k * LW + tx + 16 * u];
            if (TSIDE) xa[u] = conj_(xa[u]);  // G = Xc^H Xc
            else xb[u] = conj_(xb[u]);        // G = Xc Xc^H
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] += xa[u] * xb[v];
        }
        lds_barrier();
      }
      if constexpr (MD) {
        // 16x16 MFMA output: accumulator q of lane l is C[4 q + l / 16][l % 16]; the
        // lower quadrant of a diagonal tile also writes its mirror
        if (qlive) {
          const int l16 = lane & 15, l4 = lane >> 4;
#pragma unroll
          for (int ta = 0; ta < 2; ++ta)
#pragma unroll
            for (int tb = 0; tb < 2; ++tb)
#pragma unroll
              for (int q = 0; q < 4; ++q) {
                const int r = i0 + 32 * wi + 16 * ta + 4 * q + l4;
                const int c = j0 + 32 * wj + 16 * tb + l16;
                if (r < n && c < n) {
                  E x = mk<E>(dre[ta][tb][q], CPX ? dim[ta][tb][q] : 0.0);
                  if (a.scale) x = div_n(x, M - 1);
                  G[(size_t)r * n + c] = x;
                  if (bi != bj || wi != wj) G[(size_t)c * n + r] = conj_(x);
                }
              }
        }
      } else if constexpr (MF) {
        // 32x32 MFMA output: accumulator q of lane l is C[8(q/4) + 4(l/32) + q%4][l%32]
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int r = i0 + 32 * wi + 8 * (q >> 2) + 4 * lhi + (q & 3);
          const int c = j0 + 32 * wj + llo;
          if (r < n && c < n) {
            E x = mk<E>(cre[q], cim[q]);
            if (a.scale) x = div_n(x, M - 1);
            G[(size_t)r * n + c] = x;
            if (bi != bj) G[(size_t)c * n + r] = conj_(x);
          }
        }
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = i0 + ty + 16 * u, c = j0 + tx + 16 * v;
            if (r < n && c < n) {
              E x = acc[u][v];
              if (a.scale) x = div_n(x, M - 1);
              G[(size_t)r * n + c] = x;
              if (bi != bj) G[(size_t)c * n + r] = conj_(x);
            }
          }
      }
    }
    lds_barrier();
  }
}

// One-pass variant for FP32 real Grams up to 256 and complex up to 224: every
// 32-index chunk of the centred patch is staged once for the whole Gram, and
// each of the 8 waves keeps its share of the lower 32 x 32 blocks (block w,
// w + 8, ...) in MFMA accumulators (real: up to 5 blocks, complex: 4).  The
// patch is read once instead of once per 64 x 64 tile pair.  BF: split-BF16
// three products, as gram_big_kernel.
constexpr int kG1T = 512;  // threads of the one-pass kernel
template <bool CPX> struct G1 { static constexpr int NBW = CPX ? 4 : 5, NMAX = CPX ? 224 : 256; };

template <class IN, class E, bool TSIDE, bool BF>
__global__ __launch_bounds__(kG1T) void gram_big1_kernel(const IN* __restrict__ vol, GramArgs a) {
  using R = typename real_of<E>::type;
  static_assert(std::is_same<R, float>::value, "one-pass Gram: FP32 data");
  constexpr bool CPX = is_cpx<E>::value;
  constexpr int NPL = CPX ? 2 : 1;
  constexpr int NBW = G1<CPX>::NBW, NMAX = G1<CPX>::NMAX, NW = kG1T / kWave;
  constexpr int LW = NMAX + 1;
  const Geom& g = a.g;
  const int n = a.n, M = g.M, T = g.pt;
  const int K = TSIDE ? M : T;
  const int tid = threadIdx.x, wv = tid / kWave, lane = tid % kWave;
  const int lhi = lane >> 5, llo = lane & 31;
  const int nb = (n + 31) / 32, npad = 32 * nb, NB = nb * (nb + 1) / 2;
  extern __shared__ __align__(16) char smem[];
  char* cur = smem;
  IN* mean = lds_carve<IN>(cur, T);
  int64_t* roff = lds_carve<int64_t>(cur, M);
  E* S = BF ? nullptr : lds_carve<E>(cur, (size_t)kGBK * LW);  // S[k][i]
  __bf16* Pl = BF ? lds_carve<__bf16>(cur, (size_t)NPL * 2 * NMAX * kGBW) : nullptr;
  auto plane = [&](int ri, int hl) { return Pl + (size_t)(ri * 2 + hl) * NMAX * kGBW; };
  // this wave's blocks (bi >= bj), fixed for the kernel
  int bi[NBW], bj[NBW];
#pragma unroll
  for (int b = 0; b < NBW; ++b) {
    const int blk = wv + NW * b;
    int r = 0;
    while ((r + 1) * (r + 2) / 2 <= blk) ++r;
    bi[b] = r;
    bj[b] = blk - r * (r + 1) / 2;
  }
  for (int p = xcd_rank(), np_ = a.np_dev ? *a.np_dev : a.np; p < np_; p += gridDim.x) {
    int64_t sp;
    int t0;
    patch_origin(g, a.idx[p], sp, t0);
    const IN* X = vol + sp + t0;
    for (int r = tid; r < M; r += kG1T) roff[r] = row_offset(g, r);
    lds_barrier();
    for (int t = tid; t < T; t += kG1T) {
      mean[t] = div_n(seq_add(X[roff[0] + t], X + t, roff + 1, M - 1), M);
      if (TSIDE) reinterpret_cast<IN*>(a.MEAN)[(size_t)p * T + t] = mean[t];
    }
    floatx16 cre[NBW], cim[NBW];
#pragma unroll
    for (int b = 0; b < NBW; ++b)
#pragma unroll
      for (int q = 0; q < 16; ++q) cre[b][q] = cim[b][q] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += kGBK) {
      const int kc = min(kGBK, K - k0);
      lds_barrier();  // means ready / previous chunk consumed
      for (int q = tid; q < kGBK * npad; q += kG1T) {
        int i, k;
        if (TSIDE) {  // consecutive threads: consecutive frames of one row
          k = q / npad;
          i = q % npad;
        } else {
          i = q / kGBK;
          k = q % kGBK;
        }
        E v = E(R(0));
        if (k < kc && i < n) {
          const int rr = TSIDE ? k0 + k : i, tt = TSIDE ? i : k0 + k;
          v = cvt<E>(X[roff[rr] + tt] - mean[tt]);
        }
        if constexpr (BF) {
          const float comp[2] = {(float)re_(v), (float)im_(v)};
#pragma unroll
          for (int ri = 0; ri < NPL; ++ri) {
            const __bf16 hi = (__bf16)comp[ri];
            plane(ri, 0)[i * kGBW + k] = hi;
            plane(ri, 1)[i * kGBW + k] = (__bf16)(comp[ri] - (float)hi);
          }
        } else {
          S[k * LW + i] = v;
        }
      }
      lds_barrier();
#pragma unroll
      for (int b = 0; b < NBW; ++b) {
        if (wv + NW * b >= NB) continue;  // wave-uniform
        const int ra = 32 * bi[b] + llo, rb = 32 * bj[b] + llo;
        if constexpr (BF) {
          auto ld = [&](int ri, int hl, int row, int k) {
            return *reinterpret_cast<const bf16x8*>(plane(ri, hl) + row * kGBW + k + 8 * lhi);
          };
          auto prod3 = [&](floatx16 c, bf16x8 xh, bf16x8 xl, bf16x8 yh, bf16x8 yl) {
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yh, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yl, c, 0, 0, 0);
            return __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl, yh, c, 0, 0, 0);
          };
#pragma unroll
          for (int k = 0; k < kGBK; k += 16) {
            const bf16x8 arh = ld(0, 0, ra, k), arl = ld(0, 1, ra, k);
            const bf16x8 brh = ld(0, 0, rb, k), brl = ld(0, 1, rb, k);
            cre[b] = prod3(cre[b], arh, arl, brh, brl);
            if constexpr (CPX) {
              const bf16x8 aih = ld(1, 0, ra, k), ail = ld(1, 1, ra, k);
              const bf16x8 bih = ld(1, 0, rb, k), bil = ld(1, 1, rb, k);
              cre[b] = prod3(cre[b], aih, ail, bih, bil);
              if (TSIDE) {
                cim[b] = prod3(cim[b], arh, arl, bih, bil);
                cim[b] = prod3(cim[b], neg_bf16x8(aih), neg_bf16x8(ail), brh, brl);
              } else {
                cim[b] = prod3(cim[b], aih, ail, brh, brl);
                cim[b] = prod3(cim[b], neg_bf16x8(arh), neg_bf16x8(arl), bih, bil);
              }
            }
          }
        } else {
          for (int k = 0; k < kc; k += 2) {  // rows >= kc are zero
            const E xa = S[(k + lhi) * LW + ra];
            const E xb = S[(k + lhi) * LW + rb];
            if constexpr (CPX) {
              const float ar = re_(xa), ai = im_(xa), br = re_(xb), bim = im_(xb);
              cre[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, br, cre[b], 0, 0, 0);
              cre[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(ai, bim, cre[b], 0, 0, 0);
              if (TSIDE) {
                cim[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(ar, bim, cim[b], 0, 0, 0);
                cim[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(-ai, br, cim[b], 0, 0, 0);
              } else {
                cim[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(ai, br, cim[b], 0, 0, 0);
                cim[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(-ar, bim, cim[b], 0, 0, 0);
              }
            } else {
              cre[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(re_(xa), re_(xb), cre[b], 0, 0, 0);
            }
          }
        }
      }
    }
    // blocks -> G (lower blocks and their mirrors)
    E* G = reinterpret_cast<E*>(a.H) + (size_t)p * n * n;
#pragma unroll
    for (int b = 0; b < NBW; ++b) {
      if (wv + NW * b >= NB) continue;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int r = 32 * bi[b] + 8 * (q >> 2) + 4 * lhi + (q & 3);
        const int c = 32 * bj[b] + llo;
        if (r < n && c < n) {
          E x = mk<E>(cre[b][q], cim[b][q]);
          if (a.scale) x = div_n(x, M - 1);
          G[(size_t)r * n + c] = x;
          if (bi[b] != bj[b]) G[(size_t)c * n + r] = conj_(x);
        }
      }
    }
  }
}

template <class IN, class E, bool BF>
size_t gram_big1_lds_bytes(int M, int T) {
  constexpr bool CPX = is_cpx<E>::value;
  constexpr int NMAX = G1<CPX>::NMAX;
  auto al = [](size_t b) { return ((b + 15) / 16) * 16; };
  const size_t st = BF ? al((size_t)(CPX ? 2 : 1) * 2 * NMAX * kGBW * 2)
                       : al((size_t)kGBK * (NMAX + 1) * sizeof(E));
  return al((size_t)T * sizeof(IN)) + al((size_t)M * sizeof(int64_t)) + st;
}

template <class IN, class E, bool BF = false>
size_t gram_big_lds_bytes(int M, int T) {
  auto al = [](size_t b) { return ((b + 15) / 16) * 16; };
  const size_t st = BF ? al((size_t)2 * (is_cpx<E>::value ? 2 : 1) * 2 * kGBT * kGBW * 2)
                       : 2 * al((size_t)kGBK * (kGBT + 1) * sizeof(E));
  return al((size_t)T * sizeof(IN)) + al((size_t)M * sizeof(int64_t)) + st;
}

// Large-Gram Householder: steps k < k0 on the full matrix in global memory,
// then (once the trailing block A[k0:, k0:] fits) the rest on that block held
// in LDS as a packed lower triangle (column j of the m0 = n - k0 block holds
// rows j .. m0-1 from pk_col(j)).  Both phases compute, per step, the
// reflector from column k, p = tau A v, w = p - tau/2 (p^H v) v and
// A -= v w^H + w v^H on the trailing block; d, e, e^2 and tau go straight to
// DE / TAU and v_k to row k of the global block (columns k+1 .. n-1).
__device__ __forceinline__ float shfl_xor_e(float v, int m) { return shfl_xor_(v, m); }
__device__ __forceinline__ double shfl_xor_e(double v, int m) { return shfl_xor_(v, m); }
template <class R> __device__ __forceinline__ cpx<R> shfl_xor_e(cpx<R> v, int m) {
  return {shfl_xor_(v.x, m), shfl_xor_(v.y, m)};
}

constexpr int kTBT = 1024;  // threads of the large-Gram tridiagonalisation (512 at 2 per CU)
constexpr size_t kTBLds = 160 * 1024;     // its LDS
__host__ __device__ __forceinline__ int pk_col(int j, int m) { return j * m - (j * (j - 1)) / 2; }

// Smallest k0 with the packed (n - k0)-block, v and w within 1 / per_cu of
// the LDS.
inline int tridiag_big_k0(int n, size_t esz, int per_cu = 1) {
  for (int k0 = 0; k0 < n; ++k0) {
    const size_t m = (size_t)(n - k0);
    if (((m * (m + 1) / 2 + 2 * m + 8) * esz + 256) * per_cu <= kTBLds) return k0;
  }
  return n;
}

// Sum over the block for the large-Gram tridiagonalisation: one barrier; each
// call site owns its `red` (so no trailing barrier protects its reuse: the
// next use of the same array is several barriers later).
template <int NT, class T>
__device__ __forceinline__ T block_sum1(T v, T* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) red[wid] = v;
  lds_barrier();
  T s = red[0];
#pragma unroll
  for (int w = 1; w < NT / kWave; ++w) s += red[w];
  return s;
}

template <class E, int NT>
__global__ __launch_bounds__(NT) void tridiag_big_kernel(GramArgs a, int k0) {
  using R = typename real_of<E>::type;
  const int n = a.n, tid = threadIdx.x;
  const int m0 = n - k0;
  extern __shared__ __align__(16) char smem[];
  char* cur = smem;
  __shared__ R redr[NT / kWave];
  __shared__ E rede[NT / kWave];
  E* Bp = lds_carve<E>(cur, (size_t)m0 * (m0 + 1) / 2);  // packed trailing block
  E* v = lds_carve<E>(cur, m0);
  E* w = lds_carve<E>(cur, m0);
  // the global phase keeps v and w (length n) in the still unused packed area
  E* vg = Bp;
  E* wg = Bp + n;
  const E zero = E(R(0));
  for (int p = blockIdx.x, np_ = a.np_dev ? *a.np_dev : a.np; p < np_; p += gridDim.x) {
    E* A = reinterpret_cast<E*>(a.H) + (size_t)p * n * n;
    R* DEp = reinterpret_cast<R*>(a.DE) + (size_t)p * 3 * n;
    E* TAUp = reinterpret_cast<E*>(a.TAU) + (size_t)p * n;
    auto put_step = [&](int k, R d, R e, E t) {
      DEp[k] = d;
      DEp[n + k] = e;
      DEp[2 * n + k] = e * e;
      TAUp[k] = t;
    };
    // ---------------- global phase, steps k < k0: the rank-2 update of step k
    // is applied in step k + 1's pass over the trailing matrix, which also
    // accumulates that step's p (one read and one write of the matrix per
    // step); row k + 1 gets the update first, as the source of reflector k + 1
    E* vn = Bp + 2 * n;  // reflector of the current step (vg / wg: the pending update's)
    bool pending = false;
    for (int k = 0; k < k0 && k < n - 1; ++k) {
      E* rowk = A + (size_t)k * n;  // conj(row k) = column k (the matrix is kept Hermitian)
      if (pending) {
        const E vk = vg[k], wk = wg[k];
        for (int c = k + tid; c < n; c += NT) rowk[c] = rank2_update(rowk[c], vk, conj_(wg[c]), wk, conj_(vg[c]));
        __syncthreads();
      }
      const E alpha = conj_(rowk[k + 1]);
      const R dk = re_(rowk[k]);
      R part = R(0);
      for (int c = k + 2 + tid; c < n; c += NT) part += abs2_(rowk[c]);
      const R xn2 = block_sum<NT>(part, redr);  // its barriers order the reads above before the stores below
      const R ar = re_(alpha), ai = im_(alpha);
      if (!((xn2 > R(0)) || (ai != R(0)))) {  // H_k = I: the pending update alone
        if (pending) {
#pragma unroll
          for (int jj = 0; jj < (kNBig + NT - 1) / NT; ++jj) {
            const int c = k + 1 + tid + jj * NT;
            if (c < n) {
              E* col = A + c;
              const E cw = conj_(wg[c]), cv = conj_(vg[c]);
              for (int r = k + 1; r < n; ++r)
                col[(size_t)r * n] = rank2_update(col[(size_t)r * n], vg[r], cw, wg[r], cv);
            }
          }
          pending = false;
        }
        for (int c = k + 1 + tid; c < n; c += NT) rowk[c] = zero;
        if (tid == 0) put_step(k, dk, ar, zero);
        __syncthreads();
        continue;
      }
      const R beta = -copysign(sqrt(ar * ar + ai * ai + xn2), ar);
      const R ibeta = R(1) / beta;
      const E tk = mk<E>((beta - ar) * ibeta, -ai * ibeta);
      const E scal = recip_(alpha - E(beta));
      for (int c = tid; c < n; c += NT)
        vn[c] = c > k + 1 ? conj_(rowk[c]) * scal : (c == k + 1 ? E(R(1)) : zero);
      __syncthreads();
      for (int c = k + 1 + tid; c < n; c += NT) rowk[c] = vn[c];
      if (tid == 0) put_step(k, dk, beta, tk);
      // the pending update and p = tau A v in one pass over the trailing columns
      E pc[(kNBig + NT - 1) / NT];
#pragma unroll
      for (int jj = 0; jj < (kNBig + NT - 1) / NT; ++jj) {
        const int c = k + 1 + tid + jj * NT;
        E sacc = zero;
        if (c < n) {
          E* col = A + c;
          if (pending) {
            const E cw = conj_(wg[c]), cv = conj_(vg[c]);
            int r = k + 1;
            for (; r + 4 <= n; r += 4) {
              E x[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) x[u] = col[(size_t)(r + u) * n];
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                x[u] = rank2_update(x[u], vg[r + u], cw, wg[r + u], cv);
                col[(size_t)(r + u) * n] = x[u];
                sacc += conj_(x[u]) * vn[r + u];
              }
            }
            for (; r < n; ++r) {
              const E x = rank2_update(col[(size_t)r * n], vg[r], cw, wg[r], cv);
              col[(size_t)r * n] = x;
              sacc += conj_(x) * vn[r];
            }
          } else {
            int r = k + 1;
            for (; r + 4 <= n; r += 4) {
              E x[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) x[u] = col[(size_t)(r + u) * n];
#pragma unroll
              for (int u = 0; u < 4; ++u) sacc += conj_(x[u]) * vn[r + u];
            }
            for (; r < n; ++r) sacc += conj_(col[(size_t)r * n]) * vn[r];
          }
        }
        pc[jj] = tk * sacc;
      }
      E pv = zero;
#pragma unroll
      for (int jj = 0; jj < (kNBig + NT - 1) / NT; ++jj) {
        const int c = k + 1 + tid + jj * NT;
        if (c < n) pv += conj_(pc[jj]) * vn[c];
      }
      pv = block_sum<NT>(pv, rede);  // every thread is past its reads of vg / wg
      const E kc = (R(-0.5) * tk) * pv;
      for (int c = tid; c < n; c += NT) vg[c] = vn[c];
#pragma unroll
      for (int jj = 0; jj < (kNBig + NT - 1) / NT; ++jj) {
        const int c = k + 1 + tid + jj * NT;
        if (c < n) wg[c] = pc[jj] + kc * vn[c];
      }
      for (int c = tid; c <= k && c < n; c += NT) wg[c] = zero;
      pending = true;
      __syncthreads();
    }
    // ---------------- LDS phase: load the lower triangle of A[k0:, k0:]
    // (with the last global step's update, its v / w first copied out of the
    // area the triangle overwrites)
    if (pending) {
      for (int i = tid; i < m0; i += NT) {
        v[i] = vg[k0 + i];
        w[i] = wg[k0 + i];
      }
      __syncthreads();
    }
    for (int i = 0; i < m0; ++i) {
      const E* row = A + (size_t)(k0 + i) * n + k0;
      if (pending) {
        const E vi = v[i], wi = w[i];
        for (int j = tid; j <= i; j += NT)
          Bp[pk_col(j, m0) + (i - j)] = rank2_update(row[j], vi, conj_(w[j]), wi, conj_(v[j]));
      } else {
        for (int j = tid; j <= i; j += NT) Bp[pk_col(j, m0) + (i - j)] = row[j];
      }
    }
    __syncthreads();
    for (int kk = 0; kk < m0 - 1; ++kk) {
      const int k = k0 + kk;
      const E* colk = Bp + pk_col(kk, m0) - kk;  // colk[i] = B[i][kk], i >= kk
      const E alpha = colk[kk + 1];
      const R dk = re_(colk[kk]);
      R part = R(0);
      for (int i = kk + 2 + tid; i < m0; i += NT) part += abs2_(colk[i]);
      const R xn2 = block_sum1<NT>(part, redr);
      const R ar = re_(alpha), ai = im_(alpha);
      E* rowk = A + (size_t)k * n + k0;  // global row k, local columns
      if (!((xn2 > R(0)) || (ai != R(0)))) {
        for (int i = kk + 1 + tid; i < m0; i += NT) rowk[i] = zero;
        if (tid == 0) put_step(k, dk, ar, zero);
        __syncthreads();
        continue;
      }
      const R beta = -copysign(sqrt(ar * ar + ai * ai + xn2), ar);
      const R ibeta = R(1) / beta;
      const E tk = mk<E>((beta - ar) * ibeta, -ai * ibeta);
      const E scal = recip_(alpha - E(beta));
      for (int i = tid; i < m0; i += NT) {
        const E vi = i > kk + 1 ? colk[i] * scal : (i == kk + 1 ? E(R(1)) : zero);
        v[i] = vi;
        if (i > kk) rowk[i] = vi;
      }
      if (tid == 0) put_step(k, dk, beta, tk);
      __syncthreads();
      // p_c = tau (sum_{kk < r < c} B[c][r] v_r + sum_{r >= c} conj(B[r][c]) v_r),
      // the r range of column c split over S adjacent lanes (4 loads in flight
      // per lane), partial sums combined by a fixed butterfly
      const int L = m0 - kk - 1;  // trailing columns kk+1 .. m0-1
      int S = 1;
      while (S < 8 && 2 * S * L <= NT) S *= 2;
      const int g = tid / S, sub = tid % S;
      const int c = kk + 1 + g;
      E pacc = zero;
      if (g < L) {
        const int len = (L + S - 1) / S;
        const int r0 = kk + 1 + sub * len, r1 = min(m0, r0 + len);
        const int ra = min(r1, c);
        int r = r0;
        for (; r + 4 <= ra; r += 4) {
          E x[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) x[u] = Bp[pk_col(r + u, m0) + (c - r - u)];
#pragma unroll
          for (int u = 0; u < 4; ++u) pacc += x[u] * v[r + u];
        }
        for (; r < ra; ++r) pacc += Bp[pk_col(r, m0) + (c - r)] * v[r];
        const E* colc = Bp + pk_col(c, m0) - c;
        r = max(r0, c);
        for (; r + 4 <= r1; r += 4) {
          E x[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) x[u] = colc[r + u];
#pragma unroll
          for (int u = 0; u < 4; ++u) pacc += conj_(x[u]) * v[r + u];
        }
        for (; r < r1; ++r) pacc += conj_(colc[r]) * v[r];
      }
      for (int o = 1; o < S; o <<= 1) pacc += shfl_xor_e(pacc, o);
      const bool lead = g < L && sub == 0;
      const E pcv = tk * pacc;
      const E pv = block_sum1<NT>(lead ? conj_(pcv) * v[c] : zero, rede);
      const E kc = (R(-0.5) * tk) * pv;
      if (lead) w[c] = pcv + kc * v[c];
      __syncthreads();
      // B -= v w^H + w v^H on the lower triangle of the trailing block: column
      // pair (kk+1+q, m0-1-q) has L + 1 entries in all, split over S2 lanes
      const int P = (L + 1) / 2;
      int S2 = 1;
      while (S2 < 8 && 2 * S2 * P <= NT) S2 *= 2;
      const int q = tid / S2, sq = tid % S2;
      if (q < P) {
        const int ca = kk + 1 + q, cb = m0 - 1 - q;
        const int la = m0 - ca, lb = cb != ca ? m0 - cb : 0;
        const int len = (la + lb + S2 - 1) / S2;
        const int i0 = sq * len, i1 = min(la + lb, i0 + len);
        for (int h = 0; h < 2; ++h) {
          // part of [i0, i1) inside column h (column a: [0, la), column b: [la, la + lb))
          const int lo = h ? max(i0, la) : i0, hi = h ? i1 : min(i1, la);
          if (lo >= hi) continue;
          const int cc = h ? cb : ca;
          const int rbase = cc - (h ? la : 0);  // row of concatenated index i is rbase + i
          E* colc = Bp + pk_col(cc, m0) - cc;
          const E cw = conj_(w[cc]), cv = conj_(v[cc]);
          int i = lo;
          for (; i + 4 <= hi; i += 4) {
            E x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = colc[rbase + i + u];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const int rr = rbase + i + u;
              colc[rr] = rank2_update(x[u], v[rr], cw, w[rr], cv);
            }
          }
          for (; i < hi; ++i) {
            const int rr = rbase + i;
            colc[rr] = rank2_update(colc[rr], v[rr], cw, w[rr], cv);
          }
        }
      }
      __syncthreads();
    }
    if (tid == 0) {
      put_step(n - 1, re_(Bp[pk_col(m0 - 1, m0)]), R(0), zero);
      DEp[n + n - 1] = R(0);
      DEp[2 * n + n - 1] = R(0);
    }
    __syncthreads();
  }
}

template <class E>
size_t tridiag_big_lds_bytes(int n, int k0) {
  auto al = [](size_t b) { return ((b + 15) / 16) * 16; };
  const size_t m = (size_t)(n - k0);
  return al(m * (m + 1) / 2 * sizeof(E)) + 2 * al(m * sizeof(E));
}

// ===========================================================================
// 3. Tridiagonal eigen stage + threshold rules (lowrank.py)
// ===========================================================================

struct EigArgs {
  Geom g;
  const int32_t* idx;
  int np;
  int n, kmax, kb;  // Gram size, max components per patch, inverse-iteration batch
  int method, loss;
  int rule_f64;     // the reference's spectrum dtype is float64 (f64/c128 data)
  int scale_f64, tau_f64;
  double ts2, tau, mpm, beta;
  int adaptive_method, n_gamma;  // ADAPTIVE: PDAMD_ADAPT_*, candidates in gamma
  const double* gamma;
  double tau0_log;
  const double* qut_tau;         // QUT: tau by linear patch index
  const double* var_map;
  int var_f32;        // round the patch prior to float32 (float32 noise-level map)
  const double* LAM;  // optimal rule: per-linear-patch scale of the noise level c (null = 1)
  const void* H;
  const void* DE;
  const void* TAU;
  void* Y;           // np x n x kmax eigenvectors of the Gram (element type)
  void* Z;           // np x n x kmax tridiagonal eigenvectors (real)
  void* SCR;         // np x 3 x n x kb inverse-iteration LU scratch (real)
  double* WC;        // np x kmax synthesis weights
  int* NCOMP;        // np
  double* THETA;     // np
  int32_t* RANK;     // np (user array offset to chunk start) or null
  float* NOISE;      // np or null
  float* EIG;        // np x 10 or null
  int zero_fill;     // 1: eigenvector columns ncomp .. kmax-1 of Y are set to zero
  const int* np_dev; // device patch count (<= np, read by the kernel), or null: np
  // FP64 re-check of near-threshold decisions (MP-PCA / Hybrid): NEAR[p] = 1
  // when a comparison of the rule has its sides within refine_rel relative
  double refine_rel;
  uint8_t* NEAR;     // np flags (written when non-null)
};

template <class R> struct Lim;
template <> struct Lim<float> {
  __device__ static float ulp() { return 5.9604645e-8f; }
  __device__ static float tiny() { return FLT_MIN; }
};
template <> struct Lim<double> {
  __device__ static double ulp() { return 1.1102230246251565e-16; }
  __device__ static double tiny() { return DBL_MIN; }
};

// FP32 a / b in the recurrences (Sturm counts, LU and twisted factorizations
// of the tridiagonal): v_rcp_f32 (1 ulp) times a, two instructions where the
// correctly rounded quotient takes ten (div_scale x 2, rcp, 4 FMAs, div_fmas,
// div_fixup).  The recurrences are backward stable for quotients of a few ulp
// relative error; b is kept away from 0 by the callers' pivmin / guards.
__device__ __forceinline__ float fdiv_(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double fdiv_(double a, double b) { return a / b; }
// a / b from the hardware reciprocal (FP64: refined by one Newton step; a few
// ulp, no special-case handling; b is kept away from 0 by the callers' pivmin).
__device__ __forceinline__ float qdiv_(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double qdiv_(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  return a * r;
}

// Number of eigenvalues of the tridiagonal (d, e^2) strictly below x.
// d and e2 are wave-uniform global arrays (scalar loads); 8 entries are
// fetched per block so the recurrence never waits on memory.
template <class R, bool QD = false>
__device__ __forceinline__ int sturm_g(const R* __restrict__ d, const R* __restrict__ e2, int n,
                                       R x, R pivmin) {
  R q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q < R(0);
  int i = 1;
  for (; i + 8 <= n; i += 8) {
    R dv[8], ev[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      dv[u] = d[i + u];
      ev[u] = e2[i - 1 + u];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      q = (dv[u] - x) - (QD ? qdiv_(ev[u], q) : fdiv_(ev[u], q));
      if (fabs(q) < pivmin) q = -pivmin;
      c += q < R(0);
    }
  }
  for (; i < n; ++i) {
    q = (d[i] - x) - (QD ? qdiv_(e2[i - 1], q) : fdiv_(e2[i - 1], q));
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < R(0);
  }
  return c;
}

// NumPy's float mean of a[0..k) (k <= 10): 8-way pairwise blocks for k >= 8,
// sequential below; then divided by k.  NaN for k == 0.
template <class RT>
__device__ RT np_mean10(const RT* a, int k) {
  if (k <= 0) return RT(NAN);
  RT s;
  if (k < 8) {
    s = RT(-0.0);
    for (int i = 0; i < k; ++i) s += a[i];
  } else {
    s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    for (int i = 8; i < k; ++i) s += a[i];
  }
  return s / (RT)k;
}

template <class R>
struct WaveScratch {
  R* lo;
  R* hi;
  R* xs;
  int* cs;
  int* tgt;
  R* val;
};

// Eigenvalues of the tridiagonal at ascending indices tgt[0..nt) by
// multisection: the 64 lanes are split into groups of L = 64/nt points.
template <class R, bool QD = false>
__device__ void multisect(WaveScratch<R>& s, const R* __restrict__ d, const R* __restrict__ e2,
                          int n, R pivmin, R glo, R ghi, R atol, int nt) {
  const int lane = threadIdx.x;
  for (int b0 = 0; b0 < nt; b0 += kWave) {
    const int nb = min(kWave, nt - b0);
    const int L = kWave / nb;
    const int j = lane / L, sub = lane % L;
    const bool active = j < nb;
    if (active && sub == 0) {
      s.lo[j] = glo;
      s.hi[j] = ghi;
    }
    wave_sync();
    for (int it = 0; it < 200; ++it) {
      R lo = active ? s.lo[j] : R(0), hi = active ? s.hi[j] : R(0);
      const R tol = atol + R(2) * Lim<R>::ulp() * fmax(fabs(lo), fabs(hi));
      const R mid = lo + (hi - lo) * R(0.5);
      const bool done = !active || (hi - lo) <= tol || mid <= lo || mid >= hi;
      if (__all(done)) break;
      const R x = lo + (hi - lo) * (R)(sub + 1) / (R)(L + 1);
      const int c = sturm_g<R, QD>(d, e2, n, x, pivmin);
      s.xs[lane] = x;
      s.cs[lane] = c;
      wave_sync();
      if (active && !done && sub == 0) {
        const int tgt = s.tgt[b0 + j];
        R nlo = lo, nhi = hi;
        for (int q = 0; q < L; ++q) {
          const R xq = s.xs[lane + q];
          if (s.cs[lane + q] <= tgt) nlo = fmax(nlo, xq);
          else nhi = fmin(nhi, xq);
        }
        s.lo[j] = nlo;
        s.hi[j] = nhi;
      }
      wave_sync();
    }
    if (active && sub == 0) s.val[b0 + j] = s.lo[j] + (s.hi[j] - s.lo[j]) * R(0.5);
    wave_sync();
  }
}

// Eigenvalues at ascending indices tgt[0..nt), 64 < nt <= 128: plain
// bisection (the multisection above with one lane per target), lane j
// bisecting targets j and j + 64 in the same pass: two independent Sturm
// recurrences per lane hide each other's latency; the sequence of midpoints
// and the results are those of two passes of 64.
template <class R, bool QD = false>
__device__ void bisect_pairs(WaveScratch<R>& s, const R* __restrict__ d, const R* __restrict__ e2,
                             int n, R pivmin, R glo, R ghi, R atol, int nt) {
  const int lane = threadIdx.x;
  const bool a0 = lane < nt, a1 = lane + kWave < nt;
  const int t0 = a0 ? s.tgt[lane] : 0, t1 = a1 ? s.tgt[lane + kWave] : 0;
  R lo0 = glo, hi0 = ghi, lo1 = glo, hi1 = ghi;
  auto div = [](R a, R b) { return QD ? qdiv_(a, b) : fdiv_(a, b); };
  for (int it = 0; it < 200; ++it) {
    const R tol0 = atol + R(2) * Lim<R>::ulp() * fmax(fabs(lo0), fabs(hi0));
    const R tol1 = atol + R(2) * Lim<R>::ulp() * fmax(fabs(lo1), fabs(hi1));
    const R x0 = lo0 + (hi0 - lo0) * R(0.5), x1 = lo1 + (hi1 - lo1) * R(0.5);
    const bool done0 = !a0 || (hi0 - lo0) <= tol0 || x0 <= lo0 || x0 >= hi0;
    const bool done1 = !a1 || (hi1 - lo1) <= tol1 || x1 <= lo1 || x1 >= hi1;
    if (__all(done0 && done1)) break;
    R q0 = d[0] - x0, q1 = d[0] - x1;
    if (fabs(q0) < pivmin) q0 = -pivmin;
    if (fabs(q1) < pivmin) q1 = -pivmin;
    int c0 = q0 < R(0), c1 = q1 < R(0);
    int i = 1;
    for (; i + 8 <= n; i += 8) {
      R dv[8], ev[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        dv[u] = d[i + u];
        ev[u] = e2[i - 1 + u];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        q0 = (dv[u] - x0) - div(ev[u], q0);
        q1 = (dv[u] - x1) - div(ev[u], q1);
        if (fabs(q0) < pivmin) q0 = -pivmin;
        if (fabs(q1) < pivmin) q1 = -pivmin;
        c0 += q0 < R(0);
        c1 += q1 < R(0);
      }
    }
    for (; i < n; ++i) {
      q0 = (d[i] - x0) - div(e2[i - 1], q0);
      q1 = (d[i] - x1) - div(e2[i - 1], q1);
      if (fabs(q0) < pivmin) q0 = -pivmin;
      if (fabs(q1) < pivmin) q1 = -pivmin;
      c0 += q0 < R(0);
      c1 += q1 < R(0);
    }
    if (!done0) {
      if (c0 <= t0) lo0 = x0;
      else hi0 = x0;
    }
    if (!done1) {
      if (c1 <= t1) lo1 = x1;
      else hi1 = x1;
    }
  }
  if (a0) s.val[lane] = lo0 + (hi0 - lo0) * R(0.5);
  if (a1) s.val[lane + kWave] = lo1 + (hi1 - lo1) * R(0.5);
  wave_sync();
}

// Solve (T - lam I) x = b for the tridiagonal (d, e) with partial pivoting.
// d / e are wave-uniform global arrays (e[n-1] = 0); b is in rhs (stride ks)
// and the solution overwrites it; u0 / u1 / u2 (stride ks) hold U.
template <class R, bool QD = false>
__device__ void tri_solve(const R* __restrict__ d, const R* __restrict__ e, int n, R lam, R* u0,
                          R* u1, R* u2, R* rhs, int ks, R eps_piv) {
  auto dv_ = [](R a, R b) { return QD ? qdiv_(a, b) : fdiv_(a, b); };
  R c0 = d[0] - lam, c1 = e[0], c2 = R(0), r = rhs[0];
  auto step = [&](int i, R n0, R n1, R n2, R nr) {
    if (fabs(n0) > fabs(c0)) {
      R t;
      t = c0; c0 = n0; n0 = t;
      t = c1; c1 = n1; n1 = t;
      t = c2; c2 = n2; n2 = t;
      t = r; r = nr; nr = t;
    }
    if (QD ? fabs(c0) < eps_piv : c0 == R(0)) c0 = QD ? copysign(eps_piv, c0) : eps_piv;
    const R f = dv_(n0, c0);
    u0[i * ks] = c0;
    u1[i * ks] = c1;
    u2[i * ks] = c2;
    rhs[i * ks] = r;
    c0 = n1 - f * c1;
    c1 = n2 - f * c2;
    c2 = R(0);
    r = nr - f * r;
  };
  int i = 0;
  for (; i + 4 <= n - 1; i += 4) {
    R dv[4], ev[4], fv[4], bv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      dv[u] = d[i + 1 + u];
      ev[u] = e[i + u];
      fv[u] = e[i + 1 + u];
      bv[u] = rhs[(i + 1 + u) * ks];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) step(i + u, ev[u], dv[u] - lam, fv[u], bv[u]);
  }
  for (; i < n - 1; ++i) step(i, e[i], d[i + 1] - lam, e[i + 1], rhs[(i + 1) * ks]);
  if (QD ? fabs(c0) < eps_piv : c0 == R(0)) c0 = QD ? copysign(eps_piv, c0) : eps_piv;
  u0[(n - 1) * ks] = c0;
  // back substitution
  R z2 = R(0), z1 = dv_(r, c0);
  rhs[(n - 1) * ks] = z1;
  int k = n - 2;
  for (; k - 3 >= 0; k -= 4) {
    R a0[4], a1[4], a2[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a0[u] = u0[(k - u) * ks];
      a1[u] = u1[(k - u) * ks];
      a2[u] = u2[(k - u) * ks];
      b[u] = rhs[(k - u) * ks];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const R z = dv_(b[u] - a1[u] * z1 - a2[u] * z2, a0[u]);
      rhs[(k - u) * ks] = z;
      z2 = z1;
      z1 = z;
    }
  }
  for (; k >= 0; --k) {
    const R z = dv_(rhs[k * ks] - u1[k * ks] * z1 - u2[k * ks] * z2, u0[k * ks]);
    rhs[k * ks] = z;
    z2 = z1;
    z1 = z;
  }
}

// Shrinkers of lowrank.py:250-280 on y = s / c (float64).
__device__ double opt_x(double y, double beta) {
  const double t = y * y - beta - 1.0;
  const double x = sqrt(0.5 * (t + sqrt(t * t - 4.0 * beta)));
  return x * ((y >= 1.0 + sqrt(beta)) ? 1.0 : 0.0);
}
__device__ double shrink(int loss, double y, double beta) {
  if (loss == PDAMD_LOSS_OPE) {
    const double x = opt_x(y, beta);
    return isnan(x) ? x : fmax(x, 0.0);
  }
  if (loss == PDAMD_LOSS_NUC) {
    const double x = opt_x(y, beta);
    const double num = x * x * x * x - sqrt(beta) * x * y - beta;
    return (isnan(num) ? num : fmax(0.0, num)) / ((x * x) * y);
  }
  const double t = y * y - beta - 1.0;
  return sqrt(fmax(t * t - 4.0 * beta, 0.0) / y);
}

// Adaptive trace-norm shrinkage (lowrank.py:410-636) on the singular values
// sv[0..n) of one patch (rows nr, frames nc): dhat_i = s_i max(1 - (tau /
// s_i)^gamma, 0) and the SURE / GSURE cost of lowrank.py:410-442,
//   div  = sum_i [(1 + (gamma - 1)(tau/s_i)^gamma) [s_i >= tau] + |nr - nc| dhat_i / s_i]
//          + 2 sum_{i != j} s_i dhat_i / (s_i^2 - s_j^2)
//   rss  = sum_i (dhat_i - s_i)^2
//   SURE = sig2 (-nr nc + 2 div) + rss,   GSURE = rss / (1 - div / (nr nc))^2.
// Wave-parallel (every lane returns the same bits): lane-strided terms, the
// pair sum only over components with dhat_i != 0.  Zero singular values (a
// rank-deficient centred patch) contribute nothing.  Computed in float64.
__device__ double atn_cost(const double* __restrict__ sv, int n, double gamma, double tau,
                           bool gsure, double sig2, double nr, double nc) {
  const int lane = threadIdx.x & (kWave - 1);
  double part = 0.0, rss = 0.0, dsum = 0.0;
  for (int i = lane; i < n; i += kWave) {
    const double si = sv[i];
    if (!(si > 0.0)) continue;
    const double pw = pow(tau / si, gamma);
    const double dh = si * fmax(1.0 - pw, 0.0);
    part += (si >= tau ? 1.0 + (gamma - 1.0) * pw : 0.0) + fabs(nr - nc) * dh / si;
    rss += (dh - si) * (dh - si);
    const double tmp = si * dh;
    if (tmp != 0.0) {
      const double si2 = si * si;
      for (int j = 0; j < n; ++j)
        if (j != i && sv[j] > 0.0) dsum += tmp / (si2 - sv[j] * sv[j]);
    }
  }
  part = dpp_sum(part);
  rss = dpp_sum(rss);
  dsum = dpp_sum(dsum);
  const double div = part + 2.0 * dsum;
  if (gsure) {
    const double q = 1.0 - div / nr / nc;
    return rss / (q * q);
  }
  return sig2 * (-nr * nc + 2.0 * div) + rss;
}

// Local minimum of f (wave-uniform) reached from x0: downhill bracketing with
// growing steps, then golden-section search to |interval| <= 1e-10 (1 + |x|).
// NaN costs count as +inf.  The point with the lowest cost seen is returned.
template <class F>
__device__ double local_min(F f, double x0, double* fbest) {
  auto g = [&](double x) {
    const double v = f(x);
    return isnan(v) ? INFINITY : v;
  };
  double a = x0, fa = g(a), h = 0.25;
  double b = a + h, fb = g(b);
  if (fb > fa) {  // downhill is the other way, or x0 is already bracketed
    const double bm = a - h, fbm = g(bm);
    if (fbm >= fa) {
      // [a - h, a + h] brackets a minimum near x0
      double lo = bm, hi = b;
      double bx = a, bf = fa;
      const double gr = 0.6180339887498949;
      double x1 = hi - gr * (hi - lo), x2 = lo + gr * (hi - lo);
      double f1 = g(x1), f2 = g(x2);
      for (int it = 0; it < 200 && (hi - lo) > 1e-10 * (1.0 + fabs(bx)); ++it) {
        if (f1 <= f2) {
          hi = x2; x2 = x1; f2 = f1; x1 = hi - gr * (hi - lo); f1 = g(x1);
        } else {
          lo = x1; x1 = x2; f1 = f2; x2 = lo + gr * (hi - lo); f2 = g(x2);
        }
        if (f1 < bf) { bf = f1; bx = x1; }
        if (f2 < bf) { bf = f2; bx = x2; }
      }
      *fbest = bf;
      return bx;
    }
    b = bm;
    fb = fbm;
    h = -h;
  }
  // fb <= fa: walk downhill with doubling steps until the cost rises
  double c = b + h, fc = g(c);
  for (int it = 0; it < 60 && fc <= fb; ++it) {
    a = b; fa = fb;
    b = c; fb = fc;
    h *= 2.0;
    c = b + h;
    fc = g(c);
  }
  double lo = fmin(a, c), hi = fmax(a, c);
  double bx = b, bf = fb;
  const double gr = 0.6180339887498949;
  double x1 = hi - gr * (hi - lo), x2 = lo + gr * (hi - lo);
  double f1 = g(x1), f2 = g(x2);
  for (int it = 0; it < 200 && (hi - lo) > 1e-10 * (1.0 + fabs(bx)); ++it) {
    if (f1 <= f2) {
      hi = x2; x2 = x1; f2 = f1; x1 = hi - gr * (hi - lo); f1 = g(x1);
    } else {
      lo = x1; x1 = x2; f1 = f2; x2 = lo + gr * (hi - lo); f2 = g(x2);
    }
    if (f1 < bf) { bf = f1; bx = x1; }
    if (f2 < bf) { bf = f2; bx = x2; }
  }
  *fbest = bf;
  return bx;
}

// Eigenvector of the tridiagonal (d, e; e2 = e^2) for the eigenvalue lm by a
// twisted factorization (the eigenvector step of the MRRR algorithm, Dhillon
// & Parlett 2004), one lane: top-down pivots D+ of L+ D+ L+^T = T - lm I into
// z[i * ks], bottom-up pivots of U- D- U-^T (multipliers u into um[i * ks]),
// twist index r = argmin |gamma_r|, gamma_r = D+_r + D-_r - (d_r - lm); then
// z_r = 1, z_i = -(e_i / D+_i) z_{i+1} (i < r; each z_i replaces D+_i once
// read), z_i = -u_i z_{i-1} (i > r); unit 2-norm.  Pivots are kept at least
// eps_piv in magnitude.  False when the result is not finite.
template <class R>
__device__ bool twisted_vec(const R* d, const R* e, const R* e2, int n, R lm, R eps_piv, R* z,
                            R* um, int ks) {
  auto guard = [&](R x) { return fabs(x) < eps_piv ? copysign(eps_piv, x) : x; };
  R dv = guard(d[0] - lm);
  z[0] = dv;
  for (int i = 0; i + 1 < n; ++i) {
    dv = guard((d[i + 1] - lm) - fdiv_(e2[i], dv));
    z[(i + 1) * ks] = dv;
  }
  R dm = guard(d[n - 1] - lm);
  R best = fabs(dv + dm - (d[n - 1] - lm));
  int r = n - 1;
  for (int i = n - 2; i >= 0; --i) {
    const R u = fdiv_(e[i], dm);
    um[(i + 1) * ks] = u;
    dm = guard((d[i] - lm) - u * e[i]);
    const R g = fabs(z[i * ks] + dm - (d[i] - lm));
    if (g < best) {
      best = g;
      r = i;
    }
  }
  R zi = R(1), ss = R(1);
  for (int i = r - 1; i >= 0; --i) {
    zi = -fdiv_(e[i], z[i * ks]) * zi;
    z[i * ks] = zi;
    ss += zi * zi;
  }
  z[r * ks] = R(1);
  zi = R(1);
  for (int i = r + 1; i < n; ++i) {
    zi = -um[i * ks] * zi;
    z[i * ks] = zi;
    ss += zi * zi;
  }
  if (!(ss < R(INFINITY))) return false;
  const R inv = R(1) / sqrt(ss);
  for (int i = 0; i < n; ++i) z[i * ks] *= inv;
  return true;
}

// Every eigenvector of the tridiagonal (d, e), eigenvalues lam[0..n) in
// ascending order, into Z (row stride ks, column c = vector c).  Lane c of the
// wavefront handles component c (64 at a time) with a twisted factorization
// (the eigenvector step of the MRRR algorithm, Dhillon & Parlett 2004): the
// top-down L+ D+ L+^T and bottom-up U- D- U-^T factorizations of T - lam I
// (pivots kept at least eps_piv in magnitude), the twist index r minimising
// |gamma_r| = |D+_r + D-_r - (d_r - lam)|, then z_r = 1 and the two
// recurrences z_i = -L+_i z_{i+1} (i < r), z_i = -U-_i z_{i-1} (i > r).  This
// is one step of inverse iteration from e_r, with factors in scr (3 n x n);
// a non-finite result falls back to two inverse-iteration solves.  The
// wavefront then re-orthogonalises (modified Gram-Schmidt, twice) the vectors
// of eigenvalue clusters closer than clus.  For the back-transformation by a
// separate kernel.
template <class R>
__device__ void all_pairs_vectors(const R* __restrict__ d, const R* __restrict__ e, int n,
                                  const R* lam, R* Z, R* scr, int ks, R eps_piv, R clus) {
  const int lane = threadIdx.x;
  for (int c0 = 0; c0 < n; c0 += kWave) {
    const int c = c0 + lane;
    if (c < n) {
      R* z = Z + c;
      R* dp = scr + c;                   // D+ of L+ D+ L+^T = T - lam I (top-down)
      R* um = scr + (size_t)2 * n * ks + c;  // U- of U- D- U-^T (bottom-up)
      const R lm = lam[c];
      auto guard = [&](R x) { return fabs(x) < eps_piv ? copysign(eps_piv, x) : x; };
      // twisted factorization: z = (T - lam I)^-1 e_r up to scale, r = argmin |gamma_r|;
      // L+ (= e_i / D+_i) is recomputed from the stored D+ where z needs it
      R dv = guard(d[0] - lm);
      dp[0] = dv;
      for (int i = 0; i + 1 < n; ++i) {
        const R l = qdiv_(e[i], dv);
        dv = guard((d[i + 1] - lm) - l * e[i]);
        dp[(size_t)(i + 1) * ks] = dv;
      }
      // the loops below read the scratch columns back in blocks of 8 fetched
      // ahead of the recurrence, so it waits on memory once per block
      R dm = guard(d[n - 1] - lm);
      R best = fabs(dp[(size_t)(n - 1) * ks] + dm - (d[n - 1] - lm));
      int r = n - 1;
      auto bottom_up = [&](int i, R dpi) {
        const R u = qdiv_(e[i], dm);
        um[(size_t)(i + 1) * ks] = u;
        dm = guard((d[i] - lm) - u * e[i]);
        const R g = fabs(dpi + dm - (d[i] - lm));
        if (g < best) {
          best = g;
          r = i;
        }
      };
      int i = n - 2;
      for (; i >= 7; i -= 8) {
        R pf[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) pf[q] = dp[(size_t)(i - q) * ks];
#pragma unroll
        for (int q = 0; q < 8; ++q) bottom_up(i - q, pf[q]);
      }
      for (; i >= 0; --i) bottom_up(i, dp[(size_t)i * ks]);
      z[(size_t)r * ks] = R(1);
      R zi = R(1), ss = R(1);
      i = r - 1;
      for (; i >= 7; i -= 8) {
        R pf[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) pf[q] = dp[(size_t)(i - q) * ks];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          zi = -qdiv_(e[i - q], pf[q]) * zi;
          z[(size_t)(i - q) * ks] = zi;
          ss += zi * zi;
        }
      }
      for (; i >= 0; --i) {
        zi = -qdiv_(e[i], dp[(size_t)i * ks]) * zi;
        z[(size_t)i * ks] = zi;
        ss += zi * zi;
      }
      zi = R(1);
      i = r + 1;
      for (; i + 8 <= n; i += 8) {
        R pf[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) pf[q] = um[(size_t)(i + q) * ks];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          zi = -pf[q] * zi;
          z[(size_t)(i + q) * ks] = zi;
          ss += zi * zi;
        }
      }
      for (; i < n; ++i) {
        zi = -um[(size_t)i * ks] * zi;
        z[(size_t)i * ks] = zi;
        ss += zi * zi;
      }
      if (!(ss < R(INFINITY))) {
        // fallback: inverse iteration from a fixed pseudo-random start
        uint32_t hs = 0x9E3779B9u * (uint32_t)(c + 1);
        for (int i = 0; i < n; ++i) {
          hs ^= hs << 13; hs ^= hs >> 17; hs ^= hs << 5;
          z[(size_t)i * ks] = R(0.5) + (R)(hs & 0xFFFF) * R(1.0 / 65536.0);
        }
        for (int it = 0; it < 2; ++it) {
          tri_solve<R, true>(d, e, n, lm, dp, scr + (size_t)n * ks + c, um, z, ks, eps_piv);
          R mx = R(0);
          for (int i = 0; i < n; ++i) mx = fmax(mx, fabs(z[(size_t)i * ks]));
          const R inv = mx > R(0) ? R(1) / mx : R(1);
          for (int i = 0; i < n; ++i) z[(size_t)i * ks] *= inv;
        }
        ss = R(0);
        for (int i = 0; i < n; ++i) ss += z[(size_t)i * ks] * z[(size_t)i * ks];
      }
      const R inv = R(1) / sqrt(fmax(ss, Lim<R>::tiny()));
      i = 0;
      for (; i + 8 <= n; i += 8) {
        R zv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) zv[q] = z[(size_t)(i + q) * ks];
#pragma unroll
        for (int q = 0; q < 8; ++q) z[(size_t)(i + q) * ks] = zv[q] * inv;
      }
      for (; i < n; ++i) z[(size_t)i * ks] *= inv;
    }
  }
  __syncthreads();  // columns written lane-wise, read row-wise below
  const int i0 = lane, i1 = lane + kWave;
  for (int c = 1; c < n; ++c) {
    int qs = c;
    while (qs > 0 && fabs(lam[qs - 1] - lam[c]) <= clus) --qs;
    if (qs == c) continue;
    R zc0 = i0 < n ? Z[(size_t)i0 * ks + c] : R(0);
    R zc1 = i1 < n ? Z[(size_t)i1 * ks + c] : R(0);
    for (int pass = 0; pass < 2; ++pass) {
      for (int q = qs; q < c; ++q) {
        const R zq0 = i0 < n ? Z[(size_t)i0 * ks + q] : R(0);
        const R zq1 = i1 < n ? Z[(size_t)i1 * ks + q] : R(0);
        const R dot = dpp_sum(zq0 * zc0 + zq1 * zc1);
        zc0 -= dot * zq0;
        zc1 -= dot * zq1;
      }
      const R inv = R(1) / sqrt(fmax(dpp_sum(zc0 * zc0 + zc1 * zc1), Lim<R>::tiny()));
      zc0 *= inv;
      zc1 *= inv;
    }
    if (i0 < n) Z[(size_t)i0 * ks + c] = zc0;
    if (i1 < n) Z[(size_t)i1 * ks + c] = zc1;
  }
  __syncthreads();
}

constexpr int kKB = 16;    // components per eigenvector batch (registers per lane)
constexpr int kKBBig = 4;  // the same for Grams above 128 (more rows per lane)

// One wavefront per patch.  Eigenvalues by multisection on the tridiagonal,
// the method's rule, then for the kept components: inverse iteration (one
// lane per component, LU in scratch), Gram-Schmidt inside eigenvalue
// clusters, and back-transformation y = H_0 ... H_{n-2} z with the vectors
// held row-distributed in registers (lane l: rows l + 64 s, s < NR; batches
// of up to KBC components).
template <class E, class RT, int NR = 2, int KBC = kKB, bool AD = false>
__global__ __launch_bounds__(kWave) void trieig_kernel(EigArgs a) {
  using R = typename real_of<E>::type;
  const int lane = threadIdx.x;
  const int n = a.n, KB = a.kb;
  extern __shared__ __align__(16) char smem[];
  char* cur = smem;
  WaveScratch<R> s;
  s.lo = lds_carve<R>(cur, kWave);
  s.hi = lds_carve<R>(cur, kWave);
  s.xs = lds_carve<R>(cur, kWave);
  s.cs = lds_carve<int>(cur, kWave);
  s.tgt = lds_carve<int>(cur, n + 16);
  s.val = lds_carve<R>(cur, n + 16);
  R* lam = lds_carve<R>(cur, n + 16);       // eigenvalue of each component
  int* cidx = lds_carve<int>(cur, n + 16);  // ascending eigen index of each component
  double* wc = lds_carve<double>(cur, n + 16);
  double* sv = AD ? lds_carve<double>(cur, n + 16) : nullptr;  // ADAPTIVE: singular values, descending
  R* zb = lds_carve<R>(cur, (size_t)n * KB);  // z, component-major per lane: zb[i * KB + c]
  R* umb = lds_carve<R>(cur, (size_t)n * KB);  // twisted-factorization multipliers, same layout
  R* ld = lds_carve<R>(cur, n + 16);           // the patch's tridiagonal d | e | e^2 in LDS
  R* le = lds_carve<R>(cur, n + 16);
  R* le2 = lds_carve<R>(cur, n + 16);
  __shared__ int sh_i[4];
  __shared__ double sh_d[4];

  for (int p = blockIdx.x, np_ = a.np_dev ? *a.np_dev : a.np; p < np_; p += gridDim.x) {
    {  // the tridiagonal into LDS: every recurrence below reads it
      const R* src = reinterpret_cast<const R*>(a.DE) + (size_t)p * 3 * n;
      for (int i = lane; i < n; i += kWave) {
        ld[i] = src[i];
        le[i] = src[n + i];
        le2[i] = src[2 * n + i];
      }
      wave_sync();
    }
    const R* dg = ld;
    const R* eg = le;
    const R* e2g = le2;
    // Gershgorin interval
    R glo = R(INFINITY), ghi = R(-INFINITY), emax2 = R(0);
    for (int i = lane; i < n; i += kWave) {
      const R off = (i > 0 ? fabs(eg[i - 1]) : R(0)) + fabs(eg[i]);
      glo = fmin(glo, dg[i] - off);
      ghi = fmax(ghi, dg[i] + off);
      emax2 = fmax(emax2, e2g[i]);
    }
    for (int m = kWave / 2; m >= 1; m >>= 1) {
      glo = fmin(glo, shfl_xor_(glo, m));
      ghi = fmax(ghi, shfl_xor_(ghi, m));
      emax2 = fmax(emax2, shfl_xor_(emax2, m));
    }
    const R tnorm = fmax(fabs(glo), fabs(ghi));
    const R pivmin = Lim<R>::tiny() * fmax(R(1), emax2);
    const R atol = Lim<R>::ulp() * tnorm * R(0.5);
    glo -= R(4) * Lim<R>::ulp() * tnorm + R(4) * pivmin;
    ghi += R(4) * Lim<R>::ulp() * tnorm + R(4) * pivmin;

    // prior (Hybrid: mean variance; Optimal: mean std) over the patch voxels
    double prior = 0.0;
    if (a.var_map) {
      int64_t sp;
      int t0;
      patch_origin(a.g, a.idx[p], sp, t0);
      double acc = 0.0;
      for (int r = lane; r < a.g.M; r += kWave) {
        const double v = a.var_map[(sp + row_offset(a.g, r)) / a.g.T];
        acc += (a.method == PDAMD_OPTIMAL) ? sqrt(v) : v;
      }
      prior = dpp_sum(acc) / (double)a.g.M;
      if (a.var_f32) prior = (double)(float)prior;
    }

    int ncomp = 0, rank = 0;
    float noise = NAN;
    // ---------------- threshold rules
    if (a.method == PDAMD_MPPCA || a.method == PDAMD_HYBRID) {
      const int nt = min(10, n);
      for (int j = lane; j < nt; j += kWave) s.tgt[j] = n - 1 - j;
      wave_sync();
      multisect(s, dg, e2g, n, pivmin, glo, ghi, atol, nt);
      if (lane == 0) {
        RT ev[10];
        for (int q = 0; q < 10; ++q) {
          const int j = 9 - q;  // descending position
          ev[q] = j < nt ? (RT)s.val[j] : RT(0);
        }
        int k = 0;
        RT var;
        // a decision whose two sides lie within refine_rel of each other
        // (relative to the larger) is re-checked in FP64 when NEAR is set
        const double rr = a.refine_rel;
        bool near = false;
        auto close = [&](double x, double y) {
          near = near || fabs(x - y) <= rr * fmax(fabs(x), fabs(y));
        };
        if (a.method == PDAMD_MPPCA) {
          // lowrank.py:45-54
          double mv = (double)np_mean10(ev, 10) * (4.0 * sqrt((double)(10 - k + 1) / a.g.M));
          close(mv, (double)(ev[9 - k] - ev[0]));
          while (k < 10 && mv < (double)(ev[9 - k] - ev[0])) {
            ++k;
            mv = (double)np_mean10(ev, 10 - k) * (4.0 * sqrt((double)(a.g.pt - k + 1) / a.g.M));
            if (k < 10) close(mv, (double)(ev[9 - k] - ev[0]));
          }
          var = np_mean10(ev, 10 - k);
          int r = 0;
          if (a.scale_f64) {
            const double thr = (double)var * a.ts2;
            for (int q = 0; q < 10; ++q) {
              r += ((double)ev[q] > thr);
              close((double)ev[q], thr);
            }
          } else {
            const RT thr = var * (RT)a.ts2;
            for (int q = 0; q < 10; ++q) {
              r += (ev[q] > thr);
              close((double)ev[q], (double)thr);
            }
          }
          rank = r;
        } else {
          // lowrank.py:108-112
          var = np_mean10(ev, 10);
          close((double)var, prior);
          while ((double)var > prior && k < 10 - 2) {
            ++k;
            var = np_mean10(ev, 10 - k);
            close((double)var, prior);
          }
          rank = k;
        }
        if (a.NEAR) a.NEAR[p] = near ? 1 : 0;
        noise = (float)var;
        ncomp = min(rank, nt);
        for (int c = 0; c < ncomp; ++c) {
          cidx[c] = n - 1 - c;
          lam[c] = s.val[c];
          wc[c] = 1.0;
        }
        if (a.EIG)
          for (int q = 0; q < 10; ++q) a.EIG[(size_t)p * 10 + q] = (float)ev[q];
        sh_i[0] = ncomp;
        sh_i[1] = rank;
        sh_d[0] = (double)noise;
      }
      wave_sync();
      ncomp = sh_i[0];
      rank = sh_i[1];
      noise = (float)sh_d[0];
    } else if (a.method == PDAMD_HARD) {
      // lowrank.py:175-180: r = #{s > tau}; keep those components unchanged
      const double tlo = a.tau > 0 ? a.tau * (1.0 - 1e-5) : a.tau;
      int cand = n;
      if (tlo > 0) {
        const R x = (R)(tlo * tlo);
        cand = n - sturm_g(dg, e2g, n, x, pivmin);
      }
      for (int j = lane; j < cand; j += kWave) s.tgt[j] = n - 1 - j;
      wave_sync();
      multisect(s, dg, e2g, n, pivmin, glo, ghi, atol, cand);
      if (lane == 0) {
        int r = 0;
        for (int j = 0; j < cand; ++j) {
          const RT sv = (RT)sqrt(fmax((double)s.val[j], 0.0));
          const bool keep = a.tau_f64 ? ((double)sv > a.tau) : (sv > (RT)a.tau);
          if (!keep) break;
          ++r;
        }
        rank = r;
        ncomp = r;
        for (int c = 0; c < r; ++c) {
          cidx[c] = n - 1 - c;
          lam[c] = s.val[c];
          wc[c] = 1.0;
        }
        sh_i[0] = ncomp;
        sh_i[1] = rank;
      }
      wave_sync();
      ncomp = sh_i[0];
      rank = sh_i[1];
    } else if (a.method == PDAMD_ADAPTIVE) {
      // ---- adaptive trace-norm shrinkage, lowrank.py:597-636 (parameters: 444-540);
      // compiled only into the AD instantiation (its search code would cost the
      // other rules' kernel registers)
      if constexpr (AD) {
      for (int j = lane; j < n; j += kWave) s.tgt[j] = j;
      wave_sync();
      multisect(s, dg, e2g, n, pivmin, glo, ghi, atol, n);
      // singular values, descending, rounded to the data's precision (the
      // reference's SVD returns them in that precision)
      for (int j = lane; j < n; j += kWave)
        sv[j] = (double)(RT)sqrt(fmax((double)s.val[n - 1 - j], 0.0));
      wave_sync();
      const double sig = sqrt(prior), sig2 = sig * sig;
      const double nr = (double)a.g.M, ncl = (double)a.g.pt;
      double gam, tau;
      if (a.adaptive_method == PDAMD_ADAPT_QUT) {
        tau = a.qut_tau[a.idx[p]];
        if (a.n_gamma >= 1) {
          gam = a.gamma[0];
        } else {  // lowrank.py:478-490: gamma = exp(y) + 1, y from 0
          double fb;
          const double y = local_min(
              [&](double yy) { return atn_cost(sv, n, exp(yy) + 1.0, tau, false, sig2, nr, ncl); },
              0.0, &fb);
          gam = exp(y) + 1.0;
        }
      } else {  // lowrank.py:494-513: each gamma, tau from log(tau0)
        const bool gs = a.adaptive_method == PDAMD_ADAPT_GSURE;
        const bool x32 = isnan(a.tau0_log);  // default start: float32 / float64 log median
        double x0 = a.tau0_log;
        if (x32) {  // log(median(s)) in the singular values' precision
          const RT md = (n % 2) ? (RT)sv[n / 2] : (RT)(((RT)sv[n / 2 - 1] + (RT)sv[n / 2]) / RT(2));
          x0 = (double)(RT)log(md);
        }
        double best = INFINITY;
        gam = a.gamma[0];
        tau = exp(x0);
        if (!a.rule_f64) {
          // Single-precision data: the reference's cost is evaluated in float32
          // and its BFGS (finite-difference gradients) returns the start point
          // itself, so tau = exp(x0) (a float32 exponential for the default
          // start) and gamma = the first candidate of least cost there.
          const double tq = x32 ? (double)(RT)exp((RT)x0) : exp(x0);
          for (int q = 0; q < a.n_gamma; ++q) {
            const double fq = atn_cost(sv, n, a.gamma[q], tq, gs, sig2, nr, ncl);
            if (fq < best) {
              best = fq;
              gam = a.gamma[q];
            }
          }
          tau = tq;
        } else {
          for (int q = 0; q < a.n_gamma; ++q) {
            const double gq = a.gamma[q];
            double fq;
            const double xq = local_min(
                [&](double x) { return atn_cost(sv, n, gq, exp(x), gs, sig2, nr, ncl); }, x0, &fq);
            if (fq < best) {
              best = fq;
              gam = gq;
              tau = exp(xq);
            }
          }
        }
      }
      if (lane == 0) {
        int nk = 0, last = -1;
        for (int j = 0; j < n; ++j) {  // descending position j
          const double sj = sv[j];
          const double dh = sj > 0.0 ? sj * fmax(1.0 - pow(tau / sj, gam), 0.0) : 0.0;
          if (dh != 0.0) {
            last = j;
            const int ai = n - 1 - j;  // ascending eigen index
            cidx[nk] = ai;
            lam[nk] = s.val[ai];
            wc[nk] = dh / sqrt(fmax((double)s.val[ai], 0.0));
            ++nk;
          }
        }
        ncomp = nk;
        rank = last + 1;
        sh_i[0] = ncomp;
        sh_i[1] = rank;
      }
      wave_sync();
      ncomp = sh_i[0];
      rank = sh_i[1];
      }
    } else {
      // ---- optimal shrinkage, lowrank.py:387-407
      double sigma;
      if (a.var_map) {
        sigma = prior;
      } else {
        const int nt = (n % 2) ? 1 : 2;
        if (lane == 0) {
          s.tgt[0] = (n - 1) / 2;
          s.tgt[1] = n / 2;
        }
        wave_sync();
        multisect(s, dg, e2g, n, pivmin, glo, ghi, atol, nt);
        RT med;
        if (nt == 1) {
          med = (RT)sqrt(fmax((double)s.val[0], 0.0));
        } else {
          const RT s0 = (RT)sqrt(fmax((double)s.val[0], 0.0));
          const RT s1 = (RT)sqrt(fmax((double)s.val[1], 0.0));
          med = (RT)(s0 + s1) / RT(2);
        }
        sigma = (double)med / sqrt((double)a.g.pt * a.mpm);
      }
      const double cs =
          sqrt((double)a.g.pt) * sigma * (a.LAM ? a.LAM[a.idx[p]] : 1.0);
      const double sb = sqrt(a.beta);
      // candidates above the upper bulk edge, and (Frobenius) below the lower one
      int ctop = n, cbot = 0;
      {
        const double xt = cs * (1.0 + sb) * (1.0 - 1e-5);
        if (xt > 0 && isfinite(xt)) ctop = n - sturm_g(dg, e2g, n, (R)(xt * xt), pivmin);
        if (a.loss == PDAMD_LOSS_FRO && 1.0 - sb > 0) {
          const double xb = cs * (1.0 - sb) * (1.0 + 1e-5);
          if (isfinite(xb)) cbot = sturm_g(dg, e2g, n, (R)(xb * xb), pivmin);
          if (!isfinite(xb)) cbot = n;
        }
        if (!isfinite(cs) || cs <= 0) {
          ctop = n;
          cbot = 0;
        }
        if (ctop + cbot > n) cbot = n - ctop;
      }
      for (int j = lane; j < ctop; j += kWave) s.tgt[j] = n - 1 - j;
      for (int j = lane; j < cbot; j += kWave) s.tgt[ctop + j] = j;
      wave_sync();
      multisect(s, dg, e2g, n, pivmin, glo, ghi, atol, ctop + cbot);
      if (lane == 0) {
        int nc = 0, last = -1;
        for (int j = 0; j < ctop + cbot; ++j) {
          const double lamj = (double)s.val[j];
          const RT sv = (RT)sqrt(fmax(lamj, 0.0));
          double sh = cs * shrink(a.loss, (double)sv / cs, a.beta);
          if (isnan(sh)) sh = 0.0;
          if (sh != 0.0) {
            const int desc = j < ctop ? j : (n - 1 - s.tgt[j]);
            last = max(last, desc);
            cidx[nc] = s.tgt[j];
            lam[nc] = s.val[j];
            wc[nc] = sh / sqrt(fmax(lamj, 0.0));
            ++nc;
          }
        }
        ncomp = nc;
        rank = last + 1;
        sh_i[0] = ncomp;
        sh_i[1] = rank;
      }
      wave_sync();
      ncomp = sh_i[0];
      rank = sh_i[1];
    }
    ncomp = min(ncomp, a.kmax);

    // ---------------- eigenvectors of the kept components
    const E* Hp = reinterpret_cast<const E*>(a.H) + (size_t)p * n * n;
    const E* __restrict__ TAUp = reinterpret_cast<const E*>(a.TAU) + (size_t)p * n;
    E* Yp = reinterpret_cast<E*>(a.Y) + (size_t)p * n * a.kmax;
    R* Zp = reinterpret_cast<R*>(a.Z) + (size_t)p * n * a.kmax;
    R* scr = reinterpret_cast<R*>(a.SCR) + (size_t)p * 3 * n * KB;
    const R eps_piv = Lim<R>::ulp() * tnorm + pivmin;
    // Gram-Schmidt inside clusters of eigenvalues closer than clus: inverse
    // iteration from eigenvalues accurate to ~ulp * tnorm yields vectors
    // orthogonal to ~ulp * tnorm / gap, so FP64 needs it only for gaps below
    // 1e-8 tnorm (error <= ~1e-8); FP32 keeps the classical 1e-3 tnorm.
    const R clus = (sizeof(R) == 8 ? R(1e-8) : R(1e-3)) * tnorm;
    const int n_init = 3;  // initial solves per vector
    // rows held by this lane: lane + 64 s, s < NR
    auto row_of = [&](int sl) { return lane + kWave * sl; };
    for (int c0 = 0; c0 < ncomp; c0 += KB) {
      const int kb = min(KB, ncomp - c0);
      // inverse iteration, one lane per component
      auto solve = [&](int c) {
        tri_solve(dg, eg, n, lam[c0 + c], scr + c, scr + n * KB + c, scr + 2 * n * KB + c,
                  zb + c, KB, eps_piv);
        R mx = R(0);
        for (int i = 0; i < n; ++i) mx = fmax(mx, fabs(zb[i * KB + c]));
        const R inv = mx > R(0) ? R(1) / mx : R(1);
        for (int i = 0; i < n; ++i) zb[i * KB + c] *= inv;
      };
      // eigenvectors by twisted factorization, one lane per component; a
      // non-finite result falls back to inverse iteration from a fixed start
      if (lane < kb) {
        if (!twisted_vec(dg, eg, e2g, n, lam[c0 + lane], eps_piv, zb + lane, umb + lane, KB)) {
          uint32_t h = 0x9E3779B9u * (uint32_t)(c0 + lane + 1);
          for (int i = 0; i < n; ++i) {
            h ^= h << 13; h ^= h >> 17; h ^= h << 5;
            zb[i * KB + lane] = R(0.5) + (R)(h & 0xFFFF) * R(1.0 / 65536.0);
          }
          for (int it = 0; it < n_init; ++it) solve(lane);
        }
      }
      // components of this batch inside an eigenvalue cluster (closer than
      // clus to another kept component): they get a second orthogonalisation
      // pass after one more inverse-iteration step
      bool clustered = false;
      if (lane < kb) {
        for (int q = 0; q < ncomp; ++q)
          if (q != c0 + lane && fabs(lam[q] - lam[c0 + lane]) <= clus) clustered = true;
      }
      const bool any_cluster = __any(clustered);
      wave_sync();
      // rows -> registers
      R z[NR][KBC];
      auto load_z = [&]() {
#pragma unroll
        for (int sl = 0; sl < NR; ++sl)
#pragma unroll
          for (int c = 0; c < KBC; ++c)
            z[sl][c] = (c < kb && row_of(sl) < n) ? zb[row_of(sl) * KB + c] : R(0);
      };
      load_z();
      // sum over this lane's rows of x_s y_s, then over the wavefront
      auto dot_rows = [&](auto fx, auto fy) {
        R t = fx(0) * fy(0);
#pragma unroll
        for (int sl = 1; sl < NR; ++sl) t += fx(sl) * fy(sl);
        return dpp_sum(t);
      };
      // classical Gram-Schmidt, twice, inside eigenvalue clusters; one extra
      // inverse-iteration step in between
      for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
        for (int c = 0; c < KBC; ++c) {
          if (c >= kb) continue;
          const int gc = c0 + c;
          // earlier batches (stored z in global memory).  The components are
          // ordered monotonically by eigenvalue within each run (descending
          // top, ascending bottom / all), so the cluster of gc among them is
          // the contiguous run below c0 found by this scan.
          int qs = c0;
          while (qs > 0 && fabs(lam[qs - 1] - lam[gc]) <= clus) --qs;
          for (int q = qs; q < c0; ++q) {
            if (fabs(lam[q] - lam[gc]) > clus) continue;
            R zq[NR];
#pragma unroll
            for (int sl = 0; sl < NR; ++sl)
              zq[sl] = row_of(sl) < n ? Zp[(size_t)row_of(sl) * a.kmax + q] : R(0);
            const R dot = dot_rows([&](int sl) { return zq[sl]; }, [&](int sl) { return z[sl][c]; });
#pragma unroll
            for (int sl = 0; sl < NR; ++sl) z[sl][c] -= dot * zq[sl];
          }
#pragma unroll
          for (int q = 0; q < KBC; ++q) {
            if (q >= c) break;
            if (fabs(lam[c0 + q] - lam[gc]) > clus) continue;
            const R dot = dot_rows([&](int sl) { return z[sl][q]; }, [&](int sl) { return z[sl][c]; });
#pragma unroll
            for (int sl = 0; sl < NR; ++sl) z[sl][c] -= dot * z[sl][q];
          }
          const R nrm = dot_rows([&](int sl) { return z[sl][c]; }, [&](int sl) { return z[sl][c]; });
          const R inv = R(1) / sqrt(fmax(nrm, Lim<R>::tiny()));
#pragma unroll
          for (int sl = 0; sl < NR; ++sl) z[sl][c] *= inv;
        }
        if (pass == 0) {
          if (!any_cluster) break;  // isolated eigenvalues: the vectors are final
#pragma unroll
          for (int sl = 0; sl < NR; ++sl)
#pragma unroll
            for (int c = 0; c < KBC; ++c)
              if (c < kb && row_of(sl) < n) zb[row_of(sl) * KB + c] = z[sl][c];
          wave_sync();
          if (lane < kb && clustered) solve(lane);
          wave_sync();
          load_z();
        }
      }
      // keep z for clustering against later batches (read back by other lanes)
      if (c0 + kb < ncomp) {
#pragma unroll
        for (int sl = 0; sl < NR; ++sl)
#pragma unroll
          for (int c = 0; c < KBC; ++c)
            if (c < kb && row_of(sl) < n) Zp[(size_t)row_of(sl) * a.kmax + c0 + c] = z[sl][c];
        __syncthreads();
      }
      if constexpr (std::is_same<E, float>::value && NR == 2 && KBC % 2 == 0) {
        // FP32, n <= 128: components in pairs, one per half wavefront -- lane
        // 32 h + i holds rows i + 32 s (s < 4) of components 2 j + h -- so an
        // instruction serves two components and each dot product is a sum
        // over 32 lanes (the 64-lane layout above: 6 levels, one component).
        // z moves into that layout through LDS (zb).
        constexpr int KH = KBC / 2, PF2 = 4;
        const int hf = lane >> 5, il = lane & 31;
        wave_sync();
#pragma unroll
        for (int sl = 0; sl < NR; ++sl)
#pragma unroll
          for (int c = 0; c < KBC; ++c)
            if (c < kb && row_of(sl) < n) zb[row_of(sl) * KB + c] = z[sl][c];
        wave_sync();
        float y2[4][KH];
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
          for (int j = 0; j < KH; ++j) {
            const int r = il + 32 * s4, c = 2 * j + hf;
            y2[s4][j] = (c < kb && r < n) ? zb[r * KB + c] : 0.f;
          }
        // reflectors kbase, kbase - 1, ... (packed rows, indexed by column)
        float vr2[4][PF2], tv2[PF2];
        auto fetch2 = [&](int kbase) {
#pragma unroll
          for (int u = 0; u < PF2; ++u) {
            const int k = kbase - u;
            if (k >= 0) {
              const float* row = reinterpret_cast<const float*>(Hp) + hh_row(k, n) - (k + 1);
              tv2[u] = reinterpret_cast<const float*>(TAUp)[k];
#pragma unroll
              for (int s4 = 0; s4 < 4; ++s4) {
                const int r = il + 32 * s4;
                vr2[s4][u] = (r > k + 1 && r < n) ? row[r] : (r == k + 1 ? 1.f : 0.f);
              }
            } else {
              tv2[u] = 0.f;
#pragma unroll
              for (int s4 = 0; s4 < 4; ++s4) vr2[s4][u] = 0.f;
            }
          }
        };
        // sum over this lane's half wavefront: row sums of 16 lanes (DPP),
        // then the half's two rows; bit-identical in all lanes of the half
        auto half_sum = [&](float v) {
          v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
          v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
          v += dpp_f<0x141>(v);  // row_half_mirror
          v += dpp_f<0x140>(v);  // row_mirror
          const float s0 = readlane_(v, 0) + readlane_(v, 16);
          const float s1 = readlane_(v, 32) + readlane_(v, 48);
          return hf ? s1 : s0;
        };
        fetch2(n - 2);
        for (int kbase = n - 2; kbase >= 0; kbase -= PF2) {
          float cv[4][PF2], ct[PF2];
#pragma unroll
          for (int u = 0; u < PF2; ++u) {
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) cv[s4][u] = vr2[s4][u];
            ct[u] = tv2[u];
          }
          if (kbase - PF2 >= 0) fetch2(kbase - PF2);
#pragma unroll
          for (int u = 0; u < PF2; ++u) {
            const float tk = ct[u];
            if (tk == 0.f) continue;
#pragma unroll
            for (int j = 0; j < KH; ++j) {
              if (2 * j >= kb) break;
              float t = cv[0][u] * y2[0][j];
#pragma unroll
              for (int s4 = 1; s4 < 4; ++s4) t = fmaf(cv[s4][u], y2[s4][j], t);
              const float sdot = tk * half_sum(t);
#pragma unroll
              for (int s4 = 0; s4 < 4; ++s4) y2[s4][j] = fmaf(-cv[s4][u], sdot, y2[s4][j]);
            }
          }
        }
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
          for (int j = 0; j < KH; ++j) {
            const int r = il + 32 * s4, c = 2 * j + hf;
            if (c < kb && r < n) Yp[(size_t)r * a.kmax + c0 + c] = y2[s4][j];
          }
      } else {
      // back-transformation y = H_0 ... H_{n-2} z, reflectors prefetched PF ahead
      E y[NR][KBC];
#pragma unroll
      for (int sl = 0; sl < NR; ++sl)
#pragma unroll
        for (int c = 0; c < KBC; ++c) y[sl][c] = E(z[sl][c]);
      if constexpr (NR > kNMid / kWave) {
        // Grams above kNMid (one component per batch): each reflector row is
        // read once from memory into registers, no prefetch
        static_assert(KBC == 1, "one component per batch above kNMid");
        for (int k = n - 2; k >= 0; --k) {
          const E tk = TAUp[k];
          if (re_(tk) == R(0) && im_(tk) == R(0)) continue;
          const E* row = Hp + (size_t)k * n;
          E v[NR];
#pragma unroll
          for (int sl = 0; sl < NR; ++sl) {
            const int i = row_of(sl);
            v[sl] = (i > k + 1 && i < n) ? row[i] : E(R(i == k + 1 ? 1 : 0));
          }
          E t = conj_(v[0]) * y[0][0];
#pragma unroll
          for (int sl = 1; sl < NR; ++sl) t += conj_(v[sl]) * y[sl][0];
          const E sdot = tk * dpp_sum(t);
#pragma unroll
          for (int sl = 0; sl < NR; ++sl) y[sl][0] -= v[sl] * sdot;
        }
      } else {
      constexpr int PF = NR > 2 ? 2 : 8;
      E vr[NR][PF], tv[PF];
      auto fetch = [&](int kbase) {  // reflectors kbase, kbase-1, ..., kbase-PF+1
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int k = kbase - u;
          if (k >= 0) {
            // reflector k indexed by absolute column: packed rows (n <= 128) or
            // row k of the full matrix (larger Grams)
            const E* row = NR > 2 ? Hp + (size_t)k * n : Hp + hh_row(k, n) - (k + 1);
            tv[u] = TAUp[k];
#pragma unroll
            for (int sl = 0; sl < NR; ++sl) {
              const int i = row_of(sl);
              vr[sl][u] = (i > k + 1 && i < n) ? row[i] : E(R(i == k + 1 ? 1 : 0));
            }
          } else {
            tv[u] = E(R(0));
#pragma unroll
            for (int sl = 0; sl < NR; ++sl) vr[sl][u] = E(R(0));
          }
        }
      };
      fetch(n - 2);
      for (int kbase = n - 2; kbase >= 0; kbase -= PF) {
        E cv[NR][PF], ct[PF];
#pragma unroll
        for (int u = 0; u < PF; ++u) {
#pragma unroll
          for (int sl = 0; sl < NR; ++sl) cv[sl][u] = vr[sl][u];
          ct[u] = tv[u];
        }
        if (kbase - PF >= 0) fetch(kbase - PF);
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const E tk = ct[u];
          if (re_(tk) == R(0) && im_(tk) == R(0)) continue;
#pragma unroll
          for (int c = 0; c < KBC; ++c) {
            if (c >= kb) break;
            E t = conj_(cv[0][u]) * y[0][c];
#pragma unroll
            for (int sl = 1; sl < NR; ++sl) t += conj_(cv[sl][u]) * y[sl][c];
            const E sdot = tk * dpp_sum(t);
#pragma unroll
            for (int sl = 0; sl < NR; ++sl) y[sl][c] -= cv[sl][u] * sdot;
          }
        }
      }
      }
#pragma unroll
      for (int sl = 0; sl < NR; ++sl)
#pragma unroll
        for (int c = 0; c < KBC; ++c)
          if (c < kb && row_of(sl) < n) Yp[(size_t)row_of(sl) * a.kmax + c0 + c] = y[sl][c];
      }
      wave_sync();
    }
    if (a.zero_fill) {  // factor layout of the gather recombination: clean unused columns
      E* Yz = reinterpret_cast<E*>(a.Y) + (size_t)p * n * a.kmax;
      const int nz = a.kmax - ncomp;
      for (int q = lane; q < n * nz; q += kWave) Yz[(size_t)(q / nz) * a.kmax + ncomp + q % nz] = E(R(0));
    }
    // ---------------- per-patch scalars
    if (lane == 0) {
      a.NCOMP[p] = ncomp;
      a.THETA[p] = 1.0 / (2.0 + (double)rank);
      if (a.RANK) a.RANK[p] = rank;
      if (a.NOISE) a.NOISE[p] = noise;
    }
    for (int c = lane; c < ncomp; c += kWave) a.WC[(size_t)p * a.kmax + c] = wc[c];
    wave_sync();
  }
}

template <class E>
size_t trieig_lds_bytes(int n, int kb, bool ad = false) {
  using R = typename real_of<E>::type;
  auto al = [](size_t b) { return ((b + 15) / 16) * 16; };
  size_t s = 3 * al(kWave * sizeof(R)) + al(kWave * sizeof(int));
  s += al((n + 16) * sizeof(int)) + al((n + 16) * sizeof(R));
  s += al((n + 16) * sizeof(R)) + al((n + 16) * sizeof(int)) +
       (ad ? 2 : 1) * al((n + 16) * sizeof(double));
  s += 2 * al((size_t)n * kb * sizeof(R));       // zb, umb
  s += 3 * al((size_t)(n + 16) * sizeof(R));      // d, e, e^2
  return s;
}

// All eigenpairs of each patch's tridiagonal (recomputation for the backward
// pass): multisection for every eigenvalue (ascending, into WC), then
// all_pairs_vectors into Z.  A lean one-wavefront kernel: no threshold rules,
// no register-held batches, so many wavefronts share a SIMD and hide the
// latency of the Sturm and LU recurrences.
template <class R>
__global__ __launch_bounds__(kWave) void trieig_all_kernel(EigArgs a) {
  const int lane = threadIdx.x;
  const int n = a.n;
  extern __shared__ __align__(16) char smem[];
  char* cur = smem;
  WaveScratch<R> s;
  s.lo = lds_carve<R>(cur, kWave);
  s.hi = lds_carve<R>(cur, kWave);
  s.xs = lds_carve<R>(cur, kWave);
  s.cs = lds_carve<int>(cur, kWave);
  s.tgt = lds_carve<int>(cur, n + 16);
  s.val = lds_carve<R>(cur, n + 16);
  for (int p = blockIdx.x; p < a.np; p += gridDim.x) {
    const R* __restrict__ dg = reinterpret_cast<const R*>(a.DE) + (size_t)p * 3 * n;
    const R* __restrict__ eg = dg + n;
    const R* __restrict__ e2g = dg + 2 * n;
    R glo = R(INFINITY), ghi = R(-INFINITY), emax2 = R(0);
    for (int i = lane; i < n; i += kWave) {
      const R off = (i > 0 ? fabs(eg[i - 1]) : R(0)) + fabs(eg[i]);
      glo = fmin(glo, dg[i] - off);
      ghi = fmax(ghi, dg[i] + off);
      emax2 = fmax(emax2, e2g[i]);
    }
    for (int m = kWave / 2; m >= 1; m >>= 1) {
      glo = fmin(glo, shfl_xor_(glo, m));
      ghi = fmax(ghi, shfl_xor_(ghi, m));
      emax2 = fmax(emax2, shfl_xor_(emax2, m));
    }
    const R tnorm = fmax(fabs(glo), fabs(ghi));
    const R pivmin = Lim<R>::tiny() * fmax(R(1), emax2);
    const R atol = Lim<R>::ulp() * tnorm * R(0.5);
    glo -= R(4) * Lim<R>::ulp() * tnorm + R(4) * pivmin;
    ghi += R(4) * Lim<R>::ulp() * tnorm + R(4) * pivmin;
    for (int j = lane; j < n; j += kWave) s.tgt[j] = j;
    wave_sync();
    if (n > kWave && n <= 2 * kWave)
      bisect_pairs<R, true>(s, dg, e2g, n, pivmin, glo, ghi, atol, n);
    else
      multisect<R, true>(s, dg, e2g, n, pivmin, glo, ghi, atol, n);
    for (int c = lane; c < n; c += kWave) a.WC[(size_t)p * a.kmax + c] = (double)s.val[c];
    const R eps_piv = Lim<R>::ulp() * tnorm + pivmin;
    const R clus = (sizeof(R) == 8 ? R(1e-8) : R(1e-3)) * tnorm;
    all_pairs_vectors(dg, eg, n, (const R*)s.val, reinterpret_cast<R*>(a.Z) + (size_t)p * n * a.kmax,
                      reinterpret_cast<R*>(a.SCR) + (size_t)p * 3 * n * n, a.kmax, eps_piv, clus);
  }
}

template <class R>
size_t trieig_all_lds_bytes(int n) {
  auto al = [](size_t b) { return ((b + 15) / 16) * 16; };
  return 3 * al(kWave * sizeof(R)) + al(kWave * sizeof(int)) + al((n + 16) * sizeof(int)) +
         al((n + 16) * sizeof(R));
}

// ===========================================================================
// 4. Synthesis + scatter-accumulate recombination (base.py:243-256)
// ===========================================================================

struct SynthArgs {
  Geom g;
  const int32_t* idx;
  int np, n, kmax;
  int recomb;
  int w3d;  // weights accumulated per voxel (one temporal patch) instead of per element
  const void* Y;
  const double* WC;
  const int* NCOMP;
  const double* THETA;
  const void* MEAN;  // np x pt per-frame patch means written by the Gram kernels
  void* out;
  double* w;  // w3d ? voxel weights : element weights (float64 accumulators)
};

__device__ __forceinline__ void atomic_add_(float* p, float v) { atomicAdd(p, v); }
__device__ __forceinline__ void atomic_add_(double* p, double v) { atomicAdd(p, v); }
template <class R> __device__ __forceinline__ void atomic_add_(cpx<R>* p, cpx<R> v) {
  R* q = reinterpret_cast<R*>(p);
  atomicAdd(q, v.x);
  atomicAdd(q + 1, v.y);
}

constexpr int kSynSC = 32;  // frames (M-side) or rows (T-side) per staged chunk
constexpr int kSynKG = 16;  // components per group held in LDS
constexpr int kSynNI = 16;  // output items per thread per chunk (128 x 32 / 256)

// SC frames (M-side) or rows (T-side) per staged chunk and KG components per
// group: SC * n <= kSynNI * kNT (one staged chunk = kSynNI items per thread).
template <class IN, class E, bool TSIDE, int SC = kSynSC, int KG = kSynKG>
__global__ __launch_bounds__(kNT) void synth_kernel(const IN* __restrict__ vol, SynthArgs a) {
  using R = typename real_of<E>::type;
  const Geom& g = a.g;
  const int n = a.n, tid = threadIdx.x;
  const int M = g.M, T = g.pt;
  extern __shared__ __align__(16) char smem[];
  char* cur = smem;
  E* Ys = lds_carve<E>(cur, (size_t)n * KG);  // component group, Ys[i * KG + c]
  R* wcs = lds_carve<R>(cur, KG);
  const int ldw = TSIDE ? (T + 1) : (M + 1);
  IN* Sraw = lds_carve<IN>(cur, (size_t)SC * ldw);
  E* Sc = lds_carve<E>(cur, (size_t)SC * ldw);
  IN* mean = lds_carve<IN>(cur, TSIDE ? T : SC);
  E* B = lds_carve<E>(cur, (size_t)KG * SC);  // M-side: B[c][t]; T-side: C[r][c]
  int64_t* roff = lds_carve<int64_t>(cur, TSIDE ? SC : M);

  // Patches are visited in a permuted order (it -> it * 1000003 mod np, a
  // bijection since the prime exceeds any chunk) so that workgroups running at
  // the same time add into distant parts of the output: neighbouring patches
  // overlap, and their concurrent float atomics would contend for the same
  // cache lines.
  for (int it = blockIdx.x; it < a.np; it += gridDim.x) {
    const int p = (int)(((int64_t)it * 1000003) % a.np);
    int64_t sp;
    int t0;
    patch_origin(g, a.idx[p], sp, t0);
    const IN* X = vol + sp + t0;
    IN* O = reinterpret_cast<IN*>(a.out) + sp + t0;
    const int K = a.NCOMP[p];
    const double theta = a.recomb == PDAMD_WEIGHTED ? a.THETA[p] : 1.0;
    const E* Yp = reinterpret_cast<const E*>(a.Y) + (size_t)p * n * a.kmax;
    const double* WCp = a.WC + (size_t)p * a.kmax;
    const IN* MEANp = reinterpret_cast<const IN*>(a.MEAN) + (size_t)p * T;
    const bool center = a.recomb == PDAMD_CENTER;
    const int ctr_r = ((g.p0 / 2) * g.p1 + g.p1 / 2) * g.p2 + g.p2 / 2;
    const int ctr_t = T / 2;
    E acc[kSynNI];

    if (!TSIDE) {
      for (int r = tid; r < M; r += kNT) roff[r] = row_offset(g, r);
      lds_barrier();
      for (int c0 = 0; c0 < T; c0 += SC) {
        const int tc = min(SC, T - c0);
        if (center && (ctr_t < c0 || ctr_t >= c0 + tc)) continue;
        {
          constexpr int NL = kSynNI;
          IN buf[NL];
#pragma unroll
          for (int u = 0; u < NL; ++u) {
            const int q = tid + u * kNT;
            const int t = q % SC, r = q / SC;
            buf[u] = (q < SC * M && t < tc) ? X[roff[r] + c0 + t] : IN(0);
          }
#pragma unroll
          for (int u = 0; u < NL; ++u) {
            const int q = tid + u * kNT;
            if (q < SC * M) Sraw[(q % SC) * ldw + q / SC] = buf[u];
          }
        }
        lds_barrier();
        if (tid < tc)  // sequential over rows, as numpy's axis-0 mean
          mean[tid] = div_n(seq_add(Sraw[tid * ldw], Sraw + tid * ldw + 1, M - 1, 1), M);
        lds_barrier();
        for (int q = tid; q < tc * M; q += kNT) {
          const int t = q / M, r = q % M;
          Sc[t * ldw + r] = cvt<E>(Sraw[t * ldw + r] - mean[t]);
        }
#pragma unroll
        for (int j = 0; j < kSynNI; ++j) {
          const int q = tid + j * kNT;
          acc[j] = q < M * SC ? cvt<E>(mean[q % SC]) : E(R(0));
        }
        lds_barrier();
        for (int k0 = 0; k0 < K; k0 += KG) {
          const int kg = min(KG, K - k0);
          for (int q = tid; q < n * kg; q += kNT) {
            const int i = q / kg, c = q % kg;
            Ys[i * KG + c] = Yp[(size_t)i * a.kmax + k0 + c];
          }
          for (int c = tid; c < kg; c += kNT) wcs[c] = (R)WCp[k0 + c];
          lds_barrier();
          // B[c][t] = w_c * sum_r conj(Y[r][c]) Xc[r][t]
          for (int q = tid; q < kg * SC; q += kNT) {
            const int c = q / SC, t = q % SC;
            E b = E(R(0));
            if (t < tc)
              for (int r = 0; r < M; ++r) b += conj_(Ys[r * KG + c]) * Sc[t * ldw + r];
            B[c * SC + t] = wcs[c] * b;
          }
          lds_barrier();
#pragma unroll
          for (int j = 0; j < kSynNI; ++j) {
            const int q = tid + j * kNT;
            if (q < M * SC) {
              const int r = q / SC, t = q % SC;
              E v = acc[j];
              for (int c = 0; c < kg; ++c) v += Ys[r * KG + c] * B[c * SC + t];
              acc[j] = v;
            }
          }
          lds_barrier();
        }
#pragma unroll
        for (int j = 0; j < kSynNI; ++j) {
          const int q = tid + j * kNT;
          if (q >= M * SC) continue;
          const int r = q / SC, t = q % SC;
          if (t >= tc) continue;
          IN* dst = O + roff[r] + c0 + t;
          if (center) {
            if (r == ctr_r && c0 + t == ctr_t) *dst = cvt<IN>(acc[j]);
          } else {
            atomic_add_(dst, cvt<IN>((R)theta * acc[j]));
          }
        }
        lds_barrier();
      }
    } else {
      // per-frame means over all rows: the Gram kernel's (numpy order)
      for (int t = tid; t < T; t += kNT) mean[t] = MEANp[t];
      lds_barrier();
      for (int r0 = 0; r0 < M; r0 += SC) {
        const int rc = min(SC, M - r0);
        if (center && (ctr_r < r0 || ctr_r >= r0 + rc)) continue;
        for (int r = tid; r < rc; r += kNT) roff[r] = row_offset(g, r0 + r);
        lds_barrier();
        {
          constexpr int NL = kSynNI;
          IN buf[NL];
#pragma unroll
          for (int u = 0; u < NL; ++u) {
            const int q = tid + u * kNT;
            buf[u] = q < rc * T ? X[roff[q / T] + q % T] : IN(0);
          }
#pragma unroll
          for (int u = 0; u < NL; ++u) {
            const int q = tid + u * kNT;
            if (q < rc * T) Sc[(q / T) * ldw + q % T] = cvt<E>(buf[u] - mean[q % T]);
          }
        }
#pragma unroll
        for (int j = 0; j < kSynNI; ++j) {
          const int q = tid + j * kNT;
          acc[j] = q < rc * T ? cvt<E>(mean[q % T]) : E(R(0));
        }
        lds_barrier();
        for (int k0 = 0; k0 < K; k0 += KG) {
          const int kg = min(KG, K - k0);
          for (int q = tid; q < n * kg; q += kNT) {
            const int i = q / kg, c = q % kg;
            Ys[i * KG + c] = Yp[(size_t)i * a.kmax + k0 + c];
          }
          for (int c = tid; c < kg; c += kNT) wcs[c] = (R)WCp[k0 + c];
          lds_barrier();
          // C[r][c] = w_c * sum_t Xc[r][t] Y[t][c]; with few (r, c) pairs each sum
          // is split over 2 or 4 adjacent lanes (frames t = h mod split),
          // combined by a fixed butterfly
          const int npair = rc * kg;
          const int split = npair * 4 <= kNT ? 4 : (npair * 2 <= kNT ? 2 : 1);
          for (int q0 = tid; q0 < npair * split; q0 += kNT) {
            const int q = q0 / split, h = q0 % split;
            const int r = q / kg, c = q % kg;
            E b = E(R(0));
            for (int t = h; t < T; t += split) b += Sc[r * ldw + t] * Ys[t * KG + c];
            for (int o = 1; o < split; o <<= 1) b += shfl_xor_e(b, o);
            if (h == 0) B[r * KG + c] = wcs[c] * b;
          }
          lds_barrier();
#pragma unroll
          for (int j = 0; j < kSynNI; ++j) {
            const int q = tid + j * kNT;
            if (q < rc * T) {
              const int r = q / T, t = q % T;
              E v = acc[j];
              for (int c = 0; c < kg; ++c) v += B[r * KG + c] * conj_(Ys[t * KG + c]);
              acc[j] = v;
            }
          }
          lds_barrier();
        }
#pragma unroll
        for (int j = 0; j < kSynNI; ++j) {
          const int q = tid + j * kNT;
          if (q >= rc * T) continue;
          const int r = q / T, t = q % T;
          IN* dst = O + roff[r] + t;
          if (center) {
            if (r0 + r == ctr_r && t == ctr_t) *dst = cvt<IN>(acc[j]);
          } else {
            atomic_add_(dst, cvt<IN>((R)theta * acc[j]));
          }
        }
        lds_barrier();
      }
    }
    // weights
    if (!center) {
      if (a.w3d) {
        for (int r = tid; r < M; r += kNT)
          atomicAdd(a.w + (sp + row_offset(g, r)) / g.T, theta);
      } else {
        for (int q = tid; q < M * T; q += kNT) {
          const int r = q / T, t = q % T;
          atomicAdd(a.w + sp + row_offset(g, r) + t0 + t, theta);
        }
      }
    }
    lds_barrier();
  }
}

template <class IN, class E, bool TSIDE, int SC = kSynSC, int KG = kSynKG>
size_t synth_lds_bytes(int n, int kmax, int M, int T) {
  using R = typename real_of<E>::type;
  auto al = [](size_t b) { return ((b + 15) / 16) * 16; };
  const int ldw = TSIDE ? (T + 1) : (M + 1);
  size_t s = al((size_t)n * KG * sizeof(E)) + al(KG * sizeof(R));
  s += al((size_t)SC * ldw * sizeof(IN)) + al((size_t)SC * ldw * sizeof(E));
  s += al((TSIDE ? T : SC) * sizeof(IN)) + al((size_t)KG * SC * sizeof(E));
  s += al((TSIDE ? SC : M) * sizeof(int64_t));
  return s;
}

// ===========================================================================
// 5. Normalisation and mask (base.py:262-270)
// ===========================================================================

__device__ __forceinline__ float div_w(float a, float w) { return a / w; }
__device__ __forceinline__ double div_w(double a, float w) { return a / (double)w; }
template <class R> __device__ __forceinline__ cpx<R> div_w(cpx<R> a, float w) {
  return {a.x / (R)w, a.y / (R)w};
}

// The weights are summed in float64 and rounded to float32 once here.  The
// atomic summation order changes the float64 sum by ~1e-16 relative, which
// survives the rounding only for sums that close to a float32 rounding
// midpoint, so the float32 weights (and the normalised output) repeat from
// call to call; atomic float32 sums would differ by an ulp between calls.
template <class IN>
__global__ void finalize_kernel(IN* __restrict__ out, float* __restrict__ w4,
                                const double* __restrict__ wacc, int w3d,
                                const uint8_t* __restrict__ mask, int mask_kind, int recomb,
                                int64_t nvox, int64_t T) {
  const int64_t total = nvox * T;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = q / T;
    IN o = out[q];
    if (recomb != PDAMD_CENTER) {
      const float wv = (float)wacc[w3d ? v : q];
      w4[q] = wv;
      o = div_w(o, wv);
    }
    bool keep = true;
    if (mask_kind == PDAMD_MASK_SPATIAL) keep = mask[v] != 0;
    else if (mask_kind == PDAMD_MASK_FULL) keep = mask[q] != 0;
    if (!keep) o = IN(0);
    out[q] = o;
  }
}

// ===========================================================================
// Host side: error handling, dispatch, workspace, driver
// ===========================================================================


__global__ void zero_kernel(uint32_t* __restrict__ p, int64_t nwords) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

int zero_async(void* p, size_t bytes, hipStream_t st) {
  const int64_t nw = (int64_t)(bytes / 4);
  if (nw == 0) return PDAMD_OK;
  zero_kernel<<<(int)std::min<int64_t>((nw + 255) / 256, 65536), 256, 0, st>>>(
      reinterpret_cast<uint32_t*>(p), nw);
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
}

// LDS a workgroup occupies: static + dynamic bytes in 1280-byte granules.
// (An eigen launch of 16,240 bytes per workgroup measured as if nine fit in
// the 160 KiB of a CU, not ten: 10 x 13 granules exceed it.)
constexpr size_t kLdsGranule = 1280;
inline int lds_per_cu(size_t bytes) {
  const size_t g = (bytes + kLdsGranule - 1) / kLdsGranule * kLdsGranule;
  return g ? (int)((160 * 1024) / g) : 1 << 20;
}

// Workgroups of `kern` (`threads` each, `lds` dynamic LDS bytes) that one CU
// holds at once -- registers, LDS and wave slots together -- or 0 if the
// runtime cannot say.  Cached per kernel and shape.
inline int resident_per_cu(const void* kern, int threads, size_t lds) {
  struct Entry {
    const void* k;
    int t;
    size_t l;
    int v;
  };
  thread_local Entry cache[32];
  thread_local int used = 0;
  for (int i = 0; i < used; ++i)
    if (cache[i].k == kern && cache[i].t == threads && cache[i].l == lds) return cache[i].v;
  int nb = 0;
  hipFuncAttributes fa{};
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, threads, lds) != hipSuccess ||
      hipFuncGetAttributes(&fa, kern) != hipSuccess) {
    (void)hipGetLastError();
    nb = 0;
  } else {
    nb = std::min(nb, lds_per_cu(lds + fa.sharedSizeBytes));
  }
  if (used < 32) cache[used++] = {kern, threads, lds, std::max(nb, 0)};
  return std::max(nb, 0);
}

// Workgroups per CU of a persistent launch (patches strided over the grid):
// the caller's override (autotuner) when set, else the default capped by what
// stays resident -- workgroups beyond that would start only when the first
// ones have finished their whole share, a second pass at a fraction of the
// occupancy.  Either way at most what the LDS holds.
inline int pick_per_cu(int override_, int dflt, size_t lds, const void* kern, int threads) {
  const int cap = std::max(1, (int)((160 * 1024) / std::max<size_t>(lds, 1)));
  int v = override_ > 0 ? override_ : dflt;
  if (override_ <= 0) {
    const int res = resident_per_cu(kern, threads, lds);
    if (res > 0) v = std::min(v, res);
  }
  return std::max(1, std::min(v, cap));
}

int grid_for(int64_t np, int per_cu) {
  static int n_cu = 0;
  if (!n_cu) {
    int dev = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    if (n_cu <= 0) n_cu = 256;
  }
  return (int)std::max<int64_t>(1, std::min<int64_t>(np, (int64_t)n_cu * per_cu));
}

template <class IN, class E, bool TSIDE, int TILE, bool BF = false>
int launch_gram(const Plan& pl, const void* vol, GramArgs ga, hipStream_t st) {
  const size_t lds = gram_lds_bytes<IN, E, TSIDE, TILE, BF>(pl.n, pl.g.M);
  if (lds > 160 * 1024) return fail(PDAMD_EUNSUPPORTED, "gram LDS %zu bytes exceeds 160 KiB", lds);
  auto kern = gram_kernel<IN, E, TSIDE, TILE, BF>;
  HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int per_cu = pick_per_cu(pl.per_cu[0], 8, lds, (const void*)kern, kNT);
  kern<<<grid_for(ga.np, per_cu), kNT, lds, st>>>(reinterpret_cast<const IN*>(vol), ga);
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
}

template <class E, int RJ>
int launch_tridiag(const Plan& pl, GramArgs ga, hipStream_t st) {
  const size_t lds = tridiag_lds_bytes<E>(pl.n);
  if (lds > 160 * 1024)
    return fail(PDAMD_EUNSUPPORTED, "tridiagonalisation LDS %zu bytes exceeds 160 KiB", lds);
  auto kern = tridiag_kernel<E, RJ>;
  HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int per_cu = pick_per_cu(pl.per_cu[1], 8, lds, (const void*)kern, kNT);
  kern<<<grid_for(ga.np, per_cu), kNT, lds, st>>>(ga);
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
}

template <class IN, class E, bool TSIDE>
int launch_gram_big(const Plan& pl, const void* vol, GramArgs ga, hipStream_t st) {
  constexpr bool F32 = std::is_same<typename real_of<E>::type, float>::value;
  const bool bf = F32 && pl.bf16x3;
  const size_t lds = bf ? gram_big_lds_bytes<IN, E, true>(pl.g.M, pl.g.pt)
                        : gram_big_lds_bytes<IN, E>(pl.g.M, pl.g.pt);
  if (lds > 160 * 1024) return fail(PDAMD_EUNSUPPORTED, "gram LDS %zu bytes exceeds 160 KiB", lds);
  if constexpr (F32) {
    constexpr bool CPX = is_cpx<E>::value;
    if (pl.n <= G1<CPX>::NMAX) {
      const size_t lds1 = bf ? gram_big1_lds_bytes<IN, E, true>(pl.g.M, pl.g.pt)
                             : gram_big1_lds_bytes<IN, E, false>(pl.g.M, pl.g.pt);
      if (lds1 <= 160 * 1024) {
        auto k1 = bf ? gram_big1_kernel<IN, E, TSIDE, true> : gram_big1_kernel<IN, E, TSIDE, false>;
        HIP_TRY(hipFuncSetAttribute((const void*)k1, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds1));
        const int per_cu = pick_per_cu(pl.per_cu[0], 2, lds1, (const void*)k1, kG1T);
        k1<<<grid_for(ga.np, per_cu), kG1T, lds1, st>>>(reinterpret_cast<const IN*>(vol), ga);
        HIP_TRY(hipGetLastError());
        return PDAMD_OK;
      }
    }
  }
  auto kern = bf ? gram_big_kernel<IN, E, TSIDE, F32> : gram_big_kernel<IN, E, TSIDE>;
  HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int per_cu = pick_per_cu(pl.per_cu[0], 4, lds, (const void*)kern, kNT);
  kern<<<grid_for(ga.np, per_cu), kNT, lds, st>>>(reinterpret_cast<const IN*>(vol), ga);
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
}

template <class E>
int launch_tridiag_big(const Plan& pl, GramArgs ga, hipStream_t st) {
  // two workgroups per CU (one's barriers overlap the other's work) when that
  // costs at most a few more global-memory steps
  int per_cu = 1, k0 = tridiag_big_k0(pl.n, sizeof(E));
  const int k0b = tridiag_big_k0(pl.n, sizeof(E), 2);
  // the global phase keeps three length-n vectors in the packed triangle's area
  auto room = [&](int kk) { const size_t m = (size_t)(pl.n - kk); return kk == 0 || m * (m + 1) / 2 >= 3 * (size_t)pl.n; };
  if (k0b <= k0 + std::max(2, pl.n / 32) && room(k0b)) {
    per_cu = 2;
    k0 = k0b;
  }
  if (!room(k0)) return fail(PDAMD_EUNSUPPORTED, "tridiagonalisation LDS too small for n = %d", pl.n);
  const size_t lds = tridiag_big_lds_bytes<E>(pl.n, k0);
  if (k0 > pl.n - 2 || (lds + 256) * per_cu > kTBLds)
    return fail(PDAMD_EUNSUPPORTED, "tridiagonalisation LDS %zu bytes (k0 %d)", lds, k0);
  // one workgroup per CU: 1024 threads (16 waves hide the per-step barriers);
  // two per CU: 512 threads each (the register file holds both)
  auto kern = per_cu == 1 ? tridiag_big_kernel<E, kTBT> : tridiag_big_kernel<E, kTBT / 2>;
  HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  kern<<<grid_for(ga.np, per_cu), per_cu == 1 ? kTBT : kTBT / 2, lds, st>>>(ga, k0);
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
}

template <class IN, class E, bool TSIDE>
int launch_gram_tile(const Plan& pl, const void* vol, GramArgs ga, hipStream_t st) {
  switch (pl.tile) {
    case 0: return launch_gram_big<IN, E, TSIDE>(pl, vol, ga, st);
    case 2: return launch_gram<IN, E, TSIDE, 2>(pl, vol, ga, st);
    case 4: return launch_gram<IN, E, TSIDE, 4>(pl, vol, ga, st);
    default:
      if constexpr (std::is_same<IN, float>::value && std::is_same<E, float>::value)
        if (pl.bf16x3) return launch_gram<IN, E, TSIDE, 8, true>(pl, vol, ga, st);
      return launch_gram<IN, E, TSIDE, 8>(pl, vol, ga, st);
  }
}

template <class E>
int launch_tridiag_tile(const Plan& pl, GramArgs ga, hipStream_t st) {
  // real Grams that fit one wavefront's registers: the one-wave kernel
  constexpr bool F64 = std::is_same<E, double>::value;
  if constexpr (std::is_same<E, float>::value || F64) {
    if (pl.tile && tridiag_wave_nb(pl.n, F64)) {
      TridiagWaveArgs ta;
      ta.H = ga.H;
      ta.DE = ga.DE;
      ta.TAU = ga.TAU;
      ta.np = ga.np;
      ta.n = pl.n;
      ta.np_dev = ga.np_dev;
      if (launch_tridiag_wave(ta, F64, st))
        return fail(PDAMD_EHIP, "tridiagonalisation launch failed: %s",
                    hipGetErrorString(hipGetLastError()));
      return PDAMD_OK;
    }
  }
  // Grams up to 208 (real FP32 above 128, real FP64 and complex FP32 above
  // 64): four wavefronts per patch, tiles in registers
  constexpr int WK = std::is_same<E, float>::value ? kWgF32
                     : F64                          ? kWgF64
                     : std::is_same<E, cpx<float>>::value ? kWgC32 : -1;
  if constexpr (WK >= 0) {
    if (tridiag_wg_supported(pl.n, WK)) {
      TridiagWaveArgs ta;
      ta.H = ga.H;
      ta.DE = ga.DE;
      ta.TAU = ga.TAU;
      ta.np = ga.np;
      ta.n = pl.n;
      ta.np_dev = ga.np_dev;
      if (launch_tridiag_wg(ta, WK, st))
        return fail(PDAMD_EHIP, "tridiagonalisation launch failed: %s",
                    hipGetErrorString(hipGetLastError()));
      return PDAMD_OK;
    }
  }
  switch (pl.tile) {
    case 0: return launch_tridiag_big<E>(pl, ga, st);
    case 2: return launch_tridiag<E, 8>(pl, ga, st);
    case 4: return launch_tridiag<E, 16>(pl, ga, st);
    default: return launch_tridiag<E, 32>(pl, ga, st);
  }
}

template <class E, class RT>
int launch_eig(const Plan& pl, EigArgs ea, hipStream_t st) {
  const bool ad = ea.method == PDAMD_ADAPTIVE;
  const size_t lds = trieig_lds_bytes<E>(pl.n, pl.kb, ad);
  if (lds > 160 * 1024) return fail(PDAMD_EUNSUPPORTED, "eigen LDS %zu bytes exceeds 160 KiB", lds);
  auto kern = pl.tile ? (ad ? trieig_kernel<E, RT, 2, kKB, true> : trieig_kernel<E, RT>)
              : pl.n <= 4 * kWave
                  ? (ad ? trieig_kernel<E, RT, 4, kKBBig, true> : trieig_kernel<E, RT, 4, kKBBig>)
              : pl.n <= kNMid
                  ? (ad ? trieig_kernel<E, RT, kNMid / kWave, kKBBig, true>
                        : trieig_kernel<E, RT, kNMid / kWave, kKBBig>)
                  : (ad ? trieig_kernel<E, RT, kNBig / kWave, 1, true>
                        : trieig_kernel<E, RT, kNBig / kWave, 1>);
  HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int per_cu = pick_per_cu(pl.per_cu[2], 16, lds, (const void*)kern, kWave);
  kern<<<grid_for(ea.np, per_cu), kWave, lds, st>>>(ea);
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
}

template <class R>
int launch_eig_all(const Plan& pl, EigArgs ea, hipStream_t st) {
  const size_t lds = trieig_all_lds_bytes<R>(pl.n);
  auto kern = trieig_all_kernel<R>;
  kern<<<grid_for(ea.np, 32), kWave, lds, st>>>(ea);
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
}

constexpr int kSynSCBig = 8, kSynKGBig = 4;  // synth chunking for Grams above 128
constexpr int kSynSCHuge = 2;                // ... and for Grams above kNMid
static_assert(kSynSC * 128 <= kSynNI * kNT && kSynSCBig * kNMid <= kSynNI * kNT &&
                  2 * kSynSCBig * 256 <= kSynNI * kNT && kSynSCHuge * kNBig <= kSynNI * kNT,
              "one staged synthesis chunk must fit kSynNI items per thread");

template <class IN, class E, bool TSIDE>
int launch_synth(const Plan& pl, const void* vol, SynthArgs sa, hipStream_t st) {
  const bool big = pl.tile == 0;
  // Grams of 129 .. 256: 16-row / -frame chunks and groups of 8 components
  // when they fit the LDS, else 8 and 4
  const size_t lds_mid = synth_lds_bytes<IN, E, TSIDE, 2 * kSynSCBig, 2 * kSynKGBig>(
      pl.n, pl.kmax, pl.g.M, pl.g.pt);
  const bool mid = big && pl.n <= 256 && lds_mid <= 160 * 1024;
  // Grams above kNMid: 2-row / -frame chunks, groups of 4 components (2 when
  // 4 overflow the LDS, as for complex float64 near the maximum size)
  const bool huge = big && pl.n > kNMid;
  const size_t lds_h4 = synth_lds_bytes<IN, E, TSIDE, kSynSCHuge, kSynKGBig>(
      pl.n, pl.kmax, pl.g.M, pl.g.pt);
  const bool h2 = huge && lds_h4 > 160 * 1024;
  const size_t lds = mid ? lds_mid
                     : h2 ? synth_lds_bytes<IN, E, TSIDE, kSynSCHuge, kSynKGBig / 2>(
                                pl.n, pl.kmax, pl.g.M, pl.g.pt)
                     : huge ? lds_h4
                     : big ? synth_lds_bytes<IN, E, TSIDE, kSynSCBig, kSynKGBig>(pl.n, pl.kmax, pl.g.M, pl.g.pt)
                           : synth_lds_bytes<IN, E, TSIDE>(pl.n, pl.kmax, pl.g.M, pl.g.pt);
  if (lds > 160 * 1024) return fail(PDAMD_EUNSUPPORTED, "synth LDS %zu bytes exceeds 160 KiB", lds);
  auto kern = mid ? synth_kernel<IN, E, TSIDE, 2 * kSynSCBig, 2 * kSynKGBig>
              : h2 ? synth_kernel<IN, E, TSIDE, kSynSCHuge, kSynKGBig / 2>
              : huge ? synth_kernel<IN, E, TSIDE, kSynSCHuge, kSynKGBig>
              : big ? synth_kernel<IN, E, TSIDE, kSynSCBig, kSynKGBig> : synth_kernel<IN, E, TSIDE>;
  HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int per_cu = pick_per_cu(pl.per_cu[3], 8, lds, (const void*)kern, kNT);
  kern<<<grid_for(sa.np, per_cu), kNT, lds, st>>>(reinterpret_cast<const IN*>(vol), sa);
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
}

EigArgs eig_args(const pdamd_problem* P, const Plan& pl, char* ws, const GramArgs& ga) {
  EigArgs ea;
  ea.g = pl.g;
  ea.idx = ga.idx;
  ea.np = ga.np;
  ea.n = pl.n;
  ea.kmax = pl.kmax;
  ea.kb = pl.kb;
  ea.method = P->method;
  ea.loss = P->loss;
  ea.rule_f64 = pl.rule_f64;
  ea.scale_f64 = P->scale_f64;
  ea.tau_f64 = P->tau_f64;
  ea.ts2 = P->threshold_scale2;
  ea.tau = P->tau;
  ea.mpm = P->mp_median;
  ea.beta = (double)pl.g.pt / (double)pl.g.M;
  ea.var_map = P->var_map_dev;
  ea.var_f32 = P->var_f32;
  ea.LAM = P->lam_dev;
  ea.adaptive_method = P->adaptive_method;
  ea.n_gamma = P->n_gamma;
  ea.gamma = P->gamma_dev;
  ea.tau0_log = P->tau0_log;
  ea.qut_tau = P->qut_tau_dev;
  ea.H = ga.H;
  ea.DE = ga.DE;
  ea.TAU = ga.TAU;
  ea.Y = ws + pl.off_Y;
  ea.Z = ws + pl.off_Z;
  ea.SCR = ws + pl.off_SCR;
  ea.WC = reinterpret_cast<double*>(ws + pl.off_WC);
  ea.NCOMP = reinterpret_cast<int*>(ws + pl.off_NC);
  ea.THETA = reinterpret_cast<double*>(ws + pl.off_TH);
  ea.RANK = nullptr;
  ea.NOISE = nullptr;
  ea.EIG = nullptr;
  ea.zero_fill = 0;
  ea.np_dev = nullptr;
  ea.refine_rel = 0.0;
  ea.NEAR = nullptr;
  return ea;
}

GramArgs gram_args(const pdamd_problem* P, const Plan& pl, char* ws, const int32_t* idx, int np) {
  GramArgs ga;
  ga.g = pl.g;
  ga.idx = idx;
  ga.np = np;
  ga.n = pl.n;
  ga.scale = (P->method == PDAMD_MPPCA || P->method == PDAMD_HYBRID) ? 1 : 0;
  ga.H = ws + pl.off_H;
  ga.DE = ws + pl.off_DE;
  ga.TAU = ws + pl.off_TAU;
  ga.MEAN = ws + pl.off_MEAN;
  ga.MEANW = nullptr;
  ga.np_dev = nullptr;
  return ga;
}

// ---------------------------------------------------------------------------
// FP64 re-check of near-threshold rule decisions (float32 data, FP32 eigen
// stage, MP-PCA / Hybrid).  The eigen kernel flags a patch when a comparison
// of its rule has both sides within refine_rel of each other; the flagged
// patches are listed in increasing slot order (at most rcap), recomputed with
// an FP64 Gram, tridiagonalisation and eigen stage (refine_f32_f64), and the
// FP64 result replaces the FP32 one in the buffers the synthesis reads.
// ---------------------------------------------------------------------------

// One 1024-thread workgroup: slots[q] / ridx[q] = the q-th flagged slot and
// its linear patch index; cnt[0] = listed (<= cap), cnt[1] = flagged.
__global__ __launch_bounds__(1024) void refine_compact_kernel(
    const uint8_t* __restrict__ flags, int nflags, const int32_t* __restrict__ idx, int cap,
    int* __restrict__ cnt, int32_t* __restrict__ slots, int32_t* __restrict__ ridx) {
  __shared__ int wsum[16];
  __shared__ int base_s;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
  if (tid == 0) base_s = 0;
  __syncthreads();
  for (int b0 = 0; b0 < nflags; b0 += 1024) {
    const int i = b0 + tid;
    const bool f = i < nflags && flags[i] != 0;
    const unsigned long long m = __ballot(f);
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int off = base_s;
    for (int j = 0; j < w; ++j) off += wsum[j];
    if (f) {
      const int q = off + __popcll(m & ((1ull << lane) - 1ull));
      if (q < cap) {
        slots[q] = i;
        ridx[q] = idx[i];
      }
    }
    __syncthreads();
    if (tid == 0) {
      int t = 0;
      for (int j = 0; j < 1024 / kWave; ++j) t += wsum[j];
      base_s += t;
    }
    __syncthreads();
  }
  if (tid == 0) {
    cnt[0] = min(base_s, cap);
    cnt[1] = base_s;
  }
}

struct RefineStore {
  const int* cnt;          // [0]: listed patches
  const int32_t* slots;    // slot of each listed patch in the destination arrays
  int n, kmax_src, kmax_dst;
  const double* Ysrc;      // per list entry: n x kmax_src FP64 eigenvectors
  float* Ydst;             // per slot: n x kmax_dst
  const int* NCsrc;
  int* NCdst;
  const double* THsrc;
  double* THdst;
  const double* WCsrc;
  double* WCdst;           // or null
  const int32_t* RKsrc;
  int32_t* RKdst;          // or null (and the following)
  const float* NZsrc;
  float* NZdst;
  const float* EGsrc;
  float* EGdst;
  int32_t* REFINED;        // 1 at each re-checked slot, or null
};

// One workgroup per list entry (blocks past the listed count exit).
__global__ __launch_bounds__(kNT) void refine_store_kernel(RefineStore a) {
  const int q = blockIdx.x;
  if (q >= a.cnt[0]) return;
  const int s = a.slots[q];
  const int tid = threadIdx.x;
  const int nc = a.NCsrc[q];
  const double* __restrict__ ys = a.Ysrc + (size_t)q * a.n * a.kmax_src;
  float* __restrict__ yd = a.Ydst + (size_t)s * a.n * a.kmax_dst;
  for (int e = tid; e < a.n * a.kmax_dst; e += kNT) {
    const int i = e / a.kmax_dst, c = e - i * a.kmax_dst;
    yd[e] = c < nc ? (float)ys[(size_t)i * a.kmax_src + c] : 0.f;
  }
  if (a.WCdst && tid < a.kmax_dst)
    a.WCdst[(size_t)s * a.kmax_dst + tid] = tid < nc ? a.WCsrc[(size_t)q * a.kmax_src + tid] : 0.0;
  if (a.EGdst && tid < 10) a.EGdst[(size_t)s * 10 + tid] = a.EGsrc[(size_t)q * 10 + tid];
  if (tid == 0) {
    a.NCdst[s] = nc;
    a.THdst[s] = a.THsrc[q];
    if (a.RKdst) a.RKdst[s] = a.RKsrc[q];
    if (a.NZdst) a.NZdst[s] = a.NZsrc[q];
    if (a.REFINED) a.REFINED[s] = 1;
  }
}

// List the flagged patches (flags / idx over `nflags` slots), re-check them
// in FP64 and store the results through `sto` (its slot-indexed pointers and
// kmax_dst filled in by the caller).
int run_refine(pdamd_problem* P, const Plan& pl, char* ws, const uint8_t* flags, int nflags,
               const int32_t* idx, RefineStore sto, hipStream_t st) {
  int* cnt = reinterpret_cast<int*>(ws + pl.off_RCNT);
  int32_t* slots = reinterpret_cast<int32_t*>(ws + pl.off_RSLOT);
  refine_compact_kernel<<<1, 1024, 0, st>>>(flags, nflags, idx, pl.rcap, cnt, slots,
                                            reinterpret_cast<int32_t*>(ws + pl.off_RIDX));
  HIP_TRY(hipGetLastError());
  int rc = refine_f32_f64(P, pl, ws, st);
  if (rc) return rc;
  sto.cnt = cnt;
  sto.slots = slots;
  sto.n = pl.n;
  sto.kmax_src = pl.kmax;
  sto.Ysrc = reinterpret_cast<const double*>(ws + pl.off_RY);
  sto.NCsrc = reinterpret_cast<const int*>(ws + pl.off_RNC);
  sto.THsrc = reinterpret_cast<const double*>(ws + pl.off_RTH);
  sto.WCsrc = reinterpret_cast<const double*>(ws + pl.off_RWC);
  sto.RKsrc = reinterpret_cast<const int32_t*>(ws + pl.off_RRK);
  sto.NZsrc = reinterpret_cast<const float*>(ws + pl.off_RNZ);
  sto.EGsrc = reinterpret_cast<const float*>(ws + pl.off_REG);
  refine_store_kernel<<<pl.rcap, kNT, 0, st>>>(sto);
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
}

#include "pd_adjoint.inl"

// Gather recombination (float32 M-side eigen rules): per chunk the Gram,
// Householder and eigen stages write the per-call means / eigenvectors /
// counts, the factor kernel the per-patch rows; then one pass builds the
// patch map, the float32 weight sums and every output element.
// Patches per chunk for `count` selected patches: the plan's maximum, evened
// out so that the last chunk is not a small remainder (every launch then
// fills the GPU alike).
inline int64_t even_chunk(const Plan& pl, int64_t count) {
  if (count <= 0) return pl.chunk;
  const int64_t nch = (count + pl.chunk - 1) / pl.chunk;
  return (count + nch - 1) / nch;
}

int run_chunks_gather(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, hipStream_t st,
                      StageTimer& tm);

template <class IN, class E, class RT>
int run_chunks(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, hipStream_t st,
               StageTimer& tm) {
  if (P->refined_dev && count > 0) {
    const int rz = zero_async(P->refined_dev, (size_t)count * sizeof(int32_t), st);
    if (rz) return rz;
  }
  if constexpr (std::is_same<IN, float>::value && std::is_same<E, float>::value)
    if (pl.gather) return run_chunks_gather(P, pl, ws, count, st, tm);
  const int32_t* idx_all = reinterpret_cast<const int32_t*>(ws + pl.off_idx);
  double* wacc = reinterpret_cast<double*>(ws + pl.off_w3);
  const bool w3d = pl.g.gt == 1;
  uint8_t* rflags = reinterpret_cast<uint8_t*>(ws + pl.off_RFLAG);
  const int64_t chunk = even_chunk(pl, count);
  for (int64_t c0 = 0; c0 < count; c0 += chunk) {
    const int np = (int)std::min<int64_t>(chunk, count - c0);
    GramArgs ga = gram_args(P, pl, ws, idx_all + c0, np);
    tm.begin(PDAMD_STAGE_GRAM);
    int rc = pl.tside ? launch_gram_tile<IN, E, true>(pl, P->input_dev, ga, st)
                      : launch_gram_tile<IN, E, false>(pl, P->input_dev, ga, st);
    tm.end();
    if (rc) return rc;
    tm.begin(PDAMD_STAGE_TRIDIAG);
    rc = launch_tridiag_tile<E>(pl, ga, st);
    tm.end();
    if (rc) return rc;
    EigArgs ea = eig_args(P, pl, ws, ga);
    ea.RANK = P->rank_dev ? P->rank_dev + c0 : nullptr;
    ea.NOISE = P->noise_dev ? P->noise_dev + c0 : nullptr;
    ea.EIG = P->eig_dev ? P->eig_dev + c0 * 10 : nullptr;
    if (pl.refine) {
      ea.refine_rel = P->refine_rel;
      ea.NEAR = rflags;
    }
    tm.begin(PDAMD_STAGE_EIG);
    rc = launch_eig<E, RT>(pl, ea, st);
    tm.end();
    if (rc) return rc;
    if constexpr (std::is_same<IN, float>::value && std::is_same<E, float>::value) {
      if (pl.refine) {  // the chunk's near-threshold patches, in FP64
        RefineStore sto{};
        sto.kmax_dst = pl.kmax;
        sto.Ydst = reinterpret_cast<float*>(ea.Y);
        sto.NCdst = ea.NCOMP;
        sto.THdst = ea.THETA;
        sto.WCdst = ea.WC;
        sto.RKdst = ea.RANK;
        sto.NZdst = ea.NOISE;
        sto.EGdst = ea.EIG;
        sto.REFINED = P->refined_dev ? P->refined_dev + c0 : nullptr;
        tm.begin(PDAMD_STAGE_REFINE);
        rc = run_refine(P, pl, ws, rflags, np, idx_all + c0, sto, st);
        tm.end();
        if (rc) return rc;
      }
    }
    SynthArgs sa;
    sa.g = pl.g;
    sa.idx = ga.idx;
    sa.np = np;
    sa.n = pl.n;
    sa.kmax = pl.kmax;
    sa.recomb = P->recombination;
    sa.w3d = w3d ? 1 : 0;
    sa.Y = ea.Y;
    sa.WC = ea.WC;
    sa.NCOMP = ea.NCOMP;
    sa.THETA = ea.THETA;
    sa.MEAN = ga.MEAN;
    sa.out = P->output_dev;
    sa.w = wacc;
    tm.begin(PDAMD_STAGE_SYNTH);
    rc = pl.tside ? launch_synth<IN, E, true>(pl, P->input_dev, sa, st)
                  : launch_synth<IN, E, false>(pl, P->input_dev, sa, st);
    tm.end();
    if (rc) return rc;
  }
  const int64_t total = pl.nvox * pl.g.T;
  const int nb = (int)std::min<int64_t>((total + 255) / 256, 65536);
  tm.begin(PDAMD_STAGE_FINALIZE);
  finalize_kernel<IN><<<nb, 256, 0, st>>>(reinterpret_cast<IN*>(P->output_dev), P->weights_dev,
                                          wacc, w3d ? 1 : 0, P->mask_dev, P->mask_kind,
                                          P->recombination, pl.nvox, pl.g.T);
  tm.end();
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
}

#ifdef PD_RUN_GATHER
int run_chunks_gather(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, hipStream_t st,
                      StageTimer& tm) {
  const int32_t* idx_all = reinterpret_cast<const int32_t*>(ws + pl.off_idx);
  const int T = pl.g.pt;
  GatherArgs fa;
  fa.g = pl.g;
  fa.shape[0] = P->shape[0];
  fa.shape[1] = P->shape[1];
  fa.shape[2] = P->shape[2];
  const bool ts = pl.tside;
  fa.n = pl.g.M;
  fa.np = 0;
  fa.slot0 = 0;
  fa.tside = ts ? 1 : 0;
  fa.YT = ts ? reinterpret_cast<float*>(ws + pl.off_GY) : nullptr;
  fa.V = ts ? reinterpret_cast<const float*>(ws + pl.off_GV) : nullptr;
  fa.recomb = P->recombination;
  fa.theta_weak = P->method == PDAMD_HYBRID ? 1 : 0;
  fa.idx = idx_all;
  fa.vol = reinterpret_cast<const float*>(P->input_dev);
  fa.MEAN = reinterpret_cast<const float*>(ws + pl.off_GMEAN);
  fa.Y = reinterpret_cast<const float*>(ws + pl.off_GY);
  fa.WC = nullptr;
  fa.NCOMP = reinterpret_cast<const int*>(ws + pl.off_GNC);
  fa.THETA = reinterpret_cast<const double*>(ws + pl.off_GTH);
  fa.RF = reinterpret_cast<float*>(ws + pl.off_GRF);
  fa.MAP = reinterpret_cast<int32_t*>(ws + pl.off_GMAP);
  fa.W3 = reinterpret_cast<float*>(ws + pl.off_GW3);
  fa.mask = P->mask_dev;
  fa.mask_kind = P->mask_kind;
  fa.out = reinterpret_cast<float*>(P->output_dev);
  fa.wout = P->weights_dev;
  const int64_t chunk = even_chunk(pl, count);
  for (int64_t c0 = 0; c0 < count; c0 += chunk) {
    const int np = (int)std::min<int64_t>(chunk, count - c0);
    GramArgs ga = gram_args(P, pl, ws, idx_all + c0, np);
    // the per-frame means straight into the per-call buffer (T-side Grams
    // write them as the synthesis input, M-side ones as the gather's)
    void* mean_c = ws + pl.off_GMEAN + (size_t)c0 * T * sizeof(float);
    if (ts) ga.MEAN = mean_c;
    else ga.MEANW = mean_c;
    tm.begin(PDAMD_STAGE_GRAM);
    int rc = ts ? launch_gram_tile<float, float, true>(pl, P->input_dev, ga, st)
                : launch_gram_tile<float, float, false>(pl, P->input_dev, ga, st);
    tm.end();
    if (rc) return rc;
    tm.begin(PDAMD_STAGE_TRIDIAG);
    rc = launch_tridiag_tile<float>(pl, ga, st);
    tm.end();
    if (rc) return rc;
    EigArgs ea = eig_args(P, pl, ws, ga);
    ea.Y = ws + (ts ? pl.off_GV : pl.off_GY) + (size_t)c0 * pl.n * kGatherKM * sizeof(float);
    ea.NCOMP = reinterpret_cast<int*>(ws + pl.off_GNC) + c0;
    ea.THETA = reinterpret_cast<double*>(ws + pl.off_GTH) + c0;
    ea.zero_fill = 1;
    ea.RANK = P->rank_dev ? P->rank_dev + c0 : nullptr;
    ea.NOISE = P->noise_dev ? P->noise_dev + c0 : nullptr;
    ea.EIG = P->eig_dev ? P->eig_dev + c0 * 10 : nullptr;
    if (pl.refine) {
      ea.refine_rel = P->refine_rel;
      ea.NEAR = reinterpret_cast<uint8_t*>(ws + pl.off_RFLAG) + c0;
    }
    tm.begin(PDAMD_STAGE_EIG);
    rc = launch_eig<float, float>(pl, ea, st);
    tm.end();
    if (rc) return rc;
    fa.np = np;
    fa.slot0 = (int)c0;
    fa.WC = ea.WC;
    tm.begin(PDAMD_STAGE_SYNTH);
    rc = launch_gather_factors(fa, st);
    tm.end();
    if (rc) return fail(PDAMD_EHIP, "factor kernel launch failed");
  }
  if (pl.refine && count > 0) {
    // the call's near-threshold patches in FP64, then their factors again
    RefineStore sto{};
    sto.kmax_dst = kGatherKM;
    sto.Ydst = reinterpret_cast<float*>(ws + (ts ? pl.off_GV : pl.off_GY));
    sto.NCdst = reinterpret_cast<int*>(ws + pl.off_GNC);
    sto.THdst = reinterpret_cast<double*>(ws + pl.off_GTH);
    sto.WCdst = nullptr;
    sto.RKdst = P->rank_dev;
    sto.NZdst = P->noise_dev;
    sto.EGdst = P->eig_dev;
    sto.REFINED = P->refined_dev;
    tm.begin(PDAMD_STAGE_REFINE);
    int rc = run_refine(P, pl, ws, reinterpret_cast<const uint8_t*>(ws + pl.off_RFLAG),
                        (int)count, idx_all, sto, st);
    if (!rc) {
      GatherArgs fl = fa;
      fl.list = reinterpret_cast<const int32_t*>(ws + pl.off_RSLOT);
      fl.list_cnt = reinterpret_cast<const int*>(ws + pl.off_RCNT);
      fl.list_cap = pl.rcap;
      if (launch_gather_factors(fl, st)) rc = fail(PDAMD_EHIP, "factor kernel launch failed");
    }
    tm.end();
    if (rc) return rc;
  }
  tm.begin(PDAMD_STAGE_FINALIZE);
  int rc = launch_gather_map(fa, pl.npatch, count, st);
  if (!rc) rc = launch_gather_weights(fa, st);
  tm.end();
  if (!rc) {
    tm.begin(PDAMD_STAGE_RECOMBINE);
    rc = launch_gather_recombine(fa, st);
    tm.end();
  }
  if (rc) return fail(PDAMD_EHIP, "gather recombination launch failed");
  return PDAMD_OK;
}

#endif  // PD_RUN_GATHER

template <class IN, class E, class RT>
int run_backward(pdamd_problem* P, const Plan& pl, char* ws, int64_t count, const void* gout,
                 const float* w4, void* gin, double* glam, hipStream_t st, StageTimer& tm) {
  const int32_t* idx_all = reinterpret_cast<const int32_t*>(ws + pl.off_idx);
  {
    AdjArgs ha{};
    ha.g = pl.g;
    ha.recomb = P->recombination;
    ha.mask_kind = P->mask_kind;
    ha.mask = P->mask_dev;
    ha.gout = gout;
    ha.w4 = w4;
    ha.GH = ws + pl.off_GH;
    const int64_t total = pl.nvox * pl.g.T;
    tm.begin(PDAMD_STAGE_ADJ_CROSS);
    ghat_kernel<IN, E><<<(int)std::min<int64_t>((total + 255) / 256, 65536), 256, 0, st>>>(ha, total);
    tm.end();
    HIP_TRY(hipGetLastError());
  }
  const int64_t chunk = even_chunk(pl, count);
  for (int64_t c0 = 0; c0 < count; c0 += chunk) {
    const int np = (int)std::min<int64_t>(chunk, count - c0);
    // recompute the Gram eigendecomposition (all pairs)
    GramArgs ga = gram_args(P, pl, ws, idx_all + c0, np);
    tm.begin(PDAMD_STAGE_GRAM);
    int rc = launch_gram_tile<IN, E, true>(pl, P->input_dev, ga, st);
    tm.end();
    if (rc) return rc;
    tm.begin(PDAMD_STAGE_TRIDIAG);
    rc = launch_tridiag_tile<E>(pl, ga, st);
    tm.end();
    if (rc) return rc;
    EigArgs ea = eig_args(P, pl, ws, ga);
    ea.var_map = nullptr;
    ea.LAM = nullptr;
    tm.begin(PDAMD_STAGE_EIG);
    rc = launch_eig_all<E>(pl, ea, st);
    if (!rc) {
      BtArgs ba;
      ba.np = np;
      ba.n = pl.n;
      ba.H = ga.H;
      ba.TAU = ga.TAU;
      ba.Z = ea.Z;
      ba.Y = ea.Y;
      rc = launch_backtransform<E>(pl, ba, st);
    }
    tm.end();
    if (rc) return rc;
    AdjArgs aa;
    aa.g = pl.g;
    aa.idx = ga.idx;
    aa.np = np;
    aa.n = pl.n;
    aa.recomb = P->recombination;
    aa.mask_kind = P->mask_kind;
    aa.mask = P->mask_dev;
    aa.gout = gout;
    aa.w4 = w4;
    aa.GH = ws + pl.off_GH;
    aa.Y = ea.Y;
    aa.EV = ea.WC;
    aa.P1 = ws + pl.off_P1;
    aa.P2 = ws + pl.off_P2;
    aa.PE = ws + pl.off_PE;
    aa.GS = ws + pl.off_GS;
    aa.S1 = ws + pl.off_S1;
    aa.UV = ws + pl.off_UV;
    aa.THB = reinterpret_cast<double*>(ws + pl.off_THB);
    aa.loss = P->loss;
    aa.mpm = P->mp_median;
    aa.beta = (double)pl.g.pt / (double)pl.g.M;
    aa.var_map = P->var_map_dev;
    aa.var_f32 = P->var_f32;
    aa.LAM = P->lam_dev;
    aa.GLAM = glam;
    aa.gin = gin;
    tm.begin(PDAMD_STAGE_ADJ_CROSS);
    rc = launch_adj_patch_kernel<IN, E>(pl, adj_project_kernel<IN, E>, P->input_dev, aa, st);
    tm.end();
    if (rc) return rc;
    tm.begin(PDAMD_STAGE_ADJ_SPECTRAL);
    rc = launch_adj_spectral<E, RT>(pl, aa, st);
    tm.end();
    if (rc) return rc;
    tm.begin(PDAMD_STAGE_ADJ_SYNTH);
    rc = launch_adj_patch_kernel<IN, E>(pl, adj_synth_kernel<IN, E>, P->input_dev, aa, st);
    tm.end();
    if (rc) return rc;
  }
  return PDAMD_OK;
}

}  // namespace
