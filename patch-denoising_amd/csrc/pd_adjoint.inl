// Backward of optimal singular-value shrinkage with a per-patch threshold
// scale lambda (included by pd_kernels.hip, inside its anonymous namespace).
//
// Forward, per selected patch (T-side: the centred Casorati matrix Xc is
// M x T with T < M; Xc = U S V^T, n = T):
//   P = U g(S) V^T + 1 m^T,  g_i = c f(s_i / c),  c = lambda * c0,
//   c0 = sqrt(T) * sigma_noise (sigma_noise = median(s) / sqrt(T * mp_median)
//   or the patch mean of the prior standard deviation), f the loss's shrinker.
// Backward, with Gbar = dL/dP (adjoint of the recombination), H = Xc^T Gbar,
// C = U^T Gbar V = S^-1 V^T H V and h = g / s:
//   Q_ij = Gp_ij (C_ij + C_ji) / 2 + Gm_ij (C_ij - C_ji) / 2,
//     Gp_ij = (g_i - g_j) / (s_i - s_j)   (guarded: (f'_i + f'_j) / 2 for
//             |s_i - s_j| <= 1e-8 max(s); Gp_ii = f'(s_i / c)),
//     Gm_ij = (g_i + g_j) / (s_i + s_j),
//   dL/dXc = Xc R1 + Gbar R2,  R1 = V S^-1 (Q - C diag(h)) V^T,  R2 = V diag(h) V^T,
//   cbar = sum_i C_ii (f(y_i) - y_i f'(y_i)),  y = s / c,
//   dL/dlambda = cbar * c0, and when c0 comes from the median singular value
//   Q_mm += cbar * lambda / sqrt(mp_median) at the median index (half to each
//   of the two middle values for an even count).
//   Centring adjoint: dL/dX = dL/dXc - 1 (1^T dL/dXc) / M + 1 (Gbar^T 1)^T / M
//                           = Xc R1 + Gbar R2 + 1 u^T / M,  u = (I - R2) Gbar^T 1.
// U and S are recomputed (Gram -> tridiagonal -> all eigenpairs), never stored
// across the forward/backward boundary.

// ---------------------------------------------------------------------------
// Shrinker value and derivative (lowrank.py:250-280 restated; derivatives of
// the closed forms, zero where the shrinker is gated or clamped).
// ---------------------------------------------------------------------------
__device__ void shrink_fd(int loss, double y, double beta, double* f, double* fp) {
  const double t = y * y - beta - 1.0;
  if (loss == PDAMD_LOSS_FRO) {
    const double Z = t * t - 4.0 * beta;
    if (!(Z > 0.0) || !(y > 0.0)) {
      *f = 0.0;
      *fp = 0.0;
      return;
    }
    const double fv = sqrt(Z / y);
    *f = fv;
    *fp = (4.0 * t * y * y - Z) / (2.0 * fv * y * y);
    return;
  }
  // x(y) above the upper bulk edge, 0 below
  const double D2 = t * t - 4.0 * beta;
  if (!(y >= 1.0 + sqrt(beta)) || !(D2 > 0.0)) {
    *f = 0.0;
    *fp = 0.0;
    return;
  }
  const double D = sqrt(D2);
  const double x = sqrt(0.5 * (t + D));
  const double xp = x * y / D;
  if (loss == PDAMD_LOSS_OPE) {
    *f = x;
    *fp = xp;
    return;
  }
  const double sb = sqrt(beta);
  const double N = x * x * x * x - sb * x * y - beta;
  if (!(N > 0.0)) {
    *f = 0.0;
    *fp = 0.0;
    return;
  }
  const double Dn = x * x * y;
  const double Np = 4.0 * x * x * x * xp - sb * (xp * y + x);
  const double Dp = 2.0 * x * xp * y + x * x;
  *f = N / Dn;
  *fp = (Np * Dn - N * Dp) / (Dn * Dn);
}

// ---------------------------------------------------------------------------
// Block GEMM for one patch: C (m x nn, nn <= 128) = A (m x k) B (k x nn),
// 256 threads, each owning an 8 x 8 register tile of a 128 x 128 block;
// k streamed in panels of kGP through LDS.  Operands come from functors
// la(i, kk) / lb(kk, j); AKF / BKF say whether consecutive k of an operand are
// contiguous in memory (sets the staging order for coalesced loads).
// st(i, j, value) consumes the result.
// ---------------------------------------------------------------------------
constexpr int kGP = 8;
constexpr int kGLD = 128 + 4;

template <class R>
size_t gemm_lds_bytes() {
  return 2 * (size_t)kGP * kGLD * sizeof(R);
}

template <class R, bool AKF, bool BKF, class LA, class LB, class ST>
__device__ __forceinline__ void tile_gemm(int m, int nn, int k, LA la, LB lb, ST st, R* As,
                                          R* Bs) {
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  for (int m0 = 0; m0 < m; m0 += 128) {
    const int mb = min(128, m - m0);
    R acc[8][8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int v = 0; v < 8; ++v) acc[u][v] = R(0);
    for (int k0 = 0; k0 < k; k0 += kGP) {
      const int kc = min(kGP, k - k0);
      __syncthreads();
      for (int e = tid; e < kGP * 128; e += kNT) {
        const int kk = AKF ? e % kGP : e / 128, i = AKF ? e / kGP : e % 128;
        As[kk * kGLD + i] = (i < mb && kk < kc) ? la(m0 + i, k0 + kk) : R(0);
      }
      for (int e = tid; e < kGP * 128; e += kNT) {
        const int kk = BKF ? e % kGP : e / 128, j = BKF ? e / kGP : e % 128;
        Bs[kk * kGLD + j] = (j < nn && kk < kc) ? lb(k0 + kk, j) : R(0);
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < kGP; ++kk) {
        R a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          a[u] = As[kk * kGLD + tr * 8 + u];
          b[u] = Bs[kk * kGLD + tc * 8 + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int v = 0; v < 8; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int i = tr * 8 + u, j = tc * 8 + v;
        if (i < mb && j < nn) st(m0 + i, j, acc[u][v]);
      }
  }
  __syncthreads();
}

struct AdjArgs {
  Geom g;
  const int32_t* idx;
  int np, n;
  int recomb;
  int mask_kind;
  const uint8_t* mask;
  const void* gout;     // dL/d(output), input dtype, nvox x T
  const float* w4;      // forward weights, nvox x T
  void* HC;             // np x n x n   Xc^T Ghat
  void* GS;             // np x n       column sums of Ghat
  const void* Y;        // np x n x n   Gram eigenvectors V[i][c], ascending eigenvalues
  const double* EV;     // np x n       Gram eigenvalues (= s^2), ascending
  void* S1;             // np x n x n   scratch
  void* S2;             // np x n x n   scratch
  void* R1;             // np x n x n
  void* R2;             // np x n x n   theta * V diag(h) V^T
  void* UV;             // np x n       u
  double* THB;          // np           theta
  int loss;
  double mpm, beta;
  const double* var_map;
  const double* LAM;     // per linear patch index, or null (= 1)
  double* GLAM;          // per linear patch index, or null
  void* gin;            // dL/d(input), input dtype, nvox x T
};

// Ghat = dL/dP / theta at volume element q (voxel-major element index): the
// output gradient divided by the forward weights (weighted / average) or the
// gradient itself at the patch centre (center); zero where the output was
// masked or never written.
template <class IN, class E>
__device__ __forceinline__ E ghat_at(const AdjArgs& a, int64_t q, bool centre_elem) {
  using R = typename real_of<E>::type;
  const IN* go = reinterpret_cast<const IN*>(a.gout);
  const int64_t v = q / a.g.T;
  bool keep = true;
  if (a.mask_kind == PDAMD_MASK_SPATIAL) keep = a.mask[v] != 0;
  else if (a.mask_kind == PDAMD_MASK_FULL) keep = a.mask[q] != 0;
  if (!keep) return E(R(0));
  if (a.recomb == PDAMD_CENTER) return centre_elem ? cvt<E>(go[q]) : E(R(0));
  const float w = a.w4[q];
  if (!(w > 0.f)) return E(R(0));
  return cvt<E>(go[q]) / E(R(w));
}

// Stage the patch's row offsets and NumPy-order per-frame mean in LDS.
template <class IN>
__device__ __forceinline__ void stage_rows_mean(const Geom& g, const IN* X, int64_t* roff,
                                                IN* mean) {
  const int M = g.M, T = g.pt;
  for (int r = threadIdx.x; r < M; r += kNT) roff[r] = row_offset(g, r);
  __syncthreads();
  for (int t = threadIdx.x; t < T; t += kNT) {
    IN s = X[roff[0] + t];
    for (int r = 1; r < M; ++r) s += X[roff[r] + t];  // sequential, as numpy's axis-0 sum
    mean[t] = div_n(s, M);
  }
  __syncthreads();
}

template <class IN, class E>
size_t adj_patch_lds_bytes(int M, int T) {
  using R = typename real_of<E>::type;
  auto al = [](size_t b) { return ((b + 15) / 16) * 16; };
  return gemm_lds_bytes<R>() + al((size_t)M * sizeof(int64_t)) + al((size_t)T * sizeof(IN)) + 16;
}

// H = Xc^T Ghat (n x n) and the column sums of Ghat, per patch.
template <class IN, class E>
__global__ __launch_bounds__(kNT) void adj_cross_kernel(const IN* __restrict__ vol, AdjArgs a) {
  using R = typename real_of<E>::type;
  const Geom& g = a.g;
  const int n = a.n, M = g.M, T = g.pt;
  extern __shared__ __align__(16) char smem[];
  char* cur = smem;
  R* As = lds_carve<R>(cur, (size_t)kGP * kGLD);
  R* Bs = lds_carve<R>(cur, (size_t)kGP * kGLD);
  int64_t* roff = lds_carve<int64_t>(cur, M);
  IN* mean = lds_carve<IN>(cur, T);
  const int rc = ((g.p0 / 2) * g.p1 + g.p1 / 2) * g.p2 + g.p2 / 2, tcen = g.pt / 2;
  for (int p = blockIdx.x; p < a.np; p += gridDim.x) {
    int64_t sp;
    int t0;
    patch_origin(g, a.idx[p], sp, t0);
    const IN* X = vol + sp + t0;
    stage_rows_mean(g, X, roff, mean);
    const int64_t base = sp + t0;
    R* H = reinterpret_cast<R*>(a.HC) + (size_t)p * n * n;
    R* GS = reinterpret_cast<R*>(a.GS) + (size_t)p * n;
    tile_gemm<R, false, false>(
        n, n, M, [&](int i, int kk) { return cvt<E>(X[roff[kk] + i] - mean[i]); },
        [&](int kk, int j) {
          return ghat_at<IN, E>(a, base + roff[kk] + j, kk == rc && j == tcen);
        },
        [&](int i, int j, R v) { H[(size_t)i * n + j] = v; }, As, Bs);
    for (int j = threadIdx.x; j < n; j += kNT) {
      R s = R(0);
      for (int r = 0; r < M; ++r) s += ghat_at<IN, E>(a, base + roff[r] + j, r == rc && j == tcen);
      GS[j] = s;
    }
    __syncthreads();
  }
}

// Spectral part of the adjoint, per patch: theta, dL/dlambda, R1, R2, u.
template <class E, class RT>
__global__ __launch_bounds__(kNT) void adj_spectral_kernel(AdjArgs a) {
  using R = typename real_of<E>::type;
  const Geom& g = a.g;
  const int n = a.n, tid = threadIdx.x;
  extern __shared__ __align__(16) char smem[];
  char* cur = smem;
  R* As = lds_carve<R>(cur, (size_t)kGP * kGLD);
  R* Bs = lds_carve<R>(cur, (size_t)kGP * kGLD);
  double* sg = lds_carve<double>(cur, n);   // s_i
  double* gv = lds_carve<double>(cur, n);   // g_i
  double* fp = lds_carve<double>(cur, n);   // f'(y_i) = dg_i / ds_i
  double* dc = lds_carve<double>(cur, n);   // dg_i / dc
  double* hv = lds_carve<double>(cur, n);   // g_i / s_i
  double* red = lds_carve<double>(cur, 8);
  __shared__ int sh_min;
  __shared__ double sh_s[4];

  for (int p = blockIdx.x; p < a.np; p += gridDim.x) {
    const int64_t L = a.idx[p];
    const double* ev = a.EV + (size_t)p * n;
    const R* V = reinterpret_cast<const R*>(a.Y) + (size_t)p * n * n;
    const R* H = reinterpret_cast<const R*>(a.HC) + (size_t)p * n * n;
    const R* GS = reinterpret_cast<const R*>(a.GS) + (size_t)p * n;
    R* S1 = reinterpret_cast<R*>(a.S1) + (size_t)p * n * n;
    R* S2 = reinterpret_cast<R*>(a.S2) + (size_t)p * n * n;
    R* R1 = reinterpret_cast<R*>(a.R1) + (size_t)p * n * n;
    R* R2 = reinterpret_cast<R*>(a.R2) + (size_t)p * n * n;
    R* UV = reinterpret_cast<R*>(a.UV) + (size_t)p * n;
    // ---- noise scale c (same arithmetic as the forward rule)
    if (tid < kWave) {
      double sigma;
      if (a.var_map) {
        int64_t sp;
        int t0;
        patch_origin(g, L, sp, t0);
        double acc = 0.0;
        for (int r = tid; r < g.M; r += kWave) acc += sqrt(a.var_map[(sp + row_offset(g, r)) / g.T]);
        sigma = dpp_sum(acc) / (double)g.M;
      } else {
        RT med;
        if (n % 2) {
          med = (RT)sqrt(fmax(ev[(n - 1) / 2], 0.0));
        } else {
          const RT s0 = (RT)sqrt(fmax(ev[n / 2 - 1], 0.0));
          const RT s1 = (RT)sqrt(fmax(ev[n / 2], 0.0));
          med = (RT)(s0 + s1) / RT(2);
        }
        sigma = (double)med / sqrt((double)g.pt * a.mpm);
      }
      if (tid == 0) {
        const double c0 = sqrt((double)g.pt) * sigma;
        const double lam = a.LAM ? a.LAM[L] : 1.0;
        sh_s[0] = c0;
        sh_s[1] = lam;
        sh_min = n;
      }
    }
    __syncthreads();
    const double c0 = sh_s[0], lam = sh_s[1], cs = lam * c0;
    // ---- per-component shrinkage, derivatives, and the rank of the forward
    for (int i = tid; i < n; i += kNT) {
      const double s = sqrt(fmax(ev[i], 0.0));
      const RT sv = (RT)s;
      double f = 0.0, fd = 0.0;
      double sh = cs * shrink(a.loss, (double)sv / cs, a.beta);
      if (isnan(sh)) sh = 0.0;
      if (sh != 0.0 && isfinite(cs) && cs > 0) {
        shrink_fd(a.loss, s / cs, a.beta, &f, &fd);
        atomicMin(&sh_min, i);
      }
      sg[i] = s;
      gv[i] = sh != 0.0 ? cs * f : 0.0;
      fp[i] = sh != 0.0 ? fd : 0.0;
      dc[i] = sh != 0.0 ? (f - (s / cs) * fd) : 0.0;
      hv[i] = (sh != 0.0 && s > 0.0) ? cs * f / s : 0.0;
    }
    __syncthreads();
    const int rank = n - sh_min;
    const double theta = a.recomb == PDAMD_WEIGHTED ? 1.0 / (2.0 + (double)rank) : 1.0;
    const double smax = sg[n - 1];
    // ---- W = V^T (H V)
    tile_gemm<R, true, false>(
        n, n, n, [&](int i, int kk) { return H[(size_t)i * n + kk]; },
        [&](int kk, int j) { return V[(size_t)kk * n + j]; },
        [&](int i, int j, R v) { S1[(size_t)i * n + j] = v; }, As, Bs);
    tile_gemm<R, false, false>(
        n, n, n, [&](int i, int kk) { return V[(size_t)kk * n + i]; },
        [&](int kk, int j) { return S1[(size_t)kk * n + j]; },
        [&](int i, int j, R v) { S2[(size_t)i * n + j] = v; }, As, Bs);
    // ---- cbar = sum_i C_ii dg_i/dc, C = theta S^-1 W
    double part = 0.0;
    for (int i = tid; i < n; i += kNT)
      if (sg[i] > 0.0) part += theta * (double)S2[(size_t)i * n + i] / sg[i] * dc[i];
    part = dpp_sum(part);
    if ((tid % kWave) == 0) red[tid / kWave] = part;
    __syncthreads();
    const double cbar = (red[0] + red[1]) + (red[2] + red[3]);
    double dmed = 0.0;  // extra Q on the median diagonal(s)
    int m0 = -1, m1 = -1;
    if (!a.var_map && a.mpm > 0) {
      dmed = cbar * lam / sqrt(a.mpm);
      if (n % 2) {
        m0 = (n - 1) / 2;
      } else {
        m0 = n / 2 - 1;
        m1 = n / 2;
        dmed *= 0.5;
      }
    }
    // ---- D = S^-1 (Q - C diag(h)) into S1
    const double tie = 1e-8 * smax;
    for (int e = tid; e < n * n; e += kNT) {
      const int i = e / n, j = e % n;
      const double si = sg[i], sj = sg[j];
      const double ci = si > 0.0 ? theta * (double)S2[(size_t)i * n + j] / si : 0.0;
      const double cj = sj > 0.0 ? theta * (double)S2[(size_t)j * n + i] / sj : 0.0;
      double q;
      if (i == j) {
        q = fp[i] * ci;
        if (i == m0 || i == m1) q += dmed;
      } else {
        const double gp = fabs(si - sj) > tie ? (gv[i] - gv[j]) / (si - sj) : 0.5 * (fp[i] + fp[j]);
        const double gm = (si + sj) > 0.0 ? (gv[i] + gv[j]) / (si + sj) : 0.0;
        q = gp * 0.5 * (ci + cj) + gm * 0.5 * (ci - cj);
      }
      S1[(size_t)i * n + j] = si > 0.0 ? (R)((q - ci * hv[j]) / si) : R(0);
    }
    __syncthreads();
    // ---- R1 = V D V^T, R2 = theta V diag(h) V^T
    tile_gemm<R, true, false>(
        n, n, n, [&](int i, int kk) { return V[(size_t)i * n + kk]; },
        [&](int kk, int j) { return S1[(size_t)kk * n + j]; },
        [&](int i, int j, R v) { S2[(size_t)i * n + j] = v; }, As, Bs);
    tile_gemm<R, true, true>(
        n, n, n, [&](int i, int kk) { return S2[(size_t)i * n + kk]; },
        [&](int kk, int j) { return V[(size_t)j * n + kk]; },
        [&](int i, int j, R v) { R1[(size_t)i * n + j] = v; }, As, Bs);
    tile_gemm<R, true, true>(
        n, n, n, [&](int i, int kk) { return (R)(theta * hv[kk]) * V[(size_t)i * n + kk]; },
        [&](int kk, int j) { return V[(size_t)j * n + kk]; },
        [&](int i, int j, R v) { R2[(size_t)i * n + j] = v; }, As, Bs);
    // ---- u = theta * gs - R2 gs
    for (int j = tid; j < n; j += kNT) {
      R s = R(0);
      for (int t = 0; t < n; ++t) s += R2[(size_t)j * n + t] * GS[t];
      UV[j] = (R)theta * GS[j] - s;
    }
    if (tid == 0) {
      a.THB[p] = theta;
      if (a.GLAM) a.GLAM[L] = cbar * c0;
    }
    __syncthreads();
  }
}

// dL/dX = Xc R1 + Ghat R2 + 1 u^T / M, scatter-added into dL/d(input).
template <class IN, class E>
__global__ __launch_bounds__(kNT) void adj_synth_kernel(const IN* __restrict__ vol, AdjArgs a) {
  using R = typename real_of<E>::type;
  const Geom& g = a.g;
  const int n = a.n, M = g.M, T = g.pt;
  extern __shared__ __align__(16) char smem[];
  char* cur = smem;
  R* As = lds_carve<R>(cur, (size_t)kGP * kGLD);
  R* Bs = lds_carve<R>(cur, (size_t)kGP * kGLD);
  int64_t* roff = lds_carve<int64_t>(cur, M);
  IN* mean = lds_carve<IN>(cur, T);
  const int rc = ((g.p0 / 2) * g.p1 + g.p1 / 2) * g.p2 + g.p2 / 2, tcen = g.pt / 2;
  IN* gin = reinterpret_cast<IN*>(a.gin);
  for (int p = blockIdx.x; p < a.np; p += gridDim.x) {
    int64_t sp;
    int t0;
    patch_origin(g, a.idx[p], sp, t0);
    const IN* X = vol + sp + t0;
    stage_rows_mean(g, X, roff, mean);
    const int64_t base = sp + t0;
    const R* R1 = reinterpret_cast<const R*>(a.R1) + (size_t)p * n * n;
    const R* R2 = reinterpret_cast<const R*>(a.R2) + (size_t)p * n * n;
    const R* UV = reinterpret_cast<const R*>(a.UV) + (size_t)p * n;
    const R invM = R(1) / (R)M;
    tile_gemm<R, true, false>(
        M, n, 2 * n,
        [&](int i, int kk) -> R {
          if (kk < n) return cvt<E>(X[roff[i] + kk] - mean[kk]);
          const int t = kk - n;
          return ghat_at<IN, E>(a, base + roff[i] + t, i == rc && t == tcen);
        },
        [&](int kk, int j) {
          return kk < n ? R1[(size_t)kk * n + j] : R2[(size_t)(kk - n) * n + j];
        },
        [&](int i, int j, R v) { atomic_add_(gin + base + roff[i] + j, (IN)(v + UV[j] * invM)); },
        As, Bs);
  }
}

template <class IN, class E>
int launch_adj_patch_kernel(const Plan& pl, void (*kern)(const IN*, AdjArgs), const void* vol,
                            const AdjArgs& aa, hipStream_t st) {
  const size_t lds = adj_patch_lds_bytes<IN, E>(pl.g.M, pl.g.pt);
  if (lds > 160 * 1024) return fail(PDAMD_EUNSUPPORTED, "adjoint LDS %zu bytes exceeds 160 KiB", lds);
  HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  kern<<<grid_for(aa.np, 2), kNT, lds, st>>>(reinterpret_cast<const IN*>(vol), aa);
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
}

template <class E, class RT>
int launch_adj_spectral(const Plan& pl, const AdjArgs& aa, hipStream_t st) {
  using R = typename real_of<E>::type;
  const size_t lds = gemm_lds_bytes<R>() + 5 * (((size_t)pl.n * 8 + 15) / 16) * 16 + 64;
  auto kern = adj_spectral_kernel<E, RT>;
  HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  kern<<<grid_for(aa.np, 2), kNT, lds, st>>>(aa);
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
}
