// Backward of optimal singular-value shrinkage with a per-patch threshold
// scale lambda (included by pd_kernels.inl, inside its anonymous namespace).
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
//          = (P1 D + P2 diag(h)) V^T  with P1 = Xc V, P2 = Gbar V, D = S^-1 (Q - C diag(h)),
//   and C = S^-1 P1^T P2 (so H is never formed),
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
// k streamed in panels of kGP through double-buffered LDS: the next panel's
// elements are loaded into registers while the current one is multiplied,
// one barrier per panel.  Operands come from functors la(i, kk) / lb(kk, j);
// AKF / BKF say whether consecutive k of an operand are contiguous in memory
// (sets the staging order for coalesced loads).  st(i, j, value) consumes
// the result.
// ---------------------------------------------------------------------------
constexpr int kGP = 8;
constexpr int kGLD = 128 + 4;
typedef double double4_t __attribute__((ext_vector_type(4)));
constexpr int kGSE = kGP * 128 / kNT;  // staged elements per thread and operand

template <class R>
size_t gemm_lds_bytes() {
  return 4 * (size_t)kGP * kGLD * sizeof(R);  // A and B, two buffers each
}

template <class R, bool AKF, bool BKF, class LA, class LB, class ST>
__device__ __forceinline__ void tile_gemm(int m, int nn, int k, LA la, LB lb, ST st, R* As,
                                          R* Bs) {
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  constexpr int BUF = kGP * kGLD;
  // FP64 on the matrix cores (v_mfma_f64_16x16x4f64), FP32 on the VALU
  constexpr bool MF = std::is_same<R, double>::value;
  const int wv = tid / kWave, lane = tid % kWave;
  const int wr = wv >> 1, wc = wv & 1, lhi = lane >> 4, llo = lane & 15;
  for (int m0 = 0; m0 < m; m0 += 128) {
    const int mb = min(128, m - m0);
    R acc[MF ? 1 : 8][MF ? 1 : 8];
    double4_t macc[4][4];
    if constexpr (MF) {
#pragma unroll
      for (int ta = 0; ta < 4; ++ta)
#pragma unroll
        for (int tb = 0; tb < 4; ++tb) macc[ta][tb] = double4_t{0.0, 0.0, 0.0, 0.0};
    } else {
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int v = 0; v < 8; ++v) acc[u][v] = R(0);
    }
    R ra[kGSE], rb[kGSE];
    auto fetch = [&](int k0) {
      const int kc = min(kGP, k - k0);
#pragma unroll
      for (int u = 0; u < kGSE; ++u) {
        const int e = tid + u * kNT;
        const int ka = AKF ? e % kGP : e / 128, ia = AKF ? e / kGP : e % 128;
        ra[u] = (ia < mb && ka < kc) ? la(m0 + ia, k0 + ka) : R(0);
        const int kb = BKF ? e % kGP : e / 128, jb = BKF ? e / kGP : e % 128;
        rb[u] = (jb < nn && kb < kc) ? lb(k0 + kb, jb) : R(0);
      }
    };
    auto stash = [&](int buf) {
#pragma unroll
      for (int u = 0; u < kGSE; ++u) {
        const int e = tid + u * kNT;
        const int ka = AKF ? e % kGP : e / 128, ia = AKF ? e / kGP : e % 128;
        As[buf * BUF + ka * kGLD + ia] = ra[u];
        const int kb = BKF ? e % kGP : e / 128, jb = BKF ? e / kGP : e % 128;
        Bs[buf * BUF + kb * kGLD + jb] = rb[u];
      }
    };
    __syncthreads();  // previous users of the buffers are done
    fetch(0);
    stash(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < k; k0 += kGP) {
      const bool more = k0 + kGP < k;
      if (more) fetch(k0 + kGP);
      const R* Ab = As + buf * BUF;
      const R* Bb = Bs + buf * BUF;
      if constexpr (MF) {
        // wave (wr, wc) owns rows 64 wr .. + 64 and columns 64 wc .. + 64:
        // 4 x 4 MFMA tiles of 16 x 16, k in steps of 4.  Lane l supplies
        // A[16 t + l % 16][k + l / 16] and B[k + l / 16][16 t + l % 16].
#pragma unroll
        for (int kq = 0; kq < kGP; kq += 4) {
          R fa[4], fb[4];
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            fa[t] = Ab[(kq + lhi) * kGLD + 64 * wr + 16 * t + llo];
            fb[t] = Bb[(kq + lhi) * kGLD + 64 * wc + 16 * t + llo];
          }
#pragma unroll
          for (int ta = 0; ta < 4; ++ta)
#pragma unroll
            for (int tb = 0; tb < 4; ++tb)
              macc[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[ta], fb[tb], macc[ta][tb],
                                                                  0, 0, 0);
        }
      } else {
#pragma unroll 2
        for (int kk = 0; kk < kGP; ++kk) {
          R a[8], b[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            a[u] = Ab[kk * kGLD + tr * 8 + u];
            b[u] = Bb[kk * kGLD + tc * 8 + u];
          }
#pragma unroll
          for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int v = 0; v < 8; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
        }
      }
      if (more) stash(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
    if constexpr (MF) {
      // D[16 ta + 4 q + l / 16][16 tb + l % 16] = macc[ta][tb][q]
#pragma unroll
      for (int ta = 0; ta < 4; ++ta)
#pragma unroll
        for (int tb = 0; tb < 4; ++tb) {
          {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int i = 64 * wr + 16 * ta + 4 * q + lhi, j = 64 * wc + 16 * tb + llo;
              if (i < mb && j < nn) st(m0 + i, j, macc[ta][tb][q]);
            }
            // one 16 x 16 tile at a time: keeps the epilogue's operand loads
            // from being hoisted for all 64 elements at once
            __builtin_amdgcn_sched_barrier(0);
          }
        }
    } else {
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int v = 0; v < 8; ++v) {
          const int i = tr * 8 + u, j = tc * 8 + v;
          if (i < mb && j < nn) st(m0 + i, j, acc[u][v]);
        }
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
  void* GH;             // nvox x T  Ghat = dL/d(output) / weights (masked), element type
  const void* Y;        // np x n x n   Gram eigenvectors V[i][c], ascending eigenvalues
  const double* EV;     // np x n       Gram eigenvalues (= s^2), ascending
  void* P1;             // np x M x n   Xc V
  void* P2;             // np x M x n   Ghat V
  void* PE;             // np x M x n   P1 D + theta P2 diag(h)
  void* GS;             // np x n       column sums of Ghat
  void* S1;             // np x n x n   P1^T P2, then D
  void* UV;             // np x n       u
  double* THB;          // np           theta
  int loss;
  double mpm, beta;
  const double* var_map;
  int var_f32;  // round the patch prior to float32 (float32 noise-level map)
  const double* LAM;    // per linear patch index, or null (= 1)
  double* GLAM;         // per linear patch index, or null
  void* gin;            // dL/d(input), input dtype, nvox x T
};

// Ghat = dL/dP / theta over the whole volume: the output gradient divided by
// the forward weights (weighted / average) or the gradient itself (center,
// where only patch centres read it); zero where the output was masked or
// never written.  One elementwise pass per backward call.
template <class IN, class E>
__global__ void ghat_kernel(AdjArgs a, int64_t total) {
  using R = typename real_of<E>::type;
  const IN* go = reinterpret_cast<const IN*>(a.gout);
  E* gh = reinterpret_cast<E*>(a.GH);
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = q / a.g.T;
    bool keep = true;
    if (a.mask_kind == PDAMD_MASK_SPATIAL) keep = a.mask[v] != 0;
    else if (a.mask_kind == PDAMD_MASK_FULL) keep = a.mask[q] != 0;
    E r = E(R(0));
    if (keep) {
      if (a.recomb == PDAMD_CENTER) {
        r = cvt<E>(go[q]);
      } else {
        const float w = a.w4[q];
        if (w > 0.f) r = cvt<E>(go[q]) / E(R(w));
      }
    }
    gh[q] = r;
  }
}

// Ghat element of a patch (centre recombination: only the patch centre).
template <class E>
__device__ __forceinline__ E ghat_at(const AdjArgs& a, int64_t q, bool centre_elem) {
  using R = typename real_of<E>::type;
  if (a.recomb == PDAMD_CENTER && !centre_elem) return E(R(0));
  return reinterpret_cast<const E*>(a.GH)[q];
}

// Stage the patch's row offsets and NumPy-order per-frame mean in LDS.
template <class IN>
__device__ __forceinline__ void stage_rows_mean(const Geom& g, const IN* X, int64_t* roff,
                                                IN* mean) {
  const int M = g.M, T = g.pt;
  for (int r = threadIdx.x; r < M; r += kNT) roff[r] = row_offset(g, r);
  __syncthreads();
  for (int t = threadIdx.x; t < T; t += kNT)  // sequential, as numpy's axis-0 sum
    mean[t] = div_n(seq_add(X[roff[0] + t], X + t, roff + 1, M - 1), M);
  __syncthreads();
}

template <class IN, class E>
size_t adj_patch_lds_bytes(int M, int T) {
  using R = typename real_of<E>::type;
  auto al = [](size_t b) { return ((b + 15) / 16) * 16; };
  return gemm_lds_bytes<R>() + al((size_t)M * sizeof(int64_t)) + al((size_t)T * sizeof(IN)) + 16;
}

// P1 = Xc V, P2 = Ghat V (M x n each) and the column sums of Ghat, per patch.
template <class IN, class E>
__global__ __launch_bounds__(kNT) void adj_project_kernel(const IN* __restrict__ vol, AdjArgs a) {
  using R = typename real_of<E>::type;
  const Geom& g = a.g;
  const int n = a.n, M = g.M, T = g.pt;
  extern __shared__ __align__(16) char smem[];
  char* cur = smem;
  R* As = lds_carve<R>(cur, (size_t)2 * kGP * kGLD);
  R* Bs = lds_carve<R>(cur, (size_t)2 * kGP * kGLD);
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
    const R* V = reinterpret_cast<const R*>(a.Y) + (size_t)p * n * n;
    R* P1 = reinterpret_cast<R*>(a.P1) + (size_t)p * M * n;
    R* P2 = reinterpret_cast<R*>(a.P2) + (size_t)p * M * n;
    R* GS = reinterpret_cast<R*>(a.GS) + (size_t)p * n;
    tile_gemm<R, true, false>(
        M, n, n, [&](int i, int kk) { return cvt<E>(X[roff[i] + kk] - mean[kk]); },
        [&](int kk, int j) { return V[(size_t)kk * n + j]; },
        [&](int i, int j, R v) { P1[(size_t)i * n + j] = v; }, As, Bs);
    tile_gemm<R, true, false>(
        M, n, n,
        [&](int i, int kk) { return ghat_at<E>(a, base + roff[i] + kk, i == rc && kk == tcen); },
        [&](int kk, int j) { return V[(size_t)kk * n + j]; },
        [&](int i, int j, R v) { P2[(size_t)i * n + j] = v; }, As, Bs);
    for (int j = threadIdx.x; j < n; j += kNT) {
      R s = R(0);
      int r = 0;
      for (; r + 8 <= M; r += 8) {
        R gg[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) gg[u] = ghat_at<E>(a, base + roff[r + u] + j, r + u == rc && j == tcen);
#pragma unroll
        for (int u = 0; u < 8; ++u) s += gg[u];
      }
      for (; r < M; ++r) s += ghat_at<E>(a, base + roff[r] + j, r == rc && j == tcen);
      GS[j] = s;
    }
    __syncthreads();
  }
}

// Spectral part of the adjoint, per patch: theta, dL/dlambda, C, D, u and
// PE = P1 D + theta P2 diag(h).
template <class E, class RT>
__global__ __launch_bounds__(kNT) void adj_spectral_kernel(AdjArgs a) {
  using R = typename real_of<E>::type;
  const Geom& g = a.g;
  const int n = a.n, M = g.M, tid = threadIdx.x;
  extern __shared__ __align__(16) char smem[];
  char* cur = smem;
  R* As = lds_carve<R>(cur, (size_t)2 * kGP * kGLD);
  R* Bs = lds_carve<R>(cur, (size_t)2 * kGP * kGLD);
  double* sg = lds_carve<double>(cur, n);   // s_i
  double* gv = lds_carve<double>(cur, n);   // g_i
  double* fp = lds_carve<double>(cur, n);   // f'(y_i) = dg_i / ds_i
  double* dc = lds_carve<double>(cur, n);   // dg_i / dc
  double* hv = lds_carve<double>(cur, n);   // g_i / s_i
  double* tv = lds_carve<double>(cur, n);   // h o (V^T gs)
  double* red = lds_carve<double>(cur, 8);
  __shared__ int sh_min;
  __shared__ double sh_s[4];

  for (int p = blockIdx.x; p < a.np; p += gridDim.x) {
    const int64_t L = a.idx[p];
    const double* ev = a.EV + (size_t)p * n;
    const R* V = reinterpret_cast<const R*>(a.Y) + (size_t)p * n * n;
    const R* P1 = reinterpret_cast<const R*>(a.P1) + (size_t)p * M * n;
    const R* P2 = reinterpret_cast<const R*>(a.P2) + (size_t)p * M * n;
    R* PE = reinterpret_cast<R*>(a.PE) + (size_t)p * M * n;
    const R* GS = reinterpret_cast<const R*>(a.GS) + (size_t)p * n;
    R* S1 = reinterpret_cast<R*>(a.S1) + (size_t)p * n * n;
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
        if (a.var_f32) sigma = (double)(float)sigma;
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
    // ---- W = P1^T P2 (= V^T Xc^T Ghat V) into S1
    tile_gemm<R, false, false>(
        n, n, M, [&](int i, int kk) { return P1[(size_t)kk * n + i]; },
        [&](int kk, int j) { return P2[(size_t)kk * n + j]; },
        [&](int i, int j, R v) { S1[(size_t)i * n + j] = v; }, As, Bs);
    // ---- cbar = sum_i C_ii dg_i/dc, C = theta S^-1 W;  t = h o (V^T gs)
    double part = 0.0;
    for (int i = tid; i < n; i += kNT)
      if (sg[i] > 0.0) part += theta * (double)S1[(size_t)i * n + i] / sg[i] * dc[i];
    part = dpp_sum(part);
    if ((tid % kWave) == 0) red[tid / kWave] = part;
    for (int c = tid; c < n; c += kNT) {
      R t = R(0);
      int r = 0;
      for (; r + 8 <= n; r += 8) {
        R vv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) vv[u] = V[(size_t)(r + u) * n + c];
#pragma unroll
        for (int u = 0; u < 8; ++u) t += vv[u] * GS[r + u];
      }
      for (; r < n; ++r) t += V[(size_t)r * n + c] * GS[r];
      tv[c] = hv[c] * (double)t;
    }
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
    // ---- D = S^-1 (Q - C diag(h)) in place of W: one thread per pair {i, j},
    // which reads W_ij, W_ji and writes D_ij, D_ji (the divided differences
    // of g are symmetric in i, j and computed once per pair)
    const double tie = 1e-8 * smax;
    for (int e = tid; e < n * n; e += kNT) {
      const int i = e / n, j = e % n;
      if (i > j) continue;
      const double si = sg[i], sj = sg[j];
      const double ii = si > 0.0 ? 1.0 / si : 0.0;  // zero rows for s = 0
      const double ij = sj > 0.0 ? 1.0 / sj : 0.0;
      const double cij = theta * (double)S1[(size_t)i * n + j] * ii;
      const double cji = theta * (double)S1[(size_t)j * n + i] * ij;
      if (i == j) {
        double q = fp[i] * cij;
        if (i == m0 || i == m1) q += dmed;
        S1[(size_t)i * n + i] = (R)((q - cij * hv[i]) * ii);
        continue;
      }
      const double gp = fabs(si - sj) > tie ? (gv[i] - gv[j]) / (si - sj) : 0.5 * (fp[i] + fp[j]);
      const double gm = (si + sj) > 0.0 ? (gv[i] + gv[j]) / (si + sj) : 0.0;
      const double sym = 0.5 * gp * (cij + cji), skw = 0.5 * gm * (cij - cji);
      S1[(size_t)i * n + j] = (R)((sym + skw - cij * hv[j]) * ii);
      S1[(size_t)j * n + i] = (R)((sym - skw - cji * hv[i]) * ij);
    }
    __syncthreads();
    // ---- PE = P1 D + theta P2 diag(h)
    tile_gemm<R, true, false>(
        M, n, n, [&](int i, int kk) { return P1[(size_t)i * n + kk]; },
        [&](int kk, int j) { return S1[(size_t)kk * n + j]; },
        [&](int i, int j, R v) {
          PE[(size_t)i * n + j] = v + (R)(theta * hv[j]) * P2[(size_t)i * n + j];
        },
        As, Bs);
    // ---- u = theta (gs - V (h o V^T gs))
    for (int j = tid; j < n; j += kNT) {
      R s = R(0);
      int c = 0;
      for (; c + 8 <= n; c += 8) {
        R vv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) vv[u] = V[(size_t)j * n + c + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += vv[u] * (R)tv[c + u];
      }
      for (; c < n; ++c) s += V[(size_t)j * n + c] * (R)tv[c];
      UV[j] = (R)theta * (GS[j] - s);
    }
    if (tid == 0) {
      a.THB[p] = theta;
      if (a.GLAM) a.GLAM[L] = cbar * c0;
    }
    __syncthreads();
  }
}

// dL/dX = PE V^T + 1 u^T / M, scatter-added into dL/d(input).
template <class IN, class E>
__global__ __launch_bounds__(kNT) void adj_synth_kernel(const IN* __restrict__ vol, AdjArgs a) {
  using R = typename real_of<E>::type;
  const Geom& g = a.g;
  const int n = a.n, M = g.M;
  extern __shared__ __align__(16) char smem[];
  char* cur = smem;
  R* As = lds_carve<R>(cur, (size_t)2 * kGP * kGLD);
  R* Bs = lds_carve<R>(cur, (size_t)2 * kGP * kGLD);
  int64_t* roff = lds_carve<int64_t>(cur, M);
  IN* gin = reinterpret_cast<IN*>(a.gin);
  for (int p = blockIdx.x; p < a.np; p += gridDim.x) {
    int64_t sp;
    int t0;
    patch_origin(g, a.idx[p], sp, t0);
    for (int r = threadIdx.x; r < M; r += kNT) roff[r] = row_offset(g, r);
    __syncthreads();
    const int64_t base = sp + t0;
    const R* V = reinterpret_cast<const R*>(a.Y) + (size_t)p * n * n;
    const R* PE = reinterpret_cast<const R*>(a.PE) + (size_t)p * M * n;
    const R* UV = reinterpret_cast<const R*>(a.UV) + (size_t)p * n;
    const R invM = R(1) / (R)M;
    tile_gemm<R, true, true>(
        M, n, n, [&](int i, int kk) { return PE[(size_t)i * n + kk]; },
        [&](int kk, int j) { return V[(size_t)j * n + kk]; },
        [&](int i, int j, R v) { atomic_add_(gin + base + roff[i] + j, (IN)(v + UV[j] * invM)); },
        As, Bs);
  }
}

// Back-transformation of every tridiagonal eigenvector at once: Y = Q Z with
// Q = H_0 ... H_{n-2} from the tridiagonalisation (packed Householder rows +
// tau), Z the n x n eigenvectors of the tridiagonal.  Thread t owns columns
// 4 (t >> 3) .. + 3 and rows 16 (t & 7) .. + 15 of Z in registers (n <= 128),
// so each reflector entry read from LDS serves four columns; the eight lanes
// of a column group combine their partial dot products with three DPP steps,
// so no barrier is needed per reflector.  Reflectors are staged in LDS, kBtKR
// at a time.
constexpr int kBtKR = 32;

struct BtArgs {
  int np, n;
  const void* H;    // np x packed Householder rows
  const void* TAU;  // np x n
  const void* Z;    // np x n x n (row-major, column c = eigenvector c)
  void* Y;          // np x n x n
};

// Sum over the 8 lanes 8a .. 8a + 7; every lane of the group gets the same bits.
template <class R> __device__ __forceinline__ R sum8_(R v) {
  v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f<0x141>(v);  // row_half_mirror
  return v;
}

template <class R>
__global__ __launch_bounds__(kNT) void backtransform_kernel(BtArgs a) {
  const int n = a.n, tid = threadIdx.x;
  const int g = tid & 7, c0 = 4 * (tid >> 3);  // rows 16 g .., columns c0 ..
  __shared__ __align__(16) R vs[kBtKR][128];
  __shared__ R ts[kBtKR];
  for (int p = blockIdx.x; p < a.np; p += gridDim.x) {
    const R* Hp = reinterpret_cast<const R*>(a.H) + (size_t)p * n * n;
    const R* Tp = reinterpret_cast<const R*>(a.TAU) + (size_t)p * n;
    const R* Zp = reinterpret_cast<const R*>(a.Z) + (size_t)p * n * n;
    R* Yp = reinterpret_cast<R*>(a.Y) + (size_t)p * n * n;
    R y[4][16];
#pragma unroll
    for (int q = 0; q < 16; ++q)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int r = 16 * g + q, col = c0 + c;
        y[c][q] = (r < n && col < n) ? Zp[(size_t)r * n + col] : R(0);
      }
    for (int kh = n - 2; kh >= 0; kh -= kBtKR) {  // reflectors kh, kh-1, ..., kh-kBtKR+1
      __syncthreads();
      for (int e = tid; e < kBtKR * 128; e += kNT) {
        const int kk = e / 128, r = e % 128, k = kh - kk;
        R v = R(0);
        if (k >= 0 && r < n) {
          const R* row = Hp + hh_row(k, n) - (k + 1);
          v = r > k + 1 ? row[r] : (r == k + 1 ? R(1) : R(0));
        }
        vs[kk][r] = v;
      }
      if (tid < kBtKR) ts[tid] = (kh - tid >= 0) ? Tp[kh - tid] : R(0);
      __syncthreads();
      const int kcount = min(kBtKR, kh + 1);
      for (int kk = 0; kk < kcount; ++kk) {
        R v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = vs[kk][16 * g + q];
        R sd[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          R t = R(0);
#pragma unroll
          for (int q = 0; q < 16; ++q) t = fma(v[q], y[c][q], t);
          sd[c] = ts[kk] * sum8_(t);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
          for (int q = 0; q < 16; ++q) y[c][q] = fma(-sd[c], v[q], y[c][q]);
      }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int r = 16 * g + q, col = c0 + c;
        if (r < n && col < n) Yp[(size_t)r * n + col] = y[c][q];
      }
  }
}

template <class R>
int launch_backtransform(const Plan& pl, const BtArgs& ba, hipStream_t st) {
  backtransform_kernel<R><<<grid_for(ba.np, 2), kNT, 0, st>>>(ba);
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
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
  const size_t lds = gemm_lds_bytes<R>() + 6 * (((size_t)pl.n * 8 + 15) / 16) * 16 + 64;
  auto kern = adj_spectral_kernel<E, RT>;
  HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  kern<<<grid_for(aa.np, 2), kNT, lds, st>>>(aa);
  HIP_TRY(hipGetLastError());
  return PDAMD_OK;
}
