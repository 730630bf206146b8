// One-wavefront Householder tridiagonalisation of real symmetric Grams
// (n <= 8 NB), the reduction of LAPACK ?sytd2 (lower) that ssyevx / dsyevx
// apply before their eigen stage (reference: space_time/utils.py:131-135).
//
// Layout: the 64 lanes form an 8 x 8 grid, lane = 8 a + b.  The matrix is cut
// into 8 x 8 blocks; block (s, t) holds A[8s + a][8t + b] in lane (a, b).
// Only blocks with t <= s are kept, in registers (NB (NB + 1) / 2 per lane:
// 136 at n <= 128); diagonal blocks are kept whole (both triangles).
//
// Step k (reflector from column k, trailing update of A[k+1:, k+1:]):
//   * column k is block column 0 after the shifts below, coordinate b = k % 8:
//     its 8 owning lanes reduce |x|^2 over a (3 butterflies) and every lane
//     forms beta, tau from the same broadcast values (identical bits);
//   * v goes through LDS as Vt[a][s] (v at row 8s + a), so each lane reads its
//     row-side and column-side copies as contiguous 16-value runs;
//   * p = A v from the stored blocks: off-diagonal blocks add to the row side
//     (p_i += A_ij v_j) and the column side (p_j += A_ij v_i), diagonal blocks
//     to the row side only; row sides are summed over b (DPP within 8 lanes),
//     column sides over a (butterflies across rows), and both are brought to
//     row and column index space through LDS;
//   * w = tau p - (tau^2 / 2) (p . v) v, zero at eliminated indices;
//   * A -= v w^T + w v^T on the kept blocks (two FMAs per element);
//   * after 8 steps the leading block row and column are done and every kept
//     block moves one block up-left (register moves), so the active column
//     block is always block 0 and all register indices are compile-time.
// Outputs as the 4-wave kernel: v_k in packed row k of the patch's H block
// (columns k+1 .. n-1, v_k[k+1] = 1; zeros when H_k = I), d | e | e^2, tau.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>

#include "pd_common.h"
#include "pd_tridiag.h"

using namespace pd;

namespace {

__host__ __device__ constexpr int tri_ix(int s, int t) { return s * (s + 1) / 2 + t; }
__host__ __device__ inline int hh_row_w(int k, int n) { return k * (n - 1) - (k * (k - 1)) / 2; }

template <int CTRL>
__device__ __forceinline__ float dppf(float x) {
  return __builtin_bit_cast(float,
                            __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dppf(double x) {
  const long long v = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(v & 0xffffffff), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 8 lanes 8a .. 8a+7 (xor 1, xor 2, then the mirrored half of
// the 8-lane group); every lane of the group gets the same bits.
template <class R>
__device__ __forceinline__ R sum8_b(R v) {
  v += dppf<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dppf<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dppf<0x141>(v);  // row_half_mirror
  return v;
}
// Sum over the 8 lanes b, b+8, ..., b+56 (xor 8 inside each 16-lane row, then
// xor 16 and xor 32 across rows); every lane of the set gets the same bits.
template <class R>
__device__ __forceinline__ R sum8_a(R v) {
  v += dppf<0x128>(v);  // row_ror:8 == xor 8 inside a 16-lane row
  v += __shfl_xor(v, 16, kWave);
  v += __shfl_xor(v, 32, kWave);
  return v;
}

template <class R>
__device__ __forceinline__ R readlane_r(R x, int l) {
  if constexpr (sizeof(R) == 4) {
    return __builtin_bit_cast(R, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
  } else {
    const long long v = __builtin_bit_cast(long long, x);
    const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffff), l);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
    return __builtin_bit_cast(R, ((long long)hi << 32) | (unsigned int)lo);
  }
}
// Wave sum, identical bits in every lane.
template <class R>
__device__ __forceinline__ R wave_total(R v) {
  v = sum8_b(v);
  v = sum8_a(v);
  return readlane_r(v, 0);
}

// Steps [k, kend) of the reduction with NBP kept block rows / columns (a
// compile-time count: no per-block guards).  Entered with k a multiple of 8
// and block column 0 holding column k; shifts the kept blocks up-left after
// every 8 steps.  Blocks past the live part (at most the last three rows of
// the phase, left over by the shifts, and zero padding past n) only ever meet
// v = w = 0 there, so their FMAs change nothing that is read.
template <class R, int NB, int NBP>
__device__ __forceinline__ void tri_phase(R (&L)[NB * (NB + 1) / 2], int& k, int kend, int n,
                                          int& base, R* Vt, R* Pa, R* Pb, R* dd, R* ee, R* tt,
                                          R* Hp) {
  static_assert(NBP <= NB, "phase wider than the register block");
  const int lane = threadIdx.x;
  const int ra = lane >> 3, cb = lane & 7;
  int live = NBP;  // block rows of the phase not yet shifted out
  for (; k < kend; ++k) {
    const int u = k & 7;  // column k = coordinate u of block column 0
    // rows of block row s are live when s < live; in block row 0 only rows past
    // u + over (over = 1 also drops row k + 1, which is block row 1's first
    // row when u == 7)
    auto row_live = [&](int s, int over) {  // over: 0 -> il > u, 1 -> il > u + 1
      if (s == 0) return ra > u + over;
      const bool in_phase = s + 3 < NBP ? true : s < live;
      if (s == 1 && over == 1) return in_phase && (u < 7 || ra > 0);
      return in_phase;
    };
    const R dk = readlane_r(L[tri_ix(0, 0)], 9 * u);
    // ---- reflector from x = column k below the diagonal (lanes cb == u)
    R sq = R(0);
#pragma unroll
    for (int s = 0; s < NBP; ++s) {
      const R x = L[tri_ix(s, 0)];
      sq = row_live(s, 1) ? fma(x, x, sq) : sq;
    }
    sq = sum8_a(sq);
    const R xn2 = readlane_r(sq, u);
    const R alpha = (u < 7) ? readlane_r(L[tri_ix(0, 0)], 8 * (u + 1) + u)
                            : readlane_r(L[tri_ix(1, 0)], u);
    R beta = alpha, tk = R(0), scal = R(0);
    if (xn2 > R(0)) {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      tk = (beta - alpha) / beta;
      scal = R(1) / (alpha - beta);
    }
    const R one = xn2 > R(0) ? R(1) : R(0);
    if (cb == u) {
#pragma unroll
      for (int s = 0; s < NBP; ++s) {
        R v = row_live(s, 1) ? L[tri_ix(s, 0)] * scal : R(0);
        if (s == 0) v = ra == u + 1 ? one : v;
        if (s == 1) v = (u == 7 && ra == 0) ? one : v;
        Vt[ra * NB + s] = v;
      }
    }
    if (lane == 0) {
      dd[k] = dk;
      ee[k] = beta;
      tt[k] = tk;
    }
    wave_sync();
    R vr[NBP], vc[NBP];
#pragma unroll
    for (int s = 0; s < NBP; ++s) {
      vr[s] = Vt[ra * NB + s];
      vc[s] = Vt[cb * NB + s];
    }
    // ---- p = A v: row sides p_i += A_ij v_j, column sides p_j += A_ij v_i
    R pr[NBP], pc[NBP];
#pragma unroll
    for (int s = 0; s < NBP; ++s) {
      pr[s] = L[tri_ix(s, s)] * vc[s];
#pragma unroll
      for (int t = 0; t < s; ++t) {
        pr[s] = fma(L[tri_ix(s, t)], vc[t], pr[s]);
        pc[t] = (t + 1 == s) ? L[tri_ix(s, t)] * vr[s] : fma(L[tri_ix(s, t)], vr[s], pc[t]);
      }
    }
    pc[NBP - 1] = R(0);
#pragma unroll
    for (int s = 0; s < NBP; ++s) {
      pr[s] = sum8_b(pr[s]);
      pc[s] = sum8_a(pc[s]);
    }
    // column sides to row index space (lanes ra == 0 hold them for every cb)
    if (ra == 0) {
#pragma unroll
      for (int s = 0; s < NBP; ++s) Pb[cb * NB + s] = pc[s];
    }
    wave_sync();
    R q_r[NBP];
#pragma unroll
    for (int s = 0; s < NBP; ++s) {
      const R pt = pr[s] + Pb[ra * NB + s];
      q_r[s] = row_live(s, 0) ? tk * pt : R(0);
    }
    // row-index values to column index space (lanes cb == 0 hold them)
    if (cb == 0) {
#pragma unroll
      for (int s = 0; s < NBP; ++s) Pa[ra * NB + s] = q_r[s];
    }
    // K = -tau/2 (q . v) over the live indices (each index once: cb == 0)
    R dot = R(0);
    if (cb == 0) {
#pragma unroll
      for (int s = 0; s < NBP; ++s) dot = fma(q_r[s], vr[s], dot);
    }
    const R kc = R(-0.5) * tk * wave_total(dot);
    wave_sync();
    R wr[NBP], wc[NBP];
#pragma unroll
    for (int s = 0; s < NBP; ++s) {
      wr[s] = fma(kc, vr[s], q_r[s]);
      wc[s] = fma(kc, vc[s], Pa[cb * NB + s]);
    }
    // ---- A -= v w^T + w v^T on the kept blocks
#pragma unroll
    for (int s = 0; s < NBP; ++s)
#pragma unroll
      for (int t = 0; t <= s; ++t) {
        R x = L[tri_ix(s, t)];
        x = fma(-vr[s], wc[t], x);
        x = fma(-wr[s], vc[t], x);
        L[tri_ix(s, t)] = x;
      }
    // ---- v_k -> packed row k (columns k+1 .. n-1); every Gram load has returned
    {
      R* row = Hp + hh_row_w(k, n) - (k + 1);
      for (int i = k + 1 + lane; i < n; i += kWave) {
        const int il = i - base;
        row[i] = Vt[(il & 7) * NB + (il >> 3)];
      }
    }
    wave_sync();
    if (u == 7) {  // block row / column done: shift the kept blocks up-left
#pragma unroll
      for (int s = 0; s + 1 < NBP; ++s)
#pragma unroll
        for (int t = 0; t <= s; ++t) L[tri_ix(s, t)] = L[tri_ix(s + 1, t + 1)];
      base += 8;
      --live;
    }
  }
}

// WPE: waves per SIMD the register allocation must allow (2 for float at
// NB = 16: 256 VGPRs with a few spilled, instead of 264 and one wave).
template <class R, int NB, int WPE>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(WPE))) void tridiag_wave_kernel(TridiagWaveArgs args) {
  constexpr int NT = NB * (NB + 1) / 2;
  const int n = args.n;
  const int lane = threadIdx.x;
  const int ra = lane >> 3, cb = lane & 7;
  extern __shared__ __align__(16) char smem[];
  R* Vt = reinterpret_cast<R*>(smem);  // [8][NB]: v at local row 8s + a -> Vt[a * NB + s]
  R* Pa = Vt + 8 * NB;                 // [8][NB]: q (row side), same indexing
  R* Pb = Pa + 8 * NB;                 // [8][NB]: column-side sums, same indexing
  R* dd = Pb + 8 * NB;                 // [8 NB]: d_k
  R* ee = dd + 8 * NB;                 // [8 NB]: e_k (beta)
  R* tt = ee + 8 * NB;                 // [8 NB]: tau_k

  // one patch per workgroup: no patch loop, so no per-lane load offsets are
  // hoisted and kept live across patches
  const int p = blockIdx.x;
  if (args.np_dev && p >= *args.np_dev) return;
  R* Hp = reinterpret_cast<R*>(args.H) + (size_t)p * n * n;
  R L[NT];
#pragma unroll
  for (int s = 0; s < NB; ++s)
#pragma unroll
    for (int t = 0; t <= s; ++t) {
      const int i = 8 * s + ra, j = 8 * t + cb;
      L[tri_ix(s, t)] = (i < n && j < n) ? Hp[i * n + j] : R(0);
    }
  // every load of the Gram has returned before reflector rows overwrite it
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");

  // phases of 8 steps, one kept block row / column fewer each (one block
  // shift per phase), down to 1; phase widths are compile-time counts
  int k = 0, base = 0;
  const int last = n - 1;
  if constexpr (NB >= 16) tri_phase<R, NB, 16>(L, k, min(NB * 8 - 120, last), n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  if constexpr (NB >= 15) tri_phase<R, NB, 15>(L, k, min(NB * 8 - 112, last), n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  if constexpr (NB >= 14) tri_phase<R, NB, 14>(L, k, min(NB * 8 - 104, last), n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  if constexpr (NB >= 13) tri_phase<R, NB, 13>(L, k, min(NB * 8 - 96, last), n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  if constexpr (NB >= 12) tri_phase<R, NB, 12>(L, k, min(NB * 8 - 88, last), n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  if constexpr (NB >= 11) tri_phase<R, NB, 11>(L, k, min(NB * 8 - 80, last), n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  if constexpr (NB >= 10) tri_phase<R, NB, 10>(L, k, min(NB * 8 - 72, last), n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  if constexpr (NB >= 9) tri_phase<R, NB, 9>(L, k, min(NB * 8 - 64, last), n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  if constexpr (NB >= 8) tri_phase<R, NB, 8>(L, k, min(NB * 8 - 56, last), n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  if constexpr (NB >= 7) tri_phase<R, NB, 7>(L, k, min(NB * 8 - 48, last), n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  if constexpr (NB >= 6) tri_phase<R, NB, 6>(L, k, min(NB * 8 - 40, last), n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  if constexpr (NB >= 5) tri_phase<R, NB, 5>(L, k, min(NB * 8 - 32, last), n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  if constexpr (NB >= 4) tri_phase<R, NB, 4>(L, k, min(NB * 8 - 24, last), n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  if constexpr (NB >= 3) tri_phase<R, NB, 3>(L, k, min(NB * 8 - 16, last), n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  if constexpr (NB >= 2) tri_phase<R, NB, 2>(L, k, min(NB * 8 - 8, last), n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  tri_phase<R, NB, 1>(L, k, last, n, base, Vt, Pa, Pb, dd, ee, tt, Hp);
  {
    const int il = (n - 1) - base;  // 0 .. 7
    const R dl = readlane_r(L[tri_ix(0, 0)], 9 * il);
    if (lane == 0) dd[n - 1] = dl;
  }
  wave_sync();
  R* DEp = reinterpret_cast<R*>(args.DE) + (size_t)p * 3 * n;
  R* TAUp = reinterpret_cast<R*>(args.TAU) + (size_t)p * n;
  for (int q = lane; q < n; q += kWave) {
    const R ev = q < n - 1 ? ee[q] : R(0);
    DEp[q] = dd[q];
    DEp[n + q] = ev;
    DEp[2 * n + q] = ev * ev;
    TAUp[q] = q < n - 1 ? tt[q] : R(0);
  }
}

// ---------------------------------------------------------------------------
// Workgroup variant for Grams of 129 .. 16 NBT: four wavefronts per patch in a
// 2 x 2 grid (wa, wb).  The matrix is cut into 16 x 16 tiles; lane (a, b) of
// wavefront (wa, wb) holds element (8 wa + a, 8 wb + b) of every tile (s, t)
// with t <= s, in registers (NBT (NBT + 1) / 2 per lane: 91 at n <= 208).
// Diagonal tiles are kept whole, so every wavefront has the same compile-time
// register layout and no per-tile branches.
// Step k (column k in tile column P = k / 16, held by the wavefronts with
// wb = (k / 8) % 2 in lanes b = k % 8):
//   A  the owners write column k to LDS (zero at rows <= k + 1) with their
//      partial |x|^2, alpha = A[k+1][k] and d_k; barrier;
//   B  every lane forms beta, tau and v from the same LDS values (identical
//      bits); p = A v over its tiles: off-diagonal tiles add to the row side
//      (summed over b) and the column side (summed over a), diagonal tiles to
//      the row side; each index i then has two row parts (wb = 0, 1) and two
//      column parts (wa = 0, 1), written to LDS, and each wavefront its part
//      of v . p; barrier;
//   C  p_i = the four parts added in one fixed order, so the row and column
//      copies of w agree bit for bit; w = tau p - (tau^2 / 2)(v . p) v, zero at
//      eliminated indices; A -= v w^T + w v^T on the kept tiles.
// Phases of 16 steps start at tile column P (compile-time ranges).
// Outputs as the LDS kernel for large Grams: v_k in row k of the patch's H
// block (columns k+2 .. n-1, v_k[k+1] = 1 implicit), d | e | e^2, tau.
// LDS vectors hold index i at (i % 16) LD + i / 16 (LD = 4 mod 8: the 8 lanes
// of a row or column group read 8 different banks).
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int wg_ld(int nb) { return (nb & 7) == 4 ? nb : wg_ld(nb + 1); }

// Element arithmetic of the workgroup kernel, real or complex, as explicit
// FMAs (the same operations wherever a value is recomputed, so that the row
// and column copies of v, p and w agree bit for bit).
template <class R> __device__ __forceinline__ R e_fma(R a, R b, R c) { return fma(a, b, c); }
template <class R> __device__ __forceinline__ cpx<R> e_fma(cpx<R> a, cpx<R> b, cpx<R> c) {  // a b + c
  return {fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y))};
}
template <class R> __device__ __forceinline__ R e_fmac(R a, R b, R c) { return fma(a, b, c); }
template <class R> __device__ __forceinline__ cpx<R> e_fmac(cpx<R> a, cpx<R> b, cpx<R> c) {  // conj(a) b + c
  return {fma(a.x, b.x, fma(a.y, b.y, c.x)), fma(a.x, b.y, fma(-a.y, b.x, c.y))};
}
// x - v conj(w) - u conj(t)
template <class R> __device__ __forceinline__ R e_rank2(R x, R v, R w, R u, R t) {
  return fma(-u, t, fma(-v, w, x));
}
template <class R> __device__ __forceinline__ cpx<R> e_rank2(cpx<R> x, cpx<R> v, cpx<R> w, cpx<R> u, cpx<R> t) {
  const R re = fma(-u.x, t.x, fma(-u.y, t.y, fma(-v.x, w.x, fma(-v.y, w.y, x.x))));
  const R im = fma(-u.y, t.x, fma(u.x, t.y, fma(-v.y, w.x, fma(v.x, w.y, x.y))));
  return {re, im};
}
template <class R> __device__ __forceinline__ R e_mul(R a, R b) { return a * b; }
template <class R> __device__ __forceinline__ cpx<R> e_mul(cpx<R> a, cpx<R> b) {
  return {fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x)};
}
template <class R> __device__ __forceinline__ R e_scale(R a, R s) { return a * s; }
template <class R> __device__ __forceinline__ cpx<R> e_scale(cpx<R> a, R s) { return {a.x * s, a.y * s}; }
template <class R> __device__ __forceinline__ R e_sum8_a(R v) { return sum8_a(v); }
template <class R> __device__ __forceinline__ cpx<R> e_sum8_a(cpx<R> v) { return {sum8_a(v.x), sum8_a(v.y)}; }
template <class R> __device__ __forceinline__ R e_sum8_b(R v) { return sum8_b(v); }
template <class R> __device__ __forceinline__ cpx<R> e_sum8_b(cpx<R> v) { return {sum8_b(v.x), sum8_b(v.y)}; }
template <class R> __device__ __forceinline__ R e_total(R v) { return wave_total(v); }
template <class R> __device__ __forceinline__ cpx<R> e_total(cpx<R> v) { return {wave_total(v.x), wave_total(v.y)}; }
template <class R> __device__ __forceinline__ R e_abs2(R a) { return a * a; }
template <class R> __device__ __forceinline__ R e_abs2(cpx<R> a) { return fma(a.x, a.x, a.y * a.y); }

template <class E>
struct WgLds {
  using R = typename real_of<E>::type;
  E* X;    // [16][LD]: column k
  E* RS;   // [2][16][LD]: row parts of p, by wb
  E* CS;   // [2][16][LD]: column parts of p, by wa
  R* NRM;  // [2]: |x|^2 parts, by wa
  E* DOT;  // [4]: p^H v parts, by wavefront
  E* AL;   // [1]: alpha
  R* dd;   // [16 NBT]
  R* ee;   // [16 NBT]
  E* tt;   // [16 NBT]
};

// PK: reflector rows packed (row k at hh_row_w(k, n), the layout of the eigen
// stage for n <= 128) instead of row k of the n x n block.
// Complex (Hermitian) steps follow ?hetd2: beta real, tau complex, p = A v,
// w = tau p - (|tau|^2 / 2)(p^H v) v, A -= v w^H + w v^H.
template <class E, int NBT, bool PK, int P>
__device__ __forceinline__ void wg_phase(E (&L)[NBT * (NBT + 1) / 2], int& k, int last, int n,
                                         int wa, int wb, int wid, const WgLds<E>& s, E* Hp) {
  using R = typename real_of<E>::type;
  constexpr bool CX = is_cpx<E>::value;
  constexpr int LD = wg_ld(NBT), SL = 16 * LD;
  const int lane = threadIdx.x & (kWave - 1);
  const int a = lane >> 3, b = lane & 7;
  const int ra = 8 * wa + a, cb = 8 * wb + b;  // coordinates inside a tile
  const int kend = min(16 * P + 16, last);
  const E zero = E(R(0));
  for (; k < kend; ++k) {
    const int u = k & 7;
    // ---- A: column k below the diagonal (wavefronts wb == (k / 8) % 2)
    if (wb == ((k >> 3) & 1)) {
      R sq = R(0);
#pragma unroll
      for (int t = P; t < NBT; ++t) {
        const int i = 16 * t + ra;
        const E x = L[tri_ix(t, P)];
        const E xv = i > k + 1 ? x : zero;
        sq += e_abs2(xv);
        if (b == u) {
          s.X[ra * LD + t] = xv;
          if (i == k + 1) *s.AL = x;
          if (i == k) s.dd[k] = re_(x);
        }
      }
      sq = sum8_a(sq);
      if (lane == u) s.NRM[wa] = sq;
    }
    lds_barrier();
    // ---- B: reflector, v, p = A v
    const R xn2 = s.NRM[0] + s.NRM[1];
    const E alpha = *s.AL;
    const R ar = re_(alpha), ai = im_(alpha);
    R beta = ar;
    E tk = zero, scal = zero;
    const bool hh = (xn2 > R(0)) || (ai != R(0));  // H_k != I
    if (hh) {
      beta = -copysign(sqrt(ar * ar + ai * ai + xn2), ar);
      const R ib = R(1) / beta;
      tk = mk<E>((beta - ar) * ib, -ai * ib);
      scal = recip_(alpha - E(beta));
    }
    const E one = E(hh ? R(1) : R(0));
    if (threadIdx.x == 0) {
      s.ee[k] = beta;
      s.tt[k] = tk;
    }
    E vr[NBT], vc[NBT];
#pragma unroll
    for (int t = P; t < NBT; ++t) {
      vr[t] = e_mul(s.X[ra * LD + t], scal);
      vc[t] = e_mul(s.X[cb * LD + t], scal);
      if (t <= P + 1) {
        vr[t] = 16 * t + ra == k + 1 ? one : vr[t];
        vc[t] = 16 * t + cb == k + 1 ? one : vc[t];
      }
    }
    if (wid == (k & 3)) {  // v_k -> row k of H (columns k+2 .. n-1; packed rows also get v_k[k+1])
      E* row = PK ? Hp + hh_row_w(k, n) - (k + 1) : Hp + (size_t)k * n;
      if (PK && lane == 0) row[k + 1] = one;
      for (int i = k + 2 + lane; i < n; i += kWave) row[i] = e_mul(s.X[(i & 15) * LD + (i >> 4)], scal);
    }
    E pc[NBT];
    E drow = zero, dcol = zero;
#pragma unroll
    for (int t = P; t < NBT; ++t) {
      E pr = e_mul(L[tri_ix(t, t)], vc[t]);  // diagonal tile: row side
#pragma unroll
      for (int q = P; q < t; ++q) {
        const E x = L[tri_ix(t, q)];
        pr = e_fma(x, vc[q], pr);
        // column side: p_j += conj(A_ij) v_i; first term of column q at t = q + 1
        pc[q] = e_fmac(x, vr[t], q + 1 == t ? zero : pc[q]);
      }
      pr = e_sum8_b(pr);
      if (b == 0) s.RS[wb * SL + ra * LD + t] = pr;
      drow = e_fmac(pr, vr[t], drow);
    }
#pragma unroll
    for (int q = P; q + 1 < NBT; ++q) {
      const E c = e_sum8_a(pc[q]);
      if (a == 0) s.CS[wa * SL + cb * LD + q] = c;
      dcol = e_fmac(c, vc[q], dcol);
    }
    if (a == 0) s.CS[wa * SL + cb * LD + NBT - 1] = zero;
    E dp = (b == 0 ? drow : zero) + (a == 0 ? dcol : zero);
    dp = e_total(dp);
    if (lane == 0) s.DOT[wid] = dp;
    lds_barrier();
    // ---- C: w, rank-2 update
    const E dsum = (s.DOT[0] + s.DOT[1]) + (s.DOT[2] + s.DOT[3]);
    const E kc = e_scale(dsum, R(-0.5) * e_abs2(tk));
    auto p_at = [&](int o) { return ((s.RS[o] + s.RS[SL + o]) + s.CS[o]) + s.CS[SL + o]; };
    E wr[NBT];
#pragma unroll
    for (int t = P; t < NBT; ++t) {
      wr[t] = e_fma(kc, vr[t], e_mul(tk, p_at(ra * LD + t)));
      if (t == P) wr[t] = 16 * t + ra > k ? wr[t] : zero;
    }
#pragma unroll
    for (int q = P; q < NBT; ++q) {
      E wc = e_fma(kc, vc[q], e_mul(tk, p_at(cb * LD + q)));
      if (q == P) wc = 16 * q + cb > k ? wc : zero;
#pragma unroll
      for (int t = q; t < NBT; ++t) L[tri_ix(t, q)] = e_rank2(L[tri_ix(t, q)], vr[t], wc, wr[t], vc[q]);
    }
    (void)CX;
  }
}

template <class E, int NBT, bool PK, int... Ps>
__device__ __forceinline__ void wg_phases(std::integer_sequence<int, Ps...>, E (&L)[NBT * (NBT + 1) / 2],
                                          int& k, int last, int n, int wa, int wb, int wid,
                                          const WgLds<E>& s, E* Hp) {
  (wg_phase<E, NBT, PK, Ps>(L, k, last, n, wa, wb, wid, s, Hp), ...);
}

template <class E, int NBT, bool PK>
__global__ __launch_bounds__(4 * kWave) void tridiag_wg_kernel(TridiagWaveArgs args) {
  using R = typename real_of<E>::type;
  constexpr int NT = NBT * (NBT + 1) / 2, LD = wg_ld(NBT), SL = 16 * LD;
  const int n = args.n;
  const int lane = threadIdx.x & (kWave - 1);
  const int a = lane >> 3, b = lane & 7;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int wa = wid >> 1, wb = wid & 1;
  __shared__ E xs[SL], rs[2 * SL], cs[2 * SL], dot[4], al[1], tt[16 * NBT];
  __shared__ R nrm[2], dd[16 * NBT], ee[16 * NBT];
  const WgLds<E> s{xs, rs, cs, nrm, dot, al, dd, ee, tt};

  const int p = blockIdx.x;
  if (args.np_dev && p >= *args.np_dev) return;
  E* Hp = reinterpret_cast<E*>(args.H) + (size_t)p * n * n;
  E L[NT];
#pragma unroll
  for (int t = 0; t < NBT; ++t)
#pragma unroll
    for (int q = 0; q <= t; ++q) {
      const int i = 16 * t + 8 * wa + a, c = 16 * q + 8 * wb + b;
      L[tri_ix(t, q)] = (i < n && c < n) ? Hp[(size_t)i * n + c] : E(R(0));
    }
  // every wavefront's Gram loads have returned before reflector rows overwrite H
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();

  int k = 0;
  wg_phases<E, NBT, PK>(std::make_integer_sequence<int, NBT>{}, L, k, n - 1, n, wa, wb, wid, s, Hp);
  {  // d_{n-1}: diagonal tile (n-1) / 16, element ((n-1) % 16, (n-1) % 16)
    const int tl = (n - 1) >> 4, cl = (n - 1) & 15;
#pragma unroll
    for (int t = 0; t < NBT; ++t)
      if (t == tl && wa == (cl >> 3) && wb == (cl >> 3) && lane == 9 * (cl & 7)) dd[n - 1] = re_(L[tri_ix(t, t)]);
  }
  __syncthreads();
  R* DEp = reinterpret_cast<R*>(args.DE) + (size_t)p * 3 * n;
  E* TAUp = reinterpret_cast<E*>(args.TAU) + (size_t)p * n;
  for (int q = threadIdx.x; q < n; q += 4 * kWave) {
    const R ev = q < n - 1 ? ee[q] : R(0);
    DEp[q] = dd[q];
    DEp[n + q] = ev;
    DEp[2 * n + q] = ev * ev;
    TAUp[q] = q < n - 1 ? tt[q] : E(R(0));
  }
}

template <class R, int NB>
int launch_nb(const TridiagWaveArgs& a, hipStream_t st) {
  const size_t lds = 6 * 8 * NB * sizeof(R);
  auto kern = tridiag_wave_kernel<R, NB, (sizeof(R) == 4 && NB == 16) ? 2 : 1>;
  if (a.np < 1) return 0;
  kern<<<a.np, kWave, lds, st>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

namespace pd {

int tridiag_wg_supported(int n, int kind) {
  return n > (kind == kWgF32 ? 128 : 64) && n <= 16 * kWgNT;
}

template <class E>
static void launch_wg(const TridiagWaveArgs& a, hipStream_t st) {
  // tiles per side: the smallest instantiation that holds the Gram
  auto kern = a.n <= 128   ? tridiag_wg_kernel<E, 8, true>
              : a.n <= 144 ? tridiag_wg_kernel<E, 9, false>
              : a.n <= 176 ? tridiag_wg_kernel<E, 11, false>
                           : tridiag_wg_kernel<E, kWgNT, false>;
  kern<<<a.np, 4 * kWave, 0, st>>>(a);
}

int launch_tridiag_wg(const TridiagWaveArgs& a, int kind, hipStream_t st) {
  if (!tridiag_wg_supported(a.n, kind)) return -2;
  if (a.np < 1) return 0;
  switch (kind) {
    case kWgF32: launch_wg<float>(a, st); break;
    case kWgF64: launch_wg<double>(a, st); break;
    default: launch_wg<cpx<float>>(a, st); break;
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int tridiag_wave_nb(int n, bool f64) {
  return n <= 32 ? 4 : (n <= 64 ? 8 : (f64 ? 0 : (n <= 96 ? 12 : (n <= 128 ? 16 : 0))));
}

int launch_tridiag_wave(const TridiagWaveArgs& a, bool f64, hipStream_t st) {
  const int nb = tridiag_wave_nb(a.n, f64);
  if (a.n < 2 || nb == 0) return -2;
  if (f64) {
    switch (nb) {
      case 4: return launch_nb<double, 4>(a, st);
      case 8: return launch_nb<double, 8>(a, st);
      default: return -2;  // 136 doubles per lane do not fit: the 4-wave kernel
    }
  }
  switch (nb) {
    case 4: return launch_nb<float, 4>(a, st);
    case 8: return launch_nb<float, 8>(a, st);
    case 12: return launch_nb<float, 12>(a, st);
    default: return launch_nb<float, 16>(a, st);
  }
}

}  // namespace pd
