// Gather-based recombination for M-side float32 patches (see pd_gather.h).
//
// Reference: the recombination loop of space_time/base.py:243-266
// (out[patch] += theta * patch_denoised; w[patch] += theta, in increasing
// patch index, then out /= w and out[~mask] = 0) with the denoised patch of
// utils.py:140-146 (projection on the kept eigenvectors plus the mean).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "patch_denoise_amd.h"
#include "pd_gather.h"

using namespace pd;

namespace {

constexpr int KM = kGatherKM;
constexpr int kSlice = 8;  // gather tile: kSlice x kSlice voxels of one dim-2 index

// Grid indices g (origins g * s) of patches of size p covering coordinate c,
// as [lo, hi] (empty when hi < lo).  s == 0: one patch at origin 0.
__device__ __forceinline__ void covering(int c, int p, int s, int gcount, int& lo, int& hi) {
  if (s == 0) {
    lo = 0;
    hi = c < p ? 0 : -1;
    return;
  }
  const int a = c - p + 1;
  lo = a <= 0 ? 0 : (a + s - 1) / s;
  hi = min(c / s, gcount - 1);
}

// Factors of one patch per wavefront (lanes = frames, NCH chunks of 64
// frames per pass, so each eigenvector row is loaded once per pass and NCH
// independent data loads are in flight per row).
template <int NCH>
__global__ __launch_bounds__(kWave) void gather_factor_kernel(GatherArgs a) {
  const int p = xcd_rank();  // chunk-local patch, or list entry (XCD-major order)
  if (a.list && p >= *a.list_cnt) return;
  const int slot = a.list ? a.list[p] : a.slot0 + p;
  const int lane = threadIdx.x;
  const Geom& g = a.g;
  const int M = g.M, T = g.pt;
  int64_t sp;
  int t0;
  patch_origin(g, a.idx[slot], sp, t0);
  const float* __restrict__ X = a.vol + sp + t0;
  const float* __restrict__ U = a.Y + (size_t)slot * a.n * KM;
  const float* __restrict__ mean = a.MEAN + (size_t)slot * T;
  const int nc = a.NCOMP[slot];
  const double theta = a.recomb == PDAMD_WEIGHTED ? a.THETA[slot] : 1.0;
  float scale[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k)
    scale[k] = k < nc ? (float)(theta * (a.list ? 1.0 : a.WC[(size_t)p * KM + k])) : 0.f;
  // patch-relative row offsets, rows lane and lane + 64 (M <= 128)
  auto roff = [&](int r) {
    const int i2 = r % g.p2, q = r / g.p2;
    return (int)((q / g.p1) * g.vs0 + (q % g.p1) * g.vs1 + i2 * g.vs2);
  };
  const int ro0 = lane < M ? roff(lane) : 0;
  const int ro1 = lane + kWave < M ? roff(lane + kWave) : 0;
  float* __restrict__ R = a.RF + (size_t)slot * (KM + 1) * T;
  for (int tb = 0; tb < T; tb += NCH * kWave) {
    int tt[NCH];
    bool tv[NCH];
    float m[NCH], acc[NCH][KM];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const int t = tb + c * kWave + lane;
      tv[c] = t < T;
      tt[c] = tv[c] ? t : 0;
      m[c] = tv[c] ? mean[t] : 0.f;
#pragma unroll
      for (int k = 0; k < KM; ++k) acc[c][k] = 0.f;
    }
    for (int i = 0; i < M; ++i) {
      const int off = __builtin_amdgcn_readlane(i < kWave ? ro0 : ro1, i & (kWave - 1));
      float x[NCH];
#pragma unroll
      for (int c = 0; c < NCH; ++c) x[c] = X[off + tt[c]] - m[c];
      float u[KM];
#pragma unroll
      for (int k = 0; k < KM; ++k) u[k] = U[i * KM + k];
#pragma unroll
      for (int c = 0; c < NCH; ++c)
#pragma unroll
        for (int k = 0; k < KM; ++k) acc[c][k] = fmaf(u[k], x[c], acc[c][k]);
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      if (!tv[c]) continue;
      const int t = tt[c];
      R[t] = (float)theta * m[c];
#pragma unroll
      for (int k = 0; k < KM; ++k) R[(size_t)(1 + k) * T + t] = scale[k] * acc[c][k];
    }
  }
}

// T-side factors of one patch per wavefront: V (T x KM) and the means in LDS
// (read as broadcasts), lanes = voxel rows, 64 per pass, each lane streaming
// its row's frames (four per load when rows are 16-byte aligned).
__global__ __launch_bounds__(kWave) void gather_factor_t_kernel(GatherArgs a) {
  extern __shared__ __align__(16) float fts[];
  const int p = xcd_rank();  // chunk-local patch, or list entry (XCD-major order)
  if (a.list && p >= *a.list_cnt) return;
  const int slot = a.list ? a.list[p] : a.slot0 + p;
  const int lane = threadIdx.x;
  const Geom& g = a.g;
  const int M = g.M, T = g.pt;
  float* vs = fts;            // [T][KM]
  float* ms = fts + T * KM;   // [T]
  int64_t sp;
  int t0;
  patch_origin(g, a.idx[slot], sp, t0);
  const float* __restrict__ X = a.vol + sp + t0;
  const float* __restrict__ V = a.V + (size_t)slot * T * KM;
  const float* __restrict__ mean = a.MEAN + (size_t)slot * T;
  const int nc = a.NCOMP[slot];
  const double theta = a.recomb == PDAMD_WEIGHTED ? a.THETA[slot] : 1.0;
  const float th = (float)theta;
  float wk[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k)
    wk[k] = k < nc ? (float)(a.list ? 1.0 : a.WC[(size_t)p * KM + k]) : 0.f;
  float* __restrict__ R = a.RF + (size_t)slot * (KM + 1) * T;
  for (int t = lane; t < T; t += kWave) {
    const float m = mean[t];
    ms[t] = m;
    R[t] = th * m;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      const float v = k < nc ? V[(size_t)t * KM + k] : 0.f;
      vs[t * KM + k] = v;
      R[(size_t)(1 + k) * T + t] = th * v;
    }
  }
  wave_sync();
  float* __restrict__ U = a.YT + (size_t)slot * M * KM;
  const bool v4 = (T & 3) == 0 && ((sp + t0) & 3) == 0 && (g.vs2 & 3) == 0 && (g.vs1 & 3) == 0 &&
                  (g.vs0 & 3) == 0;
  for (int rb = 0; rb < M; rb += kWave) {
    const int r = rb + lane;
    const bool rv = r < M;
    int off = 0;
    if (rv) {
      const int i2 = r % g.p2, q = r / g.p2;
      off = (int)((q / g.p1) * g.vs0 + (q % g.p1) * g.vs1 + i2 * g.vs2);
    }
    const float* __restrict__ xr = X + off;
    float acc[KM];
#pragma unroll
    for (int k = 0; k < KM; ++k) acc[k] = 0.f;
    auto add = [&](int t, float x) {
      const float xc = x - ms[t];
      const float* vr = vs + t * KM;
#pragma unroll
      for (int k = 0; k < KM; ++k) acc[k] = fmaf(xc, vr[k], acc[k]);
    };
    int t = 0;
    if (v4) {
      for (; t + 4 <= T; t += 4) {
        const float4 x = *reinterpret_cast<const float4*>(xr + t);
        add(t, x.x);
        add(t + 1, x.y);
        add(t + 2, x.z);
        add(t + 3, x.w);
      }
    } else {
      for (; t + 4 <= T; t += 4) {
        float x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = xr[t + u];
#pragma unroll
        for (int u = 0; u < 4; ++u) add(t + u, x[u]);
      }
    }
    for (; t < T; ++t) add(t, xr[t]);
    if (rv) {
#pragma unroll
      for (int k = 0; k < KM; ++k) U[(size_t)r * KM + k] = wk[k] * acc[k];
    }
  }
}

__global__ void gather_map_fill_kernel(int32_t* __restrict__ map, int64_t npatch) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npatch;
       i += (int64_t)gridDim.x * blockDim.x)
    map[i] = -1;
}

__global__ void gather_map_set_kernel(GatherArgs a, int64_t nsel) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nsel;
       j += (int64_t)gridDim.x * blockDim.x)
    a.MAP[a.idx[j]] = (int32_t)(j * 16 + a.NCOMP[j]);
}

// Float32 weight sum per voxel over its covering patches, in increasing patch
// index, rounded after every addition as the reference's float32 array is:
// float64 theta (NumPy scalar) -> f32(f64(w) + theta); a Python-float theta
// (Hybrid) is a float32 operand -> f32(w + f32(theta)).
__global__ void gather_weight_kernel(GatherArgs a) {
  const Geom& g = a.g;
  const int64_t X = a.shape[0], Y = a.shape[1], Z = a.shape[2];
  const int64_t nvox = X * Y * Z;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(v % Z), y = (int)((v / Z) % Y), x = (int)(v / (Y * Z));
    int xl, xh, yl, yh, zl, zh;
    covering(x, g.p0, g.s0, g.g0, xl, xh);
    covering(y, g.p1, g.s1, g.g1, yl, yh);
    covering(z, g.p2, g.s2, g.g2, zl, zh);
    float w = 0.f;
    for (int gx = xl; gx <= xh; ++gx)
      for (int gy = yl; gy <= yh; ++gy)
        for (int gz = zl; gz <= zh; ++gz) {
          const int e = a.MAP[((int64_t)gx * g.g1 + gy) * g.g2 + gz];
          if (e < 0) continue;
          const double th = a.recomb == PDAMD_WEIGHTED ? a.THETA[e >> 4] : 1.0;
          w = a.theta_weak ? w + (float)th : (float)((double)w + th);
        }
    a.W3[v] = w;
  }
}

// One wavefront per (8 x 8 voxels of one dim-2 index, 64 frames).
//   1. the covering patches (C order) are looked up 64 at a time and the
//      selected ones appended, in order, to a visit list in LDS;
//   2. per visit, lane j loads the eigenvector row of box voxel j (at most 64
//      box voxels) into LDS, every lane its frame's 11 factor rows -- for the
//      next visit while the current one is accumulated;
//   3. each box voxel's contribution (11 FMAs per lane, the eigenvector row
//      read as a broadcast) is added to its accumulator, one register per
//      slice voxel (the LDS then holds only the visit list and one visit's
//      eigenvector rows: four wavefronts per SIMD).
constexpr int kVisitCap = 768;   // visit-list entries per segment of covering patches
constexpr int kPF = 1;           // visits whose global loads are in flight ahead of the current one
constexpr int kURow = 12;        // LDS row of one eigenvector row (10 used, 16-B aligned)

__global__ __launch_bounds__(kWave, 4) void gather_recombine_kernel(GatherArgs a) {
  // per lane (frame) the 64 voxels of the slice, in registers (compile-time
  // indices; each visit adds to the voxels of its box, a wavefront-uniform test)
  float acc[kSlice * kSlice];
  // visit list (+ padding: empty visits up to a multiple of kPF + 1 and the
  // prefetch distance, so the visit loop has no bounds branches)
  __shared__ int vslot[kVisitCap + 2 * (kPF + 1)];  // slot * 16 + kept components of each visit
  __shared__ int vcode[kVisitCap + 2 * (kPF + 1)];  // (gx-gxl) << 20 | (gy-gyl) << 10 | (gz-gzl), -1: empty
  // eigenvector rows of the current visit's box voxels, after kSlice rows of
  // padding (a box row's base may lie up to kSlice - 1 rows before its first
  // voxel's row)
  __shared__ __align__(16) float urow[(kSlice + kSlice * kSlice) * kURow];
  const Geom& g = a.g;
  const int lane = threadIdx.x;
  const int X = (int)a.shape[0], Y = (int)a.shape[1], Z = (int)a.shape[2];
  const int T = g.pt;
  const int nT = (T + kWave - 1) / kWave, nty = (Y + kSlice - 1) / kSlice;
  // XCD-aware work order: consecutive workgroups land on different XCDs, so
  // each XCD is given a contiguous range of work items (neighbouring slices
  // and their shared factors stay in that XCD's L2)
  const int B = gridDim.x, bid = blockIdx.x;
  const int per = B / 8, rem = B % 8, xcd = bid % 8, kk = bid / 8;
  int w = xcd < rem ? xcd * (per + 1) + kk : rem * (per + 1) + (xcd - rem) * per + kk;
  const int tc = w % nT;
  w /= nT;
  const int z = w % Z;
  w /= Z;
  const int ty = w % nty, tx = w / nty;
  const int x0 = tx * kSlice, y0 = ty * kSlice;
  const int xe = min(x0 + kSlice, X), ye = min(y0 + kSlice, Y);
  const int t = tc * kWave + lane;
  const bool tv = t < T;
  const int tt = tv ? t : 0;
#pragma unroll
  for (int v = 0; v < kSlice * kSlice; ++v) acc[v] = 0.f;
  int gxl, gxh, gyl, gyh, gzl, gzh, lo, hi;
  covering(x0, g.p0, g.s0, g.g0, gxl, hi);
  covering(xe - 1, g.p0, g.s0, g.g0, lo, gxh);
  covering(y0, g.p1, g.s1, g.g1, gyl, hi);
  covering(ye - 1, g.p1, g.s1, g.g1, lo, gyh);
  covering(z, g.p2, g.s2, g.g2, gzl, gzh);
  const int nx = max(gxh - gxl + 1, 0), ny = max(gyh - gyl + 1, 0), nz = max(gzh - gzl + 1, 0);
  const int ncov = nx * ny * nz;
  const float* __restrict__ RF = a.RF;
  const float* __restrict__ Yb = a.Y;
  const int KR = (KM + 1) * T;
  // box of a visit: [xa, xb) x [ya, yb) of the slice, eigenvector rows at zi
  auto box = [&](int code, int& ox, int& oy, int& xa, int& xb, int& ya, int& yb, int& zi) {
    if (code < 0) {  // padding: an empty box
      ox = x0;
      oy = y0;
      xa = xb = x0;
      ya = yb = y0;
      zi = 0;
      return;
    }
    const int gx = gxl + (code >> 20), gy = gyl + ((code >> 10) & 1023), gz = gzl + (code & 1023);
    ox = gx * g.s0;
    oy = gy * g.s1;
    xa = max(x0, ox);
    xb = min(xe, ox + g.p0);
    ya = max(y0, oy);
    yb = min(ye, oy + g.p1);
    zi = z - gz * g.s2;
  };
  for (int seg = 0; seg < ncov; seg += kVisitCap) {
    const int segn = min(kVisitCap, ncov - seg);
    // 1. visit list of this segment
    int nvis = 0;
    for (int q0 = 0; q0 < segn; q0 += kWave) {
      const int q = seg + q0 + lane;
      int e = -1, code = 0;
      if (q0 + lane < segn) {
        const int dz = q % nz, r = q / nz;
        const int dy = r % ny, dx = r / ny;
        e = a.MAP[((int64_t)(gxl + dx) * g.g1 + gyl + dy) * g.g2 + gzl + dz];
        code = (dx << 20) | (dy << 10) | dz;
      }
      const unsigned long long sel = __ballot(e >= 0);
      const int pos = nvis + __popcll(sel & ((1ull << lane) - 1ull));
      if (e >= 0) {
        vslot[pos] = e;
        vcode[pos] = code;
      }
      nvis += __popcll(sel);
    }
    if (nvis == 0) continue;
    const int nvp = (nvis + kPF) / (kPF + 1) * (kPF + 1);  // visits processed (padded)
    for (int q = nvis + lane; q < nvp + kPF; q += kWave) {
      vslot[q] = vslot[0] & ~15;  // no components: loads only the mean row
      vcode[q] = -1;
    }
    wave_sync();
    // 2./3. visits, the loads of the next kPF ones in flight (rotating registers)
    float rq[kPF + 1][KM + 1], uq[kPF + 1][KM];
    // factor rows past a patch's kept components are zero: not loaded
    auto fetch = [&](int v, float (&r)[KM + 1], float (&u)[KM]) {
      const int e = vslot[v], slot = e >> 4, nc = e & 15;
      const float* Rs = RF + (size_t)slot * KR + tt;
#pragma unroll
      for (int k = 0; k <= KM; ++k) {
        if (k <= nc) r[k] = Rs[(size_t)k * T];
        else r[k] = 0.f;
      }
      int ox, oy, xa, xb, ya, yb, zi;
      box(vcode[v], ox, oy, xa, xb, ya, yb, zi);
      const int by = yb - ya, nbx = (xb - xa) * by;
      const int j = lane < nbx ? lane : 0;
      const int row =
          nbx > 0 ? ((xa - ox + j / max(by, 1)) * g.p1 + (ya - oy + j % max(by, 1))) * g.p2 + zi : 0;
      const float* Ur = Yb + ((size_t)slot * a.n + row) * KM;
#pragma unroll
      for (int k = 0; k < KM; ++k) {
        if (k < nc) u[k] = Ur[k];
        else u[k] = 0.f;
      }
    };
#pragma unroll
    for (int d = 0; d < kPF; ++d) fetch(d, rq[d], uq[d]);
    // visit v uses register slot v % (kPF + 1); the visit loop is unrolled by
    // kPF + 1 so every slot index is a constant and a visit waits only for
    // its own loads (later visits' loads stay in flight)
    auto visit_one = [&](int v, auto B) {
      constexpr int b = decltype(B)::value;
      constexpr int bn = (b + kPF) % (kPF + 1);
      {  // eigenvector rows of this visit -> LDS (lane j: box voxel j)
        float* dst = urow + (kSlice + lane) * kURow;
#pragma unroll
        for (int k = 0; k < KM; ++k) dst[k] = uq[b][k];
      }
      float rf[KM + 1];
#pragma unroll
      for (int k = 0; k <= KM; ++k) rf[k] = rq[b][k];
      wave_sync();
      fetch(v + kPF, rq[bn], uq[bn]);
      int ox, oy, xa, xb, ya, yb, zi;
      box(vcode[v], ox, oy, xa, xb, ya, yb, zi);
      const int by = yb - ya;
      // box voxels four at a time: accumulator and eigenvector-row reads
      // first (distinct voxels, so the reads need not wait on the writes)
      // a visit keeping at most 4 components skips the zero terms 5 .. 10
      // (adding exact zeros changes at most the sign of a zero sum)
      const bool wide = (vslot[v] & 15) > 4;
      auto contrib = [&](const float* ur) {
        const float4* us = reinterpret_cast<const float4*>(ur);
        const float4 u0 = us[0];
        float val = rf[0];
        val = fmaf(u0.x, rf[1], val);
        val = fmaf(u0.y, rf[2], val);
        val = fmaf(u0.z, rf[3], val);
        val = fmaf(u0.w, rf[4], val);
        if (wide) {
          const float4 u1 = us[1], u2 = us[2];
          val = fmaf(u1.x, rf[5], val);
          val = fmaf(u1.y, rf[6], val);
          val = fmaf(u1.z, rf[7], val);
          val = fmaf(u1.w, rf[8], val);
          val = fmaf(u2.x, rf[9], val);
          val = fmaf(u2.y, rf[10], val);
        }
        return val;
      };
      // the box's voxels (x, y), each with the LDS eigenvector row
      // (x - xa) * by + (y - ya): rows and columns of the tile inside the box
      // as 8-bit masks (one scalar bit test per row and per voxel), a row's
      // eigenvector rows from one base plus the column's constant offset
      auto span = [](int lo, int hi) {  // bits [lo, hi) of 0 .. kSlice
        lo = min(max(lo, 0), kSlice);
        hi = min(max(hi, 0), kSlice);
        return ((1u << hi) - 1u) & ~((1u << lo) - 1u);
      };
      const unsigned rm = span(xa - x0, xb - x0), cm = span(ya - y0, yb - y0);
      auto voxels = [&](auto f) {
#pragma unroll
        for (int rx = 0; rx < kSlice; ++rx) {
          if (!(rm & (1u << rx))) continue;
          const float* ur = urow + (kSlice + (x0 + rx - xa) * by - (ya - y0)) * kURow;
#pragma unroll
          for (int ry = 0; ry < kSlice; ++ry)
            if (cm & (1u << ry)) acc[rx * kSlice + ry] += f(ur + ry * kURow);
        }
      };
      voxels(contrib);
      wave_sync();
    };
    static_assert(kPF >= 1 && kPF <= 3, "the visit loop below is unrolled for kPF + 1 <= 4 slots");
    for (int v = 0; v < nvp; v += kPF + 1) {
      visit_one(v, std::integral_constant<int, 0>());
      visit_one(v + 1, std::integral_constant<int, 1>());
      if constexpr (kPF >= 2) visit_one(v + 2, std::integral_constant<int, 2>());
      if constexpr (kPF >= 3) visit_one(v + 3, std::integral_constant<int, 3>());
    }
    wave_sync();
  }
  if (!tv) return;
#pragma unroll
  for (int v = 0; v < kSlice * kSlice; ++v) {
    const int x = x0 + v / kSlice, y = y0 + v % kSlice;
    if (x >= X || y >= Y) continue;
    const int64_t vox = ((int64_t)x * Y + y) * Z + z;
    const float wv = a.W3[vox];
    float o = acc[v] / wv;
    if (a.mask_kind == PDAMD_MASK_SPATIAL && !a.mask[vox]) o = 0.f;
    if (a.mask_kind == PDAMD_MASK_FULL && !a.mask[vox * T + t]) o = 0.f;
    a.out[vox * T + t] = o;
    a.wout[vox * T + t] = wv;
  }
}

}  // namespace

namespace pd {

int launch_gather_factors(const GatherArgs& a, hipStream_t st) {
  const int nb = a.list ? a.list_cap : a.np;
  if (nb < 1) return 0;
  if (a.tside) {
    const size_t lds = (size_t)a.g.pt * (KM + 1) * sizeof(float);
    if (lds > 64 * 1024) return -1;
    gather_factor_t_kernel<<<nb, kWave, lds, st>>>(a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
  }
  const int nch = (a.g.pt + kWave - 1) / kWave;
  if (nch <= 2) gather_factor_kernel<2><<<nb, kWave, 0, st>>>(a);
  else if (nch <= 3) gather_factor_kernel<3><<<nb, kWave, 0, st>>>(a);
  else gather_factor_kernel<5><<<nb, kWave, 0, st>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_gather_map(const GatherArgs& a, int64_t npatch, int64_t nsel, hipStream_t st) {
  const int nb1 = (int)std::min<int64_t>((npatch + 255) / 256, 65536);
  gather_map_fill_kernel<<<std::max(nb1, 1), 256, 0, st>>>(a.MAP, npatch);
  if (nsel > 0) {
    const int nb2 = (int)std::min<int64_t>((nsel + 255) / 256, 65536);
    gather_map_set_kernel<<<nb2, 256, 0, st>>>(a, nsel);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_gather_weights(const GatherArgs& a, hipStream_t st) {
  const int64_t nvox = a.shape[0] * a.shape[1] * a.shape[2];
  const int nb = (int)std::min<int64_t>((nvox + 255) / 256, 65536);
  gather_weight_kernel<<<std::max(nb, 1), 256, 0, st>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_gather_recombine(const GatherArgs& a, hipStream_t st) {
  const int64_t T = a.g.pt;
  const int64_t nT = (T + kWave - 1) / kWave;
  const int64_t ntx = (a.shape[0] + kSlice - 1) / kSlice, nty = (a.shape[1] + kSlice - 1) / kSlice;
  const int64_t items = ntx * nty * a.shape[2] * nT;
  if (items < 1) return 0;
  if (items >= ((int64_t)1 << 31)) return -1;
  gather_recombine_kernel<<<(int)items, kWave, 0, st>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace pd
