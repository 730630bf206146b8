// Shared device helpers for the space-time low-rank patch kernels (gfx950).
//
// Element types: the Casorati data is real (float/double) or complex
// (cpx<float>/cpx<double>); the eigen stage runs in a real type R that is
// float or double.  All kernels are written for 64-wide wavefronts.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pd {

constexpr int kWave = 64;

template <class R>
struct cpx {
  R x, y;
  __host__ __device__ cpx() = default;
  __host__ __device__ constexpr cpx(R a, R b = R(0)) : x(a), y(b) {}
};

template <class R> __device__ __forceinline__ cpx<R> operator+(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <class R> __device__ __forceinline__ cpx<R> operator-(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <class R> __device__ __forceinline__ cpx<R> operator*(cpx<R> a, cpx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class R> __device__ __forceinline__ cpx<R> operator*(R s, cpx<R> a) { return {s * a.x, s * a.y}; }
template <class R> __device__ __forceinline__ cpx<R>& operator+=(cpx<R>& a, cpx<R> b) { a.x += b.x; a.y += b.y; return a; }
template <class R> __device__ __forceinline__ cpx<R>& operator-=(cpx<R>& a, cpx<R> b) { a.x -= b.x; a.y -= b.y; return a; }

// ---- uniform real/complex interface --------------------------------------
template <class T> struct real_of { using type = T; };
template <class R> struct real_of<cpx<R>> { using type = R; };
template <class T> struct is_cpx { static constexpr bool value = false; };
template <class R> struct is_cpx<cpx<R>> { static constexpr bool value = true; };

__device__ __forceinline__ float conj_(float a) { return a; }
__device__ __forceinline__ double conj_(double a) { return a; }
template <class R> __device__ __forceinline__ cpx<R> conj_(cpx<R> a) { return {a.x, -a.y}; }

__device__ __forceinline__ float re_(float a) { return a; }
__device__ __forceinline__ double re_(double a) { return a; }
template <class R> __device__ __forceinline__ R re_(cpx<R> a) { return a.x; }
__device__ __forceinline__ float im_(float) { return 0.f; }
__device__ __forceinline__ double im_(double) { return 0.0; }
template <class R> __device__ __forceinline__ R im_(cpx<R> a) { return a.y; }

__device__ __forceinline__ float abs2_(float a) { return a * a; }
__device__ __forceinline__ double abs2_(double a) { return a * a; }
template <class R> __device__ __forceinline__ R abs2_(cpx<R> a) { return a.x * a.x + a.y * a.y; }

// Convert between element types (real<->real, cpx<->cpx of another precision).
template <class D, class S> struct cvt_t;
template <> struct cvt_t<float, float> { __device__ static float f(float a) { return a; } };
template <> struct cvt_t<double, float> { __device__ static double f(float a) { return (double)a; } };
template <> struct cvt_t<float, double> { __device__ static float f(double a) { return (float)a; } };
template <> struct cvt_t<double, double> { __device__ static double f(double a) { return a; } };
template <class RD, class RS> struct cvt_t<cpx<RD>, cpx<RS>> {
  __device__ static cpx<RD> f(cpx<RS> a) { return {(RD)a.x, (RD)a.y}; }
};
template <class D, class S> __device__ __forceinline__ D cvt(S a) { return cvt_t<D, S>::f(a); }

// a * conj(b), conj(a) * b
template <class T> __device__ __forceinline__ T mulc(T a, T b) { return a * conj_(b); }
template <class T> __device__ __forceinline__ T cmul(T a, T b) { return conj_(a) * b; }

template <class T> __device__ __forceinline__ T zero_() { return T(0); }

// Build an element from (re, im); the imaginary part is dropped for reals.
template <class E> struct mk_t { __device__ static E f(typename real_of<E>::type a, typename real_of<E>::type) { return a; } };
template <class R> struct mk_t<cpx<R>> { __device__ static cpx<R> f(R a, R b) { return {a, b}; } };
template <class E> __device__ __forceinline__ E mk(typename real_of<E>::type a, typename real_of<E>::type b) {
  return mk_t<E>::f(a, b);
}
__device__ __forceinline__ float recip_(float a) { return 1.0f / a; }
__device__ __forceinline__ double recip_(double a) { return 1.0 / a; }
template <class R> __device__ __forceinline__ cpx<R> recip_(cpx<R> a) {
  const R d = a.x * a.x + a.y * a.y;
  return {a.x / d, -a.y / d};
}

// ---- synchronisation --------------------------------------------------------
// Workgroup barrier that orders LDS only: global loads and stores in flight
// (prefetches, result writes) are not waited for, unlike __syncthreads().
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
// Ordering point for LDS traffic between the lanes of one wavefront (LDS
// operations of a wave complete in issue order; this only stops the compiler
// from reordering across it).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// ---- wavefront / block reductions ----------------------------------------
__device__ __forceinline__ float shfl_xor_(float v, int m) { return __shfl_xor(v, m, kWave); }
__device__ __forceinline__ double shfl_xor_(double v, int m) { return __shfl_xor(v, m, kWave); }
__device__ __forceinline__ int shfl_xor_(int v, int m) { return __shfl_xor(v, m, kWave); }

template <class R> __device__ __forceinline__ R wave_sum(R v) {
#pragma unroll
  for (int m = kWave / 2; m >= 1; m >>= 1) v += shfl_xor_(v, m);
  return v;
}
template <class R> __device__ __forceinline__ cpx<R> wave_sum(cpx<R> v) {
#pragma unroll
  for (int m = kWave / 2; m >= 1; m >>= 1) {
    v.x += shfl_xor_(v.x, m);
    v.y += shfl_xor_(v.y, m);
  }
  return v;
}

// Wavefront sum with DPP row operations plus 4 readlanes: every lane gets the
// same, bit-identical result (so waves that compute it redundantly agree).
// Must be called with all 64 lanes active.
template <int CTRL> __device__ __forceinline__ float dpp_f(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x),
                                                               CTRL, 0xF, 0xF, false));
}
template <int CTRL> __device__ __forceinline__ double dpp_f(double x) {
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float readlane_(float x, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), lane));
}
__device__ __forceinline__ double readlane_(double x, int lane) {
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
template <class R> __device__ __forceinline__ R dpp_sum(R v) {
  v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f<0x141>(v);  // row_half_mirror
  v += dpp_f<0x140>(v);  // row_mirror
  return ((readlane_(v, 0) + readlane_(v, 16)) + readlane_(v, 32)) + readlane_(v, 48);
}
template <class R> __device__ __forceinline__ cpx<R> dpp_sum(cpx<R> v) {
  return {dpp_sum(v.x), dpp_sum(v.y)};
}

// Sum over a block of NT threads; `red` is LDS scratch of NT/64 elements.
// Every thread returns the total.  Contains two barriers.
template <int NT, class T>
__device__ __forceinline__ T block_sum(T v, T* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  if (lane == 0) red[wid] = v;
  lds_barrier();
  T s = red[0];
#pragma unroll
  for (int w = 1; w < NT / kWave; ++w) s += red[w];
  lds_barrier();
  return s;
}

// ---- patch geometry --------------------------------------------------------
// A volume of shape (d0, d1, d2, T), C order, time contiguous.  Patches of
// shape (p0, p1, p2, pt) on a grid (g0, g1, g2, gt) with steps (s0, s1, s2,
// st).  Casorati row r <-> voxel (i0, i1, i2) of the patch in C order; column
// = frame.  Batches of independent Casorati matrices use the same struct with
// shape (B, M, 1, T) and patch (1, M, 1, T).
struct Geom {
  int64_t vs0, vs1, vs2;  // spatial element strides
  int g0, g1, g2, gt;
  int s0, s1, s2, st;
  int p0, p1, p2, pt;
  int M;      // p0 * p1 * p2
  int64_t T;  // frames in the volume (vs2 and voxel index = offset / T)
};

// Element offset of the patch's first voxel at frame 0 (spatial part) and the
// patch's first frame, for linear patch index L (C order over the grid).
__device__ __forceinline__ void patch_origin(const Geom& g, int64_t L, int64_t& sp, int& t0) {
  const int64_t it = L % g.gt;
  L /= g.gt;
  const int64_t i2 = L % g.g2;
  L /= g.g2;
  const int64_t i1 = L % g.g1;
  const int64_t i0 = L / g.g1;
  sp = i0 * g.s0 * g.vs0 + i1 * g.s1 * g.vs1 + i2 * g.s2 * g.vs2;
  t0 = (int)(it * g.st);
}

__device__ __forceinline__ int64_t row_offset(const Geom& g, int r) {
  const int i2 = r % g.p2;
  const int q = r / g.p2;
  const int i1 = q % g.p1;
  const int i0 = q / g.p1;
  return (int64_t)i0 * g.vs0 + (int64_t)i1 * g.vs1 + (int64_t)i2 * g.vs2;
}

// This workgroup's rank when the grid is numbered XCD by XCD.  Workgroups are
// dispatched round-robin over the 8 XCDs, each with its own L2; with patch
// p = rank (+ a multiple of the grid), the workgroups of one XCD work on
// consecutive patches at any time -- neighbours that share most of their
// voxels, so the volume reads hit that XCD's L2 instead of going past it.
__device__ __forceinline__ int xcd_rank() {
  const int G = gridDim.x, b = blockIdx.x;
  const int per = G / 8, rem = G % 8, x = b % 8;
  return x * per + (x < rem ? x : rem) + b / 8;
}

}  // namespace pd
