// Sliding-window noise map (space_time/utils.py:207-242, estimate_noise).
//
// out[c] = population std of the block^d x T elements of the stride-1 window
// centred at voxel c.  Two passes:
//   1. voxel_moments_kernel: per voxel, S1 = sum_t x and S2 = sum_t |x|^2 in
//      float64 (16 lanes per voxel row of T contiguous frames: coalesced,
//      the whole volume read once; HBM-bound);
//   2. window_std_kernel: per window centre, box sums of S1 / S2 over the
//      block^d voxels (small 3-D arrays, L2-resident) and
//      std = sqrt(max(S2 / N - |S1 / N|^2, 0)), N = block^d T.
// float64 moments keep the one-pass variance accurate far below the float32
// rounding of the reference's np.std.
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "patch_denoise_amd.h"
#include "pd_common.h"

using namespace pd;

namespace {

template <class IN> struct Mom;
template <> struct Mom<float> {
  __device__ static void add(float v, double& s1r, double& s1i, double& s2) {
    s1r += v;
    s2 += (double)v * v;
  }
};
template <> struct Mom<double> {
  __device__ static void add(double v, double& s1r, double& s1i, double& s2) {
    s1r += v;
    s2 += v * v;
  }
};
template <class R> struct Mom<cpx<R>> {
  __device__ static void add(cpx<R> v, double& s1r, double& s1i, double& s2) {
    s1r += v.x;
    s1i += v.y;
    s2 += (double)v.x * v.x + (double)v.y * v.y;
  }
};

// 16 lanes per voxel (4 voxels per wavefront): S1 (real, imag) and S2 over
// the T frames, 16-byte loads where the rows allow it, then a 16-lane
// butterfly.
constexpr int kVL = 16;

__device__ __forceinline__ double sum16(double v) {
#pragma unroll
  for (int m = kVL / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, kVL);
  return v;
}

template <class IN>
__global__ __launch_bounds__(256) void voxel_moments_kernel(const IN* __restrict__ x, int64_t nvox,
                                                            int64_t T, double* __restrict__ m) {
  const int sub = threadIdx.x % kVL;
  const int64_t g0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kVL;
  const int64_t ng = ((int64_t)gridDim.x * blockDim.x) / kVL;
  constexpr bool VEC = std::is_same<IN, float>::value;
  const bool vec = VEC && (T % 4 == 0);
  // every lane runs the same number of iterations (the butterfly needs all lanes)
  const int64_t iters = (nvox + ng - 1) / ng;
  for (int64_t it = 0; it < iters; ++it) {
    const int64_t v = g0 + it * ng;
    double s1r = 0.0, s1i = 0.0, s2 = 0.0;
    if (v < nvox) {
      const IN* row = x + v * T;
      if constexpr (VEC) {
        if (vec) {
          const float4* r4 = reinterpret_cast<const float4*>(row);
          for (int64_t t = sub; t < T / 4; t += kVL) {
            const float4 q = r4[t];
            s1r += ((double)q.x + q.y) + ((double)q.z + q.w);
            s2 += ((double)q.x * q.x + (double)q.y * q.y) + ((double)q.z * q.z + (double)q.w * q.w);
          }
        } else {
          for (int64_t t = sub; t < T; t += kVL) Mom<IN>::add(row[t], s1r, s1i, s2);
        }
      } else {
        for (int64_t t = sub; t < T; t += kVL) Mom<IN>::add(row[t], s1r, s1i, s2);
      }
    }
    s1r = sum16(s1r);
    s1i = sum16(s1i);
    s2 = sum16(s2);
    if (sub == 0 && v < nvox) {
      m[3 * v] = s1r;
      m[3 * v + 1] = s1i;
      m[3 * v + 2] = s2;
    }
  }
}

// One thread per window (corner c = centre - b/2 in each spatial dimension).
__global__ void window_std_kernel(const double* __restrict__ m, int64_t s0, int64_t s1, int64_t s2,
                                  int b0, int b1, int b2, int64_t T, double* __restrict__ out) {
  const int64_t g0 = s0 - b0 + 1, g1 = s1 - b1 + 1, g2 = s2 - b2 + 1;
  const int64_t nwin = g0 * g1 * g2;
  const double invN = 1.0 / ((double)b0 * b1 * b2 * (double)T);
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nwin;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = q / (g1 * g2), j = (q / g2) % g1, k = q % g2;
    double a = 0.0, ai = 0.0, q2 = 0.0;
    for (int di = 0; di < b0; ++di)
      for (int dj = 0; dj < b1; ++dj)
        for (int dk = 0; dk < b2; ++dk) {
          const int64_t v = ((i + di) * s1 + (j + dj)) * s2 + (k + dk);
          a += m[3 * v];
          ai += m[3 * v + 1];
          q2 += m[3 * v + 2];
        }
    const double mr = a * invN, mi = ai * invN;
    const double var = q2 * invN - (mr * mr + mi * mi);
    out[((i + b0 / 2) * s1 + (j + b1 / 2)) * s2 + (k + b2 / 2)] = sqrt(fmax(var, 0.0));
  }
}

}  // namespace

namespace pdamd_detail {
void set_error(const char* msg);  // pdamd_last_error's message (pd_api.hip)
}

namespace {
int noise_fail(int code, const char* msg) {
  pdamd_detail::set_error(msg);
  return code;
}
}  // namespace

extern "C" {

int pdamd_noise_workspace_size(const int64_t* shape4, size_t* bytes) {
  if (!shape4 || !bytes) return noise_fail(PDAMD_EINVAL, "null argument");
  *bytes = (size_t)shape4[0] * shape4[1] * shape4[2] * 3 * sizeof(double);
  return PDAMD_OK;
}

int pdamd_estimate_noise(const void* in_dev, int dtype, const int64_t* shape4, int nspatial,
                         int block_size, double* out_dev, void* workspace_dev,
                         size_t workspace_bytes, void* stream) {
  if (!in_dev || !out_dev || !workspace_dev || !shape4)
    return noise_fail(PDAMD_EINVAL, "null pointer");
  if (nspatial < 1 || nspatial > 3) return noise_fail(PDAMD_EINVAL, "1 to 3 spatial dimensions");
  if (block_size < 1) return noise_fail(PDAMD_EINVAL, "block_size must be >= 1");
  int b[3];
  for (int d = 0; d < 3; ++d) {
    b[d] = (d >= 3 - nspatial) ? block_size : 1;  // padded leading dims have size 1
    // the reference's corner list is empty (IndexError) unless size > block
    if (d >= 3 - nspatial && shape4[d] <= block_size)
      return noise_fail(PDAMD_EINVAL, "every spatial size must exceed block_size");
    if (shape4[d] < 1 || shape4[3] < 1) return noise_fail(PDAMD_EINVAL, "bad shape");
  }
  const int64_t nvox = shape4[0] * shape4[1] * shape4[2], T = shape4[3];
  if (workspace_bytes < (size_t)nvox * 3 * sizeof(double))
    return noise_fail(PDAMD_EWORKSPACE, "workspace too small");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  double* m = reinterpret_cast<double*>(workspace_dev);
  const int blocks = (int)std::min<int64_t>((nvox * kVL + 255) / 256, 8192);
  switch (dtype) {
    case PDAMD_F32:
      voxel_moments_kernel<float><<<blocks, 256, 0, st>>>(
          reinterpret_cast<const float*>(in_dev), nvox, T, m);
      break;
    case PDAMD_F64:
      voxel_moments_kernel<double><<<blocks, 256, 0, st>>>(
          reinterpret_cast<const double*>(in_dev), nvox, T, m);
      break;
    case PDAMD_C64:
      voxel_moments_kernel<cpx<float>><<<blocks, 256, 0, st>>>(
          reinterpret_cast<const cpx<float>*>(in_dev), nvox, T, m);
      break;
    case PDAMD_C128:
      voxel_moments_kernel<cpx<double>><<<blocks, 256, 0, st>>>(
          reinterpret_cast<const cpx<double>*>(in_dev), nvox, T, m);
      break;
    default:
      return noise_fail(PDAMD_EINVAL, "bad dtype");
  }
  const int64_t nwin = (shape4[0] - b[0] + 1) * (shape4[1] - b[1] + 1) * (shape4[2] - b[2] + 1);
  window_std_kernel<<<(int)std::min<int64_t>((nwin + 255) / 256, 65536), 256, 0, st>>>(
      m, shape4[0], shape4[1], shape4[2], b[0], b[1], b[2], T, out_dev);
  if (hipGetLastError() != hipSuccess) return noise_fail(PDAMD_EHIP, "launch failed");
  return PDAMD_OK;
}

}  // extern "C"
