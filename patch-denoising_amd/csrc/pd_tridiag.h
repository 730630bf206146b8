// One-wavefront Householder tridiagonalisation of real Grams (n <= 128),
// pd_tridiag_wave.hip.  Same inputs and outputs as the 4-wave register kernel
// in pd_kernels.inl: the Gram (row-major n x n, both triangles) in each
// patch's H block on input; packed reflector rows in that block, d | e | e^2
// in DE and tau in TAU on output.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pd {

struct TridiagWaveArgs {
  void* H;    // np x n x n (float or double)
  void* DE;   // np x 3n
  void* TAU;  // np x n
  int np;     // patches
  int n;      // Gram size
  const int* np_dev = nullptr;  // device patch count (<= np): blocks past it exit
};

// Blocks of 8 per side for a Gram of size n (0: not handled here: n > 128,
// or double with n > 64).
int tridiag_wave_nb(int n, bool f64);

// Enqueue the kernel on `st` (one one-wave workgroup per patch).  Returns 0,
// or a negative value if the launch failed or n is out of range.
int launch_tridiag_wave(const TridiagWaveArgs& a, bool f64, hipStream_t st);

// Workgroup variant (four wavefronts per patch, register-resident 16 x 16
// tiles) for Grams up to 16 kWgNT: real FP32 above 128, real FP64 and complex
// FP32 (Hermitian) above 64.  Reflector k in row k of the H block (columns
// k+2 .. n-1) above 128, as the LDS kernel for large Grams writes it; in
// packed rows at n <= 128.
constexpr int kWgNT = 13;
enum { kWgF32 = 0, kWgF64 = 1, kWgC32 = 2 };
int tridiag_wg_supported(int n, int kind);
int launch_tridiag_wg(const TridiagWaveArgs& a, int kind, hipStream_t st);

}  // namespace pd
