// Shared device/host helpers for the gfx950 kernels of the coupled LBM-MPM step.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/fsi_hip.h"

namespace fsi {

// ------------------------------------------------------------ errors ----

// Message of the last failing entry point on this host thread.
void set_error(const char* fmt, ...);
const char* last_error();

// Returns 0 after a successful launch sequence, else records the HIP error.
int check_launch(const char* where);

#define FSI_TRY_LAUNCH(where)             \
  do {                                    \
    int _rc = ::fsi::check_launch(where); \
    if (_rc) return _rc;                  \
  } while (0)

#define FSI_REQUIRE(cond, ...)     \
  do {                             \
    if (!(cond)) {                 \
      ::fsi::set_error(__VA_ARGS__); \
      return FSI_ERR_ARGUMENT;     \
    }                              \
  } while (0)

constexpr int kBlock = 256;

inline unsigned int blocks_for(int64_t n, int block = kBlock) {
  return (unsigned int)((n + block - 1) / block);
}

// --------------------------------------------------------- lattice ----

// fsi/lbm3d.py:221-239 velocity set and 171-174 opposite table.
__host__ __device__ constexpr int ex(int q) {
  constexpr int t[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  return t[q];
}
__host__ __device__ constexpr int ey(int q) {
  constexpr int t[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
  return t[q];
}
__host__ __device__ constexpr int ez(int q) {
  constexpr int t[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
  return t[q];
}
__host__ __device__ constexpr int opp(int q) {
  constexpr int t[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
  return t[q];
}
// weights as exact reciprocals: 3, 18, 36
__host__ __device__ constexpr int wden(int q) { return q == 0 ? 3 : (q < 7 ? 18 : 36); }

template <typename R> __host__ __device__ constexpr R wq(int q) {
  return R(1) / R(wden(q));
}

// ------------------------------------------------- population storage ----
//
// Populations are kept as deviations d_q = F_q - b_q from a per-direction base
// b_q. FP32/FP64 storage keeps the absolute value (b_q = 0) exactly as the
// reference does. The 16-bit storages keep the shift b_q = w_q rho0
// (Lehmann et al., Phys. Rev. E 106, 015308 (2022)):
//  - FP16: (F_q - w_q rho0) * scale in IEEE binary16, a fixed power-of-two
//    scale keeping slow-flow deviations in the normal range (11-bit
//    significand relative to each value);
//  - I16: (F_q - w_q rho0) / w_q * s as signed normalised 16-bit integers
//    (round(x * 32767), |x| <= 1), with a run-time scale s that the lattice
//    kernels adapt to the largest stored value, so every population is
//    resolved to 2^-15 of the field's largest normalised deviation.
// load/store take the factor of the population's weight class (rest, axis,
// diagonal), ld = 1/(store factor).

template <typename S> struct PopStore;

template <> struct PopStore<float> {
  using real = float;
  static constexpr bool shifted = false;
  static constexpr bool adaptive = false;
  __device__ static float load(float s, float) { return s; }
  __device__ static float scaled(float d, float) { return d; }
  __device__ static float store(float x) { return x; }
};

template <> struct PopStore<double> {
  using real = double;
  static constexpr bool shifted = false;
  static constexpr bool adaptive = false;
  __device__ static double load(double s, double) { return s; }
  __device__ static double scaled(double d, double) { return d; }
  __device__ static double store(double x) { return x; }
};

template <> struct PopStore<__half> {
  using real = float;
  static constexpr bool shifted = true;
  static constexpr bool adaptive = false;
  // the factors are exact powers of two
  __device__ static float load(__half s, float ld) { return __half2float(s) * ld; }
  __device__ static float scaled(float d, float st) { return d * st; }
  __device__ static __half store(float x) { return __float2half_rn(x); }
};

template <> struct PopStore<int16_t> {
  using real = float;
  static constexpr bool shifted = true;
  static constexpr bool adaptive = true;
  // ld = w_q / (32767 s); st = s / w_q
  __device__ static float load(int16_t s, float ld) { return (float)s * ld; }
  __device__ static float scaled(float d, float st) { return d * st; }
  // v_cvt_pknorm_i16_f32: round(clamp(x, -1, 1) * 32767) to nearest
  __device__ static int16_t store(float x) { return __builtin_amdgcn_cvt_pknorm_i16(x, 0.0f)[0]; }
};

// Largest of a non-negative float over the 64 lanes (all active), the same
// value in every lane: DPP row shifts make lane 15 of each row the row's
// maximum, two row broadcasts carry rows 0-2 into lane 63, which is read.
// Non-negative floats order as their bit patterns.
template <int CTRL, int ROW_MASK> __device__ __forceinline__ int dpp_max_step(int b) {
  const int o = __builtin_amdgcn_update_dpp(b, b, CTRL, ROW_MASK, 0xf, false);
  return (int)max((uint32_t)b, (uint32_t)o);
}

__device__ __forceinline__ float wave_max_nonneg(float v) {
  int b = (int)__float_as_uint(v);
  b = dpp_max_step<0x111, 0xf>(b);  // row_shr:1
  b = dpp_max_step<0x112, 0xf>(b);  // row_shr:2
  b = dpp_max_step<0x114, 0xf>(b);  // row_shr:4
  b = dpp_max_step<0x118, 0xf>(b);  // row_shr:8
  b = dpp_max_step<0x142, 0xa>(b);  // row_bcast:15 into rows 1 and 3
  b = dpp_max_step<0x143, 0xc>(b);  // row_bcast:31 into rows 2 and 3
  return __uint_as_float((uint32_t)__builtin_amdgcn_readlane(b, 63));
}

// weight class of a D3Q19 direction: 0 rest, 1 axis, 2 diagonal
__host__ __device__ constexpr int wclass(int q) { return q == 0 ? 0 : (q < 7 ? 1 : 2); }

// Adaptive I16 storage: a control block of three slots (rotating with the
// write epoch of the populations), each the scale the slot's populations were
// written with and the largest |normalised value| of that write, as
// kCtrlParts partial maxima on separate 128-byte lines (float bits, combined
// with atomicMax on their unsigned bits), and a fourth slot for the maxima of
// a measuring pass (a step run without stores, before the first writes).
constexpr int kCtrlParts = 64;
constexpr int kCtrlLine = 32;                              // floats per 128-byte line
constexpr int kCtrlSlot = kCtrlLine * (kCtrlParts + 1);    // floats per slot
constexpr int kCtrlMeasure = 3;                            // slot of the measuring pass
constexpr int kCtrlFloats = 4 * kCtrlSlot;

}  // namespace fsi
