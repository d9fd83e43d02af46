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
// reference does; FP16 storage keeps (F_q - w_q rho0) * scale with a power of
// two scale so that slow-flow deviations stay in the normal FP16 range
// (shifted FP16 storage, Lehmann et al., Phys. Rev. E 106, 015308 (2022)).

template <typename S> struct PopStore;

template <> struct PopStore<float> {
  using real = float;
  static constexpr bool shifted = false;
  __device__ static float load(float s, float) { return s; }
  __device__ static float store(float d, float) { return d; }
};

template <> struct PopStore<double> {
  using real = double;
  static constexpr bool shifted = false;
  __device__ static double load(double s, double) { return s; }
  __device__ static double store(double d, double) { return d; }
};

template <> struct PopStore<__half> {
  using real = float;
  static constexpr bool shifted = true;
  // inv_scale / scale are exact powers of two
  __device__ static float load(__half s, float inv_scale) { return __half2float(s) * inv_scale; }
  __device__ static __half store(float d, float scale) { return __float2half_rn(d * scale); }
};

}  // namespace fsi
