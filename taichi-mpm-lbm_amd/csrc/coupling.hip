// Cell-side tail of the coupled step: occupancy clamp, lattice force reset,
// coupling-force application and immersed-boundary drag in one pass.

#include "common.h"

namespace fsi {
namespace {

template <typename R> __device__ __forceinline__ R wave_sum(R v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ void wave_count(int32_t* ctr, bool pred) {
  const unsigned long long b = __ballot(pred);
  const unsigned long long act = __ballot(1);
  const int leader = __ffsll((long long)act) - 1;
  if ((int)__lane_id() == leader && b) atomicAdd(ctr, (int)__popcll(b));
}

template <typename R> struct AsmK {
  int64_t N;
  int coupling, clamp, ib, ib_use_max, write_force;
  R drag, thr, max_force;
  R base[3];
};

// fsi/coupling.py:408-411 (clamp), lbm3d.py:505-510 (reset_force),
// coupling.py:413-419 (apply coupling), 421-461 (IB compute + apply).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_assemble(AsmK<R> a, const int32_t* __restrict__ solid,
                                                     R* __restrict__ cforce,
                                                     const R* __restrict__ rho,
                                                     const R* __restrict__ vel,
                                                     R* __restrict__ force, R* __restrict__ ibf,
                                                     R* __restrict__ ib_total,
                                                     int32_t* __restrict__ counters) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = n < a.N;
  const int64_t N = a.N;
  bool active = false, clipped = false;
  R fi[3] = {R(0), R(0), R(0)};
  if (live) {
    R* cf = cforce + 4 * n;
    R phi = cf[3];
    if (a.clamp && phi > R(1)) {
      phi = R(1);
      cf[3] = phi;
    }
    const bool fluid = solid[n] == 0;
    R f[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] = fluid ? a.base[c] : R(0);
    if (a.coupling && fluid) {
#pragma unroll
      for (int c = 0; c < 3; ++c) f[c] += cf[c];
    }
    if (a.ib) {
      if (a.drag > R(0) && fluid && phi > R(0) && phi >= a.thr) {
        const R s = -a.drag * phi * rho[n];
#pragma unroll
        for (int c = 0; c < 3; ++c) fi[c] = s * vel[c * N + n];
        const R nrm = sqrt(fi[0] * fi[0] + fi[1] * fi[1] + fi[2] * fi[2]);
        if (a.ib_use_max && nrm > a.max_force && nrm > R(0)) {
#pragma unroll
          for (int c = 0; c < 3; ++c) fi[c] = fi[c] / nrm * a.max_force;
          clipped = true;
        }
        active = true;
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ibf[c * N + n] = fi[c];
        f[c] += fi[c];
      }
    }
    if (a.write_force) {
#pragma unroll
      for (int c = 0; c < 3; ++c) force[c * N + n] = f[c];
    }
  }
  if (a.ib) {
    wave_count(counters + 5, active);
    wave_count(counters + 6, clipped);
    const unsigned long long any = __ballot(active);
    if (any) {
      const R s0 = wave_sum(fi[0]), s1 = wave_sum(fi[1]), s2 = wave_sum(fi[2]);
      const unsigned long long act = __ballot(1);
      if ((int)__lane_id() == __ffsll((long long)act) - 1) {
        atomicAdd(ib_total + 0, s0);
        atomicAdd(ib_total + 1, s1);
        atomicAdd(ib_total + 2, s2);
      }
    }
  }
}

template <typename R>
__global__ void k_reset_ib(R* __restrict__ total, int32_t* __restrict__ counters) {
  if (threadIdx.x < 3) total[threadIdx.x] = R(0);
  if (threadIdx.x == 3) counters[5] = 0;
  if (threadIdx.x == 4) counters[6] = 0;
}

template <typename R>
int assemble_t(const fsi_lbm_params* lp, const fsi_lbm_buffers* lb, const fsi_force_assembly* fa,
               void* cforce, void* ib_force, void* ib_total, int32_t* counters, int write_force,
               hipStream_t st) {
  AsmK<R> a;
  a.N = (int64_t)lp->nx * lp->ny * lp->nz;
  a.coupling = fa->coupling;
  a.clamp = fa->clamp_fraction;
  a.ib = fa->ib_enabled;
  a.ib_use_max = fa->ib_use_max;
  a.write_force = write_force;
  a.drag = R(fa->ib_drag);
  a.thr = R(fa->ib_threshold);
  a.max_force = R(fa->ib_max_force);
  for (int c = 0; c < 3; ++c) a.base[c] = R(fa->base_force[c]);
  if (a.ib) {
    hipLaunchKernelGGL(k_reset_ib<R>, dim3(1), dim3(64), 0, st, (R*)ib_total, counters);
    FSI_TRY_LAUNCH("reset_ib");
  }
  hipLaunchKernelGGL(k_assemble<R>, dim3(blocks_for(a.N)), dim3(kBlock), 0, st, a, lb->solid,
                     (R*)cforce, (const R*)lb->rho, (const R*)lb->vel, (R*)lb->force, (R*)ib_force,
                     (R*)ib_total, counters);
  FSI_TRY_LAUNCH("force_assemble");
  return 0;
}

}  // namespace
}  // namespace fsi

using namespace fsi;

extern "C" int fsi_force_assemble(const fsi_lbm_params* lp, const fsi_lbm_buffers* lb,
                                  const fsi_force_assembly* a, void* cforce, void* ib_force,
                                  void* ib_total, int32_t* counters, int32_t write_force,
                                  void* stream) {
  FSI_REQUIRE(lp && lb && a && cforce && lb->solid, "force_assemble: null argument");
  FSI_REQUIRE(!write_force || lb->force, "force_assemble: null force");
  FSI_REQUIRE(!a->ib_enabled || (ib_force && ib_total && counters && lb->rho && lb->vel),
              "force_assemble: IB needs ib_force, ib_total, counters, rho and vel");
  hipStream_t st = (hipStream_t)stream;
  if (lp->storage == FSI_POP_F64)
    return assemble_t<double>(lp, lb, a, cforce, ib_force, ib_total, counters, write_force, st);
  return assemble_t<float>(lp, lb, a, cforce, ib_force, ib_total, counters, write_force, st);
}
