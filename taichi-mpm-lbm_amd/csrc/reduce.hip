// FP64 on-device reductions behind the diagnostics of the coupled step:
// the lattice sums of fsi/lbm3d.py:449-456 and fsi/coupling.py:241-244, and
// the particle sums of fsi/mpm3d.py:353-368, fsi/coupling.py:238-239 and
// 263-270. The reference copies whole fields to the host and reduces them in
// float32 numpy; here one fixed grid of workgroups writes FP64 partials in a
// fixed order and one workgroup combines them, so a result is reproducible
// bit for bit and carries no FP32 accumulation error.
#include <algorithm>

#include "common.h"

namespace fsi {
namespace {

constexpr int kSlots = FSI_SUM_SLOTS;
constexpr int kMaxBlocks = 1024;

enum Kind { kSum = 0, kMax = 1, kMin = 2 };

// max/min propagate NaN, as numpy's max/min of the reference's host
// reductions do (a blown-up field must show in the diagnostics)
__device__ __forceinline__ double combine(int kind, double a, double b) {
  if (kind == kSum) return a + b;
  if (a != a) return a;
  if (b != b) return b;
  return kind == kMax ? (a > b ? a : b) : (a < b ? a : b);
}

__device__ __forceinline__ double identity(int kind) {
  return kind == kSum ? 0.0 : (kind == kMax ? -HUGE_VAL : HUGE_VAL);
}

// Reduces v[slot] over the workgroup (fixed shuffle/LDS tree) and writes the
// result of thread 0 to dst[slot].
template <int NS>
__device__ void block_reduce(double (&v)[NS], const int (&kind)[NS], double* dst) {
  __shared__ double part[NS][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[s] = combine(kind[s], v[s], __shfl_xor(v[s], off));
    if (lane == 0) part[s][wid] = v[s];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      double r = part[s][0];
      for (int w = 1; w < kBlock / 64; ++w) r = combine(kind[s], r, part[s][w]);
      dst[s] = r;
    }
  }
}

// Lattice partials. Slots: 0 fluid mass, 1 max |u|^2 (all cells), 2..4
// coupling force summed over fluid cells (cforce AoS: xyz + phi per cell).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_lattice_partials(int64_t N, const int32_t* __restrict__ solid,
                                                             const R* __restrict__ rho,
                                                             const R* __restrict__ vel,
                                                             const R* __restrict__ cforce,
                                                             double* __restrict__ partials) {
  constexpr int NS = 5;
  const int kind[NS] = {kSum, kMax, kSum, kSum, kSum};
  double v[NS] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += stride) {
    const bool fluid = solid[n] == 0;
    const double ux = vel[n], uy = vel[N + n], uz = vel[2 * N + n];
    v[1] = combine(kMax, v[1], ux * ux + uy * uy + uz * uz);
    if (fluid) {
      v[0] += (double)rho[n];
      if (cforce) {
        v[2] += (double)cforce[4 * n + 0];
        v[3] += (double)cforce[4 * n + 1];
        v[4] += (double)cforce[4 * n + 2];
      }
    }
  }
  block_reduce<NS>(v, kind, partials + (int64_t)blockIdx.x * kSlots);
}

// Particle partials over the first `num` particles (active or not, as the
// reference's count-trimmed accessors). Slots: 0 sum m, 1..3 sum m x,
// 4 max |v|^2, 5..7 sum of particle force, 8 min valid weight, 9 sum of valid
// weights.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_particle_partials(int num, int cap, const R* __restrict__ x,
                                                              const R* __restrict__ v,
                                                              const R* __restrict__ mass,
                                                              const R* __restrict__ pforce,
                                                              const R* __restrict__ wvalid,
                                                              double* __restrict__ partials) {
  constexpr int NS = 10;
  const int kind[NS] = {kSum, kSum, kSum, kSum, kMax, kSum, kSum, kSum, kMin, kSum};
  double s[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) s[k] = identity(kind[k]);
  const int stride = gridDim.x * blockDim.x;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < num; p += stride) {
    const double m = mass[p];
    s[0] += m;
#pragma unroll
    for (int a = 0; a < 3; ++a) s[1 + a] += m * (double)x[a * cap + p];
    const double vx = v[p], vy = v[cap + p], vz = v[2 * cap + p];
    s[4] = combine(kMax, s[4], vx * vx + vy * vy + vz * vz);
#pragma unroll
    for (int a = 0; a < 3; ++a) s[5 + a] += (double)pforce[a * cap + p];
    if (wvalid) {
      const double w = wvalid[p];
      s[8] = combine(kMin, s[8], w);
      s[9] += w;
    }
  }
  block_reduce<NS>(s, kind, partials + (int64_t)blockIdx.x * kSlots);
}

struct LatticeKinds {
  static constexpr int n = 5;
  static __device__ int kind(int s) { return s == 1 ? kMax : kSum; }
};
struct ParticleKinds {
  static constexpr int n = 10;
  static __device__ int kind(int s) { return s == 4 ? kMax : (s == 8 ? kMin : kSum); }
};

// Combines `blocks` rows of partials in block order into out[0..n): one
// thread per slot walks the rows in order, so the result is deterministic.
template <typename K>
__global__ __launch_bounds__(64) void k_final(int blocks, const double* __restrict__ partials,
                                              double* __restrict__ out) {
  const int s = threadIdx.x;
  if (s >= K::n) return;
  const int kind = K::kind(s);
  double r = identity(kind);
  for (int b = 0; b < blocks; ++b) r = combine(kind, r, partials[(int64_t)b * kSlots + s]);
  out[s] = r;
}

int workspace_blocks(int64_t n, int64_t workspace_doubles) {
  int64_t blocks = std::min<int64_t>(std::max<int64_t>(blocks_for(n), 1), kMaxBlocks);
  return (int)std::min<int64_t>(blocks, workspace_doubles / kSlots);
}

template <typename R>
int lattice_t(int64_t N, const fsi_lbm_buffers* b, const void* cforce, double* ws, int blocks,
              double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_lattice_partials<R>, dim3(blocks), dim3(kBlock), 0, st, N, b->solid, (const R*)b->rho,
                     (const R*)b->vel, (const R*)cforce, ws);
  FSI_TRY_LAUNCH("reduce_lattice partials");
  hipLaunchKernelGGL(k_final<LatticeKinds>, dim3(1), dim3(64), 0, st, blocks, (const double*)ws, out);
  FSI_TRY_LAUNCH("reduce_lattice final");
  return 0;
}

template <typename R>
int particle_t(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const void* wvalid, double* ws,
               int blocks, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_particle_partials<R>, dim3(blocks), dim3(kBlock), 0, st, p->num_particles, p->capacity,
                     (const R*)b->x, (const R*)b->v, (const R*)b->mass, (const R*)b->pforce,
                     (const R*)wvalid, ws);
  FSI_TRY_LAUNCH("reduce_particles partials");
  hipLaunchKernelGGL(k_final<ParticleKinds>, dim3(1), dim3(64), 0, st, blocks, (const double*)ws, out);
  FSI_TRY_LAUNCH("reduce_particles final");
  return 0;
}

}  // namespace
}  // namespace fsi

extern "C" {

int fsi_reduce_lattice(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const void* cforce,
                       void* workspace, int64_t workspace_doubles, double* out, void* stream) {
  using namespace fsi;
  FSI_REQUIRE(p && b && b->solid && b->rho && b->vel && out && workspace, "reduce_lattice: null argument");
  FSI_REQUIRE(workspace_doubles >= kSlots, "reduce_lattice: workspace holds fewer than %d doubles", kSlots);
  const int64_t N = (int64_t)p->nx * p->ny * p->nz;
  const int blocks = workspace_blocks(N, workspace_doubles);
  hipStream_t st = (hipStream_t)stream;
  if (p->storage == FSI_POP_F64) return lattice_t<double>(N, b, cforce, (double*)workspace, blocks, out, st);
  return lattice_t<float>(N, b, cforce, (double*)workspace, blocks, out, st);
}

int fsi_reduce_particles(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const void* valid_weight,
                         void* workspace, int64_t workspace_doubles, double* out, void* stream) {
  using namespace fsi;
  FSI_REQUIRE(p && b && b->x && b->v && b->mass && b->pforce && out && workspace,
              "reduce_particles: null argument");
  FSI_REQUIRE(p->num_particles >= 0 && p->num_particles <= p->capacity, "reduce_particles: bad particle count");
  FSI_REQUIRE(workspace_doubles >= kSlots, "reduce_particles: workspace holds fewer than %d doubles", kSlots);
  const int blocks = workspace_blocks(p->num_particles, workspace_doubles);
  hipStream_t st = (hipStream_t)stream;
  if (p->real == FSI_REAL_F64)
    return particle_t<double>(p, b, valid_weight, (double*)workspace, blocks, out, st);
  return particle_t<float>(p, b, valid_weight, (double*)workspace, blocks, out, st);
}

}  // extern "C"
