// D3Q19 MRT lattice Boltzmann kernels for gfx950.
//
// State layout in HBM (SoA, one plane per population):
//   pop[q * N + n], n = (i*ny + j)*nz + k   (k fastest, as the reference)
// so a wavefront of 64 consecutive cells reads and writes 64 consecutive
// entries of each population plane.
//
// One step is one fused kernel (collide + push-stream with half-way
// bounce-back) plus a small x-face boundary kernel. The populations a cell
// reads are its own, so the reads are perfectly coalesced; the pushes go to
// the 18 neighbours and are coalesced along k.
//
// Collision: f* = F + M^-1 r with r_k = -S_k (m_k - meq_k) + (1 - S_k/2) G_k,
// algebraically the reference's m = M F; m -= S (m - meq); m += (1 - S/2) G;
// f = M^-1 m (fsi/lbm3d.py:314-333). The moments use the even/odd structure of
// M under e -> -e (pair sums and differences of opposite populations), the
// Guo moments G_k use their closed form in (u, F), and M^-1 is the exact
// rational inverse; conserved moments (S_k = 0) keep r_k = G_k.

#include "common.h"
#include "mrt_tables.h"

namespace fsi {
namespace {

// relaxation class per moment, fsi/lbm3d.py:145-168: 0 conserved, 1 s_v, 2 s_o
__host__ __device__ constexpr int rate_kind(int k) {
  constexpr int t[19] = {0, 1, 1, 0, 2, 0, 2, 0, 2, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2};
  return t[k];
}

template <typename R> struct LbmK {
  int nx, ny, nz;
  int64_t N;
  R s_v, s_o;
  R rho0;
  R rho_base;   // rho0 for shifted storage, 0 for absolute storage
  R scale, inv_scale;
  R uf[3];      // uniform force when no field is given
};

template <typename R> LbmK<R> make_consts(const fsi_lbm_params* p, const fsi_lbm_force* f) {
  LbmK<R> c;
  c.nx = p->nx;
  c.ny = p->ny;
  c.nz = p->nz;
  c.N = (int64_t)p->nx * p->ny * p->nz;
  c.s_v = R(p->s_v);
  c.s_o = R(p->s_o);
  c.rho0 = R(p->rho0);
  const bool shifted = p->storage == FSI_POP_F16;
  c.rho_base = shifted ? R(p->rho0) : R(0);
  c.scale = shifted ? R(p->pop_scale) : R(1);
  c.inv_scale = shifted ? R(1.0 / p->pop_scale) : R(1);
  for (int a = 0; a < 3; ++a) c.uf[a] = f ? R(f->uniform[a]) : R(0);
  return c;
}

// Equilibrium minus the storage base. Absolute storage: the reference's
// feq = w rho (1 + 3eu + 4.5 eu^2 - 1.5 u.u) (fsi/lbm3d.py:250-254).
// Shifted storage: w (rho - rho0) + w rho (3eu + 4.5 eu^2 - 1.5 u.u), the same
// quantity minus w rho0 without the cancellation.
template <typename R, bool SHIFTED>
__device__ __forceinline__ R feq_dev(int q, R rho, R rho0, const R u[3]) {
  const R eu = R(ex(q)) * u[0] + R(ey(q)) * u[1] + R(ez(q)) * u[2];
  const R uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const R w = wq<R>(q);
  if constexpr (SHIFTED) {
    return w * (rho - rho0) + w * rho * (R(3) * eu + R(4.5) * eu * eu - R(1.5) * uu);
  } else {
    return w * rho * (R(1) + R(3) * eu + R(4.5) * eu * eu - R(1.5) * uu);
  }
}

template <typename R> __device__ __forceinline__ R frac(long long num, long long den) {
  return R(num) / R(den);
}

// MRT collision in place on d[19] (absolute or deviation populations).
template <typename R>
__device__ __forceinline__ void mrt_collide(R d[19], R rho, const R u[3], const R fo[3], R s_v,
                                            R s_o) {
  using namespace fsi_mrt;
  // opposite pairs are (2p+1, 2p+2)
  R ps[9], pd[9];
#pragma unroll
  for (int p = 0; p < 9; ++p) {
    ps[p] = d[2 * p + 1] + d[2 * p + 2];
    pd[p] = d[2 * p + 1] - d[2 * p + 2];
  }
  const R uf = u[0] * fo[0] + u[1] * fo[1] + u[2] * fo[2];
  const R uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  R r[19];
#pragma unroll
  for (int k = 0; k < 19; ++k) {
    // closed-form Guo moment G_k
    R g = R(0);
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (GA_NUM[k * 3 + a] != 0) g += frac<R>(GA_NUM[k * 3 + a], GA_DEN[k * 3 + a]) * fo[a];
    if (GB_NUM[k] != 0) g -= frac<R>(GB_NUM[k], GB_DEN[k]) * uf;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        if (GC_NUM[(k * 3 + a) * 3 + b] != 0)
          g += frac<R>(GC_NUM[(k * 3 + a) * 3 + b], GC_DEN[(k * 3 + a) * 3 + b]) * (u[a] * fo[b]);
    if (rate_kind(k) == 0) {
      r[k] = g;
      continue;
    }
    // moment m_k from pair sums (even rows) or differences (odd rows)
    R m = R(0);
    if (kOdd[k]) {
#pragma unroll
      for (int p = 0; p < 9; ++p)
        if (M[k][2 * p + 1] != 0) m += R(M[k][2 * p + 1]) * pd[p];
    } else {
      if (M[k][0] != 0) m += R(M[k][0]) * d[0];
#pragma unroll
      for (int p = 0; p < 9; ++p)
        if (M[k][2 * p + 1] != 0) m += R(M[k][2 * p + 1]) * ps[p];
    }
    // equilibrium moment, fsi/lbm3d.py:285-312
    R me = R(0);
    switch (k) {
      case 1: me = uu; break;
      case 9: me = R(2) * u[0] * u[0] - u[1] * u[1] - u[2] * u[2]; break;
      case 11: me = u[1] * u[1] - u[2] * u[2]; break;
      case 13: me = u[0] * u[1]; break;
      case 14: me = u[1] * u[2]; break;
      case 15: me = u[0] * u[2]; break;
      default: break;
    }
    const R s = rate_kind(k) == 1 ? s_v : s_o;
    r[k] = -s * (m - me) + (R(1) - R(0.5) * s) * g;
  }
  (void)rho;
  // d += M^-1 r, by even/odd parts per opposite pair
  {
    R e0 = R(0);
#pragma unroll
    for (int k = 0; k < 19; ++k)
      if (!kOdd[k] && INV_NUM[0][k] != 0) e0 += frac<R>(INV_NUM[0][k], INV_DEN[0][k]) * r[k];
    d[0] += e0;
  }
#pragma unroll
  for (int p = 0; p < 9; ++p) {
    const int q = 2 * p + 1;
    R ev = R(0), od = R(0);
#pragma unroll
    for (int k = 0; k < 19; ++k) {
      if (INV_NUM[q][k] == 0) continue;
      const R c = frac<R>(INV_NUM[q][k], INV_DEN[q][k]);
      if (kOdd[k])
        od += c * r[k];
      else
        ev += c * r[k];
    }
    d[q] += ev + od;
    d[q + 1] += ev - od;
  }
}

struct Nbr {
  int64_t ox[3], oy[3], oz[3];  // index offsets for e = -1, 0, +1 with periodic wrap
};

__device__ __forceinline__ Nbr neighbours(int i, int j, int k, int nx, int ny, int nz) {
  const int64_t sx = (int64_t)ny * nz, sy = nz;
  Nbr o;
  o.ox[0] = i == 0 ? (int64_t)(nx - 1) * sx : -sx;
  o.ox[1] = 0;
  o.ox[2] = i == nx - 1 ? -(int64_t)(nx - 1) * sx : sx;
  o.oy[0] = j == 0 ? (int64_t)(ny - 1) * sy : -sy;
  o.oy[1] = 0;
  o.oy[2] = j == ny - 1 ? -(int64_t)(ny - 1) * sy : sy;
  o.oz[0] = k == 0 ? (int64_t)(nz - 1) : -1;
  o.oz[1] = 0;
  o.oz[2] = k == nz - 1 ? -(int64_t)(nz - 1) : 1;
  return o;
}

__device__ __forceinline__ int64_t nbr_index(int64_t n, const Nbr& o, int q) {
  return n + o.ox[ex(q) + 1] + o.oy[ey(q) + 1] + o.oz[ez(q) + 1];
}

// fsi/lbm3d.py:256-267 (init_fields) plus the neighbour-solid bit masks.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_lbm_init(LbmK<typename PopStore<S>::real> c,
                                                     const int32_t* __restrict__ solid,
                                                     uint32_t* __restrict__ nbmask,
                                                     typename PopStore<S>::real* __restrict__ rho,
                                                     typename PopStore<S>::real* __restrict__ vel,
                                                     S* __restrict__ pa, S* __restrict__ pb,
                                                     typename PopStore<S>::real u0x,
                                                     typename PopStore<S>::real u0y,
                                                     typename PopStore<S>::real u0z) {
  using R = typename PopStore<S>::real;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= c.N) return;
  const int k = (int)(n % c.nz);
  const int64_t ij = n / c.nz;
  const int j = (int)(ij % c.ny);
  const int i = (int)(ij / c.ny);
  const Nbr o = neighbours(i, j, k, c.nx, c.ny, c.nz);
  uint32_t m = 0;
#pragma unroll
  for (int q = 0; q < 19; ++q)
    if (solid[nbr_index(n, o, q)] != 0) m |= 1u << q;
  nbmask[n] = m;
  const bool is_solid = (m & 1u) != 0;
  R u[3] = {is_solid ? R(0) : u0x, is_solid ? R(0) : u0y, is_solid ? R(0) : u0z};
  rho[n] = c.rho0;
  vel[n] = u[0];
  vel[c.N + n] = u[1];
  vel[2 * c.N + n] = u[2];
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const R v = feq_dev<R, PopStore<S>::shifted>(q, c.rho0, c.rho0, u);
    const S s = PopStore<S>::store(v, c.scale);
    pa[q * c.N + n] = s;
    pb[q * c.N + n] = s;
  }
}

// Fused collide + push stream (fsi/lbm3d.py:314-361) for one fluid cell.
template <typename S, bool FORCE_FIELD, bool MACRO_IN, bool FACE_VEL>
__global__ __launch_bounds__(kBlock) void k_lbm_step(
    LbmK<typename PopStore<S>::real> c, const S* __restrict__ pin, S* __restrict__ pout,
    const uint32_t* __restrict__ nbmask, const typename PopStore<S>::real* __restrict__ force,
    const typename PopStore<S>::real* __restrict__ rho_in,
    const typename PopStore<S>::real* __restrict__ vel_in,
    typename PopStore<S>::real* __restrict__ face_vel) {
  using R = typename PopStore<S>::real;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= c.N) return;
  const uint32_t mask = nbmask[n];
  if (mask & 1u) return;  // solid cells are never collided or streamed
  const int64_t N = c.N;

  R d[19];
#pragma unroll
  for (int q = 0; q < 19; ++q) d[q] = PopStore<S>::load(pin[q * N + n], c.inv_scale);

  R fo[3];
  if constexpr (FORCE_FIELD) {
    fo[0] = force[n];
    fo[1] = force[N + n];
    fo[2] = force[2 * N + n];
  } else {
    fo[0] = c.uf[0];
    fo[1] = c.uf[1];
    fo[2] = c.uf[2];
  }

  // macroscopic state of the previous step (fsi/lbm3d.py:409-425)
  R rho, u[3];
  if constexpr (MACRO_IN) {
    rho = rho_in[n];
    u[0] = vel_in[n];
    u[1] = vel_in[N + n];
    u[2] = vel_in[2 * N + n];
  } else {
    R sum = R(0), jx = R(0), jy = R(0), jz = R(0);
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      sum += d[q];
      if (ex(q)) jx += R(ex(q)) * d[q];
      if (ey(q)) jy += R(ey(q)) * d[q];
      if (ez(q)) jz += R(ez(q)) * d[q];
    }
    rho = c.rho_base + sum;
    const R inv = R(1) / rho;
    u[0] = jx * inv + (fo[0] * R(0.5)) * inv;
    u[1] = jy * inv + (fo[1] * R(0.5)) * inv;
    u[2] = jz * inv + (fo[2] * R(0.5)) * inv;
  }

  const int k = (int)(n % c.nz);
  const int64_t ij = n / c.nz;
  const int j = (int)(ij % c.ny);
  const int i = (int)(ij / c.ny);
  if constexpr (FACE_VEL) {
    if (i == 0 || i == c.nx - 1) {
      const int64_t plane = (int64_t)c.ny * c.nz;
      R* fv = face_vel + (i == 0 ? 0 : 3 * plane);
      const int64_t jk = (int64_t)j * c.nz + k;
      fv[jk] = u[0];
      fv[plane + jk] = u[1];
      fv[2 * plane + jk] = u[2];
    }
  }

  mrt_collide<R>(d, rho, u, fo, c.s_v, c.s_o);

  const Nbr o = neighbours(i, j, k, c.nx, c.ny, c.nz);
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const S s = PopStore<S>::store(d[q], c.scale);
    if (q == 0) {
      pout[n] = s;
    } else if (mask & (1u << q)) {
      pout[opp(q) * N + n] = s;  // half-way bounce-back into the source cell
    } else {
      pout[q * N + nbr_index(n, o, q)] = s;
    }
  }
}

// x-face boundaries on the streamed populations, fsi/lbm3d.py:363-407.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_lbm_bc(LbmK<typename PopStore<S>::real> c,
                                                   S* __restrict__ pop,
                                                   const int32_t* __restrict__ solid,
                                                   const typename PopStore<S>::real* __restrict__ face_vel,
                                                   int side, int mode,
                                                   typename PopStore<S>::real rho_bc,
                                                   typename PopStore<S>::real vx,
                                                   typename PopStore<S>::real vy,
                                                   typename PopStore<S>::real vz) {
  using R = typename PopStore<S>::real;
  constexpr bool SH = PopStore<S>::shifted;
  const int64_t plane = (int64_t)c.ny * c.nz;
  const int64_t jk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (jk >= plane) return;
  const int face = side == 0 ? 0 : c.nx - 1;
  const int inner = side == 0 ? 1 : c.nx - 2;
  const int64_t n = (int64_t)face * plane + jk;
  if (solid[n] != 0) return;
  const int64_t N = c.N;
  if (mode == FSI_BC_PRESSURE) {
    // velocity of the face cell from the previous macroscopic state, or of the
    // inner neighbour when that one is solid (a solid cell's velocity is 0)
    R u[3] = {R(0), R(0), R(0)};
    if (solid[(int64_t)inner * plane + jk] == 0) {
      const R* fv = face_vel + side * 3 * plane;
      u[0] = fv[jk];
      u[1] = fv[plane + jk];
      u[2] = fv[2 * plane + jk];
    }
#pragma unroll
    for (int q = 0; q < 19; ++q)
      pop[q * N + n] = PopStore<S>::store(feq_dev<R, SH>(q, rho_bc, c.rho0, u), c.scale);
  } else {
    // velocity: sequential in-place F_q = feq(opp q; rho0, u_bc) - F_opp + feq(q; rho0, u_bc)
    const R ub[3] = {vx, vy, vz};
    R d[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) d[q] = PopStore<S>::load(pop[q * N + n], c.inv_scale);
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      const int o = opp(q);
      d[q] = feq_dev<R, SH>(o, c.rho0, c.rho0, ub) - d[o] + feq_dev<R, SH>(q, c.rho0, c.rho0, ub);
    }
#pragma unroll
    for (int q = 0; q < 19; ++q) pop[q * N + n] = PopStore<S>::store(d[q], c.scale);
  }
}

// fsi/lbm3d.py:409-425 (update_macro) without the f = F copy.
template <typename S, bool FORCE_FIELD>
__global__ __launch_bounds__(kBlock) void k_lbm_macro(LbmK<typename PopStore<S>::real> c,
                                                      const S* __restrict__ pop,
                                                      const int32_t* __restrict__ solid,
                                                      const typename PopStore<S>::real* __restrict__ force,
                                                      typename PopStore<S>::real* __restrict__ rho,
                                                      typename PopStore<S>::real* __restrict__ vel) {
  using R = typename PopStore<S>::real;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= c.N) return;
  const int64_t N = c.N;
  if (solid[n] != 0) {
    rho[n] = c.rho0;
    vel[n] = R(0);
    vel[N + n] = R(0);
    vel[2 * N + n] = R(0);
    return;
  }
  R sum = R(0), jx = R(0), jy = R(0), jz = R(0);
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const R dq = PopStore<S>::load(pop[q * N + n], c.inv_scale);
    sum += dq;
    if (ex(q)) jx += R(ex(q)) * dq;
    if (ey(q)) jy += R(ey(q)) * dq;
    if (ez(q)) jz += R(ez(q)) * dq;
  }
  R fo[3];
  if constexpr (FORCE_FIELD) {
    fo[0] = force[n];
    fo[1] = force[N + n];
    fo[2] = force[2 * N + n];
  } else {
    fo[0] = c.uf[0];
    fo[1] = c.uf[1];
    fo[2] = c.uf[2];
  }
  const R r = c.rho_base + sum;
  const R inv = R(1) / r;
  rho[n] = r;
  vel[n] = jx * inv + (fo[0] * R(0.5)) * inv;
  vel[N + n] = jy * inv + (fo[1] * R(0.5)) * inv;
  vel[2 * N + n] = jz * inv + (fo[2] * R(0.5)) * inv;
}

// fsi/lbm3d.py:499-534
template <typename R>
__global__ __launch_bounds__(kBlock) void k_force_op(int64_t N, int op, R v0, R v1, R v2,
                                                     const int32_t* __restrict__ solid,
                                                     const R* src, R* dst) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const bool fluid = solid[n] == 0;
  const R val[3] = {v0, v1, v2};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    R out;
    switch (op) {
      case 0: out = R(0); break;
      case 1: out = fluid ? val[a] : R(0); break;
      case 2: out = fluid ? src[a * N + n] + val[a] : R(0); break;
      default: out = fluid ? src[a * N + n] : R(0); break;
    }
    dst[a * N + n] = out;
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_force_import(int64_t N, const int32_t* __restrict__ solid,
                                                         const float* __restrict__ aos,
                                                         R* __restrict__ dst) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const bool fluid = solid[n] == 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) dst[a * N + n] = fluid ? R(aos[n * 3 + a]) : R(0);
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_export_pop(LbmK<typename PopStore<S>::real> c,
                                                       const S* __restrict__ pop,
                                                       typename PopStore<S>::real* __restrict__ out) {
  using R = typename PopStore<S>::real;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= c.N) return;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const R dq = PopStore<S>::load(pop[q * c.N + n], c.inv_scale);
    out[n * 19 + q] = PopStore<S>::shifted ? wq<R>(q) * c.rho0 + dq : dq;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_soa_to_aos(int64_t n, int ncomp, const T* __restrict__ soa,
                                                       T* __restrict__ aos) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * ncomp) return;
  const int64_t i = t / ncomp;
  const int c = (int)(t % ncomp);
  aos[t] = soa[c * n + i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_aos_to_soa(int64_t n, int ncomp, const T* __restrict__ aos,
                                                       T* __restrict__ soa) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * ncomp) return;
  const int c = (int)(t / n);
  const int64_t i = t % n;
  soa[t] = aos[i * ncomp + c];
}

__global__ __launch_bounds__(kBlock) void k_fill_zero(int64_t n16, uint4* __restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n16) dst[t] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ void k_fill_zero_bytes(int64_t n, unsigned char* __restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = 0;
}

// ------------------------------------------------------- dispatch ----

int check_params(const fsi_lbm_params* p) {
  FSI_REQUIRE(p != nullptr, "lbm params are null");
  FSI_REQUIRE(p->nx > 4 && p->ny > 4 && p->nz > 4, "lattice dims must exceed 4, got %d %d %d",
              p->nx, p->ny, p->nz);
  FSI_REQUIRE(p->storage == FSI_POP_F32 || p->storage == FSI_POP_F16 || p->storage == FSI_POP_F64,
              "unknown population storage %d", p->storage);
  FSI_REQUIRE(p->left_mode >= 0 && p->left_mode <= 2 && p->right_mode >= 0 && p->right_mode <= 2,
              "unknown boundary mode");
  return 0;
}

template <typename S>
int lbm_init_t(const fsi_lbm_params* p, const fsi_lbm_buffers* b, void* pa, void* pb,
               hipStream_t st) {
  using R = typename PopStore<S>::real;
  LbmK<R> c = make_consts<R>(p, nullptr);
  hipLaunchKernelGGL(k_lbm_init<S>, dim3(blocks_for(c.N)), dim3(kBlock), 0, st, c, b->solid,
                     b->nbmask, (R*)b->rho, (R*)b->vel, (S*)pa, (S*)pb,
                     R(p->initial_velocity[0]), R(p->initial_velocity[1]),
                     R(p->initial_velocity[2]));
  FSI_TRY_LAUNCH("lbm_init");
  return 0;
}

template <typename S, bool FF, bool MI>
void launch_step(const LbmK<typename PopStore<S>::real>& c, bool face, const S* pin, S* pout,
                 const fsi_lbm_buffers* b, const void* force, hipStream_t st) {
  using R = typename PopStore<S>::real;
  if (face)
    hipLaunchKernelGGL((k_lbm_step<S, FF, MI, true>), dim3(blocks_for(c.N)), dim3(kBlock), 0, st,
                       c, pin, pout, b->nbmask, (const R*)force, (const R*)b->rho,
                       (const R*)b->vel, (R*)b->face_vel);
  else
    hipLaunchKernelGGL((k_lbm_step<S, FF, MI, false>), dim3(blocks_for(c.N)), dim3(kBlock), 0, st,
                       c, pin, pout, b->nbmask, (const R*)force, (const R*)b->rho,
                       (const R*)b->vel, (R*)b->face_vel);
}

template <typename S>
int lbm_step_t(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const void* pin, void* pout,
               const fsi_lbm_force* f, int macro_in, hipStream_t st) {
  using R = typename PopStore<S>::real;
  LbmK<R> c = make_consts<R>(p, f);
  const bool field = f && f->field;
  const bool face = p->left_mode == FSI_BC_PRESSURE || p->right_mode == FSI_BC_PRESSURE;
  FSI_REQUIRE(!face || b->face_vel, "pressure boundary needs a face velocity buffer");
  const S* in = (const S*)pin;
  S* out = (S*)pout;
  const void* fp = field ? f->field : nullptr;
  if (field && macro_in)
    launch_step<S, true, true>(c, face, in, out, b, fp, st);
  else if (field)
    launch_step<S, true, false>(c, face, in, out, b, fp, st);
  else if (macro_in)
    launch_step<S, false, true>(c, face, in, out, b, fp, st);
  else
    launch_step<S, false, false>(c, face, in, out, b, fp, st);
  FSI_TRY_LAUNCH("lbm_step");
  const int64_t plane = (int64_t)p->ny * p->nz;
  for (int side = 0; side < 2; ++side) {
    const int mode = side == 0 ? p->left_mode : p->right_mode;
    if (mode == FSI_BC_PERIODIC) continue;
    hipLaunchKernelGGL(k_lbm_bc<S>, dim3(blocks_for(plane)), dim3(kBlock), 0, st, c, out, b->solid,
                       (const R*)b->face_vel, side, mode, R(p->bc_rho[side]),
                       R(p->bc_vel[side][0]), R(p->bc_vel[side][1]), R(p->bc_vel[side][2]));
    FSI_TRY_LAUNCH("lbm_bc");
  }
  return 0;
}

template <typename S>
int lbm_macro_t(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const void* pop,
                const fsi_lbm_force* f, hipStream_t st) {
  using R = typename PopStore<S>::real;
  LbmK<R> c = make_consts<R>(p, f);
  if (f && f->field)
    hipLaunchKernelGGL((k_lbm_macro<S, true>), dim3(blocks_for(c.N)), dim3(kBlock), 0, st, c,
                       (const S*)pop, b->solid, (const R*)f->field, (R*)b->rho, (R*)b->vel);
  else
    hipLaunchKernelGGL((k_lbm_macro<S, false>), dim3(blocks_for(c.N)), dim3(kBlock), 0, st, c,
                       (const S*)pop, b->solid, (const R*)nullptr, (R*)b->rho, (R*)b->vel);
  FSI_TRY_LAUNCH("lbm_macro");
  return 0;
}

template <typename S>
int lbm_export_t(const fsi_lbm_params* p, const void* pop, void* out, hipStream_t st) {
  using R = typename PopStore<S>::real;
  LbmK<R> c = make_consts<R>(p, nullptr);
  hipLaunchKernelGGL(k_export_pop<S>, dim3(blocks_for(c.N)), dim3(kBlock), 0, st, c,
                     (const S*)pop, (R*)out);
  FSI_TRY_LAUNCH("lbm_export_populations");
  return 0;
}

}  // namespace
}  // namespace fsi

using namespace fsi;

extern "C" {

int fsi_lbm_init(const fsi_lbm_params* p, const fsi_lbm_buffers* b, void* pop_a, void* pop_b,
                 void* stream) {
  if (int rc = check_params(p)) return rc;
  FSI_REQUIRE(b && b->solid && b->nbmask && b->rho && b->vel && pop_a && pop_b,
              "lbm_init: null buffer");
  hipStream_t st = (hipStream_t)stream;
  switch (p->storage) {
    case FSI_POP_F32: return lbm_init_t<float>(p, b, pop_a, pop_b, st);
    case FSI_POP_F16: return lbm_init_t<__half>(p, b, pop_a, pop_b, st);
    default: return lbm_init_t<double>(p, b, pop_a, pop_b, st);
  }
}

int fsi_lbm_step(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const void* pop_in,
                 void* pop_out, const fsi_lbm_force* force, int32_t macro_in, void* stream) {
  if (int rc = check_params(p)) return rc;
  FSI_REQUIRE(b && b->solid && b->nbmask && pop_in && pop_out && pop_in != pop_out,
              "lbm_step: null or aliased buffer");
  FSI_REQUIRE(!macro_in || (b->rho && b->vel), "lbm_step: macro_in needs rho and vel");
  hipStream_t st = (hipStream_t)stream;
  switch (p->storage) {
    case FSI_POP_F32: return lbm_step_t<float>(p, b, pop_in, pop_out, force, macro_in, st);
    case FSI_POP_F16: return lbm_step_t<__half>(p, b, pop_in, pop_out, force, macro_in, st);
    default: return lbm_step_t<double>(p, b, pop_in, pop_out, force, macro_in, st);
  }
}

int fsi_lbm_macro(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const void* pop,
                  const fsi_lbm_force* force, void* stream) {
  if (int rc = check_params(p)) return rc;
  FSI_REQUIRE(b && b->solid && b->rho && b->vel && pop, "lbm_macro: null buffer");
  hipStream_t st = (hipStream_t)stream;
  switch (p->storage) {
    case FSI_POP_F32: return lbm_macro_t<float>(p, b, pop, force, st);
    case FSI_POP_F16: return lbm_macro_t<__half>(p, b, pop, force, st);
    default: return lbm_macro_t<double>(p, b, pop, force, st);
  }
}

int fsi_lbm_force_op(const fsi_lbm_params* p, const fsi_lbm_buffers* b, int32_t op,
                     const double* host_value, const void* src, void* dst, void* stream) {
  if (int rc = check_params(p)) return rc;
  FSI_REQUIRE(b && b->solid && dst, "lbm_force_op: null buffer");
  FSI_REQUIRE(op >= 0 && op <= 3, "lbm_force_op: unknown op %d", op);
  FSI_REQUIRE(op < 2 || src, "lbm_force_op: op %d reads src", op);
  const int64_t N = (int64_t)p->nx * p->ny * p->nz;
  const double zero[3] = {0, 0, 0};
  const double* v = host_value ? host_value : zero;
  hipStream_t st = (hipStream_t)stream;
  if (p->storage == FSI_POP_F64)
    hipLaunchKernelGGL(k_force_op<double>, dim3(blocks_for(N)), dim3(kBlock), 0, st, N, (int)op,
                       v[0], v[1], v[2], b->solid, (const double*)src, (double*)dst);
  else
    hipLaunchKernelGGL(k_force_op<float>, dim3(blocks_for(N)), dim3(kBlock), 0, st, N, (int)op,
                       (float)v[0], (float)v[1], (float)v[2], b->solid, (const float*)src,
                       (float*)dst);
  FSI_TRY_LAUNCH("lbm_force_op");
  return 0;
}

int fsi_lbm_force_import(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const float* aos,
                         void* stream) {
  if (int rc = check_params(p)) return rc;
  FSI_REQUIRE(b && b->solid && b->force && aos, "lbm_force_import: null buffer");
  const int64_t N = (int64_t)p->nx * p->ny * p->nz;
  hipStream_t st = (hipStream_t)stream;
  if (p->storage == FSI_POP_F64)
    hipLaunchKernelGGL(k_force_import<double>, dim3(blocks_for(N)), dim3(kBlock), 0, st, N,
                       b->solid, aos, (double*)b->force);
  else
    hipLaunchKernelGGL(k_force_import<float>, dim3(blocks_for(N)), dim3(kBlock), 0, st, N,
                       b->solid, aos, (float*)b->force);
  FSI_TRY_LAUNCH("lbm_force_import");
  return 0;
}

int fsi_lbm_export_populations(const fsi_lbm_params* p, const void* pop, void* out_aos,
                               void* stream) {
  if (int rc = check_params(p)) return rc;
  FSI_REQUIRE(pop && out_aos, "lbm_export_populations: null buffer");
  hipStream_t st = (hipStream_t)stream;
  switch (p->storage) {
    case FSI_POP_F32: return lbm_export_t<float>(p, pop, out_aos, st);
    case FSI_POP_F16: return lbm_export_t<__half>(p, pop, out_aos, st);
    default: return lbm_export_t<double>(p, pop, out_aos, st);
  }
}

int fsi_soa_to_aos(int32_t real, int64_t n, int32_t ncomp, const void* soa, void* aos,
                   void* stream) {
  FSI_REQUIRE(soa && aos && n >= 0 && ncomp > 0, "soa_to_aos: bad arguments");
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (real == FSI_REAL_F64)
    hipLaunchKernelGGL(k_soa_to_aos<double>, dim3(blocks_for(n * ncomp)), dim3(kBlock), 0, st, n,
                       (int)ncomp, (const double*)soa, (double*)aos);
  else
    hipLaunchKernelGGL(k_soa_to_aos<float>, dim3(blocks_for(n * ncomp)), dim3(kBlock), 0, st, n,
                       (int)ncomp, (const float*)soa, (float*)aos);
  FSI_TRY_LAUNCH("soa_to_aos");
  return 0;
}

int fsi_aos_to_soa(int32_t real, int64_t n, int32_t ncomp, const void* aos, void* soa,
                   void* stream) {
  FSI_REQUIRE(soa && aos && n >= 0 && ncomp > 0, "aos_to_soa: bad arguments");
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (real == FSI_REAL_F64)
    hipLaunchKernelGGL(k_aos_to_soa<double>, dim3(blocks_for(n * ncomp)), dim3(kBlock), 0, st, n,
                       (int)ncomp, (const double*)aos, (double*)soa);
  else
    hipLaunchKernelGGL(k_aos_to_soa<float>, dim3(blocks_for(n * ncomp)), dim3(kBlock), 0, st, n,
                       (int)ncomp, (const float*)aos, (float*)soa);
  FSI_TRY_LAUNCH("aos_to_soa");
  return 0;
}

int fsi_fill_zero(int64_t nbytes, void* dst, void* stream) {
  FSI_REQUIRE(dst && nbytes >= 0, "fill_zero: bad arguments");
  FSI_REQUIRE(((uintptr_t)dst & 15u) == 0, "fill_zero: destination must be 16-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n16 = nbytes / 16;
  if (n16 > 0)
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for(n16)), dim3(kBlock), 0, st, n16, (uint4*)dst);
  const int64_t tail = nbytes - n16 * 16;
  if (tail > 0)
    hipLaunchKernelGGL(k_fill_zero_bytes, dim3(1), dim3(64), 0, st, tail,
                       (unsigned char*)dst + n16 * 16);
  FSI_TRY_LAUNCH("fill_zero");
  return 0;
}

}  // extern "C"
