// D3Q19 MRT lattice Boltzmann kernels for gfx950.
//
// State layout in HBM (SoA, one plane per population):
//   pop[q * N + n], n = (i*ny + j)*nz + k   (k fastest, as the reference)
// so a wavefront of 64 consecutive cells touches 64 consecutive entries of
// every population plane.
//
// Pull scheme. Between steps the lattice is in one of two states:
//   post: pop holds post-collision populations f* of the last step;
//   mat:  pop holds the reference's post-stream/post-boundary populations f,
//         and rho/vel hold the macroscopic fields (fsi/lbm3d.py:409-425).
// One kernel, k_lbm_pull, serves three phases:
//   STREAM_COLLIDE (post -> post): gather f* from the neighbours (half-way
//       bounce-back from the cell's own opposite slot when the neighbour is
//       solid), x-face boundaries, macroscopic fields, collision with the
//       current force, write the cell's own 19 slots;
//   STREAM_MACRO (post -> mat): the same gather and boundaries, then write f,
//       rho and u (accessors, the coupler, and any force change need them);
//   COLLIDE (mat -> post): collide the stored f with the stored rho/u.
// Every lane, solid ones included, executes the same instruction stream and
// writes only its own cell's slots (solid cells rewrite their constant rest
// populations): exec-mask divergence around memory instructions was measured
// to halve the achieved bandwidth on porous masks.
//
// Collision: f* = F + M^-1 r with r_k = -S_k (m_k - meq_k) + (1 - S_k/2) G_k,
// algebraically the reference's m = M F; m -= S (m - meq); m += (1 - S/2) G;
// f = M^-1 m (fsi/lbm3d.py:314-333). The moments use the even/odd structure of
// M under e -> -e (pair sums and differences of opposite populations), the
// Guo moments G_k their closed form in (u, F), and M^-1 = M^T (M M^T)^-1 with
// the block-diagonal exact Gram inverse; conserved moments (S_k = 0) keep
// r_k = G_k.

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
  R cv, co;     // 1 - s/2 for the two relaxation rates
  R rho0;
  R rho_base;   // rho0 for shifted storage, 0 for absolute storage
  R scale, inv_scale;
  R uf[3];      // uniform force when no field is given
  int mode[2];  // x-face boundary modes (left, right)
  R bc_rho[2];
  R bc_vel[2][3];
};

template <typename R> LbmK<R> make_consts(const fsi_lbm_params* p, const fsi_lbm_force* f) {
  LbmK<R> c;
  c.nx = p->nx;
  c.ny = p->ny;
  c.nz = p->nz;
  c.N = (int64_t)p->nx * p->ny * p->nz;
  c.s_v = R(p->s_v);
  c.s_o = R(p->s_o);
  c.cv = R(1) - R(0.5) * c.s_v;
  c.co = R(1) - R(0.5) * c.s_o;
  c.rho0 = R(p->rho0);
  const bool shifted = p->storage == FSI_POP_F16;
  c.rho_base = shifted ? R(p->rho0) : R(0);
  c.scale = shifted ? R(p->pop_scale) : R(1);
  c.inv_scale = shifted ? R(1.0 / p->pop_scale) : R(1);
  for (int a = 0; a < 3; ++a) c.uf[a] = f ? R(f->uniform[a]) : R(0);
  c.mode[0] = p->left_mode;
  c.mode[1] = p->right_mode;
  for (int sd = 0; sd < 2; ++sd) {
    c.bc_rho[sd] = R(p->bc_rho[sd]);
    for (int a = 0; a < 3; ++a) c.bc_vel[sd][a] = R(p->bc_vel[sd][a]);
  }
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

// MRT collision in place on d[19] (absolute or deviation populations), given
// the opposite-pair sums ps[p] = d[2p+1] + d[2p+2] and differences pd[p].
// cv/co are 1 - s_v/2 and 1 - s_o/2.
template <typename R>
__device__ __forceinline__ void mrt_collide(R d[19], const R ps[9], const R pd[9], const R u[3],
                                            const R fo[3], R s_v, R s_o, R cv, R co) {
  using namespace fsi_mrt;
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
    const bool sv = rate_kind(k) == 1;
    r[k] = (sv ? cv : co) * g - (sv ? s_v : s_o) * (m - me);
  }
  // M^-1 r = M^T (M M^T)^-1 r: block-diagonal Gram inverse, then integer M^T
  R y[19];
#pragma unroll
  for (int a = 0; a < 19; ++a) {
    R acc = R(0);
#pragma unroll
    for (int b = 0; b < 19; ++b)
      if (GINV_NUM[a][b] != 0) acc += frac<R>(GINV_NUM[a][b], GINV_DEN[a][b]) * r[b];
    y[a] = acc;
  }
  {
    R e0 = R(0);
#pragma unroll
    for (int k = 0; k < 19; ++k)
      if (!kOdd[k] && M[k][0] != 0) e0 += R(M[k][0]) * y[k];
    d[0] += e0;
  }
#pragma unroll
  for (int p = 0; p < 9; ++p) {
    const int q = 2 * p + 1;
    R ev = R(0), od = R(0);
#pragma unroll
    for (int k = 0; k < 19; ++k) {
      if (M[k][q] == 0) continue;
      if (kOdd[k])
        od += R(M[k][q]) * y[k];
      else
        ev += R(M[k][q]) * y[k];
    }
    d[q] += ev + od;
    d[q + 1] += ev - od;
  }
}

template <typename R> __device__ __forceinline__ R recip(R x) { return R(1) / x; }
template <> __device__ __forceinline__ float recip<float>(float x) {
  return __builtin_amdgcn_rcpf(x);  // v_rcp_f32, 1 ulp
}

// 32-bit byte offsets from a uniform base: the compiler emits SGPR-base +
// VGPR-offset addressing instead of 64-bit per-lane address arithmetic.
template <typename T> __device__ __forceinline__ T ld_off(const T* base, uint32_t byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}
template <typename T> __device__ __forceinline__ void st_off(T* base, uint32_t byte_off, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off) = v;
}
// Streaming store of a population: the next step reads it from HBM, not
// from this step's caches (a 19-plane copy measured 3-4% faster with
// nontemporal stores).
template <typename T> __device__ __forceinline__ void st_stream(T* base, uint32_t byte_off, T v) {
  __builtin_nontemporal_store(v, reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off));
}
__device__ __forceinline__ void st_stream(__half* base, uint32_t byte_off, __half v) {
  __builtin_nontemporal_store(__half_as_ushort(v),
                              reinterpret_cast<unsigned short*>(reinterpret_cast<char*>(base) + byte_off));
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
                                                     typename PopStore<S>::real* __restrict__ face_vel,
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
  if (face_vel && (i == 0 || i == c.nx - 1)) {
    // slot 0 holds the velocity the first step's pressure faces read
    const int64_t plane = (int64_t)c.ny * c.nz;
    R* fv = face_vel + (i == 0 ? 0 : 3 * plane);
    const int64_t jk = (int64_t)j * c.nz + k;
    for (int a = 0; a < 3; ++a) fv[a * plane + jk] = u[a];
  }
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const R v = feq_dev<R, PopStore<S>::shifted>(q, c.rho0, c.rho0, u);
    const S s = PopStore<S>::store(v, c.scale);
    pa[q * c.N + n] = s;
    pb[q * c.N + n] = s;
  }
}

enum { PHASE_STREAM_COLLIDE = 0, PHASE_COLLIDE = 1, PHASE_STREAM_MACRO = 2 };

// Rest population of a solid cell in storage units: w_q rho0 (absolute) or 0
// (shifted). Solid cells keep it for ever (the reference never updates them).
template <typename S>
__device__ __forceinline__ S solid_value(int q, typename PopStore<S>::real rho0,
                                         typename PopStore<S>::real scale) {
  using R = typename PopStore<S>::real;
  return PopStore<S>::store(PopStore<S>::shifted ? R(0) : wq<R>(q) * rho0, scale);
}

// x-face boundary of one face cell on its streamed populations d[19]
// (fsi/lbm3d.py:363-407). prev_u: the face cell's macroscopic velocity of the
// previous step; inner_solid: the inner neighbour is solid (its velocity is 0).
template <typename R, bool SH>
__device__ __forceinline__ void x_face_boundary(R d[19], int mode, R rho_bc, const R ubc[3],
                                                R rho0, const R prev_u[3], bool inner_solid) {
  if (mode == FSI_BC_PRESSURE) {
    R u[3] = {prev_u[0], prev_u[1], prev_u[2]};
    if (inner_solid) u[0] = u[1] = u[2] = R(0);
#pragma unroll
    for (int q = 0; q < 19; ++q) d[q] = feq_dev<R, SH>(q, rho_bc, rho0, u);
  } else if (mode == FSI_BC_VELOCITY) {
    // sequential in-place F_q = feq(opp q; rho0, u_bc) - F_opp + feq(q; rho0, u_bc)
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      const int o = opp(q);
      d[q] = feq_dev<R, SH>(o, rho0, rho0, ubc) - d[o] + feq_dev<R, SH>(q, rho0, rho0, ubc);
    }
  }
}

// One lattice update of one cell in the given phase (see the file comment).
// FACES: some x face is not periodic. Requires 19 * N * sizeof(S) < 2^32.
// Workgroup size of the lattice step. A workgroup reads 19 contiguous runs
// whose z-shifted ones straddle one extra cache line at its ends; 1024 lanes
// amortise that for FP32 (measured 2% faster at 256^3), while FP16 runs best
// at 256.
template <typename S> struct PullBlock { static constexpr int value = 256; };
template <> struct PullBlock<float> { static constexpr int value = 1024; };

template <typename S, int PHASE, bool FORCE_FIELD, bool FACES>
__global__ __launch_bounds__(PullBlock<S>::value) void k_lbm_pull(
    LbmK<typename PopStore<S>::real> c, const S* __restrict__ pin, S* __restrict__ pout,
    const uint32_t* __restrict__ nbmask, const typename PopStore<S>::real* __restrict__ force,
    typename PopStore<S>::real* __restrict__ rho_io, typename PopStore<S>::real* __restrict__ vel_io,
    const typename PopStore<S>::real* __restrict__ fv_prev,
    typename PopStore<S>::real* __restrict__ fv_next) {
  using R = typename PopStore<S>::real;
  constexpr bool SH = PopStore<S>::shifted;
  const uint32_t N = (uint32_t)c.N;
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  constexpr uint32_t B = sizeof(S);
  const uint32_t plane = N * B;  // byte size of one population plane
  const uint32_t own = n * B;
  const uint32_t mask = nbmask[n];
  const bool solid = (mask & 1u) != 0;

  const uint32_t nz = (uint32_t)c.nz, ny = (uint32_t)c.ny, nx = (uint32_t)c.nx;
  const uint32_t k = n % nz;
  const uint32_t ij = n / nz;
  const uint32_t j = ij % ny;
  const uint32_t i = ij / ny;

  R d[19];
  if constexpr (PHASE == PHASE_COLLIDE) {
#pragma unroll
    for (int q = 0; q < 19; ++q) d[q] = PopStore<S>::load(ld_off(pin, q * plane + own), c.inv_scale);
  } else {
    // gather f*(c - e_q, q), or f*(c, opp q) when c - e_q = c + e_opp(q) is
    // solid (fsi/lbm3d.py:352-361 read backwards). One-cell periodic wrap on
    // every axis (lbm3d.py:335-350). A solid lane's values are discarded; it
    // reads the plain pull sources, which lie in the cache lines its fluid
    // neighbours in the wavefront read anyway (its own slots would be extra
    // lines for the 16 planes shifted in x or y).
    const uint32_t sx = ny * nz;
    const uint32_t oxm = i == 0 ? (nx - 1) * sx : (uint32_t)(-(int32_t)sx);
    const uint32_t oxp = i == nx - 1 ? (uint32_t)(-(int32_t)((nx - 1) * sx)) : sx;
    const uint32_t oym = j == 0 ? (ny - 1) * nz : (uint32_t)(-(int32_t)nz);
    const uint32_t oyp = j == ny - 1 ? (uint32_t)(-(int32_t)((ny - 1) * nz)) : nz;
    const uint32_t ozm = k == 0 ? nz - 1 : (uint32_t)(-1);
    const uint32_t ozp = k == nz - 1 ? (uint32_t)(-(int32_t)(nz - 1)) : 1u;
    const uint32_t all = solid ? 0u : mask;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      uint32_t src = n;  // c - e_q
      if (ex(q) > 0) src += oxm;
      if (ex(q) < 0) src += oxp;
      if (ey(q) > 0) src += oym;
      if (ey(q) < 0) src += oyp;
      if (ez(q) > 0) src += ozm;
      if (ez(q) < 0) src += ozp;
      const uint32_t a0 = q * plane + src * B;
      const uint32_t a1 = opp(q) * plane + own;
      const uint32_t sel = 0u - ((all >> opp(q)) & 1u);
      d[q] = PopStore<S>::load(ld_off(pin, a0 ^ ((a0 ^ a1) & sel)), c.inv_scale);
    }
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

  bool face = false;
  int side = 0;
  uint32_t jk = 0;
  const uint32_t fplane = ny * nz;
  if constexpr (FACES && PHASE != PHASE_COLLIDE) {
    // x-face boundaries of the streamed populations, with the previous
    // macroscopic velocity of the face cell
    side = i == 0 ? 0 : 1;
    face = (i == 0 && c.mode[0] != FSI_BC_PERIODIC) || (i == nx - 1 && c.mode[1] != FSI_BC_PERIODIC);
    jk = j * nz + k;
    if (face && !solid) {
      const R* fp = fv_prev + side * 3 * fplane;
      const R pu[3] = {fp[jk], fp[fplane + jk], fp[2 * fplane + jk]};
      const bool inner_solid = (mask >> (side == 0 ? 1 : 2)) & 1u;
      x_face_boundary<R, SH>(d, c.mode[side], c.bc_rho[side], c.bc_vel[side], c.rho0, pu, inner_solid);
    }
  }

  R ps[9], pd[9];
#pragma unroll
  for (int p = 0; p < 9; ++p) {
    ps[p] = d[2 * p + 1] + d[2 * p + 2];
    pd[p] = d[2 * p + 1] - d[2 * p + 2];
  }

  // macroscopic state (fsi/lbm3d.py:409-425): stored, or from the streamed
  // populations with the force the previous collision used
  R rho, u[3];
  if constexpr (PHASE == PHASE_COLLIDE) {
    rho = rho_io[n];
    u[0] = vel_io[n];
    u[1] = vel_io[N + n];
    u[2] = vel_io[2 * N + n];
  } else {
    R sum = d[0];
    R jm[3] = {R(0), R(0), R(0)};
#pragma unroll
    for (int p = 0; p < 9; ++p) {
      sum += ps[p];
      const int q = 2 * p + 1;
      if (ex(q)) jm[0] += R(ex(q)) * pd[p];
      if (ey(q)) jm[1] += R(ey(q)) * pd[p];
      if (ez(q)) jm[2] += R(ez(q)) * pd[p];
    }
    rho = c.rho_base + sum;
    const R inv = recip(rho);
#pragma unroll
    for (int a = 0; a < 3; ++a) u[a] = (jm[a] + fo[a] * R(0.5)) * inv;
    if (solid) {
      rho = c.rho0;
      u[0] = u[1] = u[2] = R(0);
    }
    if constexpr (FACES) {
      if (face) {
        R* fn = fv_next + side * 3 * fplane;
        fn[jk] = u[0];
        fn[fplane + jk] = u[1];
        fn[2 * fplane + jk] = u[2];
      }
    }
  }

  if constexpr (PHASE == PHASE_STREAM_MACRO) {
    rho_io[n] = rho;
    vel_io[n] = u[0];
    vel_io[N + n] = u[1];
    vel_io[2 * N + n] = u[2];
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      const S v = solid ? solid_value<S>(q, c.rho0, c.scale) : PopStore<S>::store(d[q], c.scale);
      st_stream(pout, q * plane + own, v);
    }
    return;
  }

  mrt_collide<R>(d, ps, pd, u, fo, c.s_v, c.s_o, c.cv, c.co);
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const S v = solid ? solid_value<S>(q, c.rho0, c.scale) : PopStore<S>::store(d[q], c.scale);
    st_stream(pout, q * plane + own, v);
  }
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
                     b->nbmask, (R*)b->rho, (R*)b->vel, (S*)pa, (S*)pb, (R*)b->face_vel,
                     R(p->initial_velocity[0]), R(p->initial_velocity[1]),
                     R(p->initial_velocity[2]));
  FSI_TRY_LAUNCH("lbm_init");
  return 0;
}

template <typename S, int PHASE, bool FF>
void launch_pull(const LbmK<typename PopStore<S>::real>& c, bool faces, const void* pin, void* pout,
                 const fsi_lbm_buffers* b, const void* force, const void* fv_prev, void* fv_next,
                 hipStream_t st) {
  using R = typename PopStore<S>::real;
  const dim3 grid(blocks_for(c.N, PullBlock<S>::value)), block(PullBlock<S>::value);
  if (faces)
    hipLaunchKernelGGL((k_lbm_pull<S, PHASE, FF, true>), grid, block, 0, st, c, (const S*)pin,
                       (S*)pout, b->nbmask, (const R*)force, (R*)b->rho, (R*)b->vel,
                       (const R*)fv_prev, (R*)fv_next);
  else
    hipLaunchKernelGGL((k_lbm_pull<S, PHASE, FF, false>), grid, block, 0, st, c, (const S*)pin,
                       (S*)pout, b->nbmask, (const R*)force, (R*)b->rho, (R*)b->vel,
                       (const R*)fv_prev, (R*)fv_next);
}

template <typename S>
int lbm_step_t(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const void* pin, void* pout,
               const fsi_lbm_force* f, int phase, int face_slot, hipStream_t st) {
  using R = typename PopStore<S>::real;
  LbmK<R> c = make_consts<R>(p, f);
  FSI_REQUIRE(19.0 * (double)c.N * sizeof(S) < 4294967296.0,
              "lbm_step: lattice of %lld cells too large for 32-bit plane offsets", (long long)c.N);
  const bool faces = p->left_mode != FSI_BC_PERIODIC || p->right_mode != FSI_BC_PERIODIC;
  FSI_REQUIRE(!faces || b->face_vel, "lbm_step: non-periodic x faces need the face velocity buffer");
  FSI_REQUIRE(phase != FSI_PHASE_COLLIDE || (b->rho && b->vel), "lbm_step: collide needs rho and vel");
  FSI_REQUIRE(phase != FSI_PHASE_STREAM_MACRO || (b->rho && b->vel), "lbm_step: macro needs rho and vel");
  const int64_t fsize = 6LL * p->ny * p->nz;
  const R* fvp = faces ? (const R*)b->face_vel + (face_slot ? fsize : 0) : nullptr;
  R* fvn = faces ? (R*)b->face_vel + (face_slot ? 0 : fsize) : nullptr;
  const bool field = f && f->field;
  const void* fp = field ? f->field : nullptr;
#define FSI_PULL(PH)                                                      \
  do {                                                                    \
    if (field)                                                            \
      launch_pull<S, PH, true>(c, faces, pin, pout, b, fp, fvp, fvn, st); \
    else                                                                  \
      launch_pull<S, PH, false>(c, faces, pin, pout, b, fp, fvp, fvn, st); \
  } while (0)
  switch (phase) {
    case FSI_PHASE_STREAM_COLLIDE: FSI_PULL(PHASE_STREAM_COLLIDE); break;
    case FSI_PHASE_COLLIDE: FSI_PULL(PHASE_COLLIDE); break;
    case FSI_PHASE_STREAM_MACRO: FSI_PULL(PHASE_STREAM_MACRO); break;
    default: FSI_REQUIRE(false, "lbm_step: unknown phase %d", phase);
  }
#undef FSI_PULL
  FSI_TRY_LAUNCH("lbm_step");
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
                 void* pop_out, const fsi_lbm_force* force, int32_t phase, int32_t face_slot,
                 void* stream) {
  if (int rc = check_params(p)) return rc;
  FSI_REQUIRE(b && b->nbmask && pop_in && pop_out && pop_in != pop_out,
              "lbm_step: null or aliased buffer");
  hipStream_t st = (hipStream_t)stream;
  switch (p->storage) {
    case FSI_POP_F32: return lbm_step_t<float>(p, b, pop_in, pop_out, force, phase, face_slot, st);
    case FSI_POP_F16: return lbm_step_t<__half>(p, b, pop_in, pop_out, force, phase, face_slot, st);
    default: return lbm_step_t<double>(p, b, pop_in, pop_out, force, phase, face_slot, st);
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
