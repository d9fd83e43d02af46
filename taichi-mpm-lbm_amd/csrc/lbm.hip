// D3Q19 MRT lattice Boltzmann kernels for gfx950.
//
// State layout in HBM: populations blocked by lattice row. Cell n =
// (i*ny + j)*nz + k (k fastest, as the reference) keeps population q at
//   pop[((i*ny + j) * 19 + q) * nz + k]
// i.e. each z-row holds its 19 population runs of nz values back to back, so
// a wavefront of 64 consecutive cells touches 64 consecutive entries of each
// run, and a workgroup's stores fill one contiguous 19*nz block per row. On
// the pull pattern this layout moves its bytes 3% (FP32) to 9% (16-bit)
// faster than 19 lattice-sized planes: the workgroups in flight stream from
// ~9 row blocks instead of 19 planes far apart. The macroscopic fields, the
// force and the neighbour mask stay plain SoA planes indexed by n.
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
// Collision (fsi/lbm3d.py:314-333: m = M F; m -= S (m - meq); m += (1 - S/2) G;
// f = M^-1 m) in population space. Every even row of the reference's M
// except the density relaxes at s_v and every odd row except the momentum at
// s_o, and M maps even (odd) population vectors to even (odd) moments, so the
// MRT update is exactly a two-relaxation-time one:
//   f* = F - s_v (F - g)^+ - s_o (F - g)^- + w (e.F)
//          + c_v w ((e.u)(e.F) - u.F) - (1 - c_v) (2/3) w u.F
// with ^+ / ^- the halves (h_q +- h_opp(q))/2, c_v = 1 - s_v/2, and g the
// populations whose moments are the reference's equilibrium moments with F's
// own density and momentum: g = w (rho_F + 3 e.j + 4.5 (e.u)^2 - 1.5 u.u),
// rho_F = sum F, j = sum e F (M^-1 applied to meq; the forcing terms are
// M^-1 (1 - S/2) G with G = M s, s = w ((e - u).F + (e.u)(e.F))). A CPU test
// checks the identity against the reference's matrix form. Per opposite pair
// it is ~10 arithmetic instructions instead of a 19 x 19 transform.

#include "common.h"

namespace fsi {
namespace {

template <typename R> struct LbmK {
  int nx, ny, nz;
  int64_t N;
  R s_v, s_o;
  R cv;         // 1 - s_v/2
  R ke;         // (2 + c_v)/3: weight of u.F in the even forcing term
  R rho0;
  R scale, inv_scale;  // FP16 storage factors; I16: the scale fsi_lbm_init writes with
  R uf[3];      // uniform force when no field is given
  int mode[2];  // x-face boundary modes (left, right)
  R bc_rho[2];
  R bc_vel[2][3];
  float* ctrl;  // I16: control block of the adaptive scale
  int epoch;    // I16: write epoch of this launch
  int pass;     // I16: 0 normal, 1 measuring (no stores), 2 scaled from the measurement
};

template <typename R> LbmK<R> make_consts(const fsi_lbm_params* p, const fsi_lbm_force* f,
                                          const fsi_lbm_buffers* b = nullptr) {
  LbmK<R> c;
  c.nx = p->nx;
  c.ny = p->ny;
  c.nz = p->nz;
  c.N = (int64_t)p->nx * p->ny * p->nz;
  c.s_v = R(p->s_v);
  c.s_o = R(p->s_o);
  c.cv = R(1) - R(0.5) * c.s_v;
  c.ke = (R(2) + c.cv) / R(3);
  c.rho0 = R(p->rho0);
  const bool shifted = p->storage == FSI_POP_F16 || p->storage == FSI_POP_I16;
  c.scale = shifted ? R(p->pop_scale) : R(1);
  c.inv_scale = shifted ? R(1.0 / p->pop_scale) : R(1);
  c.ctrl = b ? b->store_ctrl : nullptr;
  c.epoch = p->store_epoch;
  c.pass = p->storage == FSI_POP_I16 ? p->store_pass : 0;
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

// Two-relaxation-time form of the reference's MRT collision (file comment),
// in place on d[19] (absolute populations, or deviations from w rho0: g and
// the forcing are the same expressions in both, with sigma = sum d the
// density part). ps/pd: opposite-pair sums d_q + d_opp(q) and differences
// for q = 1, 3, ..., 17; j = sum e d; u the collision velocity; fo the force.
template <typename R>
__device__ __forceinline__ void trt_collide(R d[19], const R ps[9], const R pd[9], R sigma, const R j[3],
                                            const R u[3], const R fo[3], const LbmK<R>& c) {
  const R uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const R uf = u[0] * fo[0] + u[1] * fo[1] + u[2] * fo[2];
  // w * base = s_v g^+(e.u = 0) - w ke u.F: the part of the even increment
  // common to a weight class
  const R base = c.s_v * (sigma - R(1.5) * uu) - c.ke * uf;
  const R hv = R(0.5) * c.s_v, ho = R(0.5) * c.s_o;
  d[0] += wq<R>(0) * base - c.s_v * d[0];
#pragma unroll
  for (int p = 0; p < 9; ++p) {
    const int q = 2 * p + 1;
    const R w = wq<R>(q);
    R eu = R(0), ej = R(0), ef = R(0);
    if (ex(q)) { eu += R(ex(q)) * u[0]; ej += R(ex(q)) * j[0]; ef += R(ex(q)) * fo[0]; }
    if (ey(q)) { eu += R(ey(q)) * u[1]; ej += R(ey(q)) * j[1]; ef += R(ey(q)) * fo[1]; }
    if (ez(q)) { eu += R(ez(q)) * u[2]; ej += R(ez(q)) * j[2]; ef += R(ez(q)) * fo[2]; }
    const R even = w * base - hv * ps[p] + (w * eu) * (R(4.5) * c.s_v * eu + c.cv * ef);
    const R odd = w * (R(3) * c.s_o * ej + ef) - ho * pd[p];
    d[q] += even + odd;
    d[q + 1] += even - odd;
  }
}

// I16 storage: writes that record their largest value. Every one of the
// first 64 (a flow starting from rest can grow fast), then every fourth (a
// developed flow changes slowly; the maximum costs ~50 instructions of ~600
// per wavefront). An untracked write leaves its maxima at zero, and the next
// write keeps the scale.
__device__ __forceinline__ bool i16_tracked(int epoch) { return epoch < 64 || epoch % 4 == 0; }

// I16 storage: the scale of this launch's writes from the previous write's
// scale and largest normalised value (every workgroup computes the same
// value). The first writes aim low (a flow starting from rest can double
// per step), later ones at 0.6 of the range, each rescale when the largest
// value leaves [lo, hi]; at most x8 per step upwards.
__device__ __forceinline__ float i16_next_scale(float s_in, float a, int epoch) {
  const bool early = epoch < 16;
  const float lo = early ? 0.1f : 0.4f, hi = early ? 0.5f : 0.8f, target = early ? 0.25f : 0.6f;
  if (!(a > 0.0f)) return s_in;  // nothing stored yet (or a NaN): keep
  // (1-ulp reciprocals: a scale only has to be the same where it is written
  // and where it is read, and both take it from the control block)
  const float r = target * __builtin_amdgcn_rcpf(a);
  if (a > hi) return fmaxf(s_in * r, 1.0e-6f);
  if (a < lo) return fminf(s_in * fminf(r, 8.0f), 1.0e30f);
  return s_in;
}

// Storage factors of one launch per weight class (rest, axis, diagonal): ld
// turns a stored value of the previous write into storage units, st a value
// in storage units into the stored representation of this write.
template <typename S> struct Factors {
  using R = typename PopStore<S>::real;
  R ld[3], st[3];
};

template <typename S>
__device__ __forceinline__ Factors<S> launch_factors(const LbmK<typename PopStore<S>::real>& c, bool init) {
  using R = typename PopStore<S>::real;
  Factors<S> f;
  if constexpr (PopStore<S>::adaptive) {
    const int e = c.epoch;
    float s_in, s_out;
    const int lane = threadIdx.x & 63;
    if (init) {
      s_in = s_out = (float)c.scale;
    } else if (c.pass == 1) {
      // measuring pass: values at the previous write's scale, nothing stored
      s_in = s_out = c.ctrl[kCtrlSlot * ((e + 2) % 3)];
    } else if (c.pass == 2) {
      // the largest value this step stores, measured: put it at 0.8 of the range
      const float* in = c.ctrl + kCtrlSlot * ((e + 2) % 3);
      s_in = in[0];
      const float a = wave_max_nonneg(c.ctrl[kCtrlSlot * kCtrlMeasure + kCtrlLine * (1 + (lane % kCtrlParts))]);
      s_out = a > 0.0f ? fminf(s_in * 0.8f * __builtin_amdgcn_rcpf(a), 1.0e30f) : s_in;
    } else {
      const float* in = c.ctrl + kCtrlSlot * ((e + 2) % 3);
      s_in = in[0];
      // the previous write recorded its maximum only if it was tracked
      const float a = i16_tracked(e - 1) ? wave_max_nonneg(in[kCtrlLine * (1 + (lane % kCtrlParts))]) : 0.0f;
      s_out = i16_next_scale(s_in, a, e);
    }
    if (blockIdx.x == 0 && c.pass != 1) {
      // this write's scale, and the next write's slot cleared for its maxima
      // when that write is tracked (the last tracked write's maxima stay
      // readable: its slot is next cleared three tracked writes later)
      if (threadIdx.x == 0) c.ctrl[kCtrlSlot * (e % 3)] = s_out;
      float* nxt = c.ctrl + kCtrlSlot * ((e + 1) % 3);
      if (i16_tracked(e + 1) && threadIdx.x < kCtrlParts) nxt[kCtrlLine * (1 + threadIdx.x)] = 0.0f;
    }
    const float inv_in = __builtin_amdgcn_rcpf(32767.0f * s_in);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float w = wq<float>(k == 0 ? 0 : (k == 1 ? 1 : 7));
      f.ld[k] = w * inv_in;
      f.st[k] = s_out * float(wden(k == 0 ? 0 : (k == 1 ? 1 : 7)));  // s / w, w = 1 / wden
    }
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      f.ld[k] = c.inv_scale;
      f.st[k] = c.scale;
    }
  }
  return f;
}

// Adds the workgroup's largest |stored value| (fraction of the 16-bit range)
// to this write's maxima: one atomic per workgroup, spread over kCtrlParts
// lines. Every thread of the workgroup calls it.
template <int BLOCK>
__device__ __forceinline__ void record_amax(float* ctrl, int slot, float amax) {
  __shared__ uint32_t red[BLOCK / 64];
  const float m = wave_max_nonneg(amax);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = __float_as_uint(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t b = red[0];
#pragma unroll
    for (int i = 1; i < BLOCK / 64; ++i) b = max(b, red[i]);
    float* out = ctrl + kCtrlSlot * slot + kCtrlLine * (1 + blockIdx.x % kCtrlParts);
    atomicMax(reinterpret_cast<unsigned int*>(out), b);
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
  // (I16: writes this epoch's scale before any thread returns)
  const Factors<S> fac = launch_factors<S>(c, true);
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= c.N) return;
  const int k = (int)(n % c.nz);
  const int64_t ij = n / c.nz;
  const int j = (int)(ij % c.ny);
  const int i = (int)(ij / c.ny);
  const Nbr o = neighbours(i, j, k, c.nx, c.ny, c.nz);
  const int64_t rb = ij * 19 * c.nz + k;  // population q at rb + q * nz
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
    const S s = PopStore<S>::store(PopStore<S>::scaled(v, fac.st[wclass(q)]));
    pa[rb + q * c.nz] = s;
    pb[rb + q * c.nz] = s;
  }
}

enum { PHASE_STREAM_COLLIDE = 0, PHASE_COLLIDE = 1, PHASE_STREAM_MACRO = 2 };

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

// Force assembly of the coupled step (fsi/coupling.py:408-461: phi clamp,
// reset_force, coupling force, immersed-boundary drag) evaluated by the
// collide pass that uses the force (ASM), instead of a pass of its own.
template <typename R> struct AsmArgs {
  R* cforce;         // 4N AoS coupling buffer (force density xyz, phi)
  R* ib_force;       // 3N SoA IB drag (ib)
  R* ib_total;       // 3, summed with atomics (ib)
  int32_t* counters; // [5] active IB cells, [6] clipped (ib)
  R* force_out;      // 3N SoA: the assembled lattice force
  int coupling, clamp, ib, ib_use_max;
  R drag, thr, max_force;
  R base[3];
};

template <typename R> __device__ __forceinline__ R wave_sum(R v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// One lattice update of one cell in the given phase (see the file comment).
// FACES: some x face is not periodic. Requires 19 * N * sizeof(S) < 2^32.
// Workgroup size of the lattice step. A workgroup reads 19 contiguous runs
// whose z-shifted ones straddle one extra cache line at its ends; 1024 lanes
// amortise that for FP32 (measured at 256^3 with the row bitmap: 423 us
// against 430 us at 512 and 437 us at 256), while the 16-bit storages run
// best at 256 (FP16 215 us against 226 us at 512 and 237 us at 128).
template <typename S> struct PullBlock { static constexpr int value = 256; };
template <> struct PullBlock<float> { static constexpr int value = 1024; };

// ROWS: nz is a multiple of 64, so every wavefront lies in one z-row
// (blocks are multiples of 64 lanes): the row indices are wave-uniform and
// the row offsets of the pull are scalar arithmetic. BITS (with ROWS): the
// solid flags come from the row bitmap instead of the per-cell mask.
template <typename S, int PHASE, bool FORCE_FIELD, bool FACES, bool ASM = false, bool ROWS = false,
          bool BITS = false>
__global__ __launch_bounds__(PullBlock<S>::value) void k_lbm_pull(
    LbmK<typename PopStore<S>::real> c, const S* __restrict__ pin, S* __restrict__ pout,
    const uint32_t* __restrict__ nbmask, const uint32_t* __restrict__ sbits,
    const typename PopStore<S>::real* __restrict__ force,
    typename PopStore<S>::real* __restrict__ rho_io, typename PopStore<S>::real* __restrict__ vel_io,
    const typename PopStore<S>::real* __restrict__ fv_prev,
    typename PopStore<S>::real* __restrict__ fv_next, AsmArgs<typename PopStore<S>::real> asm_args) {
  static_assert(!ASM || PHASE == PHASE_COLLIDE, "the force assembly runs in the collide phase");
  using R = typename PopStore<S>::real;
  // the arithmetic runs on deviations d = F - w rho0 for every storage: the
  // 16-bit storages keep them, the absolute FP32/FP64 populations subtract
  // and add back w rho0 at the load and the store (exact for F within a
  // factor 2 of w rho0), so the collision's FP32 rounding is relative to the
  // deviations, not to the ~w-sized populations
  constexpr bool ABS = !PopStore<S>::shifted;
  constexpr bool ADAPT = PopStore<S>::adaptive;
  const Factors<S> fac = launch_factors<S>(c, false);
  R base[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) base[k] = ABS ? wq<R>(k == 0 ? 0 : (k == 1 ? 1 : 7)) * c.rho0 : R(0);
  const uint32_t N = (uint32_t)c.N;
  const uint32_t n_lane = blockIdx.x * blockDim.x + threadIdx.x;
  // adaptive storage: every thread reaches the workgroup's maximum at the
  // end, so lanes past the lattice compute the last cell and store nothing
  const bool live = n_lane < N;
  if constexpr (!ADAPT) {
    if (!live) return;
  }
  // a measuring pass computes the step and its largest stored value only
  const bool write = live && (!ADAPT || c.pass != 1);
  const uint32_t n = live ? n_lane : N - 1;
  constexpr uint32_t B = sizeof(S);
  const uint32_t nz = (uint32_t)c.nz, ny = (uint32_t)c.ny, nx = (uint32_t)c.nx;
  const uint32_t k = n % nz;
  uint32_t ij = n / nz;
  if constexpr (ROWS) ij = __builtin_amdgcn_readfirstlane(ij);
  const uint32_t j = ij % ny;
  const uint32_t i = ij / ny;
  // neighbour-solid bits (nb(q): the cell c + e_q is solid) from the row
  // bitmap when there is one (a wavefront lies in one z-row and starts at a
  // multiple of 64: its 9 neighbour rows' words are uniform scalar loads and
  // a lane's 3 bits k-1..k+1 of each row come from one funnel shift), else
  // from the per-cell mask. BITS (16-bit storages): the bounce-back selects
  // are extracted from the rows' bits directly. FP32/FP64: the 19-bit mask
  // is assembled from them once (measured 2% faster there than the direct
  // extracts, which are 3-4% faster for FP16).
  static_assert(!BITS || (ROWS && sizeof(S) == 2), "direct row-bit extracts: 16-bit storages");
  uint32_t mask = 0u;
  uint32_t win[3][3];  // bits k-1, k, k+1 of row (i + a - 1, j + b - 1)
  auto row_bits = [&](bool uniform_rows) {
    const uint32_t words = nz / 32u + 2u;
    const uint32_t ijw = __builtin_amdgcn_readfirstlane(ij);
    const uint32_t kw = __builtin_amdgcn_readfirstlane(k);
    const uint32_t iw = ijw / ny, jw = ijw % ny;
    uint32_t ii[3] = {iw == 0 ? nx - 1 : iw - 1, iw, iw == nx - 1 ? 0u : iw + 1};
    uint32_t jj[3] = {jw == 0 ? ny - 1 : jw - 1, jw, jw == ny - 1 ? 0u : jw + 1};
    if (uniform_rows) {  // keep the neighbour rows in scalar registers
      ii[0] = __builtin_amdgcn_readfirstlane(ii[0]);
      ii[2] = __builtin_amdgcn_readfirstlane(ii[2]);
      jj[0] = __builtin_amdgcn_readfirstlane(jj[0]);
      jj[2] = __builtin_amdgcn_readfirstlane(jj[2]);
    }
    const uint32_t kk = k - kw;
    const uint32_t sh = kk & 31u;
    const bool upper = kk >= 32u;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const uint32_t* rp = sbits + (ii[a] * ny + jj[b]) * words + kw / 32u;
        const uint32_t w0 = rp[0], w1 = rp[1], w2 = rp[2];
        win[a][b] = __builtin_amdgcn_alignbit(upper ? w2 : w1, upper ? w1 : w0, sh);
      }
  };
  if constexpr (BITS) {
    row_bits(true);
  } else {
    bool from_bits = false;
    if constexpr (sizeof(S) >= 4 && ROWS) from_bits = sbits != nullptr;
    if (from_bits) {
      row_bits(false);
#pragma unroll
      for (int q = 0; q < 19; ++q) mask |= ((win[ex(q) + 1][ey(q) + 1] >> (1 + ez(q))) & 1u) << q;
    } else {
      mask = nbmask[n];
    }
  }
  auto nb = [&](int q) -> uint32_t {
    if constexpr (BITS)
      return (win[ex(q) + 1][ey(q) + 1] >> (1 + ez(q))) & 1u;
    else
      return (mask >> q) & 1u;
  };
  const bool solid = BITS ? nb(0) != 0 : (mask & 1u) != 0;
  const uint32_t run = nz * B;                // byte size of one population run
  const uint32_t rstride = 19 * nz;           // elements per row block
  const uint32_t own = (ij * rstride + k) * B;  // population q at own + q * run

  // 16-bit storages: all 19 loads are issued before any is used (a barrier
  // to the instruction scheduler: interleaving each conversion with its load
  // put a wait for every second load and cost ~25% of the step; FP32 runs
  // ~3% faster without the barrier)
  S raw[19];
  if constexpr (PHASE == PHASE_COLLIDE) {
#pragma unroll
    for (int q = 0; q < 19; ++q) raw[q] = ld_off(pin, q * run + own);
  } else {
    // gather f*(c - e_q, q), or f*(c, opp q) when c - e_q = c + e_opp(q) is
    // solid (fsi/lbm3d.py:352-361 read backwards). One-cell periodic wrap on
    // every axis (lbm3d.py:335-350). A solid lane's values are discarded; it
    // reads the plain pull sources, which lie in the cache lines its fluid
    // neighbours in the wavefront read anyway (its own slots would be extra
    // lines for the 16 planes shifted in x or y).
    // element offsets of the neighbouring row blocks and z entries (modulo 2^32)
    const uint32_t sx = ny * rstride, sy = rstride;
    const uint32_t oxm = i == 0 ? (nx - 1) * sx : (uint32_t)(-(int32_t)sx);
    const uint32_t oxp = i == nx - 1 ? (uint32_t)(-(int32_t)((nx - 1) * sx)) : sx;
    const uint32_t oym = j == 0 ? (ny - 1) * sy : (uint32_t)(-(int32_t)sy);
    const uint32_t oyp = j == ny - 1 ? (uint32_t)(-(int32_t)((ny - 1) * sy)) : sy;
    const uint32_t ozm = k == 0 ? nz - 1 : (uint32_t)(-1);
    const uint32_t ozp = k == nz - 1 ? (uint32_t)(-(int32_t)(nz - 1)) : 1u;
    // all ones when the source c - e_q = c + e_opp(q) of a fluid cell is
    // solid (BITS: its row bit sign-extended by one bit-field extract)
    const uint32_t all = solid ? 0u : mask;
    uint32_t winf[3][3];
    if constexpr (BITS) {
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) winf[a][b] = solid ? 0u : win[a][b];
    }
    auto bounce = [&](int q) -> uint32_t {
      const int o = opp(q);
      if constexpr (BITS)
        return (uint32_t)__builtin_amdgcn_sbfe((int)winf[ex(o) + 1][ey(o) + 1], 1 + ez(o), 1);
      else
        return 0u - ((all >> o) & 1u);
    };
    if constexpr (ROWS) {
      // source byte offset = (uniform row block + q's run) + the lane's z
      // entry k - e_z(q), wrapped
      const uint32_t km = k == 0 ? nz - 1 : k - 1, kp = k == nz - 1 ? 0u : k + 1;
      const uint32_t kbyte[3] = {kp * B, k * B, km * B};  // e_z = -1, 0, +1
      const uint32_t rowb = ij * rstride;
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        uint32_t u = rowb;
        if (ex(q) > 0) u += oxm;
        if (ex(q) < 0) u += oxp;
        if (ey(q) > 0) u += oym;
        if (ey(q) < 0) u += oyp;
        const uint32_t a0 = kbyte[ez(q) + 1] + (q * run + u * B);
        const uint32_t a1 = opp(q) * run + own;
        const uint32_t sel = bounce(q);
        raw[q] = ld_off(pin, a0 ^ ((a0 ^ a1) & sel));
      }
    } else {
      const uint32_t ownel = ij * rstride + k;
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        uint32_t src = ownel;  // c - e_q
        if (ex(q) > 0) src += oxm;
        if (ex(q) < 0) src += oxp;
        if (ey(q) > 0) src += oym;
        if (ey(q) < 0) src += oyp;
        if (ez(q) > 0) src += ozm;
        if (ez(q) < 0) src += ozp;
        const uint32_t a0 = q * run + src * B;
        const uint32_t a1 = opp(q) * run + own;
        const uint32_t sel = bounce(q);
        raw[q] = ld_off(pin, a0 ^ ((a0 ^ a1) & sel));
      }
    }
  }
  if constexpr (sizeof(S) == 2) __builtin_amdgcn_sched_barrier(0);
  R d[19];
#pragma unroll
  for (int q = 0; q < 19; ++q) d[q] = PopStore<S>::load(raw[q], fac.ld[wclass(q)]) - base[wclass(q)];

  R fo[3];
  if constexpr (ASM) {
    // the force of this collision, assembled from the coupled step's
    // buffers (fsi/coupling.py:408-461), written to the lattice force
    const AsmArgs<R>& a = asm_args;
    R* cf = a.cforce + 4 * (int64_t)n;
    R phi = cf[3];
    if (a.clamp && phi > R(1)) {
      phi = R(1);
      if (write) cf[3] = phi;
    }
    const R cx = cf[0], cy = cf[1], cz = cf[2];
    fo[0] = solid ? R(0) : a.base[0];
    fo[1] = solid ? R(0) : a.base[1];
    fo[2] = solid ? R(0) : a.base[2];
    if (a.coupling && !solid) {
      fo[0] += cx;
      fo[1] += cy;
      fo[2] += cz;
    }
    if (a.ib) {
      R fi[3] = {R(0), R(0), R(0)};
      bool active = false, clipped = false;
      if (a.drag > R(0) && !solid && phi > R(0) && phi >= a.thr) {
        const R sc = -a.drag * phi * rho_io[n];
#pragma unroll
        for (int k = 0; k < 3; ++k) fi[k] = sc * vel_io[k * N + n];
        const R nrm = sqrt(fi[0] * fi[0] + fi[1] * fi[1] + fi[2] * fi[2]);
        if (a.ib_use_max && nrm > a.max_force && nrm > R(0)) {
#pragma unroll
          for (int k = 0; k < 3; ++k) fi[k] = fi[k] / nrm * a.max_force;
          clipped = true;
        }
        active = true;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (write) a.ib_force[k * N + n] = fi[k];
        fo[k] += fi[k];
      }
      if (!ADAPT || c.pass != 1) {  // uniform: a measuring pass counts nothing
        active = active && live;
        clipped = clipped && live;
        const unsigned long long ab = __ballot(active), cb = __ballot(clipped), lanes = __ballot(1);
        const int leader = __ffsll((long long)lanes) - 1;
        if (ab) {
          const R s0 = wave_sum(fi[0]), s1 = wave_sum(fi[1]), s2 = wave_sum(fi[2]);
          if ((int)__lane_id() == leader) {
            atomicAdd(a.ib_total + 0, s0);
            atomicAdd(a.ib_total + 1, s1);
            atomicAdd(a.ib_total + 2, s2);
            atomicAdd(a.counters + 5, (int)__popcll(ab));
          }
        }
        if (cb && (int)__lane_id() == leader) atomicAdd(a.counters + 6, (int)__popcll(cb));
      }
    }
    if (write) {
#pragma unroll
      for (int k = 0; k < 3; ++k) a.force_out[k * N + n] = fo[k];
    }
  } else if constexpr (FORCE_FIELD) {
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
      const bool inner_solid = BITS ? (side == 0 ? nb(1) != 0 : nb(2) != 0) : ((mask >> (side == 0 ? 1 : 2)) & 1u) != 0;
      x_face_boundary<R, true>(d, c.mode[side], c.bc_rho[side], c.bc_vel[side], c.rho0, pu, inner_solid);
    }
  }

  // opposite-pair sums and differences; density part and momentum of d
  R ps[9], pd[9];
  R sigma = d[0];
  R jm[3] = {R(0), R(0), R(0)};
#pragma unroll
  for (int p = 0; p < 9; ++p) {
    const int q = 2 * p + 1;
    ps[p] = d[q] + d[q + 1];
    pd[p] = d[q] - d[q + 1];
    sigma += ps[p];
    if (ex(q)) jm[0] += R(ex(q)) * pd[p];
    if (ey(q)) jm[1] += R(ey(q)) * pd[p];
    if (ez(q)) jm[2] += R(ez(q)) * pd[p];
  }

  // macroscopic velocity (fsi/lbm3d.py:409-425): stored, or from the
  // streamed populations with the force the previous collision used
  R u[3];
  if constexpr (PHASE == PHASE_COLLIDE) {
    u[0] = vel_io[n];
    u[1] = vel_io[N + n];
    u[2] = vel_io[2 * N + n];
  } else {
    R rho = c.rho0 + sigma;
    const R inv = recip(rho);
#pragma unroll
    for (int a = 0; a < 3; ++a) u[a] = (jm[a] + fo[a] * R(0.5)) * inv;
    if (solid) {
      rho = c.rho0;
      u[0] = u[1] = u[2] = R(0);
    }
    if (write) {
      if constexpr (FACES) {
        if (face) {
          R* fn = fv_next + side * 3 * fplane;
          fn[jk] = u[0];
          fn[fplane + jk] = u[1];
          fn[2 * fplane + jk] = u[2];
        }
      }
      if constexpr (PHASE == PHASE_STREAM_MACRO) {
        rho_io[n] = rho;
        vel_io[n] = u[0];
        vel_io[N + n] = u[1];
        vel_io[2 * N + n] = u[2];
      }
    }
  }

  if constexpr (PHASE != PHASE_STREAM_MACRO) trt_collide<R>(d, ps, pd, sigma, jm, u, fo, c);
  // store the post-collision (or, for STREAM_MACRO, post-stream) populations;
  // solid cells store their rest value, deviation 0 (a zero store factor in
  // the 16-bit storages, where it saves a select per population)
  R st[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) st[k] = (!ABS && solid) ? R(0) : fac.st[k];
  R x[19];
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    if constexpr (ABS)
      x[q] = (solid ? R(0) : d[q]) + base[wclass(q)];
    else
      x[q] = PopStore<S>::scaled(d[q], st[wclass(q)]);
  }
  if (write) {
#pragma unroll
    for (int q = 0; q < 19; ++q) st_stream(pout, q * run + own, PopStore<S>::store(x[q]));
  }
  if constexpr (ADAPT) {
    if (c.pass == 1 || i16_tracked(c.epoch)) {  // uniform
      // largest |x| from the bit patterns (no NaN canonicalisation per value)
      uint32_t m = __float_as_uint((float)x[0]) & 0x7fffffffu;
#pragma unroll
      for (int q = 1; q < 19; q += 2)
        m = max(m, max(__float_as_uint((float)x[q]) & 0x7fffffffu, __float_as_uint((float)x[q + 1]) & 0x7fffffffu));
      record_amax<PullBlock<S>::value>(c.ctrl, c.pass == 1 ? kCtrlMeasure : c.epoch % 3,
                                       live ? __uint_as_float(m) : 0.0f);
    }
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
  const int64_t rb = (n / c.nz) * 19 * c.nz + n % c.nz;
  R ld[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if constexpr (PopStore<S>::adaptive) {
      // the scale of the current populations (write epoch c.epoch)
      const float w = wq<float>(k == 0 ? 0 : (k == 1 ? 1 : 7));
      ld[k] = w * __builtin_amdgcn_rcpf(32767.0f * c.ctrl[kCtrlSlot * (c.epoch % 3)]);
    } else {
      ld[k] = c.inv_scale;
    }
  }
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const R dq = PopStore<S>::load(pop[rb + q * c.nz], ld[wclass(q)]);
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
  FSI_REQUIRE(p->storage == FSI_POP_F32 || p->storage == FSI_POP_F16 || p->storage == FSI_POP_F64 ||
                  p->storage == FSI_POP_I16,
              "unknown population storage %d", p->storage);
  FSI_REQUIRE(p->storage != FSI_POP_I16 || (p->pop_scale > 0.0 && p->store_epoch >= 0),
              "I16 storage needs a positive initial scale and a write epoch");
  FSI_REQUIRE(p->left_mode >= 0 && p->left_mode <= 2 && p->right_mode >= 0 && p->right_mode <= 2,
              "unknown boundary mode");
  return 0;
}

// Row bitmap of the solid flags (fsi_lbm_buffers.solid_bits): bit b of row
// (i, j) is solid(i, j, b - 1), with the periodic neighbours k = nz - 1 at
// b = 0 and k = 0 at b = nz + 1; one thread per word.
__global__ __launch_bounds__(kBlock) void k_solid_bits(int64_t rows, int nz, const int32_t* __restrict__ solid,
                                                      uint32_t* __restrict__ bits) {
  const int words = nz / 32 + 2;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * words) return;
  const int64_t row = t / words;
  const int w = (int)(t % words);
  uint32_t v = 0;
  for (int tb = 0; tb < 32; ++tb) {
    const int b = 32 * w + tb;
    if (b > nz + 1) break;
    const int k = b == 0 ? nz - 1 : (b == nz + 1 ? 0 : b - 1);
    if (solid[row * nz + k] != 0) v |= 1u << tb;
  }
  bits[t] = v;
}

template <typename S>
int lbm_init_t(const fsi_lbm_params* p, const fsi_lbm_buffers* b, void* pa, void* pb,
               hipStream_t st) {
  using R = typename PopStore<S>::real;
  LbmK<R> c = make_consts<R>(p, nullptr, b);
  hipLaunchKernelGGL(k_lbm_init<S>, dim3(blocks_for(c.N)), dim3(kBlock), 0, st, c, b->solid,
                     b->nbmask, (R*)b->rho, (R*)b->vel, (S*)pa, (S*)pb, (R*)b->face_vel,
                     R(p->initial_velocity[0]), R(p->initial_velocity[1]),
                     R(p->initial_velocity[2]));
  FSI_TRY_LAUNCH("lbm_init");
  if (b->solid_bits) {
    const int64_t words = FSI_SOLID_BITS_WORDS(p->nx, p->ny, p->nz);
    hipLaunchKernelGGL(k_solid_bits, dim3(blocks_for(words)), dim3(kBlock), 0, st, (int64_t)p->nx * p->ny, p->nz,
                       b->solid, b->solid_bits);
    FSI_TRY_LAUNCH("solid_bits");
  }
  return 0;
}

template <typename S, int PHASE, bool FF, bool ASM = false>
void launch_pull(const LbmK<typename PopStore<S>::real>& c, bool faces, const void* pin, void* pout,
                 const fsi_lbm_buffers* b, const void* force, const void* fv_prev, void* fv_next,
                 hipStream_t st, AsmArgs<typename PopStore<S>::real> a = {}) {
  using R = typename PopStore<S>::real;
  const dim3 grid(blocks_for(c.N, PullBlock<S>::value)), block(PullBlock<S>::value);
  // z-rows of whole wavefronts: wave-uniform rows, and the row bitmap of the
  // solid flags when the caller gave one (16-bit storages: the BITS kernels;
  // FP32/FP64: the bitmap pointer is tested in the kernel)
  const bool rows = c.nz % 64 == 0;
  const bool bits = rows && b->solid_bits && sizeof(S) == 2;  // FP32/FP64 read the bitmap at run time
#define FSI_PULL_LAUNCH(FACES, ROWS, BITS)                                                                     \
  hipLaunchKernelGGL((k_lbm_pull<S, PHASE, FF, FACES, ASM, ROWS, BITS>), grid, block, 0, st, c, (const S*)pin, \
                     (S*)pout, b->nbmask, b->solid_bits, (const R*)force, (R*)b->rho, (R*)b->vel,             \
                     (const R*)fv_prev, (R*)fv_next, a)
  if constexpr (sizeof(S) == 2) {
    if (bits) {
      if (faces)
        FSI_PULL_LAUNCH(true, true, true);
      else
        FSI_PULL_LAUNCH(false, true, true);
      return;
    }
  }
  if (faces && rows)
    FSI_PULL_LAUNCH(true, true, false);
  else if (faces)
    FSI_PULL_LAUNCH(true, false, false);
  else if (rows)
    FSI_PULL_LAUNCH(false, true, false);
  else
    FSI_PULL_LAUNCH(false, false, false);
#undef FSI_PULL_LAUNCH
}

template <typename S>
int lbm_step_t(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const void* pin, void* pout,
               const fsi_lbm_force* f, int phase, int face_slot, hipStream_t st) {
  using R = typename PopStore<S>::real;
  LbmK<R> c = make_consts<R>(p, f, b);
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

template <typename R> __global__ void k_reset_ib_sums(R* __restrict__ total, int32_t* __restrict__ counters) {
  if (threadIdx.x < 3) total[threadIdx.x] = R(0);
  if (threadIdx.x == 3) counters[5] = 0;
  if (threadIdx.x == 4) counters[6] = 0;
}

template <typename S>
int lbm_collide_assembled_t(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const fsi_force_assembly* fa,
                            const void* pin, void* pout, void* cforce, void* ib_force, void* ib_total,
                            int32_t* counters, hipStream_t st) {
  using R = typename PopStore<S>::real;
  LbmK<R> c = make_consts<R>(p, nullptr, b);
  FSI_REQUIRE(19.0 * (double)c.N * sizeof(S) < 4294967296.0,
              "lbm_collide_assembled: lattice of %lld cells too large for 32-bit plane offsets", (long long)c.N);
  AsmArgs<R> a{};
  a.cforce = (R*)cforce;
  a.ib_force = (R*)ib_force;
  a.ib_total = (R*)ib_total;
  a.counters = counters;
  a.force_out = (R*)b->force;
  a.coupling = fa->coupling;
  a.clamp = fa->clamp_fraction;
  a.ib = fa->ib_enabled;
  a.ib_use_max = fa->ib_use_max;
  a.drag = R(fa->ib_drag);
  a.thr = R(fa->ib_threshold);
  a.max_force = R(fa->ib_max_force);
  for (int k = 0; k < 3; ++k) a.base[k] = R(fa->base_force[k]);
  if (a.ib && p->store_pass != 1) {
    hipLaunchKernelGGL(k_reset_ib_sums<R>, dim3(1), dim3(64), 0, st, (R*)ib_total, counters);
    FSI_TRY_LAUNCH("lbm_collide_assembled reset");
  }
  const bool faces = p->left_mode != FSI_BC_PERIODIC || p->right_mode != FSI_BC_PERIODIC;
  launch_pull<S, PHASE_COLLIDE, false, true>(c, faces, pin, pout, b, nullptr, nullptr, nullptr, st, a);
  FSI_TRY_LAUNCH("lbm_collide_assembled");
  return 0;
}

template <typename S>
int lbm_export_t(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const void* pop, void* out,
                 hipStream_t st) {
  using R = typename PopStore<S>::real;
  LbmK<R> c = make_consts<R>(p, nullptr, b);
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
  FSI_REQUIRE(p->storage != FSI_POP_I16 || b->store_ctrl, "lbm_init: I16 storage needs store_ctrl");
  hipStream_t st = (hipStream_t)stream;
  switch (p->storage) {
    case FSI_POP_F32: return lbm_init_t<float>(p, b, pop_a, pop_b, st);
    case FSI_POP_F16: return lbm_init_t<__half>(p, b, pop_a, pop_b, st);
    case FSI_POP_I16: return lbm_init_t<int16_t>(p, b, pop_a, pop_b, st);
    default: return lbm_init_t<double>(p, b, pop_a, pop_b, st);
  }
}

int fsi_lbm_step(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const void* pop_in,
                 void* pop_out, const fsi_lbm_force* force, int32_t phase, int32_t face_slot,
                 void* stream) {
  if (int rc = check_params(p)) return rc;
  FSI_REQUIRE(b && b->nbmask && pop_in && pop_out && pop_in != pop_out,
              "lbm_step: null or aliased buffer");
  FSI_REQUIRE(p->storage != FSI_POP_I16 || b->store_ctrl, "lbm_step: I16 storage needs store_ctrl");
  hipStream_t st = (hipStream_t)stream;
  switch (p->storage) {
    case FSI_POP_F32: return lbm_step_t<float>(p, b, pop_in, pop_out, force, phase, face_slot, st);
    case FSI_POP_F16: return lbm_step_t<__half>(p, b, pop_in, pop_out, force, phase, face_slot, st);
    case FSI_POP_I16: return lbm_step_t<int16_t>(p, b, pop_in, pop_out, force, phase, face_slot, st);
    default: return lbm_step_t<double>(p, b, pop_in, pop_out, force, phase, face_slot, st);
  }
}

int fsi_lbm_collide_assembled(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const fsi_force_assembly* a,
                              const void* pop_in, void* pop_out, void* cforce, void* ib_force, void* ib_total,
                              int32_t* counters, void* stream) {
  if (int rc = check_params(p)) return rc;
  FSI_REQUIRE(b && a && b->nbmask && b->rho && b->vel && b->force && cforce && pop_in && pop_out &&
                  pop_in != pop_out,
              "lbm_collide_assembled: null or aliased buffer");
  FSI_REQUIRE(!a->ib_enabled || (ib_force && ib_total && counters),
              "lbm_collide_assembled: IB needs ib_force, ib_total and counters");
  FSI_REQUIRE(p->storage != FSI_POP_I16 || b->store_ctrl, "lbm_collide_assembled: I16 storage needs store_ctrl");
  hipStream_t st = (hipStream_t)stream;
  switch (p->storage) {
    case FSI_POP_F32:
      return lbm_collide_assembled_t<float>(p, b, a, pop_in, pop_out, cforce, ib_force, ib_total, counters, st);
    case FSI_POP_F16:
      return lbm_collide_assembled_t<__half>(p, b, a, pop_in, pop_out, cforce, ib_force, ib_total, counters, st);
    case FSI_POP_I16:
      return lbm_collide_assembled_t<int16_t>(p, b, a, pop_in, pop_out, cforce, ib_force, ib_total, counters, st);
    default:
      return lbm_collide_assembled_t<double>(p, b, a, pop_in, pop_out, cforce, ib_force, ib_total, counters, st);
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

int fsi_lbm_export_populations(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const void* pop,
                               void* out_aos, void* stream) {
  if (int rc = check_params(p)) return rc;
  FSI_REQUIRE(pop && out_aos, "lbm_export_populations: null buffer");
  FSI_REQUIRE(p->storage != FSI_POP_I16 || (b && b->store_ctrl),
              "lbm_export_populations: I16 storage needs store_ctrl");
  hipStream_t st = (hipStream_t)stream;
  switch (p->storage) {
    case FSI_POP_F32: return lbm_export_t<float>(p, b, pop, out_aos, st);
    case FSI_POP_F16: return lbm_export_t<__half>(p, b, pop, out_aos, st);
    case FSI_POP_I16: return lbm_export_t<int16_t>(p, b, pop, out_aos, st);
    default: return lbm_export_t<double>(p, b, pop, out_aos, st);
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
