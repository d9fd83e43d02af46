// MLS-MPM particle passes with the penalty coupling fused in, for gfx950.
//
// Particles are SoA planes of `capacity` entries (x[a*cap + p], F[(r*3+s)*cap + p]),
// so a wavefront loads 64 consecutive values per component. The grid is AoS
// (momentum xyz, mass) per node, so the 4 atomics of one node hit one cache
// line; the coupling buffer uses the same (force density xyz, phi) layout.
//
// Per substep: clear grid -> P2G (+ coupling) -> grid update + G2P (+ contact).

#include <mutex>
#include <type_traits>

#include "common.h"

namespace fsi {
namespace {

template <typename R> struct MpmK {
  int nx, ny, nz;
  int num;
  int cap;
  R dx, inv_dx, dt;
  R two_mu, lam;
  const R* material;  // optional per-particle (2 mu, lambda) planes, 2 x cap
  R g[3];
  int bw;
  R damping;
  int apic;
  R four_inv_dx;
  R lower, upper[3];
};

template <typename R> MpmK<R> make_mpm(const fsi_mpm_params* p, const fsi_mpm_buffers* b = nullptr) {
  MpmK<R> m;
  m.material = b ? (const R*)b->material : nullptr;
  m.nx = p->nx;
  m.ny = p->ny;
  m.nz = p->nz;
  m.num = p->num_particles;
  m.cap = p->capacity;
  m.dx = R(p->dx);
  m.inv_dx = R(p->inv_dx);
  m.dt = R(p->dt);
  m.two_mu = R(p->two_mu);
  m.lam = R(p->lam);
  for (int a = 0; a < 3; ++a) m.g[a] = R(p->gravity[a]);
  m.bw = p->boundary_width;
  m.damping = R(p->boundary_damping);
  m.apic = p->apic;
  m.four_inv_dx = R(4.0 * p->inv_dx);
  m.lower = R(p->clamp_lower);
  for (int a = 0; a < 3; ++a) m.upper[a] = R(p->clamp_upper[a]);
  return m;
}

template <typename R> struct CplK {
  int enabled, build;
  R gamma, dt_ratio, cell_volume, min_w;
  int use_rel;
  R rel_limit;
  int use_force;
  R force_limit;
  int contact;
  R contact_thr, contact_damp;
  R* force_mean;  // optional step-averaged coupling force accumulator (3 x cap)
  // the lattice the coupling and contact stencils address: its dims and
  // 1 / its cell size (the MPM grid's unless refine > 1)
  int refine;
  int lnx, lny, lnz;
  R inv_dx_c;
};

template <typename R> CplK<R> make_cpl(const fsi_coupling_params* c) {
  CplK<R> k{};
  if (!c) return k;
  k.enabled = c->enabled;
  k.build = c->build_fraction;
  k.gamma = R(c->gamma_eff);
  k.dt_ratio = R(c->dt_ratio);
  k.cell_volume = R(c->cell_volume);
  k.min_w = R(c->min_valid_weight);
  k.use_rel = c->use_relative_velocity_limit;
  k.rel_limit = R(c->relative_velocity_limit);
  k.use_force = c->use_force_limit;
  k.force_limit = R(c->force_limit);
  k.contact = c->contact_enabled;
  k.contact_thr = R(c->contact_threshold);
  k.contact_damp = R(c->contact_damping);
  k.refine = c->refine > 1 ? c->refine : 1;
  return k;
}

// CplK with the lattice geometry of an MPM grid p (MPM dims = refine x
// lattice dims)
template <typename R> CplK<R> make_cpl(const fsi_coupling_params* c, const fsi_mpm_params* p) {
  CplK<R> k = make_cpl<R>(c);
  k.refine = (c && c->refine > 1) ? c->refine : 1;
  k.lnx = p->nx / k.refine;
  k.lny = p->ny / k.refine;
  k.lnz = p->nz / k.refine;
  k.inv_dx_c = R(p->inv_dx / k.refine);
  return k;
}

// Adds the number of lanes with `pred` to *ctr with one atomic per wavefront.
__device__ __forceinline__ void wave_count(int32_t* ctr, bool pred) {
  const unsigned long long b = __ballot(pred);
  const unsigned long long act = __ballot(1);
  const int leader = __ffsll((long long)act) - 1;
  if ((int)__lane_id() == leader && b) atomicAdd(ctr, (int)__popcll(b));
}

// fsi/mpm3d.py:227-232, quadratic B-spline weights per axis
template <typename R> __device__ __forceinline__ void bspline(const R fx[3], R w[3][3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const R t0 = R(1.5) - fx[a];
    const R t1 = fx[a] - R(1);
    const R t2 = fx[a] - R(0.5);
    w[0][a] = R(0.5) * (t0 * t0);
    w[1][a] = R(0.75) - t1 * t1;
    w[2][a] = R(0.5) * (t2 * t2);
  }
}

// base node (truncating cast, fsi/mpm3d.py:239) and fractional offset
// No contraction: x*inv_dx is rounded before the -0.5, exactly as written, so
// the cell key of a particle on a cell face is the oracle's bit for bit.
template <typename R>
__device__ __forceinline__ void locate(const R x[3], R inv_dx, int base[3], R fx[3]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const R gp = x[a] * inv_dx;
    base[a] = (int)(gp - R(0.5));
    fx[a] = gp - R(base[a]);
  }
}

__device__ __forceinline__ bool inside(int nx, int ny, int nz, int i, int j, int k) {
  return (unsigned)i < (unsigned)nx && (unsigned)j < (unsigned)ny && (unsigned)k < (unsigned)nz;
}

// Node index for a gather that is always in bounds: nodes outside the
// lattice map to node 0 and report in = false, so all 27 loads of a particle
// can be issued together instead of one dependent round trip per node.
__device__ __forceinline__ int64_t gather_node(int nx, int ny, int nz, int i, int j, int k, bool& in) {
  in = inside(nx, ny, nz, i, j, k);
  return in ? ((int64_t)i * ny + j) * nz + k : 0;
}

// out[s] = (a0, a1, a2)[(s + r) % 3] for r in {0, 1, 2}, with selects
template <typename T>
__device__ __forceinline__ void rotate3(int r, T a0, T a1, T a2, T out[3]) {
  out[0] = r == 0 ? a0 : (r == 1 ? a1 : a2);
  out[1] = r == 0 ? a1 : (r == 1 ? a2 : a0);
  out[2] = r == 0 ? a2 : (r == 1 ? a0 : a1);
}

template <typename R> __device__ __forceinline__ R pick3(int i, R a, R b, R c) {
  return i == 0 ? a : (i == 1 ? b : c);
}

template <typename R> __device__ __forceinline__ R pick4(int i, R a, R b, R c, R d) {
  return i < 2 ? (i == 0 ? a : b) : (i == 2 ? c : d);
}

// ------------------------------------------------------------- SVD ----
//
// Rotation factor R = U V^T of the rotation-variant SVD F = U S V^T (U, V
// proper rotations), the quantity fsi/mpm3d.py:245-246 builds from ti.svd:
// cyclic Jacobi on the symmetric F^T F gives V, a Givens QR of F V gives U
// (its diagonal carries the singular values, the last with the sign of det F).

template <typename R>
__device__ __forceinline__ void jacobi_rotate(R A[3][3], R V[3][3], int p, int q) {
  const R apq = A[p][q];
  if (apq == R(0)) return;
  const R theta = (A[q][q] - A[p][p]) / (R(2) * apq);
  const R at = fabs(theta);
  R t;
  if (at > R(1e15))
    t = R(0.5) / theta;
  else
    t = (theta >= R(0) ? R(1) : R(-1)) / (at + sqrt(theta * theta + R(1)));
  const R c = R(1) / sqrt(t * t + R(1));
  const R s = t * c;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const R akp = A[k][p], akq = A[k][q];
    A[k][p] = c * akp - s * akq;
    A[k][q] = s * akp + c * akq;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const R apk = A[p][k], aqk = A[q][k];
    A[p][k] = c * apk - s * aqk;
    A[q][k] = s * apk + c * aqk;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const R vkp = V[k][p], vkq = V[k][q];
    V[k][p] = c * vkp - s * vkq;
    V[k][q] = s * vkp + c * vkq;
  }
}

template <typename R> __device__ __forceinline__ void swap_cols(R V[3][3], R lam[3], int a, int b) {
  const R t = lam[a];
  lam[a] = lam[b];
  lam[b] = t;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const R v = V[k][a];
    V[k][a] = V[k][b];
    V[k][b] = -v;  // keep det V = +1
  }
}

template <typename R>
__device__ __forceinline__ void rotation_core(const R F[9], R Rot[9], R* lam_min_over_max) {
  R A[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int s = 0; s < 3; ++s)
      A[r][s] = F[0 * 3 + r] * F[0 * 3 + s] + F[1 * 3 + r] * F[1 * 3 + s] + F[2 * 3 + r] * F[2 * 3 + s];
  R V[3][3] = {{R(1), R(0), R(0)}, {R(0), R(1), R(0)}, {R(0), R(0), R(1)}};
  constexpr int kSweeps = sizeof(R) == 8 ? 12 : 8;
  const R eps = sizeof(R) == 8 ? R(1e-30) : R(1e-14);
  for (int sweep = 0; sweep < kSweeps; ++sweep) {
    const R off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    const R dg = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
    if (off <= eps * dg) break;
    jacobi_rotate(A, V, 0, 1);
    jacobi_rotate(A, V, 0, 2);
    jacobi_rotate(A, V, 1, 2);
  }
  R lam[3] = {A[0][0], A[1][1], A[2][2]};
  // sort descending (three compare-swaps)
  if (lam[0] < lam[1]) swap_cols(V, lam, 0, 1);
  if (lam[0] < lam[2]) swap_cols(V, lam, 0, 2);
  if (lam[1] < lam[2]) swap_cols(V, lam, 1, 2);
  if (lam_min_over_max) *lam_min_over_max = lam[0] > R(0) ? lam[2] / lam[0] : R(0);
  // B = F V and its Givens QR; Q accumulates the rotations
  R B[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int s = 0; s < 3; ++s)
      B[r][s] = F[r * 3 + 0] * V[0][s] + F[r * 3 + 1] * V[1][s] + F[r * 3 + 2] * V[2][s];
  R Qm[3][3] = {{R(1), R(0), R(0)}, {R(0), R(1), R(0)}, {R(0), R(0), R(1)}};
  constexpr int ra[3] = {0, 0, 1}, rb[3] = {1, 2, 2}, cc[3] = {0, 0, 1};
#pragma unroll
  for (int gi = 0; gi < 3; ++gi) {
    const int a = ra[gi], b = rb[gi], col = cc[gi];
    const R x = B[a][col], y = B[b][col];
    const R r = sqrt(x * x + y * y);
    R c = R(1), s = R(0);
    if (r > R(0)) {
      c = x / r;
      s = y / r;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const R ba = B[a][k], bb = B[b][k];
      B[a][k] = c * ba + s * bb;
      B[b][k] = -s * ba + c * bb;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const R qa = Qm[k][a], qb = Qm[k][b];
      Qm[k][a] = c * qa + s * qb;
      Qm[k][b] = -s * qa + c * qb;
    }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int s = 0; s < 3; ++s)
      Rot[r * 3 + s] = Qm[r][0] * V[s][0] + Qm[r][1] * V[s][1] + Qm[r][2] * V[s][2];
}

// Rotation factor of F: FP64 builds use the Jacobi SVD above; FP32 uses the
// Newton polar iteration and recomputes in FP64 (Jacobi) for inverted or
// ill-conditioned F.
// Orthogonal polar factor of F by the determinant-scaled Newton iteration
// X <- (g X + (g X)^-T) / 2, g = |det X|^(-1/3) (N. J. Higham, "Computing
// the polar decomposition - with applications", SIAM J. Sci. Stat. Comput. 7
// (1986) 1160-1174). For det F > 0 its limit is the polar factor, which is
// R = U V^T of the rotation-variant SVD. X^-T = cof(X) / det X. Returns false
// (and leaves Rot unspecified) when det F <= 0 or the Frobenius condition
// number of F exceeds 300 (the squared singular values may then span more
// than 1e5); the caller then takes the FP64 Jacobi path.
__device__ __forceinline__ bool polar_newton(const float F[9], float X[9]) {
  constexpr int kMaxIter = 10;
#pragma unroll
  for (int i = 0; i < 9; ++i) X[i] = F[i];
  for (int it = 0; it < kMaxIter; ++it) {
    float c[9];
    c[0] = X[4] * X[8] - X[5] * X[7];
    c[1] = X[5] * X[6] - X[3] * X[8];
    c[2] = X[3] * X[7] - X[4] * X[6];
    c[3] = X[2] * X[7] - X[1] * X[8];
    c[4] = X[0] * X[8] - X[2] * X[6];
    c[5] = X[1] * X[6] - X[0] * X[7];
    c[6] = X[1] * X[5] - X[2] * X[4];
    c[7] = X[2] * X[3] - X[0] * X[5];
    c[8] = X[0] * X[4] - X[1] * X[3];
    const float det = X[0] * c[0] + X[1] * c[1] + X[2] * c[2];
    if (it == 0) {
      float nf = 0.0f, nc = 0.0f;
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        nf += X[i] * X[i];
        nc += c[i] * c[i];
      }
      // kappa_F = |F|_F |F^-1|_F = |F|_F |cof F|_F / det
      if (!(det > 0.0f) || !(nf * nc <= 9.0e4f * det * det)) return false;
    }
    const float g = exp2f(-log2f(det) * (1.0f / 3.0f));
    const float a = 0.5f * g, b = 0.5f / (g * det);
    float diff = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const float xn = a * X[i] + b * c[i];
      diff = fmaxf(diff, fabsf(xn - X[i]));
      X[i] = xn;
    }
    if (diff <= 1.0e-6f) break;
  }
  return true;
}

template <typename R> __device__ __forceinline__ void rotation_factor(const R F[9], R Rot[9], R J) {
  if constexpr (sizeof(R) == 8) {
    rotation_core<R>(F, Rot, nullptr);
  } else {
    (void)J;
    if (!polar_newton(F, Rot)) {
      double Fd[9], Rd[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) Fd[i] = double(F[i]);
      rotation_core<double>(Fd, Rd, nullptr);
#pragma unroll
      for (int i = 0; i < 9; ++i) Rot[i] = R(Rd[i]);
    }
  }
}

template <typename R> __device__ __forceinline__ R det3(const R m[9]) {
  return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) +
         m[6] * (m[1] * m[5] - m[4] * m[2]);
}

// ------------------------------------------------------- P2G pass ----

// F <- (I + dt C) F (fsi/mpm3d.py:243)
template <typename R>
__device__ __forceinline__ void deformation_update(R dt, const R Cm[9], const R Fo[9], R Fn[9]) {
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int s2 = 0; s2 < 3; ++s2) {
      R acc = R(0);
#pragma unroll
      for (int t = 0; t < 3; ++t) acc += ((r == t ? R(1) : R(0)) + dt * Cm[r * 3 + t]) * Fo[t * 3 + s2];
      Fn[r * 3 + s2] = acc;
    }
}

// Affine momentum matrix and the particle's momentum (fsi/mpm3d.py:244-262):
// P = 2 mu (F - R) F^T + lam J (J - 1) I, aff = -dt vol 4/dx^2 P + m C,
// mom0 = m v + dt f (the coupling force of this substep)
template <typename R>
__device__ __forceinline__ void affine_momentum(const MpmK<R>& m, int p, const R Fn[9], const R Rot[9],
                                                const R Cm[9], const R vp[3], const R pf[3], R mp, R vp_vol,
                                                R aff[9], R mom0[3]) {
  // Lame parameters: the material's (fsi/mpm3d.py:33-38), or the particle's
  // own when a per-particle material plane is given (no counterpart in the
  // reference, which has one global E and nu)
  R two_mu = m.two_mu, lam = m.lam;
  if (m.material) {
    two_mu = m.material[p];
    lam = m.material[m.cap + p];
  }
  const R J = det3(Fn);
  const R scale = -m.dt * vp_vol * R(4) * m.inv_dx * m.inv_dx;
  const R lj = lam * J * (J - R(1));
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int s2 = 0; s2 < 3; ++s2) {
      R st = R(0);
#pragma unroll
      for (int t = 0; t < 3; ++t) st += (two_mu * (Fn[r * 3 + t] - Rot[r * 3 + t])) * Fn[s2 * 3 + t];
      if (r == s2) st += lj;
      aff[r * 3 + s2] = scale * st + mp * Cm[r * 3 + s2];
    }
#pragma unroll
  for (int a = 0; a < 3; ++a) mom0[a] = mp * vp[a] + m.dt * pf[a];
}

// P2G scatter of one particle straight to the global grid (in-lattice nodes
// only), for workgroups whose particles do not fit an LDS tile
template <typename R>
__device__ __forceinline__ void scatter_global(const MpmK<R>& m, R* __restrict__ grid, const int base[3],
                                               const R fx[3], const R w[3][3], const R mom0[3], const R aff[9],
                                               R mp) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int ni = base[0] + i, nj = base[1] + j, nk = base[2] + k;
        if (!inside(m.nx, m.ny, m.nz, ni, nj, nk)) continue;
        const R dp[3] = {(R(i) - fx[0]) * m.dx, (R(j) - fx[1]) * m.dx, (R(k) - fx[2]) * m.dx};
        const R wt = w[i][0] * w[j][1] * w[k][2];
        const int64_t node = ((int64_t)ni * m.ny + nj) * m.nz + nk;
#pragma unroll
        for (int a = 0; a < 3; ++a)
          atomicAdd(grid + 4 * node + a,
                    wt * (mom0[a] + (aff[a * 3 + 0] * dp[0] + aff[a * 3 + 1] * dp[1] + aff[a * 3 + 2] * dp[2])));
        atomicAdd(grid + 4 * node + 3, wt * mp);
      }
}

// grid clear (fsi/mpm3d.py:205-209) and, when given, the per-substep
// particle-diagnostic counters (fsi/coupling.py:291-297)
template <typename R>
__global__ __launch_bounds__(kBlock) void k_clear(int64_t n, R* __restrict__ buf,
                                                  int32_t* __restrict__ counters) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) buf[t] = R(0);
  if (counters && t < 5) counters[t] = 0;
}

// Plane split (PL = 3): a workgroup takes 64 particle slots (a quarter of a
// work-list chunk) with three wavefronts, one per x plane of the stencil, so
// the passes of a small particle set spread over four times as many CUs and
// each wavefront walks 9 of the 27 nodes.
template <int PL> struct Slots {
  static constexpr int value = PL > 1 ? 64 : kBlock;  // particle slots per workgroup
  static constexpr int per_chunk = kBlock / value;    // workgroups per work-list chunk
};

// Batched substeps record the node box each P2G workgroup scattered into
// (8 ints per workgroup: x0, y0, z0, tx, ty, tz, 0, 0), so the G2P pass that
// clears that grid for its next use zeroes only those boxes. A workgroup
// whose box exceeds this many nodes flags the pass instead, and the whole
// grid is zeroed.
constexpr int kZeroBoxMaxNodes = 8192;

// Per-workgroup P2G accumulation tile in LDS: 8 values per node (grid
// momentum xyz + mass, coupling force density xyz + phi), value-major with a
// fixed stride of kNodes 8-byte words, so the value index is an immediate
// offset of the LDS atomic. The word type is TileAcc<R>::word.
struct Tile {
  static constexpr int kWordBytes = 8;
  static constexpr int kValues = 8;
  // a work-list chunk (two z-adjacent 4x4x4 cell blocks) spans 6x6x10 = 360
  // nodes plus drift since the last sort; 496 nodes keep the tile (31 KB)
  // and the coupling's staged lattice velocity (7.75 KB, float) under 40 KB,
  // so four workgroups fit the 160 KB of a CU
  static constexpr int kNodes = 496;
};

// Grid nodes a G2P workgroup stages in LDS (the velocity after the grid
// update, 3 values per node): a work-list chunk spans at most 360.
constexpr int kG2PNodes = 512;
template <typename R> constexpr int g2p_lds_bytes() { return kG2PNodes * 3 * (int)sizeof(R); }

// The G2P velocity stage and the P2G tile (with the coupling's staged lattice
// velocity) share one LDS buffer: in the fused launch the P2G part writes its
// tile only after the workgroup barrier of its node box, which every thread
// reaches after its last read of the G2P stage. Sharing takes the fused FP32
// coupling kernel from 46 KB to 39.9 KB per workgroup, so four fit a CU.
template <typename R, bool COUPLE> constexpr int p2g_lds_bytes() {
  return Tile::kNodes * Tile::kValues * Tile::kWordBytes + (COUPLE ? Tile::kNodes : 1) * 4 * (int)sizeof(R);
}

// Tile words. FP64 build: FP64 sums (ds_add_f64). FP32 build: 64-bit fixed
// point (ds_add_u64). On gfx950 an LDS atomic wave instruction on scattered
// addresses costs ~28 clocks as ds_add_f64, ~16 as ds_add_u64 and ~196 as
// ds_add_f32 (processed lane by lane), and the tile atomics are most of the
// fused substep. Each workgroup scales its FP32 contributions by a power of
// two 2^S chosen from the largest magnitude any of its lanes adds (TileScale)
// so that every scaled contribution is below 2^30 in magnitude; it is rounded
// to an integer (v_cvt_rpi_i32_f32: floor(c + 1/2)) and added as a signed
// 64-bit word, so a node sum of any number of contributions below 2^33 is
// exact and independent of the order of the atomics. A contribution keeps
// 2^-30 of the workgroup's largest contribution, below the rounding of the
// FP32 product itself (2^-24 of its own magnitude) wherever the product is
// within 2^6 of that largest value. The flush converts each sum back once
// per node and value.
template <typename R> struct TileAcc;
template <> struct TileAcc<double> {
  using word = double;
  using P = double;  // type of the per-node products
  static constexpr bool fixed = false;
  __device__ static void add(word* a, double c) { atomicAdd(a, c); }
  __device__ static double value(word s, double) { return s; }
};
template <> struct TileAcc<float> {
  using word = unsigned long long;
  using P = float;
  static constexpr bool fixed = true;
  // c: the contribution times 2^S, |c| < 2^30: one rounding conversion and
  // the sign extension
  __device__ static void add(word* a, float c) {
    int32_t r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(c));
    atomicAdd(a, (word)(long long)r);
  }
  // sum * 2^-S (inv = 2^-S; NaN when the workgroup added a non-finite value)
  __device__ static float value(word s, double inv) {
    return (float)((double)(long long)s * inv);
  }
};


template <typename R, bool COUPLE> constexpr int pass_lds_bytes() {
  return p2g_lds_bytes<R, COUPLE>() > g2p_lds_bytes<R>() ? p2g_lds_bytes<R, COUPLE>() : g2p_lds_bytes<R>();
}
// one buffer per size: the passes of one kernel that name the same size
// address the same LDS
template <int BYTES> __device__ __forceinline__ unsigned char* pass_lds() {
  __shared__ __align__(16) unsigned char buf[BYTES];
  return buf;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
  return v;
}

// Power-of-two scales of a workgroup's tile contributions, one per tile
// value (momentum xyz and mass, or coupling force xyz and phi: quantities of
// different magnitudes): every contribution to value a has |c| <= the
// workgroup maximum of bound[a]. mul[a] = 2^S_a is applied to each FP32
// contribution before TileAcc::add, inv[a] = 2^-S_a at the flush. Every
// thread calls it (it synchronises the workgroup).
struct TileScale {
  float mul[4];
  double inv[4];
};
template <int NW>
__device__ __forceinline__ TileScale tile_scale(const float bound[4], uint32_t (*red)[NW]) {
  // nonnegative floats order as their bit patterns; NaN (0x7fc00000) and
  // inf order above every finite value
  uint32_t m[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const float b = bound[a] >= 0.0f ? bound[a] : (bound[a] == bound[a] ? -bound[a] : __int_as_float(0x7fc00000));
    m[a] = __float_as_uint(wave_max_nonneg(b));
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int a = 0; a < 4; ++a) red[a][threadIdx.x >> 6] = m[a];
  }
  __syncthreads();
  TileScale s;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    uint32_t w = red[a][0];
#pragma unroll
    for (int i = 1; i < NW; ++i) w = max(w, red[a][i]);
    const float mx = __uint_as_float(w);
    if (!(mx < __builtin_huge_valf())) {  // inf or NaN: sums become NaN at the flush
      s.mul[a] = 0.0f;
      s.inv[a] = __builtin_nan("");
      continue;
    }
    // mx < 2^e; S = 30 - e keeps each scaled contribution below 2^30 (and
    // its rounding in int32); 27 contributions per particle of a workgroup
    // of <= 256 particles sum far below 2^63
    int e = 0;
    (void)frexpf(mx, &e);
    e = min(max(e, -96), 100);  // 2^(30 - e) stays a finite float
    s.mul[a] = ldexpf(1.0f, 30 - e);
    s.inv[a] = ldexp(1.0, e - 30);
  }
  return s;
}

// ----------------------------------------------- workgroup node box ----

// Bounding box of the base cells of the workgroup's active particles, grown
// by the 3-node stencil: nodes [x0, x0+tx) x [y0, y0+ty) x [z0, z0+tz). Every
// thread must call it (it synchronises the workgroup).
struct NodeBox {
  int x0, y0, z0, tx, ty, tz;
  long long nodes;  // 0 when the workgroup has no active particle
  __device__ int index(int ni, int nj, int nk) const { return ((ni - x0) * ty + (nj - y0)) * tz + (nk - z0); }
  __device__ void coords(int t, int& ni, int& nj, int& nk) const {
    nk = z0 + t % tz;
    nj = y0 + (t / tz) % ty;
    ni = x0 + t / (tz * ty);
  }
};

template <int NW>
__device__ NodeBox workgroup_box(bool act, const int base[3], int (*red)[NW]) {
  constexpr int kBig = 1 << 30;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int mn[3], mx[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    mn[a] = wave_min(act ? base[a] : kBig);
    mx[a] = wave_max(act ? base[a] : -kBig);
  }
  if (lane == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      red[a][wid] = mn[a];
      red[3 + a][wid] = mx[a];
    }
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    mn[a] = red[a][0];
    mx[a] = red[3 + a][0];
#pragma unroll
    for (int wv = 1; wv < NW; ++wv) {
      mn[a] = min(mn[a], red[a][wv]);
      mx[a] = max(mx[a], red[3 + a][wv]);
    }
  }
  NodeBox b;
  const bool any = mn[0] <= mx[0];
  b.x0 = mn[0];
  b.y0 = mn[1];
  b.z0 = mn[2];
  // extents clamped so that the product cannot overflow for far-apart cells
  b.tx = any ? (int)min((long long)mx[0] - mn[0] + 3, (long long)kBig) : 0;
  b.ty = any ? (int)min((long long)mx[1] - mn[1] + 3, (long long)kBig) : 0;
  b.tz = any ? (int)min((long long)mx[2] - mn[2] + 3, (long long)kBig) : 0;
  b.nodes = (long long)b.tx * b.ty * b.tz;
  return b;
}


// fsi/coupling.py:332-406 (COUPLE) fused with fsi/mpm3d.py:234-266 (P2G).
// The grid and coupling accumulations go through a per-workgroup LDS tile
// (LDS float atomics) and are flushed with one global atomic per non-zero
// tile entry; the flush skips nodes outside the lattice, which receive nothing.
// WIDE: gather all 27 nodes in one batch (one memory round trip, more
// registers: for grids that fit on the chip at once); otherwise one i-plane
// of 9 nodes per batch (fewer registers, more wavefronts per SIMD).
// P2G pass of one workgroup (fsi/mpm3d.py:234-266 with the coupling kernel
// of fsi/coupling.py:332-406 fused in front when COUPLE). Lane particle p
// (live: a particle of this workgroup). REGS: x, v and C come from the
// caller's registers (the fused G2P+P2G launch) instead of memory.
// REF: the MPM grid is cp.refine times finer than the lattice; the coupling
// interpolates and spreads on the lattice's cells (its own stencil, node box
// and tile rows), the P2G on the MPM grid's.
template <typename R, bool COUPLE, bool P2G, bool WIDE, bool REGS, int PL, bool REF = false>
__device__ __forceinline__ void p2g_pass(const MpmK<R>& m, const CplK<R>& cp, int p, bool live, R* __restrict__ x,
                                         R* __restrict__ v, const R* __restrict__ C, R* __restrict__ Fd,
                                         const R* __restrict__ mass, const R* __restrict__ vol,
                                         const int32_t* __restrict__ active, R* __restrict__ pforce,
                                         R* __restrict__ grid, const int32_t* __restrict__ solid,
                                         const R* __restrict__ lbm_vel, R* __restrict__ cforce,
                                         R* __restrict__ valid_weight, int32_t* __restrict__ cmask,
                                         int32_t* __restrict__ counters, int clear_force,
                                         int32_t* __restrict__ boxes_out, int32_t* __restrict__ big_flag,
                                         int big_value, const R* __restrict__ closure, const R* rx, const R* rv,
                                         const R* rC) {
  // the plane split (small particle sets, bound by one wavefront's latency,
  // not by the LDS atomics) keeps FP64 tile sums: the fixed-point scale's
  // extra workgroup reductions would lengthen its critical path
  using TA = TileAcc<std::conditional_t<PL == 1, R, double>>;
  using A = typename TA::word;
  using P = typename TA::P;
  static_assert(sizeof(A) == Tile::kWordBytes, "tile word size");
  constexpr int kN = Tile::kNodes;
  A* tile = reinterpret_cast<A*>(pass_lds<pass_lds_bytes<R, COUPLE>()>());
  // coupling: lattice velocity and a fluid flag per node of the box (zero
  // velocity and flag on solid and out-of-lattice nodes)
  R(*ctile)[4] = reinterpret_cast<R(*)[4]>(pass_lds<pass_lds_bytes<R, COUPLE>()>() +
                                           kN * Tile::kValues * sizeof(A));
  constexpr int kS = Slots<PL>::value;
  __shared__ int red[6][PL * kS / 64];
  // fixed-point tile: maxima of the coupling and the P2G contributions
  __shared__ uint32_t sred[2][4][PL * kS / 64];
  TileScale cscale{{1.0f, 1.0f, 1.0f, 1.0f}, {1.0, 1.0, 1.0, 1.0}}, pscale = cscale;
  // plane split (PL = 3): partial support and interpolated velocity
  __shared__ R cpart[PL > 1 && COUPLE ? PL : 1][PL > 1 && COUPLE ? kS : 1][4];
  static_assert(PL == 1 || (PL == 3 && sizeof(R) == 4), "plane split: FP32 only");
  const int pln = PL > 1 ? (int)threadIdx.x / kS : 0;
  const int slot = PL > 1 ? (int)threadIdx.x % kS : (int)threadIdx.x;
  const bool writer = pln == 0;  // the wavefront that writes the particle
  const bool act = live && active[p] != 0;
  const int cap = m.cap;
  const int64_t N = (int64_t)m.nx * m.ny * m.nz;

  bool partial = false, unsupported = false, clipped = false;
  R pf[3] = {R(0), R(0), R(0)};
  R xp[3] = {R(0), R(0), R(0)}, vp[3] = {R(0), R(0), R(0)}, fx[3] = {R(0), R(0), R(0)}, w[3][3];
  R Cm[9], Fo[9], mp = R(0), vp_vol = R(0);
  int base[3] = {0, 0, 0};
  // per-particle state used by both parts (C and F are loaded where the
  // P2G part needs them, keeping registers free during the coupling gather)
  if (act) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      xp[a] = REGS ? rx[a] : x[a * cap + p];
      vp[a] = REGS ? rv[a] : v[a * cap + p];
    }
    mp = mass[p];
    vp_vol = vol[p];
    locate(xp, m.inv_dx, base, fx);
  }
  bspline(fx, w);
  // coupling stencil: the lattice's (REF), else the MPM stencil itself
  int cbase[3] = {base[0], base[1], base[2]};
  R cw[3][3];
  if constexpr (REF) {
    R cfx[3] = {R(0), R(0), R(0)};
    if (act) locate(xp, cp.inv_dx_c, cbase, cfx);
    bspline(cfx, cw);
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) cw[a][b] = w[a][b];
  }
  const int lnx = REF ? cp.lnx : m.nx, lny = REF ? cp.lny : m.ny, lnz = REF ? cp.lnz : m.nz;
  const int64_t NL = REF ? (int64_t)lnx * lny * lnz : N;  // lattice cells

  // bounding box of the workgroup's base cells -> tile of (box + 2) nodes
  constexpr int kBig = 1 << 30;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __shared__ int cred[REF ? 6 : 1][REF ? PL * kS / 64 : 1];
  int mn[3], mx[3], cmn[3], cmx[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    mn[a] = wave_min(act ? base[a] : kBig);
    mx[a] = wave_max(act ? base[a] : -kBig);
    if constexpr (REF) {
      cmn[a] = wave_min(act ? cbase[a] : kBig);
      cmx[a] = wave_max(act ? cbase[a] : -kBig);
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      red[a][wid] = mn[a];
      red[3 + a][wid] = mx[a];
      if constexpr (REF) {
        cred[a][wid] = cmn[a];
        cred[3 + a][wid] = cmx[a];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    mn[a] = red[a][0];
    mx[a] = red[3 + a][0];
#pragma unroll
    for (int wv = 1; wv < PL * kS / 64; ++wv) {
      mn[a] = min(mn[a], red[a][wv]);
      mx[a] = max(mx[a], red[3 + a][wv]);
    }
    if constexpr (REF) {
      cmn[a] = cred[a][0];
      cmx[a] = cred[3 + a][0];
#pragma unroll
      for (int wv = 1; wv < PL * kS / 64; ++wv) {
        cmn[a] = min(cmn[a], cred[a][wv]);
        cmx[a] = max(cmx[a], cred[3 + a][wv]);
      }
    } else {
      cmn[a] = mn[a];
      cmx[a] = mx[a];
    }
  }
  const bool any = mn[0] <= mx[0];
  const long long tx = any ? (long long)mx[0] - mn[0] + 3 : 0;
  const long long ty = any ? (long long)mx[1] - mn[1] + 3 : 0;
  const long long tz = any ? (long long)mx[2] - mn[2] + 3 : 0;
  const long long tnodes = tx * ty * tz;
  // the coupling's node box on the lattice (REF; else the same box)
  const long long ctx = any ? (long long)cmx[0] - cmn[0] + 3 : 0;
  const long long cty = any ? (long long)cmx[1] - cmn[1] + 3 : 0;
  const long long ctz = any ? (long long)cmx[2] - cmn[2] + 3 : 0;
  const long long ctnodes = ctx * cty * ctz;
  // scatter target of the workgroup: its LDS tile when the cells of its
  // particles fit, otherwise the global grid/coupling buffers directly
  const bool use_lds = any && tnodes <= kN && (!REF || ctnodes <= kN);
  const int x0 = mn[0], y0 = mn[1], z0 = mn[2];
  const int cx0 = cmn[0], cy0 = cmn[1], cz0 = cmn[2];
  if (boxes_out && threadIdx.x == 0) {
    // every node this workgroup writes (tile flush, untiled and deferred
    // scatters) lies in the box of its particles' stencils
    int32_t* bo = boxes_out + 8 * blockIdx.x;
    bo[0] = x0;
    bo[1] = y0;
    bo[2] = z0;
    bo[3] = (int)tx;
    bo[4] = (int)ty;
    bo[5] = (int)tz;
    if (tnodes > kZeroBoxMaxNodes) *big_flag = big_value;
  }
  const int tyi = use_lds ? (int)ty : 0, tzi = use_lds ? (int)tz : 0;
  // the tile spans base..base+2 of every particle of the workgroup, so the
  // scatter needs no per-node bounds test; out-of-lattice entries, and the
  // coupling values of solid nodes, are dropped at the flush
  const int tn_all = use_lds ? (int)tnodes : 0;
  const int ctyi = use_lds ? (int)cty : 0, ctzi = use_lds ? (int)ctz : 0;
  const int ctn_all = use_lds ? (int)ctnodes : 0;
#define FSI_GLOBAL(node, slot, val)                      \
  do {                                                   \
    if ((slot) < 4)                                      \
      atomicAdd(grid + 4 * (node) + (slot), (val));      \
    else                                                 \
      atomicAdd(cforce + 4 * (node) + ((slot) - 4), (val)); \
  } while (0)
  for (int t = threadIdx.x; t < (REF ? max(tn_all, ctn_all) : tn_all); t += blockDim.x) {
    if (!REF || t < tn_all) {
#pragma unroll
      for (int s = 0; s < (COUPLE && !REF ? Tile::kValues : 4); ++s) tile[s * kN + t] = A(0);
    }
    if constexpr (COUPLE) {
      if (REF && t >= ctn_all) continue;
      if constexpr (REF) {
#pragma unroll
        for (int s = 4; s < Tile::kValues; ++s) tile[s * kN + t] = A(0);
      }
      // stage the box's lattice velocity and fluid flag (in the lattice and
      // not solid) once per node; the 27-node gathers then read LDS
      const int ni = cx0 + t / (ctzi * ctyi), nj = cy0 + (t / ctzi) % ctyi, nk = cz0 + t % ctzi;
      bool in;
      const int64_t node = gather_node(lnx, lny, lnz, ni, nj, nk, in);
      const int32_t sd = solid[node];
      const R u0 = lbm_vel[node], u1 = lbm_vel[NL + node], u2 = lbm_vel[2 * NL + node];
      const bool ok = in && sd == 0;
      ctile[t][0] = ok ? u0 : R(0);
      ctile[t][1] = ok ? u1 : R(0);
      ctile[t][2] = ok ? u2 : R(0);
      ctile[t][3] = ok ? R(1) : R(0);
    }
  }
  __syncthreads();
  // lane-rotated node order of the LDS scatters: lane walks i, j and k from
  // offsets (lane / 9 % 3, lane % 3, lane / 3 % 3), so up to 27 particles
  // in consecutive lanes add to 27 different nodes on every instruction, and
  // the lanes of neighbouring cells spread over three x planes (LDS atomics
  // to one address serialise, ~12 clk per extra lane on gfx950); the order
  // is static, the rotated weights, offsets and tile rows are formed once
  // per particle
  const int ri = (lane / 9) % 3, rj = lane % 3, rk = (lane / 3) % 3;
  int io[3], jo[3], ko[3];
  rotate3(ri, 0, 1, 2, io);
  rotate3(rj, 0, 1, 2, jo);
  rotate3(rk, 0, 1, 2, ko);
  // tile word of node (base + (io[ii], jo[jj], ko[kk])):
  // plane[ii] + row[jj] + ko[kk]
  const int tyz = tyi * tzi;
  const int b0 = ((base[0] - x0) * tyi + (base[1] - y0)) * tzi + (base[2] - z0);
  int row[3], plane[3];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    row[s] = b0 + jo[s] * tzi;
    plane[s] = io[s] * tyz;
  }
  P wir[3], wjr[3], wkr[3];
  rotate3<P>(ri, w[0][0], w[1][0], w[2][0], wir);
  rotate3<P>(rj, w[0][1], w[1][1], w[2][1], wjr);
  rotate3<P>(rk, w[0][2], w[1][2], w[2][2], wkr);
  // plane split: this wavefront's x plane of the stencil is the first (and
  // only) one of the static order
  constexpr int kPlanes = PL > 1 ? 1 : 3;
  if constexpr (PL > 1) {
    io[0] = pln;
    plane[0] = pln * tyz;
    wir[0] = P(pick3(pln, w[0][0], w[1][0], w[2][0]));
  }
  // the same for the coupling's box and stencil (REF: on the lattice)
  const int ctyz = REF ? ctyi * ctzi : tyz;
  const int cb0 = REF ? ((cbase[0] - cx0) * ctyi + (cbase[1] - cy0)) * ctzi + (cbase[2] - cz0) : b0;
  int crow[3], cplane[3];
  P cwir[3], cwjr[3], cwkr[3];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    crow[s] = REF ? cb0 + jo[s] * ctzi : row[s];
    cplane[s] = REF ? io[s] * ctyz : plane[s];
  }
  if constexpr (REF) {
    rotate3<P>(ri, cw[0][0], cw[1][0], cw[2][0], cwir);
    rotate3<P>(rj, cw[0][1], cw[1][1], cw[2][1], cwjr);
    rotate3<P>(rk, cw[0][2], cw[1][2], cw[2][2], cwkr);
    // plane split: this wavefront's x plane of the lattice stencil
    if constexpr (PL > 1) cwir[0] = P(pick3(pln, cw[0][0], cw[1][0], cw[2][0]));
  } else {
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      cwir[s] = wir[s];
      cwjr[s] = wjr[s];
      cwkr[s] = wkr[s];
    }
  }

  if constexpr (COUPLE) {
    if (live && !act && writer) {
      valid_weight[p] = R(0);
      cmask[p] = 0;
    }
    // support of the valid (in-lattice, fluid) nodes and the interpolated
    // lattice velocity, in i, j, k order (with the plane split: this
    // wavefront's x plane, the three planes' sums added in plane order)
    R wsum = R(0), u[3] = {R(0), R(0), R(0)};
    uint32_t valid = 0;  // global path: bit q = (i*3+j)*3+k
    if (act && use_lds) {
      // staged values are zero on invalid nodes and the flag is 0 or 1, so
      // the products are the masked sums without selects
      if constexpr (PL > 1) {
        const R wi = pick3(pln, cw[0][0], cw[1][0], cw[2][0]);
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const R* c = ctile[cb0 + pln * ctyz + j * ctzi + k];
            const R c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
            const R wt = wi * cw[j][1] * cw[k][2];
            wsum += wt * c3;
            u[0] += wt * c0;
            u[1] += wt * c1;
            u[2] += wt * c2;
          }
      } else {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              const R* c = ctile[cb0 + i * ctyz + j * ctzi + k];
              const R c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
              const R wt = cw[i][0] * cw[j][1] * cw[k][2];
              wsum += wt * c3;
              u[0] += wt * c0;
              u[1] += wt * c1;
              u[2] += wt * c2;
            }
      }
    }
    if constexpr (PL > 1) {
      if (use_lds) {  // uniform over the workgroup
        cpart[pln][slot][0] = wsum;
        cpart[pln][slot][1] = u[0];
        cpart[pln][slot][2] = u[1];
        cpart[pln][slot][3] = u[2];
        __syncthreads();
        wsum = (cpart[0][slot][0] + cpart[1][slot][0]) + cpart[2][slot][0];
#pragma unroll
        for (int a = 0; a < 3; ++a) u[a] = (cpart[0][slot][1 + a] + cpart[1][slot][1 + a]) + cpart[2][slot][1 + a];
      }
    }
    // the reaction values this lane spreads on the tile (scaled by 1/wsum)
    bool spreads = false;
    P rfn[4] = {P(0), P(0), P(0), P(0)};
    if (act) {
      // global path (box too large for the tile): the nodes are gathered in
      // batches of NB i-planes, all loads of a batch in flight together;
      // every plane of a split gathers all 27
      constexpr int NB = WIDE ? 3 : 1;  // i-planes per gather batch (global path)
#pragma unroll 1
      for (int i0 = use_lds ? 3 : 0; i0 < 3; i0 += NB) {
        int32_t sn[NB * 9];
        R un[NB * 9][3];
        uint32_t inmask = 0;
#pragma unroll
        for (int ii = 0; ii < NB; ++ii)
#pragma unroll
          for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              bool in;
              const int64_t node = gather_node(lnx, lny, lnz, cbase[0] + i0 + ii, cbase[1] + j, cbase[2] + k, in);
              const int q = (ii * 3 + j) * 3 + k;
              sn[q] = solid[node];
              un[q][0] = lbm_vel[node];
              un[q][1] = lbm_vel[NL + node];
              un[q][2] = lbm_vel[2 * NL + node];
              inmask |= (uint32_t)in << q;
            }
#pragma unroll
        for (int ii = 0; ii < NB; ++ii) {
          const int i = i0 + ii;
          const R wi = pick3(i, cw[0][0], cw[1][0], cw[2][0]);
#pragma unroll
          for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              const int q = (ii * 3 + j) * 3 + k;
              const bool ok = ((inmask >> q) & 1u) && sn[q] == 0;
              valid |= (uint32_t)ok << (i * 9 + j * 3 + k);
              const R wt = wi * cw[j][1] * cw[k][2];
              wsum += ok ? wt : R(0);
              u[0] += ok ? wt * un[q][0] : R(0);
              u[1] += ok ? wt * un[q][1] : R(0);
              u[2] += ok ? wt * un[q][2] : R(0);
            }
        }
      }
      if (writer) valid_weight[p] = wsum;
      if (wsum > cp.min_w) {
        partial = wsum < R(1) - R(1e-5);
        R dv[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) dv[a] = u[a] / wsum - vp[a];
        if (cp.use_rel) {
          const R nrm = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
          if (nrm > cp.rel_limit && nrm > R(0)) {
#pragma unroll
            for (int a = 0; a < 3; ++a) dv[a] = dv[a] / nrm * cp.rel_limit;
            clipped = true;
          }
        }
        const R gm = cp.gamma * mp;
#pragma unroll
        for (int a = 0; a < 3; ++a) pf[a] = gm * dv[a];
        // learned closure: its correction joins the penalty force before the
        // force-norm limit (and so also the fluid's reaction)
        if (closure) {
#pragma unroll
          for (int a = 0; a < 3; ++a) pf[a] += closure[a * cap + p];
        }
        if (cp.use_force) {
          const R nrm = sqrt(pf[0] * pf[0] + pf[1] * pf[1] + pf[2] * pf[2]);
          if (nrm > cp.force_limit && nrm > R(0)) {
#pragma unroll
            for (int a = 0; a < 3; ++a) pf[a] = pf[a] / nrm * cp.force_limit;
            clipped = true;
          }
        }
        if (writer) cmask[p] = 1;
        // equal and opposite reaction (and occupancy) spread on the valid nodes
        const R inv_wsum = R(1) / wsum;  // one division, 27 products below
        const R fscale = cp.dt_ratio / cp.cell_volume;
        const R phi = vp_vol / cp.cell_volume;
        const R rf[3] = {-pf[0] * fscale, -pf[1] * fscale, -pf[2] * fscale};
        if (use_lds) {
          // spread below, after the workgroup's scale
          spreads = true;
          const P iw = P(inv_wsum);
          rfn[0] = P(rf[0]) * iw;
          rfn[1] = P(rf[1]) * iw;
          rfn[2] = P(rf[2]) * iw;
          rfn[3] = P(phi) * iw;
        } else if (writer) {
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
              for (int k = 0; k < 3; ++k) {
                if (!((valid >> ((i * 3 + j) * 3 + k)) & 1u)) continue;
                const int ni = cbase[0] + i, nj = cbase[1] + j, nk = cbase[2] + k;
                const R nw = (cw[i][0] * cw[j][1] * cw[k][2]) * inv_wsum;
                const int64_t node = ((int64_t)ni * lny + nj) * lnz + nk;
                FSI_GLOBAL(node, 4, nw * rf[0]);
                FSI_GLOBAL(node, 5, nw * rf[1]);
                FSI_GLOBAL(node, 6, nw * rf[2]);
                if (cp.build) FSI_GLOBAL(node, 7, nw * phi);
              }
        }
      } else {
        unsupported = true;
        if (writer) cmask[p] = 0;
      }
    }
    if (use_lds) {  // uniform over the workgroup
      if constexpr (TA::fixed) {
        float bnd[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) bnd[k] = spreads && (k < 3 || cp.build) ? fabsf((float)rfn[k]) : 0.0f;
        cscale = tile_scale<PL * kS / 64>(bnd, sred[0]);
#pragma unroll
        for (int k = 0; k < 4; ++k) rfn[k] *= cscale.mul[k];
      }
      if (spreads) {
        // all 27 nodes in the lane-rotated order; the values of invalid
        // (solid or out-of-lattice) nodes are dropped at the flush, which
        // leaves the valid nodes' sums unchanged. The value order is
        // rotated by lane as well (instruction k adds value (k + r) % NV),
        // so lanes that meet on one node in the same instruction mostly
        // write different values, i.e. different words
        auto spread = [&](auto with_phi) {
          constexpr int NV = decltype(with_phi)::value ? 4 : 3;
          const int r = NV == 4 ? (lane & 3) : (lane >> 2) % 3;
          P val[NV];
          int off[NV];
#pragma unroll
          for (int k = 0; k < NV; ++k) {
            const int c = (k + r) % NV;
            val[k] = NV == 4 ? pick4(c, rfn[0], rfn[1], rfn[2], rfn[3]) : pick3(c, rfn[0], rfn[1], rfn[2]);
            off[k] = (4 + c) * kN;
          }
#pragma unroll
          for (int ii = 0; ii < kPlanes; ++ii)
#pragma unroll
            for (int jj = 0; jj < 3; ++jj) {
              const P wij = cwir[ii] * cwjr[jj];
              A* tr = tile + crow[jj] + cplane[ii];
#pragma unroll
              for (int kk = 0; kk < 3; ++kk) {
                const P wt = wij * cwkr[kk];
                A* tn = tr + ko[kk];
#pragma unroll
                for (int k = 0; k < NV; ++k) TA::add(tn + off[k], wt * val[k]);
              }
            }
        };
        if (cp.build)
          spread(std::true_type{});
        else
          spread(std::false_type{});
      }
    }
    if (writer) {
      wave_count(counters + 0, unsupported);
      wave_count(counters + 1, partial);
      wave_count(counters + 2, clipped);
    }
    if (live && writer) {
#pragma unroll
      for (int a = 0; a < 3; ++a) pforce[a * cap + p] = pf[a];
      if (cp.force_mean) {
#pragma unroll
        for (int a = 0; a < 3; ++a) cp.force_mean[a * cap + p] += pf[a] * cp.dt_ratio;
      }
    }
  } else {
    if (clear_force) {
      if (live && writer) {
#pragma unroll
        for (int a = 0; a < 3; ++a) pforce[a * cap + p] = R(0);
      }
    } else if (act) {
#pragma unroll
      for (int a = 0; a < 3; ++a) pf[a] = pforce[a * cap + p];
    }
  }

  // FP32 particles whose F needs the FP64 SVD are scattered after the flush,
  // where little else is live (keeps the kernel at four waves per SIMD)
  bool deferred = false;
  // plane split: every plane reads F before plane 0 stores the update, so
  // the store waits for the flush barrier
  R Fn[9];
  // P2G values of this lane for the tile: momentum at the particle and the
  // affine matrix (scattered after the workgroup's scale)
  bool scatters = false;
  R aff[9], mom0[3];
  if constexpr (P2G) {
    if (act) {
#pragma unroll
      for (int r = 0; r < 9; ++r) {
        Cm[r] = REGS ? rC[r] : C[r * cap + p];
        Fo[r] = Fd[r * cap + p];
      }
      deformation_update(m.dt, Cm, Fo, Fn);
      if constexpr (PL == 1) {
#pragma unroll
        for (int r = 0; r < 9; ++r) Fd[r * cap + p] = Fn[r];
      }
      R Rot[9];
      bool have_rot = true;
      if constexpr (sizeof(R) == 8)
        rotation_core<R>(Fn, Rot, nullptr);
      else
        have_rot = polar_newton(Fn, Rot);
      deferred = !have_rot;
      if (have_rot) {
        affine_momentum(m, p, Fn, Rot, Cm, vp, pf, mp, vp_vol, aff, mom0);
        if (use_lds)
          scatters = true;
        else if (writer)
          scatter_global(m, grid, base, fx, w, mom0, aff, mp);
      }
    }
    if (use_lds) {  // uniform over the workgroup
      if constexpr (TA::fixed) {
        // largest |mom0 + aff dp| per component over the stencil (|dp| <=
        // 1.5 dx per axis), and the mass
        float bnd[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (scatters) {
#pragma unroll
          for (int a = 0; a < 3; ++a)
            bnd[a] = fabsf((float)mom0[a]) + 1.5f * (float)m.dx *
                                                 (fabsf((float)aff[a * 3]) + fabsf((float)aff[a * 3 + 1]) +
                                                  fabsf((float)aff[a * 3 + 2]));
          bnd[3] = fabsf((float)mp);
        }
        pscale = tile_scale<PL * kS / 64>(bnd, sred[1]);
      }
      if (scatters) {
        // lane-rotated order as for the coupling scatter; the affine
        // momentum is built up separably: mom0 + aff (dp_i, dp_j, dp_k) per
        // i plane, (i, j) row and node. The four values (momentum xyz, mass)
        // are written in a lane-rotated order too: instruction k adds value
        // (k + lane) % 4, with the mass as a fourth component of zero affine
        // part
        const P dx = P(m.dx);
        P dpj[3], dpk[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          dpj[s] = (P(jo[s]) - P(fx[1])) * dx;
          dpk[s] = (P(ko[s]) - P(fx[2])) * dx;
        }
        P cm[4], cx[4], cy[4], cz[4];
        int off[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int c = (k + lane) & 3;
          cm[k] = P(pick4(c, mom0[0], mom0[1], mom0[2], mp));
          cx[k] = P(pick4(c, aff[0], aff[3], aff[6], R(0)));
          cy[k] = P(pick4(c, aff[1], aff[4], aff[7], R(0)));
          cz[k] = P(pick4(c, aff[2], aff[5], aff[8], R(0)));
          if constexpr (TA::fixed) {
            const P mul = P(pick4(c, pscale.mul[0], pscale.mul[1], pscale.mul[2], pscale.mul[3]));
            cm[k] *= mul;
            cx[k] *= mul;
            cy[k] *= mul;
            cz[k] *= mul;
          }
          off[k] = c * kN;
        }
#pragma unroll
        for (int ii = 0; ii < kPlanes; ++ii) {
          const P dpi = (P(io[ii]) - P(fx[0])) * dx;
          P mi[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) mi[k] = cm[k] + cx[k] * dpi;
#pragma unroll
          for (int jj = 0; jj < 3; ++jj) {
            const P wij = wir[ii] * wjr[jj];
            P mij[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) mij[k] = mi[k] + cy[k] * dpj[jj];
            A* tr = tile + row[jj] + plane[ii];
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
              const P wt = wij * wkr[kk];
              A* tn = tr + ko[kk];
#pragma unroll
              for (int k = 0; k < 4; ++k) TA::add(tn + off[k], wt * (mij[k] + cz[k] * dpk[kk]));
            }
          }
        }
      }
    }
  }

#undef FSI_GLOBAL
  // flush the tile: one global atomic per non-zero entry of an in-lattice
  // node; coupling entries only on fluid nodes
  if (use_lds) {
    __syncthreads();
    if constexpr (REF && COUPLE) {
      // the coupling values, on the lattice's box
      for (int t = threadIdx.x; t < ctn_all; t += blockDim.x) {
        const int ni = cx0 + t / ctyz, nj = cy0 + (t / ctzi) % ctyi, nk = cz0 + t % ctzi;
        if (!inside(lnx, lny, lnz, ni, nj, nk) || ctile[t][3] == R(0)) continue;
        const int64_t node = ((int64_t)ni * lny + nj) * lnz + nk;
#pragma unroll
        for (int s = 4; s < Tile::kValues; ++s) {
          const A sum = tile[s * kN + t];
          if (sum != A(0)) atomicAdd(cforce + 4 * node + (s - 4), R(TA::value(sum, cscale.inv[s - 4])));
        }
      }
    }
    for (int t = threadIdx.x; t < tn_all; t += blockDim.x) {
      const int ni = x0 + t / tyz, nj = y0 + (t / tzi) % tyi, nk = z0 + t % tzi;
      if (!inside(m.nx, m.ny, m.nz, ni, nj, nk)) continue;
      const int64_t node = ((int64_t)ni * m.ny + nj) * m.nz + nk;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const A sum = tile[s * kN + t];
        if (sum != A(0)) atomicAdd(grid + 4 * node + s, R(TA::value(sum, pscale.inv[s])));
      }
      if constexpr (COUPLE && !REF) {
        if (ctile[t][3] != R(0)) {
#pragma unroll
          for (int s = 4; s < Tile::kValues; ++s) {
            const A sum = tile[s * kN + t];
            if (sum != A(0)) atomicAdd(cforce + 4 * node + (s - 4), R(TA::value(sum, cscale.inv[s - 4])));
          }
        }
      }
    }
  }
  if constexpr (P2G && PL > 1) {
    if (!use_lds) __syncthreads();  // (the tiled path synchronised at the flush)
    if (act && writer) {
#pragma unroll
      for (int r = 0; r < 9; ++r) Fd[r * cap + p] = Fn[r];
    }
  }
  if constexpr (P2G && sizeof(R) == 4) {
    // inverted or ill-conditioned F: reload the particle (F is the updated
    // one stored above, pforce holds this substep's coupling force) and use
    // the FP64 Jacobi rotation; global atomics, as on the untiled path
    if (deferred && writer) {
      R xr[3], vr[3], fr[3], wr[3][3], Cr[9], Fr[9], pr[3];
      int br[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        xr[a] = x[a * cap + p];
        vr[a] = v[a * cap + p];
        pr[a] = pforce[a * cap + p];
      }
#pragma unroll
      for (int r = 0; r < 9; ++r) {
        Cr[r] = C[r * cap + p];
        Fr[r] = Fd[r * cap + p];
      }
      const R mr = mass[p], volr = vol[p];
      locate(xr, m.inv_dx, br, fr);
      bspline(fr, wr);
      double Fdd[9], Rdd[9];
#pragma unroll
      for (int r = 0; r < 9; ++r) Fdd[r] = double(Fr[r]);
      rotation_core<double>(Fdd, Rdd, nullptr);
      R Rot[9], aff[9], mom0[3];
#pragma unroll
      for (int r = 0; r < 9; ++r) Rot[r] = R(Rdd[r]);
      affine_momentum(m, p, Fr, Rot, Cr, vr, pr, mr, volr, aff, mom0);
      scatter_global(m, grid, br, fr, wr, mom0, aff, mr);
    }
  }
}

template <typename R, bool COUPLE, bool P2G, bool WIDE, int PL, bool REF = false>
__global__ __launch_bounds__(PL * Slots<PL>::value) void k_p2g(MpmK<R> m, CplK<R> cp, R* __restrict__ x,
                                                R* __restrict__ v, const R* __restrict__ C,
                                                R* __restrict__ Fd, const R* __restrict__ mass,
                                                const R* __restrict__ vol,
                                                const int32_t* __restrict__ active,
                                                R* __restrict__ pforce, R* __restrict__ grid,
                                                const int32_t* __restrict__ solid,
                                                const R* __restrict__ lbm_vel, R* __restrict__ cforce,
                                                R* __restrict__ valid_weight,
                                                int32_t* __restrict__ cmask,
                                                int32_t* __restrict__ counters, int clear_force,
                                                int32_t* __restrict__ zero_ctr,
                                                const int32_t* __restrict__ chunks,
                                                int32_t* __restrict__ boxes_out,
                                                int32_t* __restrict__ big_flag, int big_value,
                                                const R* __restrict__ closure) {
  // batched substeps: the G2P-pass counters of this substep start at zero
  if (zero_ctr && blockIdx.x == 0 && threadIdx.x < 2) zero_ctr[threadIdx.x] = 0;
  // particle of this lane: from the work list (one chunk of at most 256
  // particles of two z-adjacent 4x4x4 cell blocks per workgroup), else slot
  // order
  // PL wavefront groups share the particles of the workgroup (plane split)
  constexpr int kS = Slots<PL>::value, kSub = Slots<PL>::per_chunk;
  const int slot = (int)threadIdx.x % kS;
  int p;
  bool live;
  if (chunks) {
    const int ch = blockIdx.x / kSub, sub = (blockIdx.x % kSub) * kS + slot;
    p = chunks[2 * ch] + sub;
    live = sub < chunks[2 * ch + 1];
  } else {
    p = blockIdx.x * kS + slot;
    live = p < m.num;
  }
  p2g_pass<R, COUPLE, P2G, WIDE, false, PL, REF>(m, cp, p, live, x, v, C, Fd, mass, vol, active, pforce, grid, solid, lbm_vel,
                                        cforce, valid_weight, cmask, counters, clear_force, boxes_out, big_flag,
                                        big_value, closure, nullptr, nullptr, nullptr);
}

// ------------------------------------------------- grid + G2P pass ----

// fsi/mpm3d.py:268-288 (grid_op) for one node, evaluated where G2P reads it:
// v = p/m + dt g on nodes with mass, then the sticky/separating walls.
template <typename R>
__device__ __forceinline__ void grid_velocity(const MpmK<R>& m, const R g[4], int i, int j, int k,
                                              R out[3]) {
  // branch-free: nodes without mass keep their (zero) momentum
  const R mass = g[3];
  const bool has = mass > R(0);
  const R inv_mass = R(1) / mass;  // one division per node (massless nodes: unused)
  const int idx[3] = {i, j, k};
  const int n[3] = {m.nx, m.ny, m.nz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    R va = g[a] * inv_mass + m.dt * m.g[a];
    const bool flip = (idx[a] < m.bw && va < R(0)) || (idx[a] >= n[a] - m.bw && va > R(0));
    va = flip ? -m.damping * va : va;
    out[a] = has ? va : g[a];
  }
}

// fsi/coupling.py:463-496 for one particle at position xp (REF: on the
// lattice's cells of a refined MPM grid)
template <typename R, bool REF = false>
__device__ __forceinline__ bool contact_test(const MpmK<R>& m, const CplK<R>& cp, const R xp[3],
                                             const int32_t* __restrict__ solid,
                                             const R* __restrict__ cforce) {
  int base[3];
  R fx[3], w[3][3];
  locate(xp, REF ? cp.inv_dx_c : m.inv_dx, base, fx);
  bspline(fx, w);
  const int lnx = REF ? cp.lnx : m.nx, lny = REF ? cp.lny : m.ny, lnz = REF ? cp.lnz : m.nz;
  R dyn = R(0), stat = R(0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        bool in;
        const int64_t node = gather_node(lnx, lny, lnz, base[0] + i, base[1] + j, base[2] + k, in);
        const R phi = cforce[4 * node + 3];
        const int32_t sflag = solid[node];
        const R wt = w[i][0] * w[j][1] * w[k][2];
        dyn += in ? wt * phi : R(0);
        stat += (in && sflag != 0) ? wt : R(0);
      }
  return dyn >= cp.contact_thr || stat >= cp.contact_thr;
}

// fsi/mpm3d.py:268-328: grid update (per gathered node) + g2p + clamp, and
// the contact pass of fsi/coupling.py:463-496 when CONTACT

// Batched substeps: zero the grid the next P2G pass fills (read by nobody
// in this pass): only the node boxes the P2G pass that last wrote it
// recorded, unless one of them was too large or there are none; then the
// whole grid (16-byte words), spread over the launch.
template <typename R>
__device__ __forceinline__ void zero_grid_pass(const MpmK<R>& m, uint4* __restrict__ zero_grid, int64_t zero_n16,
                                               const int32_t* __restrict__ boxes_in,
                                               const int32_t* __restrict__ big_flag, int big_expect,
                                               int num_boxes) {
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  if (zero_grid && boxes_in && *big_flag != big_expect) {
    R* zg = reinterpret_cast<R*>(zero_grid);
    for (int c = blockIdx.x; c < num_boxes; c += gridDim.x) {
      const int32_t* bx = boxes_in + 8 * c;
      const int x0 = bx[0], y0 = bx[1], z0 = bx[2], ty = bx[4], tz = bx[5];
      const int n = bx[3] * ty * tz;
      for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int ni = x0 + t / (ty * tz), nj = y0 + (t / tz) % ty, nk = z0 + t % tz;
        if (!inside(m.nx, m.ny, m.nz, ni, nj, nk)) continue;
        R* gp = zg + 4 * (((int64_t)ni * m.ny + nj) * m.nz + nk);
#pragma unroll
        for (int a = 0; a < 4; ++a) gp[a] = R(0);
      }
    }
  } else {
    for (int64_t t = gt; t < zero_n16; t += (int64_t)gridDim.x * blockDim.x) zero_grid[t] = make_uint4(0, 0, 0, 0);
  }
}

// Grid update + G2P pass of one workgroup (fsi/mpm3d.py:268-328, + the
// contact pass of fsi/coupling.py:463-496 when CONTACT) for lane particle p.
// Stores x, v and C and returns them in xo, vo, Co (zeros for lanes without
// an active particle) for a P2G pass in the same launch. PL = 3: the
// workgroup has three wavefronts per particle slot; the one of plane
// `plane` gathers the 9 nodes of x offset `plane`, the partial sums of v and
// C are added in plane order through LDS, every plane continues with the
// same totals and plane 0 alone writes the particle.
template <typename R, bool CONTACT, bool WIDE, int PL, int LDS = g2p_lds_bytes<R>(), bool REF = false>
__device__ __forceinline__ void g2p_pass(const MpmK<R>& m, const CplK<R>& cp, int p, bool live, R* __restrict__ x,
                                         R* __restrict__ v, R* __restrict__ C, const int32_t* __restrict__ active,
                                         const R* __restrict__ grid, const int32_t* __restrict__ solid,
                                         const R* __restrict__ cforce, int32_t* __restrict__ contact_mask,
                                         int32_t* __restrict__ counters, R xo[3], R vo[3], R Co[9]) {
  static_assert(PL == 1 || (PL == 3 && sizeof(R) == 4), "plane split: FP32 only");
  static_assert(LDS >= g2p_lds_bytes<R>(), "G2P stage does not fit its LDS buffer");
  R(*vtile)[3] = reinterpret_cast<R(*)[3]>(pass_lds<LDS>());
  constexpr int kS = Slots<PL>::value;
  __shared__ int red[6][PL * kS / 64];
  __shared__ R gpart[PL > 1 ? PL : 1][PL > 1 ? kS : 1][12];
  const int plane = PL > 1 ? (int)threadIdx.x / kS : 0;
  const int slot = PL > 1 ? (int)threadIdx.x % kS : (int)threadIdx.x;
  const bool writer = plane == 0;
  const bool act = live && active[p] != 0;
  const int cap = m.cap;
  bool cand = false, damped = false;
  R xp[3] = {R(0), R(0), R(0)}, fx[3] = {R(0), R(0), R(0)}, w[3][3];
  int base[3] = {0, 0, 0};
  if (act) {
#pragma unroll
    for (int a = 0; a < 3; ++a) xp[a] = x[a * cap + p];
    locate(xp, m.inv_dx, base, fx);
  }
  bspline(fx, w);

  // stage the grid velocity of the workgroup's node box in LDS: the grid
  // update (fsi/mpm3d.py:268-288) runs once per node instead of once per
  // particle and node, and the 27-node gathers read LDS
  const NodeBox box = workgroup_box<PL * kS / 64>(act, base, red);
  const bool staged = box.nodes > 0 && box.nodes <= kG2PNodes;
  if (staged) {
    for (int t = threadIdx.x; t < (int)box.nodes; t += blockDim.x) {
      int ni, nj, nk;
      box.coords(t, ni, nj, nk);
      bool in;
      const int64_t node = gather_node(m.nx, m.ny, m.nz, ni, nj, nk, in);
      const R* gp = grid + 4 * node;
      const R g4[4] = {gp[0], gp[1], gp[2], gp[3]};
      R gv[3];
      grid_velocity(m, g4, ni, nj, nk, gv);
      // out-of-lattice nodes contribute exact zeros
#pragma unroll
      for (int a = 0; a < 3; ++a) vtile[t][a] = in ? gv[a] : R(0);
    }
    __syncthreads();
  }

  R nv[3] = {R(0), R(0), R(0)};
  R nc[9] = {R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0)};
  if (act) {
    // one node's contribution, in the reference's i, j, k order
    auto accumulate = [&](int i, R wi, int j, int k, const R gv[3]) {
      const R dp[3] = {R(i) - fx[0], R(j) - fx[1], R(k) - fx[2]};
      const R wt = wi * w[j][1] * w[k][2];
#pragma unroll
      for (int a = 0; a < 3; ++a) nv[a] += wt * gv[a];
      const R s = m.four_inv_dx * wt;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) nc[r * 3 + c] += s * (gv[r] * dp[c]);
    };
    if (staged) {
      // the 27-node sums in separable form: per (i, j) row of the stencil
      // the k-sums A = sum_k w_k v and B = sum_k w_k (k - fx_z) v, then
      // v += w_i w_j A and T = sum w v dp^T by columns: w_i w_j (i - fx_x) A,
      // w_i w_j (j - fx_y) A, w_i w_j B; C = 4 inv_dx T. A third of the
      // per-node products of the direct form.
      R wkz[3], T[9];
#pragma unroll
      for (int k = 0; k < 3; ++k) wkz[k] = w[k][2] * (R(k) - fx[2]);
#pragma unroll
      for (int r = 0; r < 9; ++r) T[r] = R(0);
      auto row = [&](int i, R wi, int j) {
        const R(*tv)[3] = vtile + box.index(base[0] + i, base[1] + j, base[2]);
        R A[3] = {R(0), R(0), R(0)}, B[3] = {R(0), R(0), R(0)};
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            const R g = tv[k][a];
            A[a] += w[k][2] * g;
            B[a] += wkz[k] * g;
          }
        const R wij = wi * w[j][1];
        const R ax = wij * (R(i) - fx[0]), ay = wij * (R(j) - fx[1]);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          nv[a] += wij * A[a];
          T[a * 3 + 0] += ax * A[a];
          T[a * 3 + 1] += ay * A[a];
          T[a * 3 + 2] += wij * B[a];
        }
      };
      if constexpr (PL > 1) {
        const R wi = pick3(plane, w[0][0], w[1][0], w[2][0]);
#pragma unroll
        for (int j = 0; j < 3; ++j) row(plane, wi, j);
      } else {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) row(i, w[i][0], j);
      }
#pragma unroll
      for (int r = 0; r < 9; ++r) nc[r] = m.four_inv_dx * T[r];
    } else {
      // box too large for LDS (particles far apart): gather from the grid in
      // batches of NB i-planes, all loads of a batch in flight together (one
      // plane per wavefront with the plane split)
      constexpr int NB = PL > 1 ? 1 : (WIDE ? 3 : 1);
#pragma unroll 1
      for (int i0 = PL > 1 ? plane : 0; i0 < (PL > 1 ? plane + 1 : 3); i0 += NB) {
        R gn[NB * 9][4];
        uint32_t inmask = 0;
#pragma unroll
        for (int ii = 0; ii < NB; ++ii)
#pragma unroll
          for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              bool in;
              const int64_t node = gather_node(m.nx, m.ny, m.nz, base[0] + i0 + ii, base[1] + j, base[2] + k, in);
              const R* gp = grid + 4 * node;
              const int q = (ii * 3 + j) * 3 + k;
#pragma unroll
              for (int a = 0; a < 4; ++a) gn[q][a] = gp[a];
              inmask |= (uint32_t)in << q;
            }
#pragma unroll
        for (int ii = 0; ii < NB; ++ii)
#pragma unroll
          for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              const int i = i0 + ii;
              const int q = (ii * 3 + j) * 3 + k;
              R gv[3];
              grid_velocity(m, gn[q], base[0] + i, base[1] + j, base[2] + k, gv);
              const bool in = (inmask >> q) & 1u;
#pragma unroll
              for (int a = 0; a < 3; ++a) gv[a] = in ? gv[a] : R(0);
              // w[i][0] with i dynamic when NB == 1
              accumulate(i, pick3(i, w[0][0], w[1][0], w[2][0]), j, k, gv);
            }
      }
    }
  }
  if constexpr (PL > 1) {
    // the three planes' partial sums, added in plane order
#pragma unroll
    for (int a = 0; a < 3; ++a) gpart[plane][slot][a] = nv[a];
#pragma unroll
    for (int r = 0; r < 9; ++r) gpart[plane][slot][3 + r] = nc[r];
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 3; ++a) nv[a] = (gpart[0][slot][a] + gpart[1][slot][a]) + gpart[2][slot][a];
#pragma unroll
    for (int r = 0; r < 9; ++r)
      nc[r] = (gpart[0][slot][3 + r] + gpart[1][slot][3 + r]) + gpart[2][slot][3 + r];
  }
  if (act) {
#pragma unroll
    for (int r = 0; r < 9; ++r)
      if (writer) C[r * cap + p] = m.apic ? nc[r] : R(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const R xn = xp[a] + m.dt * nv[a];
      xp[a] = fmin(fmax(xn, m.lower), m.upper[a]);
      if (writer) x[a * cap + p] = xp[a];
    }
    if constexpr (CONTACT) {
      cand = contact_test<R, REF>(m, cp, xp, solid, cforce);
      if (cand && cp.contact_damp > R(0)) {
        damped = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) nv[a] *= R(1) - cp.contact_damp;
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (writer) v[a * cap + p] = nv[a];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      xo[a] = xp[a];
      vo[a] = nv[a];
    }
#pragma unroll
    for (int r = 0; r < 9; ++r) Co[r] = m.apic ? nc[r] : R(0);
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a) xo[a] = vo[a] = R(0);
#pragma unroll
    for (int r = 0; r < 9; ++r) Co[r] = R(0);
  }
  if constexpr (CONTACT) {
    if (writer) {
      if (live) contact_mask[p] = cand ? 1 : 0;
      wave_count(counters + 3, cand);
      wave_count(counters + 4, damped);
    }
  }
}

template <typename R, bool CONTACT, bool WIDE, int PL, bool REF = false>
__global__ __launch_bounds__(PL * Slots<PL>::value) void k_g2p(MpmK<R> m, CplK<R> cp, R* __restrict__ x,
                                                R* __restrict__ v, R* __restrict__ C,
                                                const int32_t* __restrict__ active,
                                                const R* __restrict__ grid,
                                                const int32_t* __restrict__ solid,
                                                const R* __restrict__ cforce,
                                                int32_t* __restrict__ contact_mask,
                                                int32_t* __restrict__ counters,
                                                uint4* __restrict__ zero_grid, int64_t zero_n16,
                                                int32_t* __restrict__ zero_ctr,
                                                const int32_t* __restrict__ chunks, int num_chunks,
                                                const int32_t* __restrict__ boxes_in,
                                                const int32_t* __restrict__ big_flag, int big_expect,
                                                int num_boxes) {
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  zero_grid_pass<R>(m, zero_grid, zero_n16, boxes_in, big_flag, big_expect, num_boxes);
  if (zero_ctr && gt < 3) zero_ctr[gt] = 0;
  // particle of this lane: the work list of the particle sort (workgroups
  // beyond it only zero), else slot order
  constexpr int kS = Slots<PL>::value, kSub = Slots<PL>::per_chunk;
  const int slot = (int)threadIdx.x % kS;
  int p;
  bool live;
  if (chunks) {
    const int ch = blockIdx.x / kSub, sub = (blockIdx.x % kSub) * kS + slot;
    const bool has = ch < num_chunks;
    p = has ? chunks[2 * ch] + sub : 0;
    live = has && sub < chunks[2 * ch + 1];
  } else {
    p = blockIdx.x * kS + slot;
    live = p < m.num;
  }
  R xo[3], vo[3], Co[9];
  g2p_pass<R, CONTACT, WIDE, PL, g2p_lds_bytes<R>(), REF>(m, cp, p, live, x, v, C, active, grid, solid, cforce, contact_mask, counters, xo, vo,
                                 Co);
}

// Substeps s-1 and s back to back in one launch: the grid update + G2P of
// substep s-1 (from grid_src), then, in the same workgroup and with the
// particle's new x, v and C still in registers, the P2G of substep s (into
// grid_dst, zeroed before the launch). The work list is the same for both
// passes; the zeroing part prepares the grid of substep s+1. Workgroups
// beyond the particle blocks only zero. The FP32 one-wavefront kernel is
// held to 128 VGPRs (4 wavefronts per SIMD, with the shared LDS buffer four
// workgroups per CU): measured 4.7% faster per coupled step on a 448,861
// particle suspension than 146 VGPRs at 3, despite a 14-byte spill; the wide
// variants would spill ~190 bytes and keep their registers.
template <typename R, bool COUPLE, bool CONTACT, bool WIDE, int PL, bool REF = false>
__global__ __launch_bounds__(PL * Slots<PL>::value)
__attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 && PL == 1 && !WIDE ? 4 : 1))) void k_g2p_p2g(
    MpmK<R> m, CplK<R> cp, R* __restrict__ x, R* __restrict__ v, R* __restrict__ C, R* __restrict__ Fd,
    const R* __restrict__ mass, const R* __restrict__ vol, const int32_t* __restrict__ active,
    R* __restrict__ pforce, const R* __restrict__ grid_src, R* __restrict__ grid_dst,
    const int32_t* __restrict__ solid, const R* __restrict__ lbm_vel, R* __restrict__ cforce,
    R* __restrict__ valid_weight, int32_t* __restrict__ cmask, int32_t* __restrict__ contact_mask,
    int32_t* __restrict__ g2p_counters, int32_t* __restrict__ p2g_counters, int clear_force,
    const int32_t* __restrict__ chunks, int particle_blocks, uint4* __restrict__ zero_grid, int64_t zero_n16,
    const int32_t* __restrict__ boxes_in, const int32_t* __restrict__ big_in, int big_expect, int num_boxes,
    int32_t* __restrict__ boxes_out, int32_t* __restrict__ big_out, int big_value, const R* __restrict__ closure) {
  zero_grid_pass<R>(m, zero_grid, zero_n16, boxes_in, big_in, big_expect, num_boxes);
  const bool has = (int)blockIdx.x < particle_blocks;
  constexpr int kS = Slots<PL>::value, kSub = Slots<PL>::per_chunk;
  const int slot = (int)threadIdx.x % kS;
  int p;
  bool live;
  if (chunks) {
    const int ch = blockIdx.x / kSub, sub = (blockIdx.x % kSub) * kS + slot;
    p = has ? chunks[2 * ch] + sub : 0;
    live = has && sub < chunks[2 * ch + 1];
  } else {
    p = blockIdx.x * kS + slot;
    live = p < m.num;
  }
  R xo[3], vo[3], Co[9];
  g2p_pass<R, CONTACT, WIDE, PL, pass_lds_bytes<R, COUPLE>(), REF>(m, cp, p, live, x, v, C, active, grid_src, solid, cforce,
                                                          contact_mask, g2p_counters, xo, vo, Co);
  // the P2G stage reuses the G2P stage's LDS buffer: every read of the
  // staged grid velocities completes before any thread writes its tile
  __syncthreads();
  if (has)  // uniform over the workgroup
    p2g_pass<R, COUPLE, true, WIDE, true, PL, REF>(m, cp, p, live, x, v, C, Fd, mass, vol, active, pforce, grid_dst, solid,
                                          lbm_vel, cforce, valid_weight, cmask, p2g_counters, clear_force, boxes_out,
                                          big_out, big_value, closure, xo, vo, Co);
}

__global__ void k_fill_i32(int n, int32_t* __restrict__ a, int32_t value) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = value;
}

__global__ void k_copy_i32(int n, const int32_t* __restrict__ src, int32_t* __restrict__ dst) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[t];
}

// Learned-closure input features of one particle (no counterpart in the
// reference): the penalty coupling's support and interpolated lattice
// velocity (fsi/coupling.py:352-372, valid = in lattice and fluid), the
// velocity gradient over the same nodes, the interpolated occupancy phi of
// the last coupled step and its normalised negative gradient (an outward
// surface normal of the solid), all in lattice units.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_closure_features(MpmK<R> m, const R* __restrict__ x,
                                                             const R* __restrict__ v,
                                                             const int32_t* __restrict__ active,
                                                             const int32_t* __restrict__ solid,
                                                             const R* __restrict__ lbm_vel,
                                                             const R* __restrict__ cforce,
                                                             float* __restrict__ X) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m.num) return;
  float* out = X + (int64_t)p * FSI_MLP_IN;
  if (active[p] == 0) {
    for (int k = 0; k < FSI_MLP_IN; ++k) out[k] = 0.0f;
    return;
  }
  const int cap = m.cap;
  const int64_t N = (int64_t)m.nx * m.ny * m.nz;
  R xp[3], vp[3], fx[3], w[3][3], dw[3][3];
  int base[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    xp[a] = x[a * cap + p];
    vp[a] = v[a * cap + p];
  }
  locate(xp, m.inv_dx, base, fx);
  bspline(fx, w);
  // d w / d x of the quadratic B-spline (derivative of fsi/mpm3d.py:227-232)
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    dw[0][a] = (fx[a] - R(1.5)) * m.inv_dx;
    dw[1][a] = R(-2) * (fx[a] - R(1)) * m.inv_dx;
    dw[2][a] = (fx[a] - R(0.5)) * m.inv_dx;
  }
  R W = R(0), u[3] = {R(0), R(0), R(0)}, G[9] = {}, phi = R(0), gphi[3] = {R(0), R(0), R(0)};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        bool in;
        const int64_t node = gather_node(m.nx, m.ny, m.nz, base[0] + i, base[1] + j, base[2] + k, in);
        if (!in) continue;
        const R wt = w[i][0] * w[j][1] * w[k][2];
        const R gw[3] = {dw[i][0] * w[j][1] * w[k][2], w[i][0] * dw[j][1] * w[k][2], w[i][0] * w[j][1] * dw[k][2]};
        const R ph = cforce[4 * node + 3];
        phi += wt * ph;
#pragma unroll
        for (int b = 0; b < 3; ++b) gphi[b] += gw[b] * ph;
        if (solid[node] != 0) continue;
        const R un[3] = {lbm_vel[node], lbm_vel[N + node], lbm_vel[2 * N + node]};
        W += wt;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          u[a] += wt * un[a];
#pragma unroll
          for (int b = 0; b < 3; ++b) G[a * 3 + b] += gw[b] * un[a];
        }
      }
  R rel[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) rel[a] = (W > R(0) ? u[a] / W : R(0)) - vp[a];
  const R gn = sqrt(gphi[0] * gphi[0] + gphi[1] * gphi[1] + gphi[2] * gphi[2]);
#pragma unroll
  for (int a = 0; a < 3; ++a) out[a] = (float)rel[a];
#pragma unroll
  for (int a = 0; a < 9; ++a) out[3 + a] = (float)G[a];
  out[12] = (float)phi;
#pragma unroll
  for (int a = 0; a < 3; ++a) out[13 + a] = gn > R(1e-6) ? (float)(-gphi[a] / gn) : 0.0f;
  out[16] = (float)W;
  out[17] = (float)sqrt(rel[0] * rel[0] + rel[1] * rel[1] + rel[2] * rel[2]);
}

template <typename R, bool REF = false>
__global__ __launch_bounds__(kBlock) void k_contact(MpmK<R> m, CplK<R> cp, const R* __restrict__ x,
                                                    R* __restrict__ v,
                                                    const int32_t* __restrict__ active,
                                                    const int32_t* __restrict__ solid,
                                                    const R* __restrict__ cforce,
                                                    int32_t* __restrict__ contact_mask,
                                                    int32_t* __restrict__ counters) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = p < m.num;
  const bool act = live && active[p] != 0;
  bool cand = false, damped = false;
  if (act) {
    const R xp[3] = {x[p], x[m.cap + p], x[2 * m.cap + p]};
    cand = contact_test<R, REF>(m, cp, xp, solid, cforce);
    if (cand && cp.contact_damp > R(0)) {
      damped = true;
#pragma unroll
      for (int a = 0; a < 3; ++a) v[a * m.cap + p] *= R(1) - cp.contact_damp;
    }
  }
  if (live) contact_mask[p] = cand ? 1 : 0;
  wave_count(counters + 3, cand);
  wave_count(counters + 4, damped);
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_polar(int n, const R* __restrict__ F, R* __restrict__ Rot) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  R f[9], r[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) f[i] = F[t * 9 + i];
  rotation_factor<R>(f, r, det3(f));
#pragma unroll
  for (int i = 0; i < 9; ++i) Rot[t * 9 + i] = r[i];
}

__global__ void k_zero_counters(int32_t* __restrict__ c, int first, int count) {
  const int t = threadIdx.x;
  if (t < count) c[first + t] = 0;
}

int check_mpm(const fsi_mpm_params* p, const fsi_mpm_buffers* b) {
  FSI_REQUIRE(p && b, "mpm: null params or buffers");
  FSI_REQUIRE(p->nx > 0 && p->ny > 0 && p->nz > 0, "mpm: bad grid dims");
  FSI_REQUIRE(p->num_particles >= 0 && p->num_particles <= p->capacity,
              "mpm: %d particles exceed capacity %d", p->num_particles, p->capacity);
  FSI_REQUIRE(p->real == FSI_REAL_F32 || p->real == FSI_REAL_F64, "mpm: unknown real type");
  return 0;
}

// An MPM grid finer than the lattice (coupling refine > 1) takes the REF
// variants of the particle passes.
inline bool refined(const fsi_coupling_params* c) { return c && c->refine > 1; }

// Plane split of the particle passes (three wavefronts per particle, one per
// x plane of the stencil, of the lattice stencil too on a refined grid):
// FP32, a work list of at most split_max_chunks chunks (one-batch gathers).
template <typename R>
bool plane_split(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_params*) {
  return sizeof(R) == 4 && b->chunks && p->num_chunks > 0 && p->num_chunks <= p->split_max_chunks &&
         (!b->zero_boxes || (int64_t)Slots<3>::per_chunk * p->num_chunks <= (int64_t)b->zero_boxes_cap);
}

// Particle workgroups of one P2G or G2P pass (one node box each).
template <typename R>
int64_t particle_workgroups(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_params* c) {
  if (p->num_particles == 0) return 0;
  if (b->chunks && p->num_chunks > 0)
    return (int64_t)p->num_chunks * (plane_split<R>(p, b, c) ? Slots<3>::per_chunk : 1);
  return blocks_for(p->num_particles);
}

template <typename R>
int p2g_t(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_params* c,
          const fsi_coupling_buffers* cb, int clear_force, int do_p2g, bool zero_counters,
          R* grid_target, int32_t* zero_ctr, hipStream_t st, int32_t* boxes_out = nullptr,
          int32_t* big_flag = nullptr, int big_value = 0, int32_t* counters_override = nullptr) {
  MpmK<R> m = make_mpm<R>(p, b);
  int32_t* ctrs = counters_override ? counters_override : (cb ? cb->counters : nullptr);
  CplK<R> cp = make_cpl<R>(c, p);
  cp.force_mean = cb ? (R*)cb->force_mean : nullptr;
  const bool couple = c && c->enabled;
  if (couple && zero_counters) {
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, cb->counters, 0, 5);
    FSI_TRY_LAUNCH("zero_counters");
  }
  if (m.num == 0) return 0;
  // with a work list from the particle sort, one workgroup per chunk
  const int32_t* chunks = (b->chunks && p->num_chunks > 0) ? b->chunks : nullptr;
  const dim3 grid((unsigned)particle_workgroups<R>(p, b, c)), block(kBlock);
#define FSI_P2G_ARGS                                                                            \
  m, cp, (R*)b->x, (R*)b->v, (const R*)b->C, (R*)b->F, (const R*)b->mass, (const R*)b->volume, \
      b->active, (R*)b->pforce, grid_target, cb ? cb->solid : nullptr,                          \
      cb ? (const R*)cb->lbm_vel : nullptr, cb ? (R*)cb->cforce : nullptr,                      \
      cb ? (R*)cb->valid_weight : nullptr, cb ? cb->coupling_mask : nullptr,                    \
      ctrs, clear_force, zero_ctr, chunks, boxes_out, big_flag, big_value,                          \
      cb ? (const R*)cb->closure_force : nullptr
  // a grid that fits on the chip at once (2 workgroups per CU with the wide
  // gathers' registers) takes the one-batch gathers: its time is one
  // workgroup's latency
  const bool wide = (int64_t)grid.x <= (int64_t)p->wide_max_workgroups;
  const bool split = plane_split<R>(p, b, c);
  const bool ref = refined(c);
#define FSI_P2G_LAUNCH(C, P)                                                                          \
  do {                                                                                                \
    if constexpr (sizeof(R) == 4) {                                                                   \
      if (split) {                                                                                    \
        if (ref && C)                                                                                 \
          hipLaunchKernelGGL((k_p2g<R, C, P, false, 3, true>), grid, dim3(3 * Slots<3>::value), 0, st, \
                             FSI_P2G_ARGS);                                                           \
        else                                                                                          \
          hipLaunchKernelGGL((k_p2g<R, C, P, false, 3>), grid, dim3(3 * Slots<3>::value), 0, st, FSI_P2G_ARGS); \
        break;                                                                                        \
      }                                                                                               \
    }                                                                                                 \
    if (ref && C) {                                                                                   \
      if (wide)                                                                                       \
        hipLaunchKernelGGL((k_p2g<R, C, P, true, 1, true>), grid, block, 0, st, FSI_P2G_ARGS);        \
      else                                                                                            \
        hipLaunchKernelGGL((k_p2g<R, C, P, false, 1, true>), grid, block, 0, st, FSI_P2G_ARGS);       \
    } else if (wide)                                                                                  \
      hipLaunchKernelGGL((k_p2g<R, C, P, true, 1>), grid, block, 0, st, FSI_P2G_ARGS);                \
    else                                                                                              \
      hipLaunchKernelGGL((k_p2g<R, C, P, false, 1>), grid, block, 0, st, FSI_P2G_ARGS);               \
  } while (0)
  if (couple && do_p2g)
    FSI_P2G_LAUNCH(true, true);
  else if (couple)
    FSI_P2G_LAUNCH(true, false);
  else if (do_p2g)
    FSI_P2G_LAUNCH(false, true);
  else
    FSI_P2G_LAUNCH(false, false);
#undef FSI_P2G_LAUNCH
#undef FSI_P2G_ARGS
  FSI_TRY_LAUNCH("mpm_p2g");
  return 0;
}

template <typename R>
int grid_g2p_t(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_params* c,
               const fsi_coupling_buffers* cb, bool zero_counters, const R* grid_src,
               void* zero_grid, int32_t* zero_ctr, hipStream_t st, const int32_t* boxes_in = nullptr,
               const int32_t* big_flag = nullptr, int big_expect = 0, int num_boxes = 0,
               int32_t* counters_override = nullptr) {
  MpmK<R> m = make_mpm<R>(p, b);
  int32_t* ctrs = counters_override ? counters_override : (cb ? cb->counters : nullptr);
  CplK<R> cp = make_cpl<R>(c, p);
  const bool contact = c && c->contact_enabled;
  if (contact && zero_counters) {
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, cb->counters, 3, 2);
    FSI_TRY_LAUNCH("zero_counters");
  }
  // the zeroing of the other grid (16-byte words) is spread over at most
  // 1024 workgroups; particle workgroups beyond that also take part
  const int64_t n16 = zero_grid ? (4LL * p->nx * p->ny * p->nz * (int64_t)sizeof(R)) / 16 : 0;
  const int64_t zero_blocks = n16 ? std::min<int64_t>(blocks_for(n16), 1024) : 0;
  // with a work list from the particle sort, one workgroup per chunk (as P2G)
  const int32_t* chunks = (b->chunks && p->num_chunks > 0) ? b->chunks : nullptr;
  const int num_chunks = chunks ? p->num_chunks : 0;
  const int64_t particle_blocks = particle_workgroups<R>(p, b, c);
  const int64_t blocks = std::max<int64_t>(particle_blocks, std::max<int64_t>(zero_blocks, zero_ctr ? 1 : 0));
  if (blocks == 0) return 0;
  // particle workgroups that fit on the chip at once take the one-batch
  // gathers (as for P2G)
  const bool wide = particle_blocks <= (int64_t)p->wide_max_workgroups;
#define FSI_G2P_ARGS(CONTACT_ON)                                                                   \
  m, cp, (R*)b->x, (R*)b->v, (R*)b->C, b->active, grid_src, CONTACT_ON ? cb->solid : nullptr,      \
      CONTACT_ON ? (const R*)cb->cforce : nullptr, CONTACT_ON ? cb->contact_mask : nullptr,        \
      CONTACT_ON ? ctrs : nullptr, (uint4*)zero_grid, n16, zero_ctr, chunks, num_chunks,            \
      boxes_in, big_flag, big_expect, num_boxes
  const dim3 g2p_grid((unsigned)blocks), g2p_block(kBlock);
  bool launched = false;
  if constexpr (sizeof(R) == 4) {
    if (plane_split<R>(p, b, c)) {
      if (contact && refined(c))
        hipLaunchKernelGGL((k_g2p<R, true, false, 3, true>), g2p_grid, dim3(3 * Slots<3>::value), 0, st,
                           FSI_G2P_ARGS(true));
      else if (contact)
        hipLaunchKernelGGL((k_g2p<R, true, false, 3>), g2p_grid, dim3(3 * Slots<3>::value), 0, st,
                           FSI_G2P_ARGS(true));
      else
        hipLaunchKernelGGL((k_g2p<R, false, false, 3>), g2p_grid, dim3(3 * Slots<3>::value), 0, st,
                           FSI_G2P_ARGS(false));
      launched = true;
    }
  }
  if (launched) {
  } else if (contact && refined(c) && wide)
    hipLaunchKernelGGL((k_g2p<R, true, true, 1, true>), g2p_grid, g2p_block, 0, st, FSI_G2P_ARGS(true));
  else if (contact && refined(c))
    hipLaunchKernelGGL((k_g2p<R, true, false, 1, true>), g2p_grid, g2p_block, 0, st, FSI_G2P_ARGS(true));
  else if (contact && wide)
    hipLaunchKernelGGL((k_g2p<R, true, true, 1>), g2p_grid, g2p_block, 0, st, FSI_G2P_ARGS(true));
  else if (contact)
    hipLaunchKernelGGL((k_g2p<R, true, false, 1>), g2p_grid, g2p_block, 0, st, FSI_G2P_ARGS(true));
  else if (wide)
    hipLaunchKernelGGL((k_g2p<R, false, true, 1>), g2p_grid, g2p_block, 0, st, FSI_G2P_ARGS(false));
  else
    hipLaunchKernelGGL((k_g2p<R, false, false, 1>), g2p_grid, g2p_block, 0, st, FSI_G2P_ARGS(false));
#undef FSI_G2P_ARGS
  FSI_TRY_LAUNCH("mpm_g2p");
  return 0;
}

template <typename R>
int substeps_t(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_params* c,
               const fsi_coupling_buffers* cb, int32_t* ctr, int substeps, int build_first,
               int clear_force, hipStream_t st);

template <typename R>
int clear_t(const fsi_mpm_params* p, const fsi_mpm_buffers* b, int32_t* counters, hipStream_t st) {
  const int64_t n = 4LL * p->nx * p->ny * p->nz;
  hipLaunchKernelGGL(k_clear<R>, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, (R*)b->grid, counters);
  FSI_TRY_LAUNCH("mpm_clear_grid");
  return 0;
}

// Full clear of one grid.
template <typename R>
int clear_grid_t(const fsi_mpm_params* p, R* grid, hipStream_t st) {
  const int64_t n = 4LL * p->nx * p->ny * p->nz;
  hipLaunchKernelGGL(k_clear<R>, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, grid, nullptr);
  FSI_TRY_LAUNCH("mpm_clear_grid");
  return 0;
}

// Substeps with three grids, S + 1 launches: P2G of substep 0, then for
// s = 1..S-1 one launch of G2P(s-1) + P2G(s) (k_g2p_p2g), then G2P(S-1).
// Substep s scatters into grid s%3; the launch of substep s zeroes grid
// (s+1)%3, which the launch of s-1 finished reading, from the boxes the P2G
// of substep s-2 recorded (whole grids on the first launches). Counters of
// substep s go to row s of cb->counter_rows (zeroed once); the last row is
// copied to cb->counters at the end.
template <typename R>
int substeps_fused_t(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_params* c,
                     const fsi_coupling_buffers* cb, int32_t* ctr, int substeps, int build_first, int clear_force,
                     hipStream_t st, int64_t num_wg) {
  fsi_coupling_params local{};
  if (c) local = *c;
  R* G[3] = {(R*)b->grid, (R*)b->grid_alt, (R*)b->grid_alt2};
  int32_t* rows = ctr ? cb->counter_rows : nullptr;
  auto row = [&](int s) { return rows ? rows + 8 * s : nullptr; };
  auto flag_of = [&](int q) { return b->zero_boxes + q; };
  auto boxes_of = [&](int q) { return b->zero_boxes + 16 + (int64_t)q * 8 * b->zero_boxes_cap; };
  if (int rc = clear_grid_t<R>(p, G[0], st)) return rc;
  if (int rc = clear_grid_t<R>(p, G[1], st)) return rc;
  if (rows) {
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(8 * substeps)), dim3(kBlock), 0, st, 8 * substeps, rows, 0);
    FSI_TRY_LAUNCH("counter_rows");
  }
  local.build_fraction = build_first ? 1 : 0;
  if (int rc = p2g_t<R>(p, b, c ? &local : nullptr, cb, clear_force, 1, false, G[0], nullptr, st, boxes_of(0),
                        flag_of(0), 1, row(0)))
    return rc;
  local.build_fraction = 0;
  const fsi_coupling_params* lc = c ? &local : nullptr;
  MpmK<R> m = make_mpm<R>(p, b);
  CplK<R> cp = make_cpl<R>(lc, p);
  cp.force_mean = cb ? (R*)cb->force_mean : nullptr;
  const bool couple = lc && lc->enabled;
  const bool contact = lc && lc->contact_enabled;
  const int32_t* chunks = (b->chunks && p->num_chunks > 0) ? b->chunks : nullptr;
  const int64_t n16 = (4LL * p->nx * p->ny * p->nz * (int64_t)sizeof(R)) / 16;
  const int64_t zero_blocks = std::min<int64_t>(blocks_for(n16), 1024);
  const dim3 grid((unsigned)std::max<int64_t>(num_wg, zero_blocks)), block(kBlock);
  const bool wide = num_wg <= (int64_t)p->wide_max_workgroups;
  const bool split = plane_split<R>(p, b, c);
  const bool ref = refined(c);
  for (int s = 1; s < substeps; ++s) {
    const int q = (s + 1) % 3;
    const bool by_boxes = s >= 2;
#define FSI_FUSED_ARGS                                                                                          \
  m, cp, (R*)b->x, (R*)b->v, (R*)b->C, (R*)b->F, (const R*)b->mass, (const R*)b->volume, b->active,            \
      (R*)b->pforce, (const R*)G[(s - 1) % 3], G[s % 3], cb ? cb->solid : nullptr,                              \
      cb ? (const R*)cb->lbm_vel : nullptr, cb ? (R*)cb->cforce : nullptr, cb ? (R*)cb->valid_weight : nullptr, \
      cb ? cb->coupling_mask : nullptr, cb ? cb->contact_mask : nullptr, row(s - 1), row(s), clear_force,      \
      chunks, (int)num_wg, (uint4*)G[q], n16, by_boxes ? boxes_of(q) : nullptr, by_boxes ? flag_of(q) : nullptr, \
      s - 1, (int)num_wg, boxes_of(s % 3), flag_of(s % 3), s + 1, cb ? (const R*)cb->closure_force : nullptr
#define FSI_FUSED_LAUNCH(CPL, CON)                                                                        \
  do {                                                                                                    \
    if constexpr (sizeof(R) == 4) {                                                                       \
      if (split) {                                                                                        \
        if (ref && (CPL || CON))                                                                          \
          hipLaunchKernelGGL((k_g2p_p2g<R, CPL, CON, false, 3, true>), grid, dim3(3 * Slots<3>::value), 0, \
                             st, FSI_FUSED_ARGS);                                                         \
        else                                                                                              \
          hipLaunchKernelGGL((k_g2p_p2g<R, CPL, CON, false, 3>), grid, dim3(3 * Slots<3>::value), 0, st,   \
                             FSI_FUSED_ARGS);                                                             \
        break;                                                                                            \
      }                                                                                                   \
    }                                                                                                     \
    if (ref && (CPL || CON)) {                                                                            \
      if (wide)                                                                                           \
        hipLaunchKernelGGL((k_g2p_p2g<R, CPL, CON, true, 1, true>), grid, block, 0, st, FSI_FUSED_ARGS);  \
      else                                                                                                \
        hipLaunchKernelGGL((k_g2p_p2g<R, CPL, CON, false, 1, true>), grid, block, 0, st, FSI_FUSED_ARGS); \
    } else if (wide)                                                                                      \
      hipLaunchKernelGGL((k_g2p_p2g<R, CPL, CON, true, 1>), grid, block, 0, st, FSI_FUSED_ARGS);          \
    else                                                                                                  \
      hipLaunchKernelGGL((k_g2p_p2g<R, CPL, CON, false, 1>), grid, block, 0, st, FSI_FUSED_ARGS);         \
  } while (0)
    if (couple && contact)
      FSI_FUSED_LAUNCH(true, true);
    else if (couple)
      FSI_FUSED_LAUNCH(true, false);
    else if (contact)
      FSI_FUSED_LAUNCH(false, true);
    else
      FSI_FUSED_LAUNCH(false, false);
#undef FSI_FUSED_LAUNCH
#undef FSI_FUSED_ARGS
    FSI_TRY_LAUNCH("mpm_g2p_p2g");
  }
  const int last = substeps - 1;
  if (int rc = grid_g2p_t<R>(p, b, lc, cb, false, G[last % 3], nullptr, nullptr, st, nullptr, nullptr, 0, 0,
                             row(last)))
    return rc;
  if (rows) {
    hipLaunchKernelGGL(k_copy_i32, dim3(1), dim3(64), 0, st, 5, row(last), cb->counters);
    FSI_TRY_LAUNCH("counter_rows_copy");
  }
  return 0;
}

// fsi/mpm3d.py:189-214 repeated `substeps` times. With one grid: clear
// (+ counters), P2G (+ coupling), grid update + G2P (+ contact) per substep.
// With a second grid the clear runs once: substep s scatters into grid s%2,
// and its G2P pass zeroes grid (s+1)%2 and counters 0..2 for substep s+1
// while P2G zeroes counters 3..4, so a substep is two launches.
template <typename R>
int substeps_t(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_params* c,
               const fsi_coupling_buffers* cb, int32_t* ctr, int substeps, int build_first,
               int clear_force, hipStream_t st) {
  fsi_coupling_params local{};
  if (c) local = *c;
  R* grids[2] = {(R*)b->grid, (R*)b->grid_alt};
  const bool two = b->grid_alt != nullptr && b->grid_alt != b->grid;
  if (two && substeps > 0)
    if (int rc = clear_t<R>(p, b, ctr, st)) return rc;
  // dirty boxes: P2G pass s records its workgroups' boxes under parity s%2;
  // G2P pass s >= 1 zeroes the other grid from the boxes of pass s-1 (pass 0
  // zeroes it whole: its contents are unknown)
  const int64_t num_wg = particle_workgroups<R>(p, b, c);
  const bool boxes = two && b->zero_boxes && num_wg <= (int64_t)b->zero_boxes_cap;
  // three grids, box scratch for three passes and counter rows: one launch
  // per substep boundary
  const bool fused = boxes && b->grid_alt2 && b->grid_alt2 != b->grid && b->grid_alt2 != b->grid_alt &&
                     substeps >= 2 && num_wg > 0 &&
                     (!ctr || (cb && cb->counter_rows && cb->counter_rows_cap >= substeps));
  if (fused) return substeps_fused_t<R>(p, b, c, cb, ctr, substeps, build_first, clear_force, st, num_wg);
  auto flag_of = [&](int q) { return b->zero_boxes + q; };
  auto boxes_of = [&](int q) { return b->zero_boxes + 16 + (int64_t)q * 8 * b->zero_boxes_cap; };
  for (int s = 0; s < substeps; ++s) {
    local.build_fraction = (s == 0 && build_first) ? 1 : 0;
    const fsi_coupling_params* lc = c ? &local : nullptr;
    R* g = two ? grids[s & 1] : grids[0];
    if (!two)
      if (int rc = clear_t<R>(p, b, ctr, st)) return rc;
    int32_t* zc_p2g = (two && ctr && s > 0) ? ctr + 3 : nullptr;
    if (int rc = p2g_t<R>(p, b, lc, cb, clear_force, 1, false, g, zc_p2g, st, boxes ? boxes_of(s & 1) : nullptr,
                          boxes ? flag_of(s & 1) : nullptr, s + 1))
      return rc;
    const bool more = s + 1 < substeps;
    void* zg = (two && more) ? (void*)grids[(s + 1) & 1] : nullptr;
    int32_t* zc_g2p = (two && more && ctr) ? ctr : nullptr;
    const bool use_boxes = boxes && s >= 1;
    if (int rc = grid_g2p_t<R>(p, b, lc, cb, false, g, zg, zc_g2p, st, use_boxes ? boxes_of((s + 1) & 1) : nullptr,
                               use_boxes ? flag_of((s + 1) & 1) : nullptr, s, (int)num_wg))
      return rc;
  }
  return 0;
}

template <typename R>
int contact_t(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_params* c,
              const fsi_coupling_buffers* cb, hipStream_t st) {
  MpmK<R> m = make_mpm<R>(p, b);
  CplK<R> cp = make_cpl<R>(c, p);
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, cb->counters, 3, 2);
  FSI_TRY_LAUNCH("zero_counters");
  if (m.num == 0) return 0;
  if (refined(c))
    hipLaunchKernelGGL((k_contact<R, true>), dim3(blocks_for(m.num)), dim3(kBlock), 0, st, m, cp,
                       (const R*)b->x, (R*)b->v, b->active, cb->solid, (const R*)cb->cforce,
                       cb->contact_mask, cb->counters);
  else
    hipLaunchKernelGGL(k_contact<R>, dim3(blocks_for(m.num)), dim3(kBlock), 0, st, m, cp,
                       (const R*)b->x, (R*)b->v, b->active, cb->solid, (const R*)cb->cforce,
                       cb->contact_mask, cb->counters);
  FSI_TRY_LAUNCH("contact");
  return 0;
}

// ------------------------------------------- allocation-free substeps ----
//
// fsi_mpm_substeps_noalloc keeps the substep loop's scratch (the node boxes of
// three P2G passes and the per-substep counter rows) in one file-scope device
// array instead of the caller's fsi_mpm_buffers.zero_boxes and
// fsi_coupling_buffers.counter_rows. Every GPU holds its own copy until the
// process ends; calls on one GPU share it, so calls that overlap on one GPU
// race. Its stale contents are harmless: each call writes the boxes before
// they are read and zeroes the counter rows with a kernel, and a box-size
// flag left over from an earlier call can only select a whole-grid clear.

constexpr int kNoAllocMaxWorkgroups = 65536;  // P2G workgroups of one pass
constexpr int kNoAllocMaxSubsteps = 4096;     // substeps of one call with counters
constexpr int64_t kNoAllocBoxInts = 16 + 24LL * kNoAllocMaxWorkgroups;
constexpr int64_t kNoAllocScratchInts = kNoAllocBoxInts + 8LL * kNoAllocMaxSubsteps;  // 6.4 MB

}  // namespace
}  // namespace fsi

static __device__ __align__(256) int32_t scratch[fsi::kNoAllocScratchInts];

namespace fsi {
namespace {

// Device address of `scratch` on the current device, looked up once per
// device (host code must not pass the symbol itself to a kernel: its host-side
// name is not a device address).
int noalloc_scratch(int32_t** out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static int32_t* ptrs[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error("mpm_substeps_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    return FSI_ERR_HIP;
  }
  FSI_REQUIRE(dev >= 0 && dev < kMaxDevices, "mpm_substeps_noalloc: device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("mpm_substeps_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
      return FSI_ERR_HIP;
    }
    ptrs[dev] = static_cast<int32_t*>(p);
  }
  *out = ptrs[dev];
  return 0;
}

// a refined MPM grid: dims a multiple of the lattice's, no learned closure
// (its features are sampled on the MPM stencil)
int check_refine(const fsi_mpm_params* p, const fsi_coupling_params* c, const fsi_coupling_buffers* cb,
                 const char* name) {
  if (!refined(c)) return 0;
  const int r = c->refine;
  FSI_REQUIRE(p->nx % r == 0 && p->ny % r == 0 && p->nz % r == 0,
              "%s: MPM grid %d x %d x %d is not %d x a lattice", name, p->nx, p->ny, p->nz, r);
  FSI_REQUIRE(!cb || !cb->closure_force, "%s: the learned closure needs co-located grids (refine 1)", name);
  return 0;
}

int check_substeps(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_params* c,
                   const fsi_coupling_buffers* cb, int32_t substeps, const char* name) {
  if (int rc = check_mpm(p, b)) return rc;
  if (int rc = check_refine(p, c, cb, name)) return rc;
  FSI_REQUIRE(substeps >= 0, "%s: negative substep count", name);
  FSI_REQUIRE(b->x && b->v && b->C && b->F && b->mass && b->volume && b->active && b->pforce && b->grid,
              "%s: null particle or grid buffer", name);
  if (c && c->enabled)
    FSI_REQUIRE(cb && cb->solid && cb->lbm_vel && cb->cforce && cb->valid_weight && cb->coupling_mask &&
                    cb->counters,
                "%s: coupling needs all coupling buffers", name);
  if (c && c->contact_enabled)
    FSI_REQUIRE(cb && cb->solid && cb->cforce && cb->contact_mask && cb->counters,
                "%s: contact needs the coupling buffers", name);
  return 0;
}

}  // namespace
}  // namespace fsi

using namespace fsi;

extern "C" {

int fsi_mpm_clear_grid(const fsi_mpm_params* p, const fsi_mpm_buffers* b, void* stream) {
  if (int rc = check_mpm(p, b)) return rc;
  FSI_REQUIRE(b->grid, "mpm_clear_grid: null grid");
  hipStream_t st = (hipStream_t)stream;
  if (p->real == FSI_REAL_F64) return clear_t<double>(p, b, nullptr, st);
  return clear_t<float>(p, b, nullptr, st);
}

int fsi_mpm_p2g(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_params* c,
                const fsi_coupling_buffers* cb, int32_t clear_force, int32_t do_p2g, void* stream) {
  if (int rc = check_mpm(p, b)) return rc;
  if (int rc = check_refine(p, c, cb, "mpm_p2g")) return rc;
  FSI_REQUIRE(b->x && b->v && b->C && b->F && b->mass && b->volume && b->active && b->pforce,
              "mpm_p2g: null particle buffer");
  FSI_REQUIRE(!do_p2g || b->grid, "mpm_p2g: null grid");
  if (c && c->enabled)
    FSI_REQUIRE(cb && cb->solid && cb->lbm_vel && cb->cforce && cb->valid_weight &&
                    cb->coupling_mask && cb->counters,
                "mpm_p2g: coupling needs all coupling buffers");
  hipStream_t st = (hipStream_t)stream;
  if (p->real == FSI_REAL_F64)
    return p2g_t<double>(p, b, c, cb, clear_force, do_p2g, true, (double*)b->grid, nullptr, st);
  return p2g_t<float>(p, b, c, cb, clear_force, do_p2g, true, (float*)b->grid, nullptr, st);
}

int fsi_mpm_grid_g2p(const fsi_mpm_params* p, const fsi_mpm_buffers* b,
                     const fsi_coupling_params* c, const fsi_coupling_buffers* cb, void* stream) {
  if (int rc = check_mpm(p, b)) return rc;
  if (int rc = check_refine(p, c, cb, "mpm_grid_g2p")) return rc;
  FSI_REQUIRE(b->x && b->v && b->C && b->active && b->grid, "mpm_grid_g2p: null buffer");
  if (c && c->contact_enabled)
    FSI_REQUIRE(cb && cb->solid && cb->cforce && cb->contact_mask && cb->counters,
                "mpm_grid_g2p: contact needs the coupling buffers");
  hipStream_t st = (hipStream_t)stream;
  if (p->real == FSI_REAL_F64)
    return grid_g2p_t<double>(p, b, c, cb, true, (const double*)b->grid, nullptr, nullptr, st);
  return grid_g2p_t<float>(p, b, c, cb, true, (const float*)b->grid, nullptr, nullptr, st);
}

int fsi_contact(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_params* c,
                const fsi_coupling_buffers* cb, void* stream) {
  if (int rc = check_mpm(p, b)) return rc;
  if (int rc = check_refine(p, c, cb, "contact")) return rc;
  FSI_REQUIRE(c && cb && cb->solid && cb->cforce && cb->contact_mask && cb->counters && b->x &&
                  b->v && b->active,
              "contact: null buffer");
  hipStream_t st = (hipStream_t)stream;
  if (p->real == FSI_REAL_F64) return contact_t<double>(p, b, c, cb, st);
  return contact_t<float>(p, b, c, cb, st);
}

int fsi_mpm_substeps(const fsi_mpm_params* p, const fsi_mpm_buffers* b,
                     const fsi_coupling_params* c, const fsi_coupling_buffers* cb,
                     int32_t substeps, int32_t build_first, int32_t clear_force, void* stream) {
  if (int rc = check_substeps(p, b, c, cb, substeps, "mpm_substeps")) return rc;
  hipStream_t st = (hipStream_t)stream;
  const bool counters = c && (c->enabled || c->contact_enabled);
  int32_t* ctr = counters ? cb->counters : nullptr;
  if (p->real == FSI_REAL_F64) return substeps_t<double>(p, b, c, cb, ctr, substeps, build_first, clear_force, st);
  return substeps_t<float>(p, b, c, cb, ctr, substeps, build_first, clear_force, st);
}

int fsi_mpm_substeps_noalloc(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_params* c,
                             const fsi_coupling_buffers* cb, int32_t substeps, int32_t build_first,
                             int32_t clear_force, void* stream) {
  if (int rc = check_substeps(p, b, c, cb, substeps, "mpm_substeps_noalloc")) return rc;
  const int64_t wg = p->real == FSI_REAL_F64 ? particle_workgroups<double>(p, b, c) : particle_workgroups<float>(p, b, c);
  FSI_REQUIRE(wg <= kNoAllocMaxWorkgroups,
              "mpm_substeps_noalloc: %lld particle workgroups per pass exceed the static scratch's %d "
              "(use fsi_mpm_substeps with caller-provided zero_boxes)",
              (long long)wg, kNoAllocMaxWorkgroups);
  const bool counters = c && (c->enabled || c->contact_enabled);
  FSI_REQUIRE(!counters || substeps <= kNoAllocMaxSubsteps,
              "mpm_substeps_noalloc: %d substeps exceed the static counter rows' %d", substeps,
              kNoAllocMaxSubsteps);
  int32_t* s = nullptr;
  if (int rc = noalloc_scratch(&s)) return rc;
  fsi_mpm_buffers lb = *b;
  lb.zero_boxes = s;
  lb.zero_boxes_cap = kNoAllocMaxWorkgroups;
  fsi_coupling_buffers lcb{};
  if (cb) {
    lcb = *cb;
    lcb.counter_rows = s + kNoAllocBoxInts;
    lcb.counter_rows_cap = kNoAllocMaxSubsteps;
  }
  const fsi_coupling_buffers* pcb = cb ? &lcb : nullptr;
  hipStream_t st = (hipStream_t)stream;
  int32_t* ctr = counters ? cb->counters : nullptr;
  if (p->real == FSI_REAL_F64)
    return substeps_t<double>(p, &lb, c, pcb, ctr, substeps, build_first, clear_force, st);
  return substeps_t<float>(p, &lb, c, pcb, ctr, substeps, build_first, clear_force, st);
}

int fsi_closure_features(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_buffers* cb,
                         float* X, void* stream) {
  if (int rc = check_mpm(p, b)) return rc;
  FSI_REQUIRE(cb && cb->solid && cb->lbm_vel && cb->cforce && X, "closure_features: null buffer");
  hipStream_t st = (hipStream_t)stream;
  const int n = p->num_particles;
  if (n == 0) return 0;
  if (p->real == FSI_REAL_F64)
    hipLaunchKernelGGL(k_closure_features<double>, dim3(blocks_for(n)), dim3(kBlock), 0, st, make_mpm<double>(p),
                       (const double*)b->x, (const double*)b->v, b->active, cb->solid, (const double*)cb->lbm_vel,
                       (const double*)cb->cforce, X);
  else
    hipLaunchKernelGGL(k_closure_features<float>, dim3(blocks_for(n)), dim3(kBlock), 0, st, make_mpm<float>(p),
                       (const float*)b->x, (const float*)b->v, b->active, cb->solid, (const float*)cb->lbm_vel,
                       (const float*)cb->cforce, X);
  FSI_TRY_LAUNCH("closure_features");
  return 0;
}

int fsi_polar_rotation(int32_t real, int32_t n, const void* F, void* R, void* stream) {
  FSI_REQUIRE(F && R && n >= 0, "polar_rotation: bad arguments");
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (real == FSI_REAL_F64)
    hipLaunchKernelGGL(k_polar<double>, dim3(blocks_for(n)), dim3(kBlock), 0, st, (int)n,
                       (const double*)F, (double*)R);
  else
    hipLaunchKernelGGL(k_polar<float>, dim3(blocks_for(n)), dim3(kBlock), 0, st, (int)n,
                       (const float*)F, (float*)R);
  FSI_TRY_LAUNCH("polar_rotation");
  return 0;
}

}  // extern "C"
