// CPU oracle for the coupled LBM-MPM time step.
//
// TEST INFRASTRUCTURE ONLY. This file restates, loop for loop and in the same
// floating-point operation order, the Taichi kernels of the reference solver
// (lizhuoh9/taichi-mpm-lbm). Tests, the smoke check and the CPU-baseline leg of
// bench.py use it to check the HIP kernels. The product library never links,
// loads or calls it.
//
// Layouts are the reference's Taichi layouts (row-major, AoS):
//   cell index c = (i*ny + j)*nz + k
//   f, F        : [c*19 + q]
//   v, force    : [c*3 + a]
//   particle x  : [p*3 + a], C and F : [p*9 + r*3 + s]
//
// Built with -ffp-contract=off so that a*b+c is never fused: the FP32
// instantiation evaluates every expression in the order written here, which is
// the order the reference's Python source writes it. Taichi's own fast-math
// code generation and its unordered float atomics are not reproducible bit for
// bit, so this is the sequential IEEE reading of the reference.
//
// Python-level float constants (1/dx, 2*mu, 4*inv_dx, clamp bounds) arrive as
// doubles and are rounded to T exactly where Taichi would embed them as
// constants of the kernel's float type.
//
// Cell loops that write disjoint entries run under OpenMP; every loop whose
// result depends on order (particle scatters, counters, sums) stays serial.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

// Argument blocks passed by pointer from Python (ctypes.Structure mirrors).
struct MpmArgs {
  int nx, ny, nz;
  int num_particles;
  double dx;            // self.dx
  double inv_dx;        // self.inv_dx = 1/dx (Python float)
  double dt;            // dt_field
  double two_mu;        // 2.0 * self.mu evaluated in Python
  double lam;           // self.lam
  double gravity[3];
  int boundary_width;
  double boundary_damping;
  int apic;
  double four_inv_dx;   // 4.0 * self.inv_dx evaluated in Python
  double clamp_lower;   // 0.5 * dx
  double clamp_upper[3];  // (n - 1.5) * dx
};

struct CouplingArgs {
  int nx, ny, nz;
  int num_particles;
  double inv_dx;        // coupler.inv_dx = 1/mpm.dx (Python float)
  double gamma;
  double dt_ratio;
  double cell_volume;   // mpm.dx ** 3
  int build_fraction;
  double min_valid_weight;
  int use_relative_velocity_limit;
  double relative_velocity_limit;
  int use_force_limit;
  double force_limit;
  int gamma_ramp_steps;
  int coupled_step;
};

namespace {

constexpr int Q = 19;

// fsi/lbm3d.py:221-239 velocity set e; 171-174 opposite-direction table LR
constexpr int E[Q][3] = {
    {0, 0, 0},  {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},  {0, -1, 0},
    {0, 0, 1},  {0, 0, -1},  {1, 1, 0},   {-1, -1, 0}, {1, -1, 0},
    {-1, 1, 0}, {1, 0, 1},   {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
    {0, 1, 1},  {0, -1, -1}, {0, 1, -1},  {0, -1, 1}};
constexpr int LR[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

// fsi/lbm3d.py:244-248: w0 = 1/3, w1..6 = 1/18, w7..18 = 1/36, each the
// division evaluated in the kernel's float type.
template <typename T> inline T wq(int q) {
  if (q == 0) return T(1.0) / T(3.0);
  if (q < 7) return T(1.0) / T(18.0);
  return T(1.0) / T(36.0);
}

struct Dims {
  int nx, ny, nz;
  inline int64_t idx(int i, int j, int k) const { return ((int64_t)i * ny + j) * nz + k; }
  inline int64_t n() const { return (int64_t)nx * ny * nz; }
};

// Taichi Vector.dot: (a*b).sum(), summed left to right.
template <typename T> inline T dot3(const T* a, const T* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// fsi/lbm3d.py:250-254
template <typename T> inline T feq(int q, T rho, const T* u) {
  T ef[3] = {T(E[q][0]), T(E[q][1]), T(E[q][2])};
  T eu = dot3(ef, u);
  T uv = dot3(u, u);
  return wq<T>(q) * rho * (T(1.0) + T(3.0) * eu + T(4.5) * eu * eu - T(1.5) * uv);
}

// fsi/lbm3d.py:285-312: equilibrium moments, no rho factor on velocity terms.
template <typename T> inline T meq_vec(int m, T rho, const T* u) {
  switch (m) {
    case 0: return rho;
    case 1: return dot3(u, u);
    case 3: return u[0];
    case 5: return u[1];
    case 7: return u[2];
    case 9: return T(2.0) * u[0] * u[0] - u[1] * u[1] - u[2] * u[2];
    case 11: return u[1] * u[1] - u[2] * u[2];
    case 13: return u[0] * u[1];
    case 14: return u[1] * u[2];
    case 15: return u[0] * u[2];
    default: return T(0.0);
  }
}

// fsi/lbm3d.py:276-283: Guo forcing projected on moment row m.
template <typename T> inline T guo_force(int m, const T* M, const T* u, const T* fo) {
  T out = T(0.0);
  for (int q = 0; q < Q; ++q) {
    T ef[3] = {T(E[q][0]), T(E[q][1]), T(E[q][2])};
    T emu[3] = {ef[0] - u[0], ef[1] - u[1], ef[2] - u[2]};
    T e_minus_u_dot_f = dot3(emu, fo);
    T eu_ef = dot3(ef, u) * dot3(ef, fo);
    out += wq<T>(q) * (e_minus_u_dot_f + eu_ef) * M[m * Q + q];
  }
  return out;
}

// fsi/lbm3d.py:335-350: a one-cell periodic wrap on every axis.
inline int wrap(int v, int n) {
  if (v < 0) return n - 1;
  if (v > n - 1) return 0;
  return v;
}

// ---------------------------------------------------------------- LBM ----

// fsi/lbm3d.py:256-267
template <typename T>
void lbm_init_fields(Dims d, double rho0, const double* u0, const int32_t* solid, T* rho, T* v,
                     T* f, T* F) {
  const int64_t n = d.n();
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n; ++c) {
    rho[c] = T(rho0);
    T u[3] = {T(u0[0]), T(u0[1]), T(u0[2])};
    if (solid[c] != 0) u[0] = u[1] = u[2] = T(0.0);
    for (int a = 0; a < 3; ++a) v[c * 3 + a] = u[a];
    for (int q = 0; q < Q; ++q) {
      T value = feq<T>(q, rho[c], u);
      f[c * Q + q] = value;
      F[c * Q + q] = value;
    }
  }
}

// fsi/lbm3d.py:314-333 (with multiply_m 269-274, meq_vec 285-312, guo_force 276-283)
template <typename T>
void lbm_collide(Dims d, const int32_t* solid, const T* F, const T* rho, const T* v,
                 const T* force, const T* S, const T* M, const T* invM, T* moment, T* f) {
  const int64_t n = d.n();
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n; ++c) {
    if (solid[c] != 0) continue;
    const T* Fc = F + c * Q;
    const T* u = v + c * 3;
    const T* fo = force + c * 3;
    T* mc = moment + c * Q;
    for (int m = 0; m < Q; ++m) {
      T mv = T(0.0);
      for (int q = 0; q < Q; ++q) mv += M[m * Q + q] * Fc[q];
      T me = meq_vec<T>(m, rho[c], u);
      mv -= S[m] * (mv - me);
      mv += (T(1.0) - T(0.5) * S[m]) * guo_force<T>(m, M, u, fo);
      mc[m] = mv;
    }
    for (int q = 0; q < Q; ++q) {
      T acc = T(0.0);
      for (int m = 0; m < Q; ++m) acc += invM[q * Q + m] * mc[m];
      f[c * Q + q] = acc;
    }
  }
}

// fsi/lbm3d.py:352-361: push from fluid sources, half-way bounce-back into the
// source cell when the target is solid. Every written entry has one writer.
template <typename T> void lbm_stream(Dims d, const int32_t* solid, const T* f, T* F) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < d.nx; ++i)
    for (int j = 0; j < d.ny; ++j)
      for (int k = 0; k < d.nz; ++k) {
        const int64_t c = d.idx(i, j, k);
        if (solid[c] != 0) continue;
        for (int q = 0; q < Q; ++q) {
          const int64_t t =
              d.idx(wrap(i + E[q][0], d.nx), wrap(j + E[q][1], d.ny), wrap(k + E[q][2], d.nz));
          if (solid[t] == 0)
            F[t * Q + q] = f[c * Q + q];
          else
            F[c * Q + LR[q]] = f[c * Q + q];
        }
      }
}

// fsi/lbm3d.py:363-407. mode: 0 periodic, 1 pressure, 2 velocity.
template <typename T>
void lbm_apply_boundary(Dims d, int left_mode, int right_mode, double rho_left, double rho_right,
                        const double* vel_left, const double* vel_right, double rho0,
                        const int32_t* solid, const T* v, T* F) {
  const T r0 = T(rho0);
  for (int side = 0; side < 2; ++side) {
    const int mode = side == 0 ? left_mode : right_mode;
    if (mode == 0) continue;
    const int face = side == 0 ? 0 : d.nx - 1;
    const int inner = side == 0 ? 1 : d.nx - 2;
    const T rho_bc = T(side == 0 ? rho_left : rho_right);
    const double* vb = side == 0 ? vel_left : vel_right;
    const T vel_bc[3] = {T(vb[0]), T(vb[1]), T(vb[2])};
    for (int j = 0; j < d.ny; ++j)
      for (int k = 0; k < d.nz; ++k) {
        const int64_t c = d.idx(face, j, k);
        if (solid[c] != 0) continue;
        T* Fc = F + c * Q;
        if (mode == 1) {
          // pressure: equilibrium at rho_bc with the previous macroscopic
          // velocity of the face cell, or of its inner neighbour when that
          // neighbour is solid (lbm3d.py:369-372)
          const int64_t ci = d.idx(inner, j, k);
          const T* u = solid[ci] > 0 ? v + ci * 3 : v + c * 3;
          for (int q = 0; q < Q; ++q) Fc[q] = feq<T>(q, rho_bc, u);
        } else {
          // velocity: sequential in-place update over q using rho0
          for (int q = 0; q < Q; ++q) {
            const int opp = LR[q];
            Fc[q] = feq<T>(opp, r0, vel_bc) - Fc[opp] + feq<T>(q, r0, vel_bc);
          }
        }
      }
  }
}

// fsi/lbm3d.py:409-425
template <typename T>
void lbm_update_macro(Dims d, double rho0, const int32_t* solid, const T* F, const T* force, T* f,
                      T* rho, T* v) {
  const int64_t n = d.n();
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n; ++c) {
    if (solid[c] == 0) {
      T r = T(0.0);
      T u[3] = {T(0.0), T(0.0), T(0.0)};
      for (int q = 0; q < Q; ++q) {
        const T fq = F[c * Q + q];
        f[c * Q + q] = fq;
        r += fq;
        for (int a = 0; a < 3; ++a) u[a] += T(E[q][a]) * fq;
      }
      for (int a = 0; a < 3; ++a) u[a] /= r;
      for (int a = 0; a < 3; ++a) u[a] += (force[c * 3 + a] * T(0.5)) / r;
      rho[c] = r;
      for (int a = 0; a < 3; ++a) v[c * 3 + a] = u[a];
    } else {
      rho[c] = T(rho0);
      for (int a = 0; a < 3; ++a) v[c * 3 + a] = T(0.0);
    }
  }
}

// fsi/lbm3d.py:499-534. op: 0 clear, 1 reset/set uniform (fluid = value,
// solid = 0), 2 add uniform (fluid += value, solid = 0), 3 zero on solid.
template <typename T>
void lbm_force_op(Dims d, int op, const double* value, const int32_t* solid, T* force) {
  const int64_t n = d.n();
  const T val[3] = {T(value[0]), T(value[1]), T(value[2])};
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n; ++c) {
    T* fo = force + c * 3;
    switch (op) {
      case 0:
        fo[0] = fo[1] = fo[2] = T(0.0);
        break;
      case 1:
        for (int a = 0; a < 3; ++a) fo[a] = solid[c] == 0 ? val[a] : T(0.0);
        break;
      case 2:
        for (int a = 0; a < 3; ++a) fo[a] = solid[c] == 0 ? fo[a] + val[a] : T(0.0);
        break;
      case 3:
        if (solid[c] != 0) fo[0] = fo[1] = fo[2] = T(0.0);
        break;
      default:
        break;
    }
  }
}

// ---------------------------------------------------------------- MPM ----


// fsi/mpm3d.py:227-232: quadratic B-spline weights per axis.
template <typename T> inline void bspline_weights(const T* fx, T w[3][3]) {
  for (int a = 0; a < 3; ++a) {
    const T t0 = T(1.5) - fx[a];
    const T t1 = fx[a] - T(1.0);
    const T t2 = fx[a] - T(0.5);
    w[0][a] = T(0.5) * (t0 * t0);
    w[1][a] = T(0.75) - t1 * t1;
    w[2][a] = T(0.5) * (t2 * t2);
  }
}

inline bool inside(int nx, int ny, int nz, int i, int j, int k) {
  return 0 <= i && i < nx && 0 <= j && j < ny && 0 <= k && k < nz;
}

// 3x3 matrix helpers on row-major arrays; Taichi's matmul sums k left to right.
template <typename T> inline void matmul3(const T* a, const T* b, T* out) {
  for (int r = 0; r < 3; ++r)
    for (int s = 0; s < 3; ++s)
      out[r * 3 + s] = a[r * 3 + 0] * b[0 * 3 + s] + a[r * 3 + 1] * b[1 * 3 + s] +
                       a[r * 3 + 2] * b[2 * 3 + s];
}

template <typename T> inline T det3(const T* m) {
  return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) +
         m[6] * (m[1] * m[5] - m[4] * m[2]);
}

// Rotation factor R = U V^T of the rotation-variant SVD F = U diag(s) V^T
// (U, V proper rotations; the smallest singular value carries the sign of
// det F), the quantity fsi/mpm3d.py:245-246 forms from ti.svd.
// Computed in double: cyclic Jacobi on F^T F for V, Givens QR of F V for U.
void polar_rotation_d(const double* Fm, double* R) {
  double A[9];
  for (int r = 0; r < 3; ++r)
    for (int s = 0; s < 3; ++s)
      A[r * 3 + s] = Fm[0 * 3 + r] * Fm[0 * 3 + s] + Fm[1 * 3 + r] * Fm[1 * 3 + s] +
                     Fm[2 * 3 + r] * Fm[2 * 3 + s];
  double V[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int sweep = 0; sweep < 60; ++sweep) {
    const double off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
    const double diag = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
    if (off <= 1e-36 * diag || off == 0.0) break;
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0;
      const int q = pq == 0 ? 1 : 2;
      const double apq = A[p * 3 + q];
      if (apq == 0.0) continue;
      const double theta = (A[q * 3 + q] - A[p * 3 + p]) / (2.0 * apq);
      const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
      const double c = 1.0 / std::sqrt(t * t + 1.0);
      const double s = t * c;
      // A <- J^T A J with J the rotation in the (p,q) plane
      for (int k = 0; k < 3; ++k) {
        const double akp = A[k * 3 + p], akq = A[k * 3 + q];
        A[k * 3 + p] = c * akp - s * akq;
        A[k * 3 + q] = s * akp + c * akq;
      }
      for (int k = 0; k < 3; ++k) {
        const double apk = A[p * 3 + k], aqk = A[q * 3 + k];
        A[p * 3 + k] = c * apk - s * aqk;
        A[q * 3 + k] = s * apk + c * aqk;
      }
      for (int k = 0; k < 3; ++k) {
        const double vkp = V[k * 3 + p], vkq = V[k * 3 + q];
        V[k * 3 + p] = c * vkp - s * vkq;
        V[k * 3 + q] = s * vkp + c * vkq;
      }
    }
  }
  // sort eigenvalues descending, keeping V a proper rotation
  double lam[3] = {A[0], A[4], A[8]};
  int order[3] = {0, 1, 2};
  std::sort(order, order + 3, [&](int a, int b) { return lam[a] > lam[b]; });
  double Vs[9];
  for (int r = 0; r < 3; ++r)
    for (int s = 0; s < 3; ++s) Vs[r * 3 + s] = V[r * 3 + order[s]];
  if (det3(Vs) < 0)
    for (int r = 0; r < 3; ++r) Vs[r * 3 + 2] = -Vs[r * 3 + 2];
  // B = F V; Givens QR B = Qm Rt, Qm a rotation
  double B[9];
  matmul3(Fm, Vs, B);
  double Qm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  const int rows[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  const int cols[3] = {0, 0, 1};
  for (int g = 0; g < 3; ++g) {
    const int a = rows[g][0], b = rows[g][1], col = cols[g];
    const double x = B[a * 3 + col], y = B[b * 3 + col];
    const double r = std::hypot(x, y);
    double c = 1.0, s = 0.0;
    if (r > 0.0) {
      c = x / r;
      s = y / r;
    }
    for (int k = 0; k < 3; ++k) {
      const double ba = B[a * 3 + k], bb = B[b * 3 + k];
      B[a * 3 + k] = c * ba + s * bb;
      B[b * 3 + k] = -s * ba + c * bb;
    }
    // Qm <- Qm G^T so that Qm * (G B) stays equal to the original F V
    for (int k = 0; k < 3; ++k) {
      const double qa = Qm[k * 3 + a], qb = Qm[k * 3 + b];
      Qm[k * 3 + a] = c * qa + s * qb;
      Qm[k * 3 + b] = -s * qa + c * qb;
    }
  }
  for (int r = 0; r < 3; ++r)
    for (int s = 0; s < 3; ++s)
      R[r * 3 + s] = Qm[r * 3 + 0] * Vs[s * 3 + 0] + Qm[r * 3 + 1] * Vs[s * 3 + 1] +
                     Qm[r * 3 + 2] * Vs[s * 3 + 2];
}

template <typename T> void polar_rotation(const T* F, T* R) {
  double Fd[9], Rd[9];
  for (int i = 0; i < 9; ++i) Fd[i] = double(F[i]);
  polar_rotation_d(Fd, Rd);
  for (int i = 0; i < 9; ++i) R[i] = T(Rd[i]);
}

// fsi/mpm3d.py:234-266. Particles are processed in index order, which fixes
// the order of the (in Taichi, atomic) grid accumulations.
template <typename T>
void mpm_p2g(const MpmArgs& a, const T* x, const T* v, const T* C, T* Fdef, const T* mass,
             const T* vol, const int32_t* active, const T* pforce, T* grid_v, T* grid_m) {
  const T dt = T(a.dt), inv_dx = T(a.inv_dx), dxT = T(a.dx);
  const T two_mu = T(a.two_mu), lam = T(a.lam);
  for (int p = 0; p < a.num_particles; ++p) {
    if (active[p] == 0) continue;
    T gp[3], fx[3];
    int base[3];
    for (int d = 0; d < 3; ++d) {
      gp[d] = x[p * 3 + d] * inv_dx;
      base[d] = (int)(gp[d] - T(0.5));  // truncation toward zero, as .cast(ti.i32)
      fx[d] = gp[d] - T(base[d]);
    }
    T w[3][3];
    bspline_weights(fx, w);
    // F <- (I + dt C) F
    T IdC[9];
    for (int r = 0; r < 9; ++r) IdC[r] = (r % 4 == 0 ? T(1.0) : T(0.0)) + dt * C[p * 9 + r];
    T Fn[9];
    matmul3(IdC, Fdef + p * 9, Fn);
    for (int r = 0; r < 9; ++r) Fdef[p * 9 + r] = Fn[r];
    const T J = det3(Fn);
    T R[9];
    polar_rotation(Fn, R);
    // stress = 2 mu (F - R) F^T + lam J (J - 1) I
    T FmR[9], FT[9], st[9];
    for (int r = 0; r < 3; ++r)
      for (int s = 0; s < 3; ++s) {
        FmR[r * 3 + s] = two_mu * (Fn[r * 3 + s] - R[r * 3 + s]);
        FT[r * 3 + s] = Fn[s * 3 + r];
      }
    matmul3(FmR, FT, st);
    const T lj = lam * J * (J - T(1.0));
    for (int r = 0; r < 9; ++r) st[r] = st[r] + lj * (r % 4 == 0 ? T(1.0) : T(0.0));
    const T scale = -dt * vol[p] * T(4.0) * inv_dx * inv_dx;
    T affine[9];
    for (int r = 0; r < 9; ++r) affine[r] = scale * st[r] + mass[p] * C[p * 9 + r];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          const int ni = base[0] + i, nj = base[1] + j, nk = base[2] + k;
          if (!inside(a.nx, a.ny, a.nz, ni, nj, nk)) continue;
          const T dpos[3] = {(T(i) - fx[0]) * dxT, (T(j) - fx[1]) * dxT, (T(k) - fx[2]) * dxT};
          const T weight = w[i][0] * w[j][1] * w[k][2];
          const int64_t node = ((int64_t)ni * a.ny + nj) * a.nz + nk;
          for (int d = 0; d < 3; ++d) {
            const T ad = affine[d * 3 + 0] * dpos[0] + affine[d * 3 + 1] * dpos[1] +
                         affine[d * 3 + 2] * dpos[2];
            const T mom = mass[p] * v[p * 3 + d] + ad + dt * pforce[p * 3 + d];
            grid_v[node * 3 + d] += weight * mom;
          }
          grid_m[node] += weight * mass[p];
        }
  }
}

// fsi/mpm3d.py:268-288
template <typename T> void mpm_grid_op(const MpmArgs& a, T* grid_v, const T* grid_m) {
  const T dt = T(a.dt);
  const T g[3] = {T(a.gravity[0]), T(a.gravity[1]), T(a.gravity[2])};
  const T damp = T(a.boundary_damping);
  const int bw = a.boundary_width;
  const int n[3] = {a.nx, a.ny, a.nz};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < a.nx; ++i)
    for (int j = 0; j < a.ny; ++j)
      for (int k = 0; k < a.nz; ++k) {
        const int64_t node = ((int64_t)i * a.ny + j) * a.nz + k;
        const T m = grid_m[node];
        if (!(m > T(0.0))) continue;
        T* gv = grid_v + node * 3;
        for (int d = 0; d < 3; ++d) gv[d] = gv[d] / m;
        for (int d = 0; d < 3; ++d) gv[d] = gv[d] + dt * g[d];
        const int idx[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) {
          if (idx[d] < bw && gv[d] < T(0.0)) gv[d] = -damp * gv[d];
          if (idx[d] >= n[d] - bw && gv[d] > T(0.0)) gv[d] = -damp * gv[d];
        }
      }
}

// fsi/mpm3d.py:290-328
template <typename T>
void mpm_g2p(const MpmArgs& a, T* x, T* v, T* C, const int32_t* active, const T* grid_v) {
  const T dt = T(a.dt), inv_dx = T(a.inv_dx), c4 = T(a.four_inv_dx);
  const T lower = T(a.clamp_lower);
  const T upper[3] = {T(a.clamp_upper[0]), T(a.clamp_upper[1]), T(a.clamp_upper[2])};
#pragma omp parallel for schedule(static)
  for (int p = 0; p < a.num_particles; ++p) {
    if (active[p] == 0) continue;
    T gp[3], fx[3];
    int base[3];
    for (int d = 0; d < 3; ++d) {
      gp[d] = x[p * 3 + d] * inv_dx;
      base[d] = (int)(gp[d] - T(0.5));
      fx[d] = gp[d] - T(base[d]);
    }
    T w[3][3];
    bspline_weights(fx, w);
    T nv[3] = {T(0.0), T(0.0), T(0.0)};
    T nc[9] = {T(0.0), T(0.0), T(0.0), T(0.0), T(0.0), T(0.0), T(0.0), T(0.0), T(0.0)};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          const int ni = base[0] + i, nj = base[1] + j, nk = base[2] + k;
          if (!inside(a.nx, a.ny, a.nz, ni, nj, nk)) continue;
          const T dpos[3] = {T(i) - fx[0], T(j) - fx[1], T(k) - fx[2]};
          const T weight = w[i][0] * w[j][1] * w[k][2];
          const int64_t node = ((int64_t)ni * a.ny + nj) * a.nz + nk;
          const T* gv = grid_v + node * 3;
          for (int d = 0; d < 3; ++d) nv[d] += weight * gv[d];
          const T s = c4 * weight;
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) nc[r * 3 + c] += s * (gv[r] * dpos[c]);
        }
    for (int d = 0; d < 3; ++d) v[p * 3 + d] = nv[d];
    for (int r = 0; r < 9; ++r) C[p * 9 + r] = a.apic ? nc[r] : T(0.0);
    for (int d = 0; d < 3; ++d) x[p * 3 + d] += dt * v[p * 3 + d];
    for (int d = 0; d < 3; ++d) x[p * 3 + d] = std::min(std::max(x[p * 3 + d], lower), upper[d]);
  }
}

// ----------------------------------------------------------- coupling ----


// counters[]: 0 unsupported, 1 partial support, 2 clipped, 3 contact
// candidates, 4 contact damped, 5 IB active cells, 6 IB clipped cells.

// fsi/coupling.py:332-406 (effective gamma 321-330). The fluid side
// accumulations are done in particle order.
template <typename T>
void coupling_forces(const CouplingArgs& a, const T* x, const T* v, const T* mass, const T* vol,
                     const int32_t* active, const int32_t* solid, const T* lbm_v, T* pforce,
                     T* valid_weight, int32_t* cmask, T* cforce, T* svf, int32_t* counters) {
  const T inv_dx = T(a.inv_dx), dt_ratio = T(a.dt_ratio), cell_volume = T(a.cell_volume);
  const T min_w = T(a.min_valid_weight), tol = T(1.0e-5);
  for (int p = 0; p < a.num_particles; ++p) {
    if (active[p] == 0) continue;
    T gp[3], fx[3];
    int base[3];
    for (int d = 0; d < 3; ++d) {
      gp[d] = x[p * 3 + d] * inv_dx;
      base[d] = (int)(gp[d] - T(0.5));
      fx[d] = gp[d] - T(base[d]);
    }
    T w[3][3];
    bspline_weights(fx, w);
    T wsum = T(0.0);
    T u[3] = {T(0.0), T(0.0), T(0.0)};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          const int ni = base[0] + i, nj = base[1] + j, nk = base[2] + k;
          if (!inside(a.nx, a.ny, a.nz, ni, nj, nk)) continue;
          const int64_t node = ((int64_t)ni * a.ny + nj) * a.nz + nk;
          if (solid[node] != 0) continue;
          const T weight = w[i][0] * w[j][1] * w[k][2];
          wsum += weight;
          for (int d = 0; d < 3; ++d) u[d] += weight * lbm_v[node * 3 + d];
        }
    valid_weight[p] = wsum;
    if (wsum > min_w) {
      if (wsum < T(1.0) - tol) counters[1] += 1;
      for (int d = 0; d < 3; ++d) u[d] /= wsum;
      T dv[3] = {u[0] - v[p * 3 + 0], u[1] - v[p * 3 + 1], u[2] - v[p * 3 + 2]};
      int clipped = 0;
      if (a.use_relative_velocity_limit) {
        const T nrm = std::sqrt(dot3(dv, dv));
        const T lim = T(a.relative_velocity_limit);
        if (nrm > lim && nrm > T(0.0)) {
          for (int d = 0; d < 3; ++d) dv[d] = dv[d] / nrm * lim;
          clipped = 1;
        }
      }
      T gamma = T(a.gamma);
      if (a.gamma_ramp_steps > 0) {
        const T ramp = (T(a.coupled_step) + T(1.0)) / T(a.gamma_ramp_steps);
        gamma *= std::min(T(1.0), ramp);
      }
      T pf[3];
      for (int d = 0; d < 3; ++d) pf[d] = gamma * mass[p] * dv[d];
      if (a.use_force_limit) {
        const T nrm = std::sqrt(dot3(pf, pf));
        const T lim = T(a.force_limit);
        if (nrm > lim && nrm > T(0.0)) {
          for (int d = 0; d < 3; ++d) pf[d] = pf[d] / nrm * lim;
          clipped = 1;
        }
      }
      for (int d = 0; d < 3; ++d) pforce[p * 3 + d] = pf[d];
      cmask[p] = 1;
      if (clipped) counters[2] += 1;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) {
            const int ni = base[0] + i, nj = base[1] + j, nk = base[2] + k;
            if (!inside(a.nx, a.ny, a.nz, ni, nj, nk)) continue;
            const int64_t node = ((int64_t)ni * a.ny + nj) * a.nz + nk;
            if (solid[node] != 0) continue;
            const T weight = w[i][0] * w[j][1] * w[k][2];
            const T nw = weight / wsum;
            for (int d = 0; d < 3; ++d) cforce[node * 3 + d] += -nw * pf[d] * dt_ratio / cell_volume;
            if (a.build_fraction) svf[node] += nw * vol[p] / cell_volume;
          }
    } else {
      for (int d = 0; d < 3; ++d) pforce[p * 3 + d] = T(0.0);
      cmask[p] = 0;
      counters[0] += 1;
    }
  }
}

// fsi/coupling.py:421-452. total[3] accumulates in cell order.
template <typename T>
void ib_compute(Dims d, int enabled, double drag_d, double thr_d, int use_max, double max_d,
                const int32_t* solid, const T* svf, const T* rho, const T* v, T* ibf,
                int32_t* counters, T* total) {
  counters[5] = 0;
  counters[6] = 0;
  total[0] = total[1] = total[2] = T(0.0);
  const T drag = T(drag_d), thr = T(thr_d), max_force = T(max_d);
  const int64_t n = d.n();
  for (int64_t c = 0; c < n; ++c) {
    T fo[3] = {T(0.0), T(0.0), T(0.0)};
    if (enabled != 0 && drag > T(0.0) && solid[c] == 0) {
      const T phi = svf[c];
      if (phi > T(0.0) && phi >= thr) {
        for (int a = 0; a < 3; ++a) fo[a] = -drag * phi * rho[c] * v[c * 3 + a];
        const T nrm = std::sqrt(dot3(fo, fo));
        if (use_max != 0 && nrm > max_force && nrm > T(0.0)) {
          for (int a = 0; a < 3; ++a) fo[a] = fo[a] / nrm * max_force;
          counters[6] += 1;
        }
        counters[5] += 1;
        for (int a = 0; a < 3; ++a) total[a] += fo[a];
      }
    }
    for (int a = 0; a < 3; ++a) ibf[c * 3 + a] = fo[a];
  }
}

// fsi/coupling.py:463-496
template <typename T>
void contact(const CouplingArgs& a, int enabled, double thr_d, double damping_d, const T* x, T* v,
             const int32_t* active, const int32_t* solid, const T* svf, int32_t* contact_mask,
             int32_t* counters) {
  counters[3] = 0;
  counters[4] = 0;
  const T inv_dx = T(a.inv_dx), thr = T(thr_d), damping = T(damping_d);
  for (int p = 0; p < a.num_particles; ++p) {
    contact_mask[p] = 0;
    if (enabled == 0 || active[p] == 0) continue;
    T gp[3], fx[3];
    int base[3];
    for (int d = 0; d < 3; ++d) {
      gp[d] = x[p * 3 + d] * inv_dx;
      base[d] = (int)(gp[d] - T(0.5));
      fx[d] = gp[d] - T(base[d]);
    }
    T w[3][3];
    bspline_weights(fx, w);
    T dyn = T(0.0), stat = T(0.0);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          const int ni = base[0] + i, nj = base[1] + j, nk = base[2] + k;
          if (!inside(a.nx, a.ny, a.nz, ni, nj, nk)) continue;
          const int64_t node = ((int64_t)ni * a.ny + nj) * a.nz + nk;
          const T weight = w[i][0] * w[j][1] * w[k][2];
          dyn += weight * svf[node];
          if (solid[node] != 0) stat += weight;
        }
    if (dyn >= thr || stat >= thr) {
      contact_mask[p] = 1;
      counters[3] += 1;
      if (damping > T(0.0)) {
        for (int d = 0; d < 3; ++d) v[p * 3 + d] *= T(1.0) - damping;
        counters[4] += 1;
      }
    }
  }
}

}  // namespace

extern "C" {

int oref_abi_version() { return 1; }

}  // extern "C"

#define OREF_INSTANTIATE(T, SUF)                                                                 \
  extern "C" {                                                                                   \
  void oref_lbm_init_fields_##SUF(int nx, int ny, int nz, double rho0, const double* u0,        \
                                  const int32_t* solid, T* rho, T* v, T* f, T* F) {              \
    lbm_init_fields<T>({nx, ny, nz}, rho0, u0, solid, rho, v, f, F);                              \
  }                                                                                              \
  void oref_lbm_collide_##SUF(int nx, int ny, int nz, const int32_t* solid, const T* F,         \
                              const T* rho, const T* v, const T* force, const T* S, const T* M,  \
                              const T* invM, T* moment, T* f) {                                  \
    lbm_collide<T>({nx, ny, nz}, solid, F, rho, v, force, S, M, invM, moment, f);                 \
  }                                                                                              \
  void oref_lbm_stream_##SUF(int nx, int ny, int nz, const int32_t* solid, const T* f, T* F) {  \
    lbm_stream<T>({nx, ny, nz}, solid, f, F);                                                     \
  }                                                                                              \
  void oref_lbm_apply_boundary_##SUF(int nx, int ny, int nz, int left_mode, int right_mode,     \
                                     double rho_left, double rho_right, const double* vel_left, \
                                     const double* vel_right, double rho0,                       \
                                     const int32_t* solid, const T* v, T* F) {                   \
    lbm_apply_boundary<T>({nx, ny, nz}, left_mode, right_mode, rho_left, rho_right, vel_left,    \
                          vel_right, rho0, solid, v, F);                                         \
  }                                                                                              \
  void oref_lbm_update_macro_##SUF(int nx, int ny, int nz, double rho0, const int32_t* solid,   \
                                   const T* F, const T* force, T* f, T* rho, T* v) {             \
    lbm_update_macro<T>({nx, ny, nz}, rho0, solid, F, force, f, rho, v);                          \
  }                                                                                              \
  void oref_lbm_force_op_##SUF(int nx, int ny, int nz, int op, const double* value,             \
                               const int32_t* solid, T* force) {                                 \
    lbm_force_op<T>({nx, ny, nz}, op, value, solid, force);                                       \
  }                                                                                              \
  void oref_mpm_p2g_##SUF(const MpmArgs* a, const T* x, const T* v, const T* C, T* Fdef,        \
                          const T* mass, const T* vol, const int32_t* active, const T* pforce,   \
                          T* grid_v, T* grid_m) {                                                \
    mpm_p2g<T>(*a, x, v, C, Fdef, mass, vol, active, pforce, grid_v, grid_m);                     \
  }                                                                                              \
  void oref_mpm_grid_op_##SUF(const MpmArgs* a, T* grid_v, const T* grid_m) {                    \
    mpm_grid_op<T>(*a, grid_v, grid_m);                                                           \
  }                                                                                              \
  void oref_mpm_g2p_##SUF(const MpmArgs* a, T* x, T* v, T* C, const int32_t* active,            \
                          const T* grid_v) {                                                     \
    mpm_g2p<T>(*a, x, v, C, active, grid_v);                                                      \
  }                                                                                              \
  void oref_polar_rotation_##SUF(const T* F, T* R) { polar_rotation<T>(F, R); }                   \
  void oref_coupling_forces_##SUF(const CouplingArgs* a, const T* x, const T* v,                \
                                  const T* mass, const T* vol, const int32_t* active,            \
                                  const int32_t* solid, const T* lbm_v, T* pforce,               \
                                  T* valid_weight, int32_t* cmask, T* cforce, T* svf,            \
                                  int32_t* counters) {                                           \
    coupling_forces<T>(*a, x, v, mass, vol, active, solid, lbm_v, pforce, valid_weight, cmask,   \
                       cforce, svf, counters);                                                   \
  }                                                                                              \
  void oref_ib_compute_##SUF(int nx, int ny, int nz, int enabled, double drag, double thr,      \
                             int use_max, double max_force, const int32_t* solid, const T* svf,  \
                             const T* rho, const T* v, T* ibf, int32_t* counters, T* total) {    \
    ib_compute<T>({nx, ny, nz}, enabled, drag, thr, use_max, max_force, solid, svf, rho, v, ibf, \
                  counters, total);                                                              \
  }                                                                                              \
  void oref_contact_##SUF(const CouplingArgs* a, int enabled, double thr, double damping,       \
                          const T* x, T* v, const int32_t* active, const int32_t* solid,        \
                          const T* svf, int32_t* contact_mask, int32_t* counters) {              \
    contact<T>(*a, enabled, thr, damping, x, v, active, solid, svf, contact_mask, counters);     \
  }                                                                                              \
  }

OREF_INSTANTIATE(float, f32)
OREF_INSTANTIATE(double, f64)
