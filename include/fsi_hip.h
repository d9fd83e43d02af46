/*
 * C ABI of the MI355X (gfx950) kernels of the coupled LBM-MPM time step.
 *
 * The reference solver (lizhuoh9/taichi-mpm-lbm) has no FFI: its boundary is
 * the Python object API of fsi/lbm3d.py, fsi/mpm3d.py and fsi/coupling.py,
 * whose @ti.kernel methods run on Taichi fields. Each entry point below
 * replaces one or a fused group of those kernels; the comment names them.
 * The Python package (fsi/_native.py) binds this header with ctypes and keeps
 * the reference's class API on top.
 *
 * Conventions
 *  - every pointer is a device pointer unless named host_*; sizes are element
 *    counts; no entry point allocates, frees or synchronises unless it says so.
 *  - every entry point enqueues on the given HIP stream and returns 0 on
 *    success or a negative FSI_ERR_* code; fsi_last_error() then holds a
 *    message for the calling thread. Nothing throws across the ABI.
 *  - fields are SoA: component c of cell n lives at [c * ncells + n]; the cell
 *    index is n = (i * ny + j) * nz + k, the reference's row-major order.
 *    Population buffers (19 * ncells entries of the storage type) are blocked
 *    by lattice row instead: population q of cell n lives at
 *    [((i * ny + j) * 19 + q) * nz + k]; fsi_lbm_export_populations returns
 *    them in the reference's (n, 19) order.
 *  - `real` buffers are float for FSI_REAL_F32 and double for FSI_REAL_F64.
 */
#ifndef FSI_HIP_H_
#define FSI_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSI_ABI_VERSION 5

enum {
  FSI_OK = 0,
  FSI_ERR_ARGUMENT = -1,
  FSI_ERR_HIP = -2,
};

enum { FSI_REAL_F32 = 0, FSI_REAL_F64 = 1 };

/* population storage: absolute FP32 (reference), shifted+scaled FP16, FP64,
 * shifted signed-normalised 16-bit integers with an adaptive scale (I16) */
enum { FSI_POP_F32 = 0, FSI_POP_F16 = 1, FSI_POP_F64 = 2, FSI_POP_I16 = 3 };

/* floats of the I16 storage's control block (fsi_lbm_buffers.store_ctrl) */
#define FSI_STORE_CTRL_FLOATS 8320

/* x-face boundary modes, fsi/lbm3d.py:16-27 */
enum { FSI_BC_PERIODIC = 0, FSI_BC_PRESSURE = 1, FSI_BC_VELOCITY = 2 };

/* -------------------------------------------------------------- LBM ---- */

typedef struct fsi_lbm_params {
  int32_t nx, ny, nz;
  int32_t storage;         /* FSI_POP_* ; FSI_POP_F64 implies real = double */
  double rho0;
  double initial_velocity[3];
  double s_v, s_o;         /* relaxation rates, fsi/lbm3d.py:141-168 */
  int32_t left_mode, right_mode;
  double bc_rho[2];        /* x_left.rho, x_right.rho */
  double bc_vel[2][3];     /* x_left.velocity, x_right.velocity */
  double pop_scale;        /* FP16 storage scale (power of two); I16: the scale of the
                              populations fsi_lbm_init writes; ignored otherwise */
  int32_t store_epoch;     /* I16: population writes so far (fsi_lbm_init writes epoch 0,
                              each fsi_lbm_step call the next one); ignored otherwise */
  int32_t store_pass;      /* I16, fsi_lbm_step: 0 a normal step (its scale adapted from the
                              previous write's largest value); 1 a measuring pass that
                              computes the step without writing anything but the largest
                              value of what it would store (its slot of store_ctrl zeroed
                              before); 2 the step after a measuring pass, written at a scale
                              from that measurement. Ignored otherwise */
} fsi_lbm_params;

typedef struct fsi_lbm_buffers {
  const int32_t* solid;    /* N, 0 fluid / 1 solid */
  uint32_t* nbmask;        /* N, bit q set when the q-neighbour (bit 0: the cell) is solid */
  void* rho;               /* N real: macroscopic density */
  void* vel;               /* 3N real: macroscopic velocity */
  void* force;             /* 3N real: per-cell body force (fsi/lbm3d.py:126) */
  void* face_vel;          /* 2 slots x 2 faces x 3 x ny*nz real: macroscopic velocity of the
                              x=0 / x=nx-1 cells, read by the next pressure boundary */
  float* store_ctrl;       /* I16 storage: FSI_STORE_CTRL_FLOATS floats, zeroed before
                              fsi_lbm_init: per write epoch (three rotating slots) the scale
                              the populations were written with and the largest |value| as a
                              fraction of the 16-bit range, and a fourth slot for a measuring
                              pass (floats [6240, 8320)); each step adapts the scale from the
                              previous write's maximum. Unused otherwise (may be null) */
  uint32_t* solid_bits;    /* optional (may be null), used when nz % 64 == 0:
                              FSI_SOLID_BITS_WORDS(nx, ny, nz) words written by
                              fsi_lbm_init, the solid flags of each z-row as bits (bit b of
                              row (i, j): cell k = b - 1, the periodic neighbours k = nz - 1 at
                              b = 0 and k = 0 at b = nz + 1). The lattice step then reads a
                              wavefront's 9 neighbour rows with scalar loads instead of the
                              4-byte nbmask of every cell */
} fsi_lbm_buffers;

/* words of fsi_lbm_buffers.solid_bits: nx * ny rows of nz / 32 + 2 words */
#define FSI_SOLID_BITS_WORDS(nx, ny, nz) ((int64_t)(nx) * (ny) * ((nz) / 32 + 2))

/* Force seen by one step. When `field` is null, every fluid cell uses
 * `uniform` and no force memory is read. */
typedef struct fsi_lbm_force {
  const void* field;       /* 3N real or null */
  double uniform[3];
} fsi_lbm_force;

/* Lattice phases of fsi_lbm_step. Between steps the populations are either
 * post-collision (after a step) or post-stream with rho/u materialised (the
 * reference's state after update_macro; after initialize). */
enum {
  FSI_PHASE_STREAM_COLLIDE = 0, /* post -> post: one full step */
  FSI_PHASE_COLLIDE = 1,        /* materialised -> post: collide the stored state */
  FSI_PHASE_STREAM_MACRO = 2,   /* post -> materialised: stream, boundaries, rho/u */
};

/* fsi/lbm3d.py:193-203 + 219-267 (initialize: static_init, init_fields):
 * neighbour masks, rho = rho0, u = u0 (0 on solid), both population
 * buffers = feq(rho0, u) (the materialised state), face velocities in slot 0. */
int fsi_lbm_init(const fsi_lbm_params* p, const fsi_lbm_buffers* b, void* pop_a, void* pop_b,
                 void* stream);

/* The lattice update of fsi/lbm3d.py:427-433 (collide 314-333, stream
 * 352-361 with half-way bounce-back, apply_boundary 363-407, update_macro
 * 409-425) as one pull-scheme kernel in the given phase: pop_in -> pop_out.
 * A full step from the post-collision state is STREAM_COLLIDE: each cell
 * gathers its streamed populations, applies the x-face boundaries, forms
 * rho/u with `force` (the force of the previous collision, unchanged), and
 * collides with `force`. When the force changed since the last step the
 * caller first runs STREAM_MACRO (writes the post-stream populations and
 * rho/u) and then COLLIDE, which collides with the new force from the
 * stored rho/u - exactly the reference's order. face_slot (0/1) selects the
 * half of b->face_vel holding the previous face velocities; phases that form
 * rho/u write the other half. */
int fsi_lbm_step(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const void* pop_in,
                 void* pop_out, const fsi_lbm_force* force, int32_t phase, int32_t face_slot,
                 void* stream);

/* fsi/lbm3d.py:499-534. op: 0 clear, 1 set (fluid = value, solid = 0; used by
 * reset_force and set_uniform_force), 2 add (fluid += value, solid = 0),
 * 3 zero on solid. `src` (3N real, may equal b->force) is read by op 2/3. */
int fsi_lbm_force_op(const fsi_lbm_params* p, const fsi_lbm_buffers* b, int32_t op,
                     const double* host_value, const void* src, void* dst, void* stream);

/* fsi/lbm3d.py:475-478 (set_force_from_numpy): AoS (N,3) float32 -> SoA
 * force, zero on solid. */
int fsi_lbm_force_import(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const float* aos,
                         void* stream);

/* fsi/lbm3d.py:444-445 (distribution_numpy): populations as the reference's
 * absolute values, AoS (N,19), real. b: the lattice buffers (I16 storage reads
 * the scale of the current write epoch from b->store_ctrl; may be null for the
 * other storages). */
int fsi_lbm_export_populations(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const void* pop,
                               void* out_aos, void* stream);

/* SoA (ncomp, n) -> AoS (n, ncomp) transposition for the *_numpy accessors. */
int fsi_soa_to_aos(int32_t real, int64_t n, int32_t ncomp, const void* soa, void* aos,
                   void* stream);
int fsi_aos_to_soa(int32_t real, int64_t n, int32_t ncomp, const void* aos, void* soa,
                   void* stream);

/* -------------------------------------------------------------- MPM ---- */

typedef struct fsi_mpm_params {
  int32_t nx, ny, nz;
  int32_t real;            /* FSI_REAL_* */
  int32_t num_particles;
  int32_t capacity;        /* max_particles: stride of the particle SoA arrays */
  double dx, inv_dx, dt;
  double two_mu, lam;      /* 2*mu and lambda, fsi/mpm3d.py:33-38 */
  double gravity[3];
  int32_t boundary_width;
  double boundary_damping;
  int32_t apic;
  double clamp_lower, clamp_upper[3];
  int32_t num_chunks;      /* entries of fsi_mpm_buffers.chunks (0: P2G takes the
                              particles in slot order, 256 per workgroup) */
  int32_t split_max_chunks; /* FP32 with a work list of at most this many chunks:
                              each particle gets three wavefronts, one per x plane
                              of its 27-node stencil, in the P2G and G2P passes
                              (small particle sets, whose passes are bound by one
                              wavefront's instruction latency); results differ
                              only in the order of FP32 sums; 0 disables */
  int32_t wide_max_workgroups; /* particle passes of at most this many workgroups gather
                              all 27 stencil nodes in one batch (one memory round trip,
                              more registers: a launch that fits on the chip at once);
                              larger ones gather one x plane of 9 nodes at a time and the
                              FP32 fused G2P+P2G kernel is held to 128 VGPRs. Results are
                              the same up to the order of float atomics; 0: always the
                              plane-batch kernels */
} fsi_mpm_params;

typedef struct fsi_mpm_buffers {
  void* x;                 /* 3*cap real */
  void* v;                 /* 3*cap real */
  void* C;                 /* 9*cap real, row-major 3x3 per particle */
  void* F;                 /* 9*cap real */
  const void* mass;        /* cap real */
  const void* volume;      /* cap real */
  const int32_t* active;   /* cap */
  void* pforce;            /* 3*cap real */
  void* grid;              /* 4N real, AoS (momentum/velocity xyz, mass) per node */
  void* grid_alt;          /* optional second 4N real grid (may be NULL): with it,
                              fsi_mpm_substeps alternates the two grids and each
                              G2P pass zeroes the one the next P2G fills, so a
                              substep needs no separate clear launch */
  const int32_t* chunks;   /* optional P2G work list from fsi_mpm_sort: (first slot,
                              count <= 256) pairs, one workgroup each; every chunk
                              lies in a 4x4x8 cell box */
  int32_t* zero_boxes;     /* optional scratch for fsi_mpm_substeps with grid_alt:
                              16 + 24 * zero_boxes_cap int32 (three passes). Each P2G
                              pass records the node box of every workgroup, and the
                              pass that clears that grid zeroes only those boxes
                              instead of the whole grid. NULL: whole-grid clears */
  int32_t zero_boxes_cap;  /* workgroups it holds per pass (>= P2G workgroups:
                              num_chunks, or ceil(num_particles / 256) unsorted) */
  void* grid_alt2;         /* optional third 4N real grid: with grid_alt, zero_boxes
                              and (when counters are kept) the coupling buffers'
                              counter rows, fsi_mpm_substeps runs G2P of substep s-1
                              and P2G of substep s in one launch (S + 1 launches for
                              S substeps) */
  const void* material;    /* optional 2*cap real: per-particle Lame parameters, 2 mu at
                              [p] and lambda at [cap + p], used by the P2G stress instead
                              of fsi_mpm_params.two_mu / lam. No counterpart in the
                              reference (one global E and nu, fsi/mpm3d.py:33-38): a
                              build-only extension for scenes of bodies of different
                              stiffness. NULL: the global parameters */
} fsi_mpm_buffers;

/* Coupling inputs/outputs read and written inside the particle passes
 * (fsi/coupling.py:332-406 and 463-496). */
typedef struct fsi_coupling_params {
  int32_t enabled;         /* run the penalty coupling in the P2G pass */
  int32_t build_fraction;  /* scatter the occupancy phi on this pass */
  double gamma_eff;        /* gamma with the ramp applied (coupling.py:321-330) */
  double dt_ratio;
  double cell_volume;
  double min_valid_weight;
  int32_t use_relative_velocity_limit;
  double relative_velocity_limit;
  int32_t use_force_limit;
  double force_limit;
  int32_t contact_enabled; /* run the contact pass after G2P */
  double contact_threshold;
  double contact_damping;
  int32_t refine;          /* MPM cells per lattice cell along each axis: 1 (the reference:
                              one co-located grid) or r > 1, an MPM grid r times finer than
                              the lattice (MPM dims = r x lattice dims, lattice cell = r x
                              MPM dx). The coupling and contact stencils then use the
                              lattice's cells; cell_volume is the lattice cell's. 0 means 1 */
} fsi_coupling_params;

typedef struct fsi_coupling_buffers {
  const int32_t* solid;    /* N lattice solid mask */
  const void* lbm_vel;     /* 3N real, macroscopic velocity of the lattice */
  void* cforce;            /* 4N real, AoS (coupling force density xyz, phi) per cell */
  void* valid_weight;      /* cap real */
  int32_t* coupling_mask;  /* cap */
  int32_t* contact_mask;   /* cap */
  int32_t* counters;       /* 8: unsupported, partial, clipped, contact, damped, ib active,
                              ib clipped, spare */
  const void* closure_force; /* optional 3*cap real (SoA): learned correction added to
                              the penalty force of every coupled particle before the
                              force-norm limit (NULL: the reference's force) */
  int32_t* counter_rows;   /* optional scratch for the fused substeps: 8 int32 per
                              substep (the last substep's row is copied to counters) */
  int32_t counter_rows_cap; /* substeps it holds */
  void* force_mean;        /* optional 3*cap real (SoA): when set, every coupled substep
                              adds its particle's coupling force times the substep's
                              fraction of the lattice step, so after a step that began
                              with it zeroed it holds the step-averaged coupling force
                              per particle slot. No counterpart in the reference (which
                              keeps only the last substep's force, fsi/coupling.py:389):
                              the learned closure's training targets use it. */
} fsi_coupling_buffers;

/* Zero the grid (fsi/mpm3d.py:205-209, clear_grid). */
int fsi_mpm_clear_grid(const fsi_mpm_params* p, const fsi_mpm_buffers* b, void* stream);

/* Particle-to-grid pass, fsi/mpm3d.py:234-266, with the penalty coupling of
 * fsi/coupling.py:332-406 fused in when c->enabled (velocity interpolation,
 * force, reaction and phi scatter) and the per-substep particle force clear
 * of fsi/mpm3d.py:211-214 when clear_force != 0. do_p2g = 0 runs only the
 * coupling part (LBMMpmCoupler.compute_coupling_forces). */
int fsi_mpm_p2g(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_params* c,
                const fsi_coupling_buffers* cb, int32_t clear_force, int32_t do_p2g, void* stream);

/* fsi/mpm3d.py:268-288 (grid_op) then 290-328 (g2p + clamp), with the contact
 * pass of fsi/coupling.py:463-496 fused after the position update when
 * c->contact_enabled. */
int fsi_mpm_grid_g2p(const fsi_mpm_params* p, const fsi_mpm_buffers* b,
                     const fsi_coupling_params* c, const fsi_coupling_buffers* cb, void* stream);

/* The substep loop of LBMMpmCoupler.step (fsi/coupling.py:147-160) and of
 * repeated MPMSolver3D.substep calls: `substeps` x (clear grid, P2G with the
 * optional fused coupling, grid update + G2P with optional contact), one call.
 * The occupancy phi is scattered on the first substep only when build_first.
 * With b->grid_alt set the grid is cleared once per call and each substep is
 * two launches (the G2P pass zeroes the other grid for the next P2G pass);
 * the last substep's P2G result is then in grid or grid_alt by parity. */
int fsi_mpm_substeps(const fsi_mpm_params* p, const fsi_mpm_buffers* b,
                     const fsi_coupling_params* c, const fsi_coupling_buffers* cb,
                     int32_t substeps, int32_t build_first, int32_t clear_force, void* stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops). Same as
 * fsi_mpm_substeps, but b->zero_boxes / b->zero_boxes_cap and
 * cb->counter_rows / cb->counter_rows_cap are ignored: the node boxes and
 * counter rows live in a static device array of the library (6.4 MB, one
 * per GPU, held until the process ends). Limits: at most 65536 particle
 * workgroups per pass (the sort's chunk count, four per chunk with the plane
 * split, or ceil(num_particles / 256) unsorted) and, when coupling or
 * contact counters are kept, at most 4096 substeps per call; larger calls
 * fail with FSI_ERR_ARGUMENT. Calls no allocating routine. Every call on one
 * GPU shares the array: calls that overlap on one GPU (other streams or
 * threads) race. */
int fsi_mpm_substeps_noalloc(const fsi_mpm_params* p, const fsi_mpm_buffers* b,
                             const fsi_coupling_params* c, const fsi_coupling_buffers* cb,
                             int32_t substeps, int32_t build_first, int32_t clear_force, void* stream);

/* Device workspace (bytes) fsi_mpm_sort needs for this grid and capacity. */
int fsi_mpm_sort_workspace(const fsi_mpm_params* p, int64_t* bytes);

/* Cell-key counting sort of the first num_particles particles (no
 * counterpart in the reference, whose particles keep their input order; it
 * changes only the order in which particles are processed). Key: the
 * truncating base cell of fsi/mpm3d.py:239 (computed as P2G computes it),
 * clamped to [-1, n] per axis, in 4x4x4-blocked order (blocks row-major,
 * an even number of blocks in z); inactive particles last. The permutation is that of a stable sort of the keys (ties keep
 * index order). Every particle array of `src` (with the material plane, when
 * src->material is set; dst->material must then be set too) is gathered into `dst` in
 * sorted order and ids_dst[t] = ids_src[perm[t]]; keys_out / perm_out (n
 * each, optional) receive the sorted keys and the permutation. With
 * `chunks` (2 * max_chunks int32, optional) it also writes the P2G work list
 * of the sorted order: the particles of each pair of z-adjacent 4x4x4 cell
 * blocks (and the inactive ones) cut into chunks of at most 256, and their
 * number to the
 * device int *chunk_count (at most num_particles/256 + occupied blocks + 1;
 * an error when max_chunks is smaller). */
int fsi_mpm_sort(const fsi_mpm_params* p, const fsi_mpm_buffers* src, const fsi_mpm_buffers* dst,
                 const int32_t* ids_src, int32_t* ids_dst, void* workspace, int64_t workspace_bytes,
                 uint32_t* keys_out, int32_t* perm_out, int32_t* chunks, int32_t max_chunks,
                 int32_t* chunk_count, void* stream);

/* fsi/coupling.py:463-496 alone (used by compute_coupling_forces). */
int fsi_contact(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_params* c,
                const fsi_coupling_buffers* cb, void* stream);

/* Learned-closure input features of the first num_particles particles, from
 * the lattice velocity and solid mask and the occupancy phi of the last
 * coupled step (slot 3 of cb->cforce): X[p * 18 + k] (float), zero rows for
 * inactive particles. k: 0-2 u - v (u interpolated over the valid nodes as
 * the coupling does, coupling.py:352-372), 3-11 grad u (row-major, over the
 * valid nodes), 12 phi, 13-15 -grad(phi)/|grad(phi)| (0 if |grad phi| <
 * 1e-6), 16 the valid support weight, 17 |u - v|. The lattice buffers have
 * the MPM grid's dims (co-located grids: the closure is not offered with
 * fsi_coupling_params.refine > 1). No counterpart in the reference. */
int fsi_closure_features(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const fsi_coupling_buffers* cb,
                         float* X, void* stream);

/* Rotation factor R = U V^T of F = U S V^T per particle (fsi/mpm3d.py:245-246)
 * for n matrices; exposed for tests of the in-kernel SVD. */
int fsi_polar_rotation(int32_t real, int32_t n, const void* F, void* R, void* stream);

/* --------------------------------------------------------- coupling ---- */

typedef struct fsi_force_assembly {
  int32_t coupling;        /* add the coupling force (coupling.py:413-419) */
  int32_t clamp_fraction;  /* phi = min(phi, 1) (coupling.py:408-411) */
  int32_t ib_enabled;      /* immersed-boundary drag (coupling.py:421-461) */
  double ib_drag, ib_threshold, ib_max_force;
  int32_t ib_use_max;
  double base_force[3];    /* LBMConfig.force: the value of reset_force */
} fsi_force_assembly;

/* LBMMpmCoupler.step tail, fsi/coupling.py:162-170: clamp phi, reset the
 * lattice force, add the coupling force, compute and add the IB drag, all in
 * one pass over the cells. ib_force (3N real) and ib_total (3 real, summed
 * with atomics) may be null when IB is off. */
int fsi_force_assemble(const fsi_lbm_params* lp, const fsi_lbm_buffers* lb,
                       const fsi_force_assembly* a, void* cforce, void* ib_force, void* ib_total,
                       int32_t* counters, int32_t write_force, void* stream);

/* LBMMpmCoupler.step tail fused with the next lattice collision,
 * fsi/coupling.py:162-170 (clamp phi, reset the force, add the coupling
 * force, compute and add the IB drag) + fsi/lbm3d.py:314-333 (collide): from
 * the materialised state (pop_in, b->rho, b->vel) each cell assembles its
 * force from cforce (and, with IB, rho and u), writes it to b->force (and the
 * drag to ib_force, summed into ib_total, counted in counters[5..6]) and
 * collides with it into pop_out, the post-collision state of
 * fsi_lbm_step. One pass instead of fsi_force_assemble + the collide
 * phase. I16 storage: a measuring pass (store_pass 1) writes and counts
 * nothing. */
int fsi_lbm_collide_assembled(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const fsi_force_assembly* a,
                              const void* pop_in, void* pop_out, void* cforce, void* ib_force, void* ib_total,
                              int32_t* counters, void* stream);

/* ------------------------------------------------------- reductions ---- */

/* doubles per workgroup row of a reduction workspace */
#define FSI_SUM_SLOTS 16

/* FP64 sums behind LBMSolver3D.total_mass (fsi/lbm3d.py:449-452),
 * max_velocity_norm (454-456) and LBMMpmCoupler.total_fluid_coupling_force
 * (fsi/coupling.py:241-244), in one pass: out[0] = sum of rho over fluid
 * cells, out[1] = max |u|^2 over all cells, out[2..4] = coupling force summed
 * over fluid cells (cforce: the 4N real AoS coupling buffer, or null for 0).
 * workspace: device doubles, at least FSI_SUM_SLOTS (up to 1024 rows of
 * FSI_SUM_SLOTS are used). Fixed grid and summation order: the result is
 * reproducible bit for bit. out: 5 device doubles. */
int fsi_reduce_lattice(const fsi_lbm_params* p, const fsi_lbm_buffers* b, const void* cforce,
                       void* workspace, int64_t workspace_doubles, double* out, void* stream);

/* FP64 sums behind MPMSolver3D.center_of_mass (fsi/mpm3d.py:353-362),
 * max_velocity_norm (364-368), LBMMpmCoupler.total_particle_coupling_force
 * (fsi/coupling.py:238-239) and the valid-weight statistics of
 * coupling_diagnostics (263-270), over the first num_particles particles:
 * out[0] sum m, out[1..3] sum m x, out[4] max |v|^2, out[5..7] particle force
 * summed, out[8] min and out[9] sum of valid_weight (cap real, or null).
 * out: 10 device doubles. Workspace as for fsi_reduce_lattice. */
int fsi_reduce_particles(const fsi_mpm_params* p, const fsi_mpm_buffers* b, const void* valid_weight,
                         void* workspace, int64_t workspace_doubles, double* out, void* stream);

/* -------------------------------------------------- learned closure ---- */

/* No counterpart in the reference (a learned correction to the penalty
 * coupling force of fsi/coupling.py:333-406). An MLP
 * 18 -> 64 -> 64 -> 3 with GELU, BF16 weights and activations, FP32
 * accumulation on the matrix cores. Parameters: one FP32 vector of
 * FSI_MLP_PARAMS floats, W1[64][32] (input columns 18..31 zero), b1[64],
 * W2[64][64], b2[64], W3[16][64] (rows 3..15 zero), b3[16]; layer l computes
 * z = W h + b. */
#define FSI_MLP_IN 18
#define FSI_MLP_HIDDEN 64
#define FSI_MLP_OUT 3
#define FSI_MLP_PARAMS 7312
/* normalisation block of fsi_mlp_infer: mean[18], inv_std[18], lo[18],
 * hi[18] (envelope, normalised units), out_mean[3], out_std[3], y_lo[3],
 * y_hi[3] (bounds of the normalised output: the training targets' range) */
#define FSI_MLP_NORM 84

/* One training batch: forward, loss = mean over rows and the 3 outputs of
 * (y - T)^2, backward. Rows idx[0..n) of X (n x 18 normalised features,
 * row-major; idx NULL: rows 0..n) against T (rows x 3). Writes
 * dLoss/dparams to grads (FSI_MLP_PARAMS floats; each wavefront's partial
 * sums go to `workspace`, then are added in a fixed order, so a batch gives
 * the same gradient bit for bit) and adds the squared-error sum into
 * *loss_sum. workspace: at least fsi_mlp_train_workspace(n) floats. */
int fsi_mlp_train_workspace(int32_t n, int64_t* floats);
int fsi_mlp_train_step(int32_t n, const float* X, const float* T, const int32_t* idx, const float* params,
                       float* grads, double* loss_sum, float* workspace, int64_t workspace_floats, void* stream);
/* Adam update of params with moments m, v (FSI_MLP_PARAMS each); step >= 1 */
int fsi_mlp_adam(float* params, const float* grads, float* m, float* v, int32_t step, float lr, float beta1,
                 float beta2, float eps, void* stream);
/* Network outputs for n rows of normalised features: Y (n x 3) */
int fsi_mlp_forward(int32_t n, const float* X, const float* params, float* Y, void* stream);
/* Closure forces for n rows of raw features: out[a * stride + r] =
 * out_mean[a] + out_std[a] * y[a], or 0 when a normalised feature of row r
 * lies outside [lo, hi] (counted into *fallbacks, a device int). */
int fsi_mlp_infer(int32_t n, const float* X, const float* params, const float* norm, float* out, int64_t stride,
                  int32_t* fallbacks, void* stream);

/* ----------------------------------------------------------- common ---- */

int fsi_abi_version(void);
const char* fsi_last_error(void);
/* fills a real buffer with zeros (a kernel, usable inside stream capture) */
int fsi_fill_zero(int64_t nbytes, void* dst, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FSI_HIP_H_ */
