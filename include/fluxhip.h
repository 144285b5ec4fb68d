/*
 * libfluxhip: C-ABI of the MI355X (gfx950) dense smoke/fire step.
 *
 * Every entry point takes plain integers, floats and pointers, returns 0 on
 * success or a negative FX_E* code, and leaves a message for fx_last_error().
 * Device pointers come from the caller (torch tensors own the memory), and the
 * dense-step, solver and network entry points allocate no device memory.  The
 * one exception is the brick pool handle (fx_bricks_*): it owns its device
 * mirror and allocates it in fx_bricks_sync (on capacity growth, which also
 * synchronises the device) and fx_bricks_transport (its second payload buffer
 * and brick map, on first use or growth); and the brick comparisons
 * (fx_projection_compare, fx_coupled_compare, fx_transport_compare,
 * fx_mac_compare), diagnostics that allocate both of their engines within a
 * byte budget and free them before returning.  `stream` is a hipStream_t
 * (NULL = default).
 *
 * Layout everywhere is the reference's flat x-fast index x + nx*(y + ny*z)
 * (physics/reference.py:3).  Face arrays are U (nx+1,ny,nz), V (nx,ny+1,nz),
 * W (nx,ny,nz+1) (physics/config.py:21-23).  The vertical (buoyancy) axis is z.
 *
 * Each declaration cites the reference interface it replaces.  In the
 * reference the boundary is a GLSL dispatch (backend/device.py:57-74) issued by
 * the Python solver classes; here it is one C call per fused pass.
 */
#ifndef FLUXHIP_H
#define FLUXHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Size in doubles of the FP64 reduction workspace `ws` taken by the norm and
 * solver entry points: 16 finalised scalars followed by 1024 per-block
 * partials for each of them (reductions use no atomics). */
#define FX_WS_DOUBLES 16400

enum {
  FX_OK = 0,
  FX_EINVAL = -1,   /* invalid argument  -> Python ValueError   */
  FX_EDEVICE = -2,  /* HIP runtime error -> GPUUnavailable      */
  FX_ESTATE = -3    /* misuse of a handle -> RuntimeError       */
};

/* Grid descriptor: cell counts and domain extent in metres
 * (physics/config.py:6-23 GridSpec.shape/extent). */
typedef struct fx_grid {
  int32_t nx, ny, nz;
  float ex, ey, ez;
} fx_grid;

/* Packed source table: `count` rows of 16 floats in the row order of
 * physics/emission.py:49-58 (centre xyz, radius, start xyz, coupling,
 * velocity xyz, density rate, heat rate, target flag, solid-profile flag,
 * fuel rate).  `multi` selects the source-table semantics of
 * shaders/multi_advect_*.glsl (density clamped to >= the max target);
 * 0 selects the single-emitter uniforms of shaders/closed_advect_scalar.glsl and
 * shaders/emitting_advect_velocity.glsl. */
typedef struct fx_sources {
  int32_t count;
  int32_t multi;
  float table[128];
} fx_sources;

const char* fx_last_error(void);
int fx_version(void);

/* --------------------------------------------------------------------------
 * masks and flags
 * ------------------------------------------------------------------------ */

/* Union one analytic collider into a uint8 solid mask (1 = solid).
 * kind 0 = sphere, 1 = box; rows = 3x4 inverse affine rows.
 * Replaces SolidDevice mask construction (backend/collision.py:24-36,
 * shaders/solid_mask.glsl:1-9). */
int fx_solid_raster(fx_grid g, int32_t kind, const float* rows12, uint8_t* solid, void* stream);

/* Per-cell open-face bits of the level-0 pressure operator (bit 2a: low face of
 * axis a open, bit 2a+1: high face open, bit 6: fluid cell).  solid may be NULL
 * (closed box).  Encodes shaders/solid_boundary.glsl:2-10 solidCell/solidFace. */
int fx_build_flags(fx_grid g, const uint8_t* solid, uint8_t* flags, void* stream);

/* out = solid ? 0 : in (shaders/solid_clear.glsl:1-6; backend/collision.py:65-66). */
int fx_solid_clear(fx_grid g, const uint8_t* solid, const float* in, float* out, void* stream);

/* Initial fields of reset(): kind 0 writes the cell field scale*w^2 of the soft
 * sphere (shaders/seed.glsl, temperature_seed.glsl); kind 1..3 writes face
 * component kind-1 of base + rotation about z (shaders/mac_seed.glsl).
 * params9 = centre xyz, radius, scale, base velocity xyz, angular speed. */
int fx_seed(fx_grid g, int32_t kind, const float* params9, float* out, void* stream);

/* out[i] = value (graph-safe clear; replaces texture.clear). */
int fx_fill(float* out, int64_t n, float value, void* stream);

/* out[i] = factor * in[i] (shaders/pressure_scale.glsl:1-6). */
int fx_scale(const float* in, float* out, int64_t n, float factor, void* stream);

/* --------------------------------------------------------------------------
 * projection (backend/pressure.py:49-98)
 * ------------------------------------------------------------------------ */

/* Zero normal velocity on wall and solid faces, copy the rest
 * (PressureProjector.enforce_walls, backend/pressure.py:49-51,
 * shaders/pressure_walls.glsl:1-15).  flags NULL = closed box. */
int fx_enforce_walls(fx_grid g, const uint8_t* flags, const float* u, const float* v, const float* w,
                     float* uo, float* vo, float* wo, void* stream);

/* div = sum_a (u_a[c+e_a]-u_a[c])/h_a with wall/solid faces read as 0; 0 in
 * solid cells (compute_divergence, backend/pressure.py:53-55,
 * shaders/pressure_divergence.glsl:1-12, applied to enforce_walls' output). */
int fx_divergence(fx_grid g, const uint8_t* flags, const float* u, const float* v, const float* w,
                  float* div, void* stream);

/* One weighted-Jacobi sweep of L p = rhs*inv_dt with Neumann walls and solids
 * (shaders/pressure_jacobi.glsl:2-34; backend/pressure.py:95-98). */
int fx_jacobi_sweep(fx_grid g, const uint8_t* flags, const float* p, const float* rhs, float* out,
                    float inv_dt, float omega, void* stream);

/* Open faces: out = in - scale*(p[c]-p[c-e])/h; wall/solid faces 0
 * (backend/pressure.py:76-80, shaders/pressure_gradient.glsl:1-21 fused with
 * pressure_walls.glsl so the bounded copies need not exist). */
int fx_gradient(fx_grid g, const uint8_t* flags, const float* p, const float* u, const float* v,
                const float* w, float* uo, float* vo, float* wo, float scale, void* stream);

/* As fx_gradient with the pressure split p + dp: the neighbour differences of
 * both parts are formed before they are added, so an impulse carried as
 * warm start p plus solver increment dp keeps its full accuracy in fp32.
 * psum (may be NULL, must not alias p or dp) receives p + dp in the same pass. */
int fx_gradient_split(fx_grid g, const uint8_t* flags, const float* p, const float* dp, const float* u,
                      const float* v, const float* w, float* uo, float* vo, float* wo, float* psum, float scale,
                      void* stream);


/* Moving colliders (MovingSolidDevice, backend/moving_collision.py:26-41 with
 * shaders/moving_boundary.glsl): as fx_enforce_walls / fx_divergence /
 * fx_gradient, but every closed face (domain wall or solid) reads or is set to
 * the wall velocity bu/bv/bw of fx_wall_velocity instead of 0. */
int fx_enforce_walls_moving(fx_grid g, const uint8_t* flags, const float* u, const float* v, const float* w,
                            const float* bu, const float* bv, const float* bw, float* uo, float* vo, float* wo,
                            void* stream);
int fx_divergence_moving(fx_grid g, const uint8_t* flags, const float* u, const float* v, const float* w,
                         const float* bu, const float* bv, const float* bw, float* div, void* stream);

/* fx_divergence (bu/bv/bw NULL) or fx_divergence_moving, plus vmax[0..2] =
 * max |U|, |V|, |W| over every face of the three arrays (device floats,
 * non-finite values as 3.4e38, as fx_max_abs): the velocity bound of the
 * active-brick plans taken from the loads the divergence makes anyway (no
 * reference counterpart).  partials: 3 * ceil(nx*ny*nz / 1024) floats of
 * scratch. */
int fx_divergence_bound(fx_grid g, const uint8_t* flags, const float* u, const float* v, const float* w,
                        const float* bu, const float* bv, const float* bw, float* div, float* partials, float* vmax,
                        void* stream);
int fx_gradient_moving(fx_grid g, const uint8_t* flags, const float* p, const float* u, const float* v,
                       const float* w, const float* bu, const float* bv, const float* bw, float* uo, float* vo,
                       float* wo, float scale, void* stream);

/* fx_gradient_split with moving walls: the gradient of p + dp on open faces,
 * the wall velocity on closed ones, and psum = p + dp (nullable).  The
 * tolerance projector's correction for moving colliders (no reference
 * counterpart: the reference projects moving colliders with masked Jacobi,
 * backend/projected.py:21-28). */
int fx_gradient_split_moving(fx_grid g, const uint8_t* flags, const float* p, const float* dp, const float* u,
                             const float* v, const float* w, const float* bu, const float* bv, const float* bw,
                             float* uo, float* vo, float* wo, float* psum, float scale, void* stream);

/* Owner raster of moving collider `owner` (1..8): mask[c] = owner where the
 * unit sphere/box under the inverse rows r (3 x 4) covers the cell centre and
 * no earlier collider owns the cell (MovingSolidDevice.update,
 * backend/moving_collision.py:53-71, shaders/moving_mask.glsl:1-11). */
int fx_owner_raster(fx_grid g, int32_t kind, const float* r, int32_t owner, uint8_t* mask, void* stream);

/* Rigid wall velocity on interior faces next to solid cells: rows (host,
 * count x 3 x 4 floats) are each collider's velocity_rows; the low cell's
 * owner, else the high cell's, is evaluated at the face centre; 0 elsewhere
 * (shaders/moving_wall.glsl:1-19). */
int fx_wall_velocity(fx_grid g, const uint8_t* owner, int32_t count, const float* rows, float* bu, float* bv,
                     float* bw, void* stream);

/* Cells solid now or before start empty (MovingSolidDevice.remap,
 * backend/moving_collision.py:73-83, shaders/moving_remap_scalar.glsl). */
int fx_remap_scalar(fx_grid g, const uint8_t* solid, const uint8_t* old_solid, const float* in, float* out,
                    void* stream);

/* Faces blocked now take the wall velocity, faces blocked before start at
 * rest, others keep their value, domain walls 0 (shaders/moving_remap_face.glsl). */
int fx_remap_faces(fx_grid g, const uint8_t* solid, const uint8_t* old_solid, const float* bu, const float* bv,
                   const float* bw, const float* u, const float* v, const float* w, float* uo, float* vo, float* wo,
                   void* stream);

/* The update and remap above restricted to a box of cells, for colliders that
 * are all moving (every solid cell before and after the move lies inside the
 * box): box = {x0, y0, z0, x1, y1, z1}, half-open cell ranges inside the grid,
 * covering the colliders' old and new poses.  fx_moving_update_box clears the
 * owners in the box and rasters the `count` colliders (kinds 0 sphere, 1 box;
 * 3 x 4 inverse rows each) in order, then rebuilds the open-face flags one
 * cell wider, the clearance (fx_build_clearance) 15 cells wider through the
 * two distinct scratch arrays, and the wall velocity (3 x 4 velocity rows per
 * collider) on the box's faces; outside those ranges none of them can change,
 * so the results equal the whole-grid passes.  fx_remap_box empties, in place,
 * the box's cells solid now or before in up to three scalar fields (s1, s2
 * nullable) and applies fx_remap_faces' rule to the box's faces of u, v, w in
 * place (MovingSolidDevice.update/remap, backend/moving_collision.py:53-83,
 * without the whole-grid passes). */
int fx_moving_update_box(fx_grid g, const int32_t* box, int32_t count, const int32_t* kinds,
                         const float* inverse_rows, const float* wall_rows, uint8_t* owner, uint8_t* flags,
                         uint8_t* clearance, uint8_t* scratch, uint8_t* scratch2, float* bu, float* bv, float* bw,
                         void* stream);
int fx_remap_box(fx_grid g, const int32_t* box, const uint8_t* solid, const uint8_t* old_solid, const float* bu,
                 const float* bv, const float* bw, float* u, float* v, float* w, float* s0, float* s1, float* s2,
                 void* stream);

/* out = a + b (n floats). */
int fx_combine(const float* a, const float* b, float* out, int64_t n, void* stream);

/* --------------------------------------------------------------------------
 * multigrid engine (backend/multigrid.py:20-95, backend/solid_multigrid.py:7-76)
 * ------------------------------------------------------------------------ */
enum { FX_F32 = 0, FX_BF16 = 1, FX_F16 = 2 };
/* Level operators.  BOX: closed box, no data.  FLAGS: uint8 open-face bits per
 * cell (bit 2a low face of axis a open, 2a+1 high face, bit 6 fluid).  EDGES:
 * float4 (wx, wy, wz, active) conductances to the +x/+y/+z neighbour.  PACKED:
 * uint32 per cell, bits [0,10) kx, [10,20) ky, [20,30) kz, bit 30 active, the
 * conductance to the +a neighbour being k_a * level.w[a] (rediscretised masked
 * levels: fractions k/4^l of open fine faces, exact in 10 bits up to l = 5). */
enum { FX_OP_BOX = 0, FX_OP_FLAGS = 1, FX_OP_EDGES = 2, FX_OP_PACKED = 3 };
enum { FX_RESTRICT_MEAN = 0, FX_RESTRICT_FULL = 1 };
enum { FX_PROLONG_TRILINEAR = 0, FX_PROLONG_RENORM = 1, FX_PROLONG_CONSTANT = 2 };

/* One level.  op BOX: Neumann box with weights w[a] = (n_a/extent_a)^2;
 * FLAGS: op_data = uint8 open-face bits (fx_build_flags), weights w[a];
 * EDGES: op_data = float4 (w+x, w+y, w+z, active) absolute conductances to the
 * positive neighbour (the RGBA32F edge records of physics/solid_hierarchy.py:11-47).
 * p/q are the ping-pong solution arrays, rhs the right-hand side, all in the
 * config's storage type. */
typedef struct fx_mg_level {
  int32_t nx, ny, nz, op;
  float w[3];
  float pad_;
  const void* op_data;
  void* p;
  void* q;
  void* rhs;
  const uint8_t* active;  /* optional, EDGES/PACKED levels (fast renormalised prolongation): bit 0 = active
                            cell; bit 1 (may be left 0) = the cell and its 26 neighbours, indices clamped
                            to the level, are all active */
} fx_mg_level;

typedef struct fx_mg_config {
  int32_t levels;
  int32_t storage;        /* FX_F32 / FX_BF16 / FX_F16 */
  int32_t pre, post;      /* smoothing sweeps per level */
  float omega;            /* Jacobi relaxation */
  int32_t restriction;    /* FX_RESTRICT_* */
  int32_t prolongation;   /* FX_PROLONG_* */
  int32_t coarse_sweeps;  /* > 0: smooth the coarsest level this many times; 0: direct (coarse_matrix);
                            < 0: FP64 conjugate gradients in one workgroup (fx_mg_coarse_cg) */
  const double* coarse_matrix; /* device, m x m FP64 pseudo-inverse, m = cells of coarsest */
  float coarse_scale;     /* direct: p = coarse_scale * M rhs */
  int32_t separate_legs;  /* 0: on levels below the finest, each V-cycle leg (the pre-sweeps from zero with
                            the restriction, the prolongation with the post-sweeps) runs as one fused launch
                            when pre = post = 2, the restriction is the 8-child mean and the prolongation
                            trilinear or renormalised (same results); 1: always separate launches */
  fx_mg_level level[16];
} fx_mg_config;

/* One V-cycle on level 0 (MultigridPressureProjector.cycle,
 * backend/multigrid.py:57-81; SolidMultigridProjector.cycle,
 * backend/solid_multigrid.py:46-64).  inv_dt scales the level-0 rhs (coarse
 * levels use 1, as the reference recurses with dt=1).  Ping-pong swaps are
 * written back into cfg->level[i].p/q. */
int fx_mg_vcycle(fx_mg_config* cfg, float inv_dt, void* stream);

/* `count` smoothing sweeps of one level (MultigridPressureProjector.smooth,
 * backend/multigrid.py:50-55).  Swaps are written back into the level. */
int fx_mg_smooth(fx_mg_config* cfg, int32_t index, int32_t count, float inv_dt, void* stream);

/* r = rhs*inv_dt - L p on one level (shaders/mg_residual.glsl:1-13). */
int fx_mg_residual(const fx_mg_config* cfg, int32_t index, float inv_dt, void* out, void* stream);

/* coarse rhs = restriction of fine residual, fused (mg_residual + mg_restrict). */
int fx_mg_residual_restrict(fx_mg_config* cfg, int32_t index, float inv_dt, void* stream);

/* coarse rhs = restriction of an arbitrary fine array (shaders/mg_restrict.glsl:2-9). */
int fx_mg_restrict(const fx_mg_config* cfg, int32_t index, const void* fine, void* coarse, void* stream);

/* fine out = fine in + prolongation(coarse) (shaders/mg_prolong.glsl:2-15,
 * shaders/solid_mg_prolong.glsl:1-7); fine_in NULL: a zero fine input (out
 * = the prolonged correction, inactive cells 0). */
int fx_mg_prolong(const fx_mg_config* cfg, int32_t index, const void* fine_in, const void* coarse,
                  void* fine_out, void* stream);

/* Coarsest direct solve p = scale * M rhs (shaders/coarse_pressure.glsl:1-12,
 * shaders/solid_mg_direct.glsl:1-12), FP64 accumulation. */
int fx_mg_direct(const fx_mg_config* cfg, void* stream);

/* Coarsest-level solve by FP64 conjugate gradients in one workgroup (coarsest
 * level of at most 1024 cells), what the V-cycle runs when coarse_sweeps < 0:
 * A p = coarse_scale * rhs from p = 0 to ||r|| <= 1e-7 ||b|| (the
 * pseudo-inverse's solution for a consistent rhs), for operators rebuilt every
 * step, where no pseudo-inverse is kept.  info (nullable, device, 2 doubles):
 * iterations taken and the final ||r|| / ||b||.  No reference counterpart. */
int fx_mg_coarse_cg(const fx_mg_config* cfg, double* info, void* stream);

/* Rebuild levels 1.. of a rediscretised masked hierarchy on the device from
 * level 0's open-face flags (FX_OP_FLAGS, from fx_build_flags): per level the
 * packed words (FX_OP_PACKED, levels 1-4) or float4 conductances
 * (FX_OP_EDGES), and the activity bytes, into cfg->level[l].op_data and
 * .active (writable device arrays of the level's cell count, l >= 1).  A coarse
 * face conducts the mean open fraction of the four finer faces it covers and a
 * coarse cell is active when a child is, the same integers and roundings as
 * the host construction of static masks.  ex/ey/ez: the domain extent.
 * coarse_solid (nullable): level 1's solid mask (1 where all eight children
 * are solid).  For moving colliders (no reference counterpart: the reference
 * projects them with masked Jacobi, backend/projected.py:21-28). */
int fx_mg_rebuild_masked(const fx_mg_config* cfg, double ex, double ey, double ez, uint8_t* coarse_solid,
                         void* stream);

/* Solve info returned by the tolerance-driven solvers (no reference
 * counterpart: the reference runs fixed cycle budgets). */
typedef struct fx_solve_info {
  int32_t iterations;   /* V-cycles or Krylov iterations taken */
  int32_t converged;    /* 1 if residual <= tol * bnorm */
  int32_t fallback;     /* learned FPCG: 1 if plain MG finished the solve */
  int32_t pad_;
  double bnorm;         /* ||b||_2 after mean projection */
  double residual0;     /* ||b - A x0||_2 */
  double residual;      /* final true ||b - A x||_2 (FP64 norm of the FP32 residual) */
} fx_solve_info;

/* Controls of the tolerance-driven solvers. */
typedef struct fx_solve_params {
  float tol;           /* stop when ||b - L x|| <= tol * reference norm */
  int32_t max_iters;   /* V-cycles / refinement steps / CG iterations */
  int32_t check_from;  /* iterations run before the first host-side convergence check */
  int32_t ref_q;       /* ws[ref_q] holds the squared reference norm; -1: ||b|| of this solve */
  int32_t x_is_zero;   /* the start iterate is zero (skips reading it in the first sweep) */
  int32_t pad_;
} fx_solve_params;

/* r0 = b - L x0 (neighbour differences first): the right-hand side of the
 * increment equation L d = r0 around a warm start x0.  With ws != NULL also
 * ws[6] = ||r0||^2 and ws[7] = ||b||^2 (FP64, device side, no sync), and a
 * start no better than zero (||r0|| >= ||b||) is dropped on the device:
 * x0 = 0, r0 = b, ws[6] = ws[7]. */
int fx_increment_rhs(fx_grid g, const uint8_t* flags, float* x0, const float* b, float* r0, double* ws,
                     void* stream);

/* Remove the fluid-cell mean from rhs (Neumann compatibility), FP64 sums.
 * ws: FX_WS_DOUBLES doubles of device scratch. */
int fx_project_mean(fx_grid g, const uint8_t* flags, float* rhs, double* ws, void* stream);

/* fx_project_mean without the ws argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid allocation).  No counterpart in the reference.  It allocates nothing:
 * its FP64 workspace is a static device array of FX_WS_DOUBLES doubles
 * (131 KB) per GPU, held until the process ends, which covers every grid the
 * library accepts (reductions fold into 1024 block partials).  Calls that
 * overlap on one GPU share that array and race; results equal fx_project_mean
 * bit for bit. */
int fx_project_mean_noalloc(fx_grid g, const uint8_t* flags, float* rhs, void* stream);

/* ||b - A x||_2 and ||b||_2 in FP64 for the level-0 operator (flags NULL = box).
 * Synchronises the stream; writes host doubles. ws: FX_WS_DOUBLES doubles. */
int fx_residual_norm(fx_grid g, const uint8_t* flags, const float* x, const float* b,
                     double* rnorm, double* bnorm, double* ws, void* stream);

/* Stationary V-cycles on level 0 of cfg (storage FX_F32) until the true
 * residual meets prm (x = level[0].p, b = level[0].rhs).  ws: FX_WS_DOUBLES
 * doubles.  Synchronises once per cycle from prm->check_from on. */
int fx_mg_solve(fx_mg_config* cfg, const fx_solve_params* prm, double* ws, fx_solve_info* info, void* stream);

/* Mixed-precision iterative refinement on level 0 of ecfg (storage BF16/FP16,
 * its own level-0 p/q/rhs hold the error and the 16-bit residual): x (fp32)
 * is refined until ||b - L x|| <= tol*||b||; the fp32 residual is recomputed
 * from x each step with the update fused in.  *x_out receives x or x_spare,
 * whichever holds the result.  ws: FX_WS_DOUBLES doubles. */
int fx_mixed_solve(fx_mg_config* ecfg, float* x, float* x_spare, const float* b, const fx_solve_params* prm,
                   double* ws, fx_solve_info* info, float** x_out, void* stream);

/* Flexible preconditioned CG (Polak-Ribiere beta, restarted from the true
 * residual if the recurrence drifts) on L x = b with one V-cycle of pcfg as
 * preconditioner.  r, p, Ap: fp32 scratch; zold: scratch in pcfg's storage.
 * ws: FX_WS_DOUBLES doubles. */
int fx_pcg_solve(fx_mg_config* pcfg, float* x, const float* b, float* r, float* p, float* Ap, void* zold,
                 const fx_solve_params* prm, double* ws, fx_solve_info* info, void* stream);

/* --------------------------------------------------------------------------
 * transport (backend/detail.py:16-37, backend/projected.py:150-195)
 * ------------------------------------------------------------------------ */

/* Chebyshev distance (cells; 255 = more than 15) from every cell to the
 * nearest solid or out-of-range cell, by a separable three-pass transform.  The
 * transport kernels use it to skip the solid trace clip (3-D DDA of
 * shaders/solid_boundary.glsl:13-36) and the 3x3x3 halo test (:39-44) where no
 * solid is within reach; results are unchanged.  scratch: n bytes. */
int fx_build_clearance(fx_grid g, const uint8_t* solid, uint8_t* clearance, uint8_t* scratch, void* stream);

/* Interior fast path of the transport kernels (no reference counterpart; for
 * tests that compare the two paths): a wave whose sites all lie at least 4
 * cells from every solid and wall, with traces within 2 cells, skips the
 * clamps and solid tests, which are identities there; results are the same
 * bits either way.  enable = 0 sends every site through the general path, 1
 * (the default) restores the fast path.  Process-wide, read when a transport
 * entry point is called.  Returns the previous setting. */
int fx_transport_fast_paths(int32_t enable);

/* Face predictor for U, V and W (DetailPasses.transport_velocity,
 * backend/detail.py:27-37, shaders/transport_face.glsl:1-20). */
int fx_face_predict(fx_grid g, float dt, const float* u, const float* v, const float* w, float* pu,
                    float* pv, float* pw, const uint8_t* solid, const uint8_t* clearance, void* stream);
/* fx_face_predict plus, per face, the min (lo) and max (hi) of the eight donors
 * of the departure sample when the departure point may be MacCormack-corrected
 * (inside the face lattice, 3x3x3 halo fluid, trace unclipped), else
 * lo = +inf, hi = -inf.  lo/hi: 3 face arrays each. */
int fx_face_predict_bounds(fx_grid g, float dt, const float* u, const float* v, const float* w, float* const* pred,
                           float* const* lo, float* const* hi, const uint8_t* solid, const uint8_t* clearance,
                           void* stream);

typedef struct fx_velocity_args {
  fx_grid g;
  float dt;
  int32_t maccormack;
  const float* u; const float* v; const float* w;      /* old velocity */
  const float* pu; const float* pv; const float* pw;   /* face predictors (maccormack) */
  float* ou; float* ov; float* ow;                     /* new velocity (provisional W) */
  const float* density; const float* temperature;      /* pre-transport, for buoyancy; may be NULL */
  float thermal_lift; float density_weight;
  const uint8_t* solid;
  const uint8_t* clearance;                            /* fx_build_clearance of solid, or NULL */
  const fx_sources* sources;                           /* host pointer */
  /* optional (maccormack): donor bounds from fx_face_predict_bounds; the
   * predictor then doubles as the semi-Lagrangian value, so the back-trace
   * is not repeated.  NULL: full recomputation. */
  const float* lu; const float* lv; const float* lw;
  const float* hu; const float* hv; const float* hw;
} fx_velocity_args;

/* Fused velocity SL/MacCormack transport + jet relaxation + buoyancy on W
 * (backend/projected.py:155-167: emitting_advect_velocity.glsl /
 * multi_advect_velocity.glsl with correct_velocity_fused.glsl, then
 * buoyancy.glsl). */
int fx_velocity_advect(const fx_velocity_args* a, void* stream);

/* Scalar predictor for up to 3 channels (DetailPasses.transport with +dt,
 * backend/detail.py:16-25, shaders/transport_scalar.glsl:1-19). */
int fx_scalar_predict(fx_grid g, float dt, const float* u, const float* v, const float* w, int32_t channels,
                      const float* const* fields, float* const* preds, const uint8_t* solid,
                      const uint8_t* clearance, void* stream);
/* fx_scalar_predict plus, per channel and cell, the min (lo) and max (hi) of
 * the eight donors of the departure sample when the departure point may be
 * MacCormack-corrected (inside [h/2, L - h/2], halo fluid, trace unclipped),
 * else lo = +inf, hi = -inf. */
int fx_scalar_predict_bounds(fx_grid g, float dt, const float* u, const float* v, const float* w, int32_t channels,
                             const float* const* fields, float* const* preds, float* const* lo, float* const* hi,
                             const uint8_t* solid, const uint8_t* clearance, void* stream);

typedef struct fx_scalar_args {
  fx_grid g;
  float dt;
  int32_t maccormack;
  int32_t channels;                /* 1..3 */
  int32_t boundary;                /* 0 closed box (clamped indices), 1 zero extension */
  const float* u; const float* v; const float* w;     /* projected velocity */
  const float* field[3];
  const float* pred[3];            /* maccormack predictors */
  float* out[3];
  float decay[3];                  /* exp(-decay*dt) factor per channel */
  int32_t channel_id[3];           /* 0 density, 1 temperature, 2 fuel */
  const uint8_t* solid;
  const uint8_t* clearance;        /* fx_build_clearance of solid, or NULL */
  const fx_sources* sources;       /* host pointer */
  /* optional (maccormack): donor bounds from fx_scalar_predict_bounds; the
   * predictor then doubles as the semi-Lagrangian value.  NULL: recompute. */
  const float* lo[3];
  const float* hi[3];
} fx_scalar_args;

/* Fused closed-box scalar transport (reverse sampling, MacCormack correction,
 * clamp, decay, sources) for up to 3 channels (backend/projected.py:178-195,
 * shaders/closed_advect_scalar.glsl, multi_advect_scalar.glsl,
 * correct_scalar.glsl, closed_sample.glsl).  boundary = 1 gives the
 * zero-extension sampling of shaders/mac_advect_density.glsl (backend/mac.py:60-76). */
int fx_scalar_advect(const fx_scalar_args* a, void* stream);

/* Fused combustion: flame = burned/dt, fuel/density/temperature updated in one
 * pass (CombustionPasses.react, backend/combustion.py:21-31,
 * shaders/combustion_burn.glsl, combustion_apply.glsl). */
int fx_combustion(int64_t n, float dt, float ignition, float burn_rate, float heat_yield, float smoke_yield,
                  const float* fuel, const float* temperature, const float* density, float* fuel_out,
                  float* temperature_out, float* density_out, float* flame, void* stream);

/* Exact active bricks of the scalar passes.  No reference counterpart: the
 * reference's sparse engine drops density below 1e-4 (native/projection.mm);
 * these masks track exact zeros instead, so the passes below give the same
 * values as their whole-grid forms.  A mask holds one byte per 8^3 brick of
 * the cell grid (ceil(n/8) bricks per axis, x fastest).  A field's live mask
 * byte 0 guarantees every cell of that brick of the field is exactly 0.
 *
 * fx_active_scan: the live mask of a field (1 where any cell is not exactly
 * 0, NaN included). */
int fx_active_scan(fx_grid g, const float* field, uint8_t* live, void* stream);

/* Plan of the next scalar pass from the live masks of its inputs (nin, 1..4)
 * and of the buffers it writes (nout, 0..4).  Plan bytes: 0 skip the brick,
 * 1 compute it, 2 store 0 (the outputs are exactly zero there but a buffer may
 * hold other values).  transport = 1 (fx_scalar_advect_active): computed are
 * the bricks within one brick of a live input or of an emitter's reach
 * (sources: the step's source table), and predict_bricks receives the bricks
 * within one brick of those (fx_scalar_predict_bounds_active); vmax is the
 * device max |U|, |V|, |W| of the transporting velocity (fx_max_abs), and a
 * plan with dt * max|u| / h above 2.5 cells on any axis computes every brick.
 * transport = 0 (fx_combustion_active): computed are the live inputs'
 * bricks.  The output masks are cleared on every brick the pass writes; the
 * pass sets them again where it stores nonzero values. */
int fx_active_plan(fx_grid g, float dt, const float* vmax, int32_t transport, int32_t nin,
                   const uint8_t* const* live_in, int32_t nout, uint8_t* const* live_out, const fx_sources* sources,
                   uint8_t* predict_bricks, uint8_t* write_bricks, void* stream);

/* fx_scalar_predict_bounds on the bricks predict_bricks computes (other
 * cells of the outputs untouched). */
int fx_scalar_predict_bounds_active(fx_grid g, float dt, const float* u, const float* v, const float* w,
                                    int32_t channels, const float* const* fields, float* const* preds,
                                    float* const* lo, float* const* hi, const uint8_t* solid,
                                    const uint8_t* clearance, const uint8_t* bricks, void* stream);

/* fx_scalar_advect under a transport plan; live[k] (the output's live mask)
 * gets 1 for every brick where channel k stores a nonzero value. */
int fx_scalar_advect_active(const fx_scalar_args* a, const uint8_t* bricks, uint8_t* const* live, void* stream);

/* fx_combustion under a pointwise plan; live[0..3]: the live masks of the
 * fuel, temperature, density and flame outputs. */
int fx_combustion_active(fx_grid g, float dt, float ignition, float burn_rate, float heat_yield, float smoke_yield,
                         const float* fuel, const float* temperature, const float* density, float* fuel_out,
                         float* temperature_out, float* density_out, float* flame, const uint8_t* bricks,
                         uint8_t* const* live, void* stream);

/* max |x| over `count` arrays into out[i] (device floats), NaN/Inf -> 3.4e38
 * (AdaptiveTimestep.bounds, backend/timestep.py:24-33, shaders/max_reduce.glsl). */
int fx_max_abs(int32_t count, const float* const* arrays, const int64_t* sizes, float* out, void* stream);

/* Cell-centred vorticity (omega, |omega|) as float4 per cell, x-fast, from the
 * face velocities (DetailPasses.confine, backend/detail.py:39-49,
 * shaders/curl.glsl:1-18).  curl4: 4 floats per cell, 16-byte aligned. */
int fx_curl(fx_grid g, const float* u, const float* v, const float* w, float* curl4, void* stream);

/* Confined faces: in + dt * the mean confinement force of the two cells of the
 * face (backend/detail.py:39-49, shaders/confinement.glsl:1-25); the force is
 * strength*h*cross(grad|w|/|grad|w||, w) capped at `limit`. */
int fx_confine(fx_grid g, const float* u, const float* v, const float* w, const float* curl4, float* uo, float* vo,
               float* wo, float dt, float strength, float limit, void* stream);
/* fx_confine with the cell forces computed once: force3 is caller scratch of
 * 3*nx*ny*nz floats (component planes); results equal fx_confine's bit for bit. */
int fx_confine_split(fx_grid g, const float* u, const float* v, const float* w, const float* curl4, float* force3,
                     float* uo, float* vo, float* wo, float dt, float strength, float limit, void* stream);


/* --------------------------------------------------------------------------
 * learned preconditioner (no reference counterpart): 3-level 3D U-Net on
 * channels-last bf16 activations act[v*C + c], v = x + nx*(y + ny*z), fp32
 * accumulation on the matrix cores.  Weights: fp32 masters w[27][cin][cout]
 * (tap = (dz+1)*9 + (dy+1)*3 + (dx+1)), packed to bf16 for the kernels.
 * ------------------------------------------------------------------------ */
typedef struct fx_conv_desc {
  int32_t nx, ny, nz;  /* output grid */
  int32_t c0;          /* channels of in0 (multiple of 8) */
  int32_t c1;          /* channels of in1, concatenated after in0 (0: none) */
  int32_t cout;        /* output channels (16..96, multiple of 16) */
  int32_t up0;         /* in0 lives on the half grid and is nearest-upsampled */
  int32_t relu;        /* ReLU after the bias */
} fx_conv_desc;

/* out[v][co] = act(bias[co] + sum_{tap,ci} in(v + off(tap), ci) wt[co][tap*(c0+c1) + ci]),
 * zero padding; in = concat(upsample?(in0), in1).  wt: bf16 [cout][kp],
 * kp = 27*(c0+c1) rounded up to 32 (fx_conv_pack_weights).  Also the data
 * gradient, with weights packed with flip = 1 and c0 = cout of the layer. */
int fx_conv3d_forward(const fx_conv_desc* conv, const void* in0, const void* in1, const void* wt, const float* bias,
                      void* out, void* stream);
/* dw[27][cin][cout] (+)= sum_v in(v + off, ci) g'(v, co), db[co] (+)= sum_v g'(v, co) with
 * g' = g * (act > 0) when act != NULL.  `work` (fx_conv3d_weight_grad_workspace
 * bytes) holds channel-major copies and per-chunk partial sums. */
int fx_conv3d_weight_grad(const fx_conv_desc* conv, const void* in0, const void* in1, const void* g,
                          const void* act, void* work, int64_t work_bytes, float* dw, float* db, int32_t accumulate,
                          void* stream);
int64_t fx_conv3d_weight_grad_workspace(const fx_conv_desc* conv);
/* fp32 w[27][cin][cout] -> bf16 kernel layout: flip = 0 -> [cout][kp] with the
 * input padded to cpack channels; flip = 1 -> data-gradient layout
 * [round16(cin)][27*cout rounded to 32] with w[26 - tap]. */
int fx_conv_pack_weights(int32_t cin, int32_t cpack, int32_t cout, int32_t flip, const float* w, void* wt,
                         void* stream);
/* g[i] = act[i] > 0 ? g[i] : 0 (bf16, in place) */
int fx_relu_mask(int64_t n, void* g, const void* act, void* stream);
/* 2x2x2 mean of a channels-last bf16 tensor on the fine grid (nx, ny, nz);
 * c a multiple of 8, 16-byte aligned arrays */
int fx_pool(int32_t nx, int32_t ny, int32_t nz, int32_t c, const void* in, void* out, void* stream);
/* out = skip + (adjoint of the 2x2x2 mean)(gpool) on the fine grid; skip may be NULL */
int fx_pool_backward_add(int32_t nx, int32_t ny, int32_t nz, int32_t c, const void* gpool, const void* skip,
                         void* out, void* stream);
/* gradient of concat(upsample(a), b) on the fine grid: gup = sum over the 8
 * children of g[:, :c0] (half grid), gskip = g[:, c0:] */
int fx_concat_backward(int32_t nx, int32_t ny, int32_t nz, int32_t c0, int32_t c1, const void* g, void* gup,
                       void* gskip, void* stream);
/* packed network input [n][8]: (r / rms_fluid(r), solid, 0...) with
 * ws[8] = sum_fluid r^2, ws[9] = fluid count (device side); n a multiple of 4,
 * r and out 16-byte aligned, solid (optional) 4-byte aligned */
int fx_unet_input(int64_t n, const float* r, const uint8_t* solid, double* ws, void* out, void* stream);
/* z[v] = fluid ? (sum_c act[v][c] w[c] + b) * s : 0, s = rms_fluid(r) * post_scale
 * from ws[8..9] (fx_unet_input), or s = post_scale when ws == NULL; n a multiple
 * of 4, act and z 16-byte aligned, solid (optional) 4-byte aligned */
int fx_unet_head(int64_t n, const void* act, const float* w, const float* b, const uint8_t* solid, const double* ws,
                 float post_scale, float* z, void* stream);
/* head backward: gact = gout (x) w; dwb[0..15] += sum gout*act, dwb[16] += sum gout */
int fx_unet_head_backward(int64_t n, const void* act, const float* gout, const float* w, void* gact, float* dwb,
                          void* stream);
/* masked MSE: gout = fluid ? 2 (out - target) * inv_count : 0; ws[10] = sum_fluid
 * (out - target)^2, ws[11] = fluid count */
int fx_masked_mse(int64_t n, const float* out, const float* target, const uint8_t* solid, float inv_count,
                  float* gout, double* ws, void* stream);
/* Adam (bias-corrected, step t >= 1) on fp32 parameters */
int fx_adam(int64_t n, float* p, const float* g, float* m, float* v, int32_t t, float lr, float beta1, float beta2,
            float eps, void* stream);
int fx_to_bf16(int64_t n, const float* a, void* b, void* stream);


/* Network buffers for fx_unet_forward / fx_learned_pcg_solve.  Level-0 grid
 * dims divisible by 4 (two poolings).  Layer order: e0a e0b e1a e1b b2a b2b
 * d1a d1b d0a d0b; act[i] is the ReLU output of layer i. */
typedef struct fx_unet {
  int32_t nx, ny, nz;
  int32_t pad_;
  const void* wt[10];    /* packed bf16 forward weights (fx_conv_pack_weights, flip = 0) */
  const float* bias[10]; /* fp32 biases */
  const float* head;     /* fp32 [17]: 1x1 head weights (16) then bias */
  void* act[10];         /* bf16 [V_level][cout] */
  void* pool1;           /* bf16 [V/8][16] */
  void* pool2;           /* bf16 [V/64][32] */
  void* input;           /* bf16 [V][8] */
} fx_unet;

/* z = U-Net(r): input (r / rms_fluid(r), solid); output on fluid cells
 * scaled by post_scale, times rms_fluid(r) when output_rms != 0.  Uses
 * ws[8], ws[9]. */
int fx_unet_forward(const fx_unet* net, const float* r, const uint8_t* solid, double* ws, float post_scale,
                    int32_t output_rms, float* z, void* stream);

/* FP32 flexible PCG on the level-0 operator of `mg` with the U-Net as the
 * preconditioner: z = U-Net(r) * post_scale followed by `sweeps` (even)
 * damped-Jacobi sweeps on L z = r (ztmp: fp32 ping-pong); when the true residual misses
 * the tolerance it continues with V-cycle preconditioned FPCG on `mg`
 * (info->fallback = 1).  Vectors: x (in/out), b, r, p, Ap, zold, z, rT fp32
 * of the level-0 size; mg->level[0] p/q/rhs are the V-cycle's work arrays and
 * mg_zold a fourth level-0 array, all in mg->storage (the fallback swaps
 * level[0].p with mg_zold as fx_pcg_solve does).
 * The network grid is level 0's, or level 1's: then the network sees the
 * 8-child mean of the residual, its output is prolonged like a V-cycle
 * correction before the sweeps, `solid` is level 1's mask (a coarse cell is
 * solid when all eight children are) and post_scale is level 1's h^2. */
int fx_learned_pcg_solve(const fx_unet* net, fx_mg_config* mg, float* x, const float* b, float* r, float* p,
                         float* Ap, float* zold, float* z, float* rT, float* ztmp, void* mg_zold,
                         const uint8_t* solid, float post_scale, int32_t learned_iters, int32_t sweeps,
                         const fx_solve_params* prm, double* ws, fx_solve_info* info, void* stream);

/* ------------------------------------------------------------------------
 * Sparse brick pool (FluxFX_Source/native/bricks.hpp, regions.hpp,
 * brick_pool.hpp; Python surface native/__init__.py:31-55 BrickPool).
 * Bricks are side^3 voxels (side 8 or 16) addressed by signed brick
 * coordinates in [-1048576, 1048575]; a slot is the brick's row in the pool
 * (stable until deactivation, then reusable).  Topology and region calls are
 * host-only; the device mirror is refreshed by fx_bricks_sync.  Errors:
 * FX_EINVAL for invalid arguments (the reference's std::invalid_argument),
 * FX_ESTATE for capacity exhaustion and closed pools (std::runtime_error).
 * ------------------------------------------------------------------------ */
typedef struct fx_bricks fx_bricks;

/* Device arrays of a synced pool: field[slot * side^3 + (lz*side + ly)*side + lx],
 * coords[4*slot + {0,1,2}] = brick x, y, z, active[0..active_count),
 * neighbors[6*slot + d] for d = -X, +X, -Y, +Y, -Z, +Z (-1: none). */
typedef struct fx_bricks_dev {
  float* field;
  int32_t* coords;
  uint32_t* active;
  int32_t* neighbors;
  int32_t side;
  uint32_t capacity;
  uint32_t active_count;
  uint32_t reallocations; /* device field reallocations by capacity growth */
} fx_bricks_dev;

/* BrickPool(side, capacity, budget_bytes) (brick_pool.hpp:16-31): capacity
 * 1..262144, side 8 or 16, capacity*(4*side^3 + 28) (256-byte rounded) <= budget. */
int fx_bricks_create(int32_t side, uint32_t capacity, uint64_t budget_bytes, fx_bricks** out);
int fx_bricks_destroy(fx_bricks* b);
/* BrickPool.close: device memory released; later calls fail with FX_ESTATE. */
int fx_bricks_close(fx_bricks* b);
/* BrickTopology::activate (bricks.hpp:55-64): idempotent; slot of the brick. */
int fx_bricks_activate(fx_bricks* b, int32_t x, int32_t y, int32_t z, uint32_t* slot);
/* BrickTopology::deactivate (bricks.hpp:65-70): *removed = 0 when absent. */
int fx_bricks_deactivate(fx_bricks* b, int32_t x, int32_t y, int32_t z, int32_t* removed);
/* BrickTopology::lookup (bricks.hpp:52-54): -1 when absent. */
int fx_bricks_lookup(const fx_bricks* b, int32_t x, int32_t y, int32_t z, int32_t* slot);
/* BrickTopology::neighbors (bricks.hpp:71-79): six slots, -1 when absent. */
int fx_bricks_neighbors(const fx_bricks* b, int32_t x, int32_t y, int32_t z, int32_t* out6);
/* BrickTopology::activate_box (bricks.hpp:85-93): all-or-nothing. */
int fx_bricks_activate_box(fx_bricks* b, int32_t x, int32_t y, int32_t z, int32_t nx, int32_t ny, int32_t nz);
/* Field-driven activation (no reference counterpart; the support scan of
 * docs/DENSITY_CAPACITY.md): activates brick (x, y, z) for every nonzero
 * flags[(z*gy + y)*gx + x] (host array), in z, y, x order, all-or-nothing. */
int fx_bricks_activate_flags(fx_bricks* b, const uint8_t* flags, int32_t gx, int32_t gy, int32_t gz);
/* RegionTracker::update (regions.hpp:14-51): sources are `count` rows of
 * (x, y, z, radius, vx, vy, vz) in domain units. */
int fx_bricks_update_regions(fx_bricks* b, const double* sources, int32_t count, int32_t grid, double dt,
                             int32_t linger, int32_t halo);
int fx_bricks_count(const fx_bricks* b, uint32_t* active, uint32_t* capacity);
/* BrickPool.snapshot (brick_bindings.inc:124-135): rows of (x, y, z, slot, age)
 * in active-list order; max_rows >= the active count. */
int fx_bricks_snapshot(const fx_bricks* b, int32_t* rows, uint32_t max_rows, uint32_t* count);
int fx_bricks_host_bytes(const fx_bricks* b, uint64_t* bytes);
/* Capacity doubling (docs/DENSITY_CAPACITY.md): the capacity doubles until it
 * holds `needed` bricks, capped at `limit` (e.g. the domain's brick count);
 * slots, coordinates and (after the next sync) field values are kept. */
int fx_bricks_reserve(fx_bricks* b, uint32_t needed, uint32_t limit, uint32_t* capacity);
/* Upload the topology and rebuild the device neighbour table on the GPU
 * (brick_pool.hpp probe preparation); synchronises the stream. */
int fx_bricks_sync(fx_bricks* b, void* stream);
int fx_bricks_device(const fx_bricks* b, fx_bricks_dev* out);
/* field of every active brick <- dense nx*ny*nz (x fastest) at brick*side
 * (voxels outside the grid read 0), and the reverse. */
int fx_bricks_gather(fx_bricks* b, int32_t nx, int32_t ny, int32_t nz, const float* dense, void* stream);
int fx_bricks_scatter(fx_bricks* b, int32_t nx, int32_t ny, int32_t nz, float* dense, void* stream);
/* flags[(bz*by + by_)*bx + bx_] = any |dense| > threshold in that brick
 * (bx = ceil(nx/side) ...), device arrays. */
int fx_brick_support(int32_t side, int32_t nx, int32_t ny, int32_t nz, const float* dense, float threshold,
                     uint8_t* flags, void* stream);
/* Transport-aware preflight (native/projection.mm:126-138 requiredDensity):
 * flags[(bz*g + by)*g + bx] (g = n/side, device array, cleared here) is 1
 * when some cell of the brick is inside the emitter or its RK2 back-trace
 * (velocity u/v/w on the n^3 MAC faces, in cells per step) has a donor cell
 * with density > 0 or inside the emitter.  density: n^3 cells; emitter:
 * host floats (cx, cy, cz, radius, rate) in cells (the reference: 0.5n,
 * 0.5n, 0.45n, n/12, 0.1). */
int fx_brick_required(int32_t side, int32_t n, const float* density, const float* u, const float* v, const float* w,
                      const float* emitter, uint8_t* flags, void* stream);
/* One semi-Lagrangian step of the brick field on the cubic n^3 domain
 * (native/projection.mm:139-147 transportDensity): each voxel of an active
 * brick samples the old field trilinearly at its RK2 back-trace through the
 * MAC velocity u/v/w (cells per step, dense n^3 face arrays); donors in
 * inactive bricks read 0.  The result replaces the field (a second pool
 * buffer and an n/side brick map are allocated on the first call). */
int fx_bricks_transport(fx_bricks* b, int32_t n, const float* u, const float* v, const float* w, void* stream);
/* BrickPool.probe (brick_bindings.inc:64-90): resets every reserved voxel to
 * a sentinel, writes a per-voxel pattern of (slot, voxel, neighbour ids) into
 * active bricks and verifies all slots on the host. */
int fx_bricks_probe(fx_bricks* b, uint32_t* mismatches, float* gpu_ms, void* stream);

/* --------------------------------------------------------------------------
 * Projection on brick storage (native/projection.mm:9-168 kernels, 169-462
 * engine and compare_projection; Python surface native/__init__.py:70-81
 * compare_projection / compare_multigrid).  Unit cells, impulse pressure,
 * a fixed analytic solve region (mode 0 whole box, 1 central half, 2 two
 * boxes, 3 zero velocity, 4 gradient of a cosine q), Neumann outer walls,
 * q = 0 outside the region and on missing bricks.  Bricks hold both faces of
 * their cells along each axis; the face at local index side along an axis
 * belongs to the next brick (canonical ownership) and is never written
 * unless the brick is the last along that axis.  The sparse engine (bricks
 * within one cell of the region) and a dense engine run the same seeded
 * problem: divergence, `iterations` damped-Jacobi sweeps (omega 2/3) or, with
 * cycles > 0, that many V-cycles (4 + 4 sweeps, 4 x child mean restriction,
 * clamped trilinear prolongation, 128 sweeps at 8^3, levels down to 8^3),
 * gradient correction, divergence after.
 * ------------------------------------------------------------------------ */
typedef struct fx_projection_report {
  uint32_t resolution, side, iterations, mode, cycles, levels, active_bricks, solve_cells;
  uint64_t sparse_bytes, dense_bytes, sparse_hierarchy_bytes, dense_hierarchy_bytes;
  double topology_ms, sparse_gpu_ms, dense_gpu_ms;
  double max_error;        /* max |brick - dense| over the projected faces */
  double pressure_error;   /* max |brick - dense| over q */
  double divergence_error; /* max |brick - dense| over the divergence after */
  double before_rms, after_rms, residual_rms;  /* over solve cells; residual from the exported q */
  double wall_error;       /* max |face| on the outer walls */
  double padding_error;    /* max |padding face - 12345| (never written) */
} fx_projection_report;

/* compare_projection / compare_multigrid (projection.mm:370-462): n 32/64/128,
 * side 8/16, iterations 0..1024, cycles 0..16, mode 0..4, budget 1 B..1 GiB per
 * engine (checked before any allocation).  Host outputs from the brick engine
 * in dense layout: faces 3*(n+1)*n^2 (axis-major, each x-fast with n+1 along
 * its axis), pressure, before (divergence) and after, n^3 each. */
int fx_projection_compare(uint32_t n, uint32_t side, uint32_t iterations, uint32_t cycles, uint32_t mode,
                          uint64_t budget, float* faces, float* pressure, float* before, float* after,
                          fx_projection_report* report, void* stream);

/* --------------------------------------------------------------------------
 * Coupled transport + projection + density on brick storage
 * (native/projection.mm:101-149 kernels, 242-366 engine; Python surface
 * native/__init__.py:84-166 compare_coupled, compare_global_coupled,
 * compare_hybrid_coupled, compare_adaptive_coupled, compare_capacity_coupled).
 * Each step self-advects the faces (RK2, dt = 1 voxel) on the faces next to
 * the solve region, projects them with `cycles` V-cycles from q = 0, adds
 * the emitted density (a soft sphere at (0.5, 0.5, 0.45) n of radius n/12,
 * schedule 0 fixed, 1 swinging in x, 2 on even steps) and advects the
 * density through the projected faces.  Options: DENSE_FULL runs the dense
 * engine on the whole box; GLOBAL solves on the whole box (the seed keeps
 * its mode); HYBRID (needs GLOBAL and DENSE_FULL) keeps pressure and
 * velocity dense and the density on bricks of its own, a fixed support
 * around the central half; ADAPTIVE (needs HYBRID) grows that support from
 * the centre brick by a full-domain preflight before each density pass,
 * reallocating the engine block; POOLED (needs ADAPTIVE) keeps the density
 * in buffers of their own whose capacity doubles when exceeded.
 * ------------------------------------------------------------------------ */
enum {
  FX_COUPLED_DENSE_FULL = 1,
  FX_COUPLED_GLOBAL = 2,
  FX_COUPLED_HYBRID = 4,
  FX_COUPLED_ADAPTIVE = 8,
  FX_COUPLED_POOLED = 16
};

typedef struct fx_coupled_report {
  fx_projection_report projection;  /* last step's projection figures */
  uint32_t steps, schedule, dense_full, global_support, hybrid, adaptive, pooled, expansions;
  uint32_t density_bricks, density_capacity, density_reallocations, density_reuses;
  uint64_t density_cells, density_bytes, preserved_bytes, global_copy_bytes, peak_sparse_bytes;
  double density_error;             /* max |brick - dense| density */
  double sparse_mass, dense_mass, injected_mass, dense_after_rms;
  double density_topology_ms, preflight_ms, sparse_wall_ms, dense_wall_ms;
} fx_coupled_report;

/* compare_coupled and its variants (projection.mm:370-462): steps 1..16,
 * cycles 1..16, schedule 0..2, other ranges and budget as
 * fx_projection_compare; density growth is held to the budget with the old
 * and new storage counted together.  Host outputs as fx_projection_compare
 * plus density (n^3) from the brick engine in dense layout. */
int fx_coupled_compare(uint32_t n, uint32_t side, uint32_t steps, uint32_t cycles, uint32_t mode, uint32_t schedule,
                       uint32_t options, uint64_t budget, float* faces, float* pressure, float* before, float* after,
                       float* density, fx_coupled_report* report, void* stream);

/* --------------------------------------------------------------------------
 * Advection diagnostics on brick storage, each next to the same work on
 * dense storage; missing bricks read 0.
 * - fx_transport_compare (native/transport.mm; native/__init__.py:69-72
 *   compare_transport): a centred cube of ones (edge 7n/16) moved `steps`
 *   times by a constant displacement of at most 4 voxels per axis, each step
 *   a trilinear sample with clamped indices; the bricks cover the cube's
 *   reach.  n 32/64/128/256, side 8/16, steps 1..32.
 * - fx_mac_compare (native/mac.mm; native/__init__.py:75-78 compare_mac):
 *   MAC faces (mode 0 localized affine, 1 affine, 2 constant, 3 signed box)
 *   self-advected `steps` times (RK2, dt = 1 voxel, coordinate-clamped
 *   trilinear samples) on canonically owned brick faces.  n 32/64/128,
 *   side 8/16, steps 0..8, mode 0..3.
 * Budget 1 B..1 GiB per engine, checked before any allocation.
 * ------------------------------------------------------------------------ */
typedef struct fx_advect_report {
  uint32_t resolution, side, steps, mode, active_bricks, pad_;
  uint64_t sparse_bytes, dense_bytes;
  uint64_t owned_faces, padding_faces;        /* MAC: stored faces by ownership */
  double max_error, rms_error, padding_error; /* brick vs dense; padding vs 12345 */
  double topology_ms, sparse_wall_ms, dense_wall_ms;
  double sparse_mass, dense_mass, initial_mass; /* scalar transport */
  double sparse_gpu_ms[32], dense_gpu_ms[32];   /* per step */
} fx_advect_report;

/* density: n^3 floats, the brick engine's field in dense layout */
int fx_transport_compare(uint32_t n, uint32_t side, uint32_t steps, float dx, float dy, float dz, uint64_t budget,
                         float* density, fx_advect_report* report, void* stream);

/* faces: 3*(n+1)*n^2 floats, the brick engine's faces in dense layout */
int fx_mac_compare(uint32_t n, uint32_t side, uint32_t steps, uint32_t mode, uint64_t budget, float* faces,
                   fx_advect_report* report, void* stream);

#ifdef __cplusplus
}
#endif
#endif
