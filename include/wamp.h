/*
 * wamp.h -- C-ABI of libwamp.so, the MI355X-native (gfx950) five-equation
 * stiffened-gas solver step.
 *
 * The reference (aschamarthi/Wave-Multiphase) has no plugin API: its
 * boundary is the set of wp.launch call sites inside main()'s closures
 * (sainath_SC_it_works.py, "SC", lines 1894-1995; Sainath_M10.py, "M10",
 * lines 1815-1907).  Each entry point below replaces one closure (or one
 * group of launches) and cites it.
 *
 * Conventions
 *  - State arrays are SoA FP64, one array per conserved field in the order
 *    [a1*rho1, a2*rho2, rho*u, rho*v, E, a1] (SC:777, 1842-1844), each of
 *    shape (nx+2g, ld) row-major with j (y) contiguous, interior cells at
 *    [g, g+nx) x [g, g+ny) (wp.array2d layout, SC:1839-1840).  ld >= ny+2g.
 *  - Pointers are device pointers (e.g. torch.Tensor.data_ptr()).  The
 *    caller owns every state, flux and primitive buffer; the context owns
 *    only scratch it allocates in wamp_ctx_create.
 *  - Every call enqueues on `stream` (a hipStream_t; NULL = default stream)
 *    and returns without synchronising.
 *  - Return value: 0 ok, > 0 a hipError_t, < 0 an argument error.  The
 *    message of the last failure on this thread is wamp_last_error().
 */
#ifndef WAMP_H
#define WAMP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wamp_ctx wamp_ctx;

/* Grid: SC:26-32.  bc: x-lo, x-hi, y-lo, y-hi; 0 = zero-gradient copy of the
   edge cell (bc_x_kernel SC:1616-1646), 1 = slip wall, the same copy with
   the wall-normal momentum negated (bc_y_kernel SC:1650-1682). */
typedef struct wamp_grid {
    int32_t nx, ny, g, ld;
    double dx, dy;
    int32_t bc[4];
} wamp_grid;

/* Two stiffened-gas phases (gamma, p_inf), SC:40-43. */
typedef struct wamp_eos {
    double gamma1, pinf1, gamma2, pinf2;
} wamp_eos;

/* Scheme switches.  base: 0 MUSCL (SC:1061-1073), 1 MP5 (SC:353-405).
   shear: 0 none (M10), 1 central4 (SC:1074-1082), 2 central6 (SC:407-412),
   Ducros-gated (SC:866-916).  vf_thinc: 0 gated (SC:1085-1092), 1 always
   (M10:971-979).  face_pp: 0 first-order fallback (SC:1102-1115), 1 Wong
   face limiter + pressure floor (M10:1007-1025).  repair: enforce_admissible
   after every stage (USE_EMERGENCY_REPAIR, SC:37).  cfl: SC:34. */
typedef struct wamp_scheme {
    int32_t base, shear, vf_thinc, face_pp, repair;
    double cfl;
} wamp_scheme;

#define WAMP_PREC_FP64 0
#define WAMP_PREC_MIXED 1

/* Create a context for one grid/EOS/scheme; allocates the step's scratch
   (face fluxes, two stage buffers, dt accumulator) on the current device:
   40 FP64 planes of the padded grid, or in WAMP_PREC_MIXED 16 FP64 planes
   and 24 FP32 face planes.  Replaces main()'s allocation of the 47 work
   arrays, SC:1839-1866. */
int wamp_ctx_create(const wamp_grid *grid, const wamp_eos *eos, const wamp_scheme *scheme,
                    int32_t precision, wamp_ctx **out);
int wamp_ctx_destroy(wamp_ctx *ctx);
const char *wamp_last_error(void);

/* Bytes of device scratch the context holds. */
int64_t wamp_ctx_scratch_bytes(const wamp_ctx *ctx);

/* apply_bc, SC:1899-1905: fill the g ghost layers of U (and of the
   primitive arrays P = [u, v, p] when P is non-NULL) from the interior. */
int wamp_apply_bc(wamp_ctx *ctx, double *const U[6], double *const P[3], void *stream);

/* update_prims, SC:1913-1917 (cons_to_prim_kernel SC:840-864) over the
   padded grid: clamp a1 to [1e-8, 1-1e-8] in place, write u, v, p. */
int wamp_cons_to_prim(wamp_ctx *ctx, double *const U[6], double *const P[3], void *stream);

/* enforce_admissible, SC:1907-1911 (enforce_admissible_kernel SC:918-958),
   over the padded grid.  A no-op when scheme.repair == 0. */
int wamp_enforce_admissible(wamp_ctx *ctx, double *const U[6], void *stream);

/* The initial sequence SC:1998-2001 (enforce_admissible, apply_bc,
   update_prims, apply_bc) on the interior: repair (if enabled) and clamp a1,
   in place.  Ghost layers are left to wamp_apply_bc. */
int wamp_prepare(wamp_ctx *ctx, double *const U[6], void *stream);

/* compute_fluxes + compute_low_fluxes, SC:1919-1947 (ducros_kernel,
   fx_kernel, gy_kernel, fx_low_kernel, gy_low_kernel).  F receives 28
   arrays of the state's shape in the reference's order: fx0..fx5, srcx,
   gy0..gy5, srcy, fxl0..fxl5, srcxl, gyl0..gyl5, srcyl.  x-face values sit
   at [i, j] for face i+1/2, i in [g-1, g+nx-1]; y-faces likewise.  The 28
   planes must be equally spaced in one allocation (e.g. a (28, nx+2g, ld)
   array): F[m] = F[0] + m * (F[1] - F[0]).  The boundary fill is applied on
   the fly; U's ghost layers are not read. */
int wamp_compute_fluxes(wamp_ctx *ctx, double *const U[6], double *const F[28], void *stream);

/* limit_fluxes + rk3_update + enforce_admissible + update_prims,
   SC:1949-1984, 1907-1917: the three Wong stages, the SSP-RK3 combine
   Uout = alpha*Un + beta*(U - dt*R(F)), repair and the a1 clamp, on the
   interior.  F as in wamp_compute_fluxes.  Uout must not alias U (halo cells of U are read by
   neighbouring tiles); it may alias Un.  dt is read from device memory.
   When dt_acc is non-NULL the CFL minimum of the new state is folded into
   *dt_acc (atomic min, timestep_kernel SC:1686-1705). */
int wamp_limit_update(wamp_ctx *ctx, double *const Un[6], double *const U[6],
                      double *const F[28], double *const Uout[6], double alpha, double beta,
                      const double *dt_dev, double *dt_acc, void *stream);

/* get_dt, SC:1986-1995 (timestep_kernel SC:1686-1705): *dt_dev =
   cfl * min over interior cells of sx*sy/(sx+sy).  Device-resident; no
   host synchronisation. */
int wamp_get_dt(wamp_ctx *ctx, double *const U[6], double *dt_dev, void *stream);

/* One SSP-RK3 step in place on U, SC:2028-2059: copy_old, three stages of
   (apply_bc, fluxes, limiter, rk3_update, repair, prims), then the next dt.
   dt is read from dt_dev; cfl * min of the new state is written to
   dt_next_dev (may equal dt_dev).  U's ghost layers are unspecified on
   return (call wamp_apply_bc for them). */
int wamp_step(wamp_ctx *ctx, double *const U[6], const double *dt_dev, double *dt_next_dev,
              void *stream);

/* No counterpart in the reference: for callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc).  wamp_step without a context: the same step (the same
   kernels, the same result bit for bit) with its scratch -- face planes,
   stage buffers, dt accumulator -- in a static device array of
   WAMP_NOALLOC_SCRATCH_BYTES per device, held until the process ends.  It
   allocates nothing.  Grids whose scratch does not fit are refused with an
   error (about 209k padded cells in FP64, e.g. 512 x 256; 299k in mixed
   precision).  Calls on one device share its array: calls that overlap on one
   device (other streams, other threads) race. */
#define WAMP_NOALLOC_SCRATCH_BYTES (64u << 20)
int wamp_step_noalloc(const wamp_grid *grid, const wamp_eos *eos, const wamp_scheme *scheme,
                      int32_t precision, double *const U[6], const double *dt_dev,
                      double *dt_next_dev, void *stream);

/* ---- The CPU device ------------------------------------------------------
   The reference's --cpu (SC:15-21, M10:14-20: wp.init() with device "cpu")
   runs its Warp kernels on the host.  These entry points run this library's
   FP64 step on the host: the same per-cell code as the gfx950 kernels (the
   face tile phases, the limiter stages and the update are host-and-device
   functions), with the IEEE divide and square root that the kernels' division
   and square-root sequences reproduce, and the same pow and tanh -- the
   result is the FP64 GPU step's, bit for bit, except in the last bits of
   subnormal values (below 2^-1022), where the kernels' shortened division
   is not the IEEE quotient.  Pointers are HOST pointers,
   with the layout of the device entry points; every call returns when done.
   FP64 only (precision is not an argument).  One context is not re-entrant:
   calls on one context from several threads race.  No HIP call is made. */
typedef struct wamp_cpu_ctx wamp_cpu_ctx;

/* threads: worker threads per call; 0: OMP_NUM_THREADS when set, else the
   CPUs of the process's affinity mask.  Allocates
   the scratch of the device context in host memory (40 planes) plus three
   limiter theta planes. */
int wamp_cpu_ctx_create(const wamp_grid *grid, const wamp_eos *eos, const wamp_scheme *scheme,
                        int32_t threads, wamp_cpu_ctx **out);
int wamp_cpu_ctx_destroy(wamp_cpu_ctx *ctx);
/* Worker threads of a context. */
int32_t wamp_cpu_threads(const wamp_cpu_ctx *ctx);
/* As wamp_prepare, wamp_apply_bc, wamp_cons_to_prim, wamp_compute_fluxes
   (SC:1998-2001, 1899-1905, 1913-1917, 1919-1947). */
int wamp_cpu_prepare(wamp_cpu_ctx *ctx, double *const U[6]);
int wamp_cpu_apply_bc(wamp_cpu_ctx *ctx, double *const U[6], double *const P[3]);
int wamp_cpu_cons_to_prim(wamp_cpu_ctx *ctx, double *const U[6], double *const P[3]);
int wamp_cpu_compute_fluxes(wamp_cpu_ctx *ctx, double *const U[6], double *const F[28]);
/* get_dt, SC:1986-1995: *dt = cfl * min over interior cells. */
int wamp_cpu_get_dt(wamp_cpu_ctx *ctx, double *const U[6], double *dt);
/* One SSP-RK3 step in place on U with time step dt (SC:2028-2059); the next
   dt, cfl * min of the new state, to *dt_next (may be NULL). */
int wamp_cpu_step(wamp_cpu_ctx *ctx, double *const U[6], double dt, double *dt_next);

/* ---- Initial conditions on the device -------------------------------------
   Replace the reference's initial-condition kernel (ic_kernel2, SC:779-837 /
   M10:736-790), a dimensional sharp-interface setup, with the nondimensional
   cases the solver is measured on, built in device memory (the host-side
   numpy versions are wamp/cases.py).  Both fill U's (nx+2g) x ld planes:
   interior cells from the case, all other cells zero.  x_i = xmin +
   (i-g+1/2) dx, y likewise (SC:1832). */

/* Planar shock at shock_x moving +x into air at rest (rho_air, p0) holding
   water droplets (rho_water): drops[4k..4k+3] = (xc, yc, r, r_cavity), a
   device array; r_cavity > 0 carves a concentric air cavity.  Volume
   fraction a1 = max over droplets of 1/2 (1 - tanh((d - r)/width)) (times
   1 - 1/2 (1 - tanh((d - r_cavity)/width)) with a cavity), clipped to
   [1e-6, 1 - 1e-6]; behind the shock (x < shock_x) the air state is
   (rho_post, p_post, u_post). */
typedef struct wamp_droplet_ic {
    double xmin, ymin;
    double shock_x, rho_post, p_post, u_post;
    double rho_water, rho_air, p0;
    double width;
    const double *drops;
    int32_t n_drops;
} wamp_droplet_ic;
int wamp_init_droplets(const wamp_grid *grid, const wamp_eos *eos, const wamp_droplet_ic *ic,
                       double *const U[6], void *stream);

/* Three constant states at rest: x < x_split (left), x >= x_split and
   y >= y_split (top) in phase 1 (a1 = 1 - 1e-6), the rest (bottom) in
   phase 2 (a1 = 1e-6) -- the gas-gas triple point. */
typedef struct wamp_triple_ic {
    double xmin, ymin, x_split, y_split;
    double rho_left, p_left, rho_top, p_top, rho_bottom, p_bottom;
} wamp_triple_ic;
int wamp_init_triple_point(const wamp_grid *grid, const wamp_eos *eos, const wamp_triple_ic *ic,
                           double *const U[6], void *stream);

/* ---- Batches ------------------------------------------------------------
   No counterpart in the reference (one case per script run).  A batch
   advances many independent cases -- any grids and EOS, one scheme and
   precision -- with one persistent launch per kernel kind per RK stage over
   a work list of every case's face tiles / limiter rows.  Per-case dt, time,
   end time and a finished flag stay on the device; a step can be captured in
   a hipGraph.  Cases are held in slots; the caller owns each case's state
   arrays, the batch owns each case's scratch (face and stage buffers). */
typedef struct wamp_batch wamp_batch;

/* scheme must be SC (base 0, shear 1, face_pp 0), M10 (0, 0, 1) or MP5 (1, 2, 0). */
int wamp_batch_create(const wamp_scheme *scheme, int32_t precision, int32_t max_cases,
                      wamp_batch **out);
int wamp_batch_destroy(wamp_batch *b);
/* Put a case in a slot (allocates its scratch); t_end <= 0: no end time. */
int wamp_batch_set_case(wamp_batch *b, int32_t slot, const wamp_grid *grid, const wamp_eos *eos,
                        double *const U[6], double t0, double t_end);
int wamp_batch_clear_case(wamp_batch *b, int32_t slot);
/* Work decomposition: tile 0 = 8x32 face tiles, 1 = 16x16; persistent blocks
   per CU for the face kernel (1..16) and the limiter kernel (0 = one wave per
   limiter item, 1..16 persistent); limiter rows per wave (4..4096, or 0:
   chosen at each commit for the resident cases, the fewest rounds x fronts).
   Results do not depend on these. */
int wamp_batch_set_tuning(wamp_batch *b, int32_t tile, int32_t faces_blocks_per_cu,
                          int32_t lu_blocks_per_cu, int32_t lu_seg);
/* Upload the case table and work lists (after set/clear/tuning). */
int wamp_batch_commit(wamp_batch *b, void *stream);
/* Initial sequence and first dt of the given slots (all occupied if n_slots < 0). */
int wamp_batch_prepare(wamp_batch *b, const int32_t *slots, int32_t n_slots, void *stream);
int wamp_batch_set_done(wamp_batch *b, int32_t slot, int32_t done, void *stream);
/* Capture `steps` steps in a hipGraph (0 drops it). */
int wamp_batch_graph(wamp_batch *b, int32_t steps);
/* One step (or one graph replay) of every unfinished case; returns the number
   of steps enqueued (> 0) or an error (< 0 / hipError_t). */
int wamp_batch_step(wamp_batch *b, void *stream);
/* Per-slot time, dt and finished flag to the host (synchronises the stream). */
int wamp_batch_status(wamp_batch *b, double *t, double *dt, int32_t *done, void *stream);
/* Work-list lengths: which = 0 face tiles, 1 limiter wave items; 2: the
   limiter rows per wave in effect. */
int64_t wamp_batch_work(const wamp_batch *b, int32_t which);

/* Measurement hook (no counterpart in the reference; for benchmarks and
   profilers).  wamp_timing_enable(ctx, n) makes the next n calls of
   wamp_step record hipEvents on their stream around each stage's face-flux
   kernel and limiter/update kernel (n = 0 disables).  wamp_timing_read
   waits for the recorded events and returns the summed milliseconds and
   launch counts: [0] face-flux kernel, [1] limiter/update kernel. */
int wamp_timing_enable(wamp_ctx *ctx, int32_t max_steps);
int wamp_timing_read(wamp_ctx *ctx, double ms[2], int64_t launches[2]);

/* Arithmetic self-test (no counterpart in the reference; for tests).  For
   i < n: out[i] = a[i] / b[i] by the FP64 kernels' division sequence (the
   hardware reciprocal, two Newton steps, one correction), out[n + i] = the
   compiler's IEEE a[i] / b[i], out[2n + i] = sqrt(b[i]) by the kernels'
   square-root sequence, out[3n + i] = the IEEE sqrt(b[i]).  a, b: n device
   doubles; out: 4n device doubles. */
int wamp_selftest_arith(const double *a, const double *b, double *out, int64_t n, void *stream);

/* Elementary-function self-test (no counterpart in the reference; for
   tests).  For i < n: out[i] = x[i]^y[i] and out[n + i] = tanh(x[i]) as the
   FP64 kernels evaluate them (the entropy's pow, cell_entropy SC:308-314;
   THINC's tanh, SC:327-349): functions built from IEEE operations only, so a
   host build of the same source gives the same bits.  x, y: n device
   doubles; out: 2n device doubles. */
int wamp_selftest_elementary(const double *x, const double *y, double *out, int64_t n,
                             void *stream);

/* ---- Three-dimensional extension (no counterpart in the reference, whose
   scripts are 2D; its README mentions 3D runs whose code it does not hold).
   The same step with a z-sweep and the 7-variable state [a1*rho1, a2*rho2,
   rho*u, rho*v, rho*w, E, a1]; each array has shape (nx+2g, ny+2g, ldz)
   with z contiguous, interior [g, g+nx) x [g, g+ny) x [g, g+nz).  A
   z-invariant state with w = 0 takes exactly the 2D step of its (x, y)
   slice, and a y-invariant state with v = 0 exactly the 2D step of its
   (x, z) slice, for the same dt.  FP64 or FP32-mixed; schemes SC, M10 and
   MP5. */
typedef struct wamp3_ctx wamp3_ctx;

/* bc: x-lo, x-hi, y-lo, y-hi, z-lo, z-hi, as in wamp_grid. */
typedef struct wamp_grid3 {
    int32_t nx, ny, nz, g, ldz;
    double dx, dy, dz;
    int32_t bc[6];
} wamp_grid3;

/* Allocates 68 FP64 planes of the padded grid: 48 face planes, two stage
   buffers, pressure / entropy / Ducros planes, three limiter theta planes.
   precision WAMP_PREC_MIXED: FP32 projection, reconstruction and HLLC in the
   face kernels (energy flux, pressure, indicators and limiters FP64), and
   the face planes as 6 FP64 energy planes + 42 FP32 planes (47 FP64-plane
   equivalents in all). */
int wamp3_ctx_create(const wamp_grid3 *grid, const wamp_eos *eos, const wamp_scheme *scheme,
                     int32_t precision, wamp3_ctx **out);
int wamp3_ctx_destroy(wamp3_ctx *ctx);
int64_t wamp3_ctx_scratch_bytes(const wamp3_ctx *ctx);
/* Interior repair + a1 clamp, as wamp_prepare. */
int wamp3_prepare(wamp3_ctx *ctx, double *const U[7], void *stream);
/* *dt_dev = cfl * min over interior cells of 1 / (1/sx + 1/sy + 1/sz),
   s_d = d_d / (|u_d| + c): the 2D sx*sy/(sx+sy) of get_dt with z added. */
int wamp3_get_dt(wamp3_ctx *ctx, double *const U[7], double *dt_dev, void *stream);
/* One SSP-RK3 step in place on U, as wamp_step. */
int wamp3_step(wamp3_ctx *ctx, double *const U[7], const double *dt_dev, double *dt_next_dev,
               void *stream);
/* The split form of one stage, as wamp_compute_fluxes / wamp_limit_update
   (the reference's compute_fluxes + compute_low_fluxes and limit_fluxes +
   rk3_update + enforce_admissible + update_prims, SC:1919-1984, with the
   z-sweep added).  F: 48 FP64 face planes of the padded grid, equally spaced
   in one allocation; direction d (x, y, z) at planes d*16 .. d*16+15: the
   high-order fluxes of the 7 conserved variables and u*, then the
   first-order ones; the face d + 1/2 of a cell is stored at the cell's
   index.  In mixed precision the face kernels' FP32 values are stored
   widened, so the limiter sees the numbers wamp3_step does. */
int wamp3_compute_fluxes(wamp3_ctx *ctx, double *const U[7], double *const F[48], void *stream);
/* Uout = alpha * Un + beta * (U - dt R(U)) with the three Wong limiter
   stages on F, repair and the a1 clamp; Uout must not alias U.  dt is read
   from dt_dev; when dt_acc is non-NULL the CFL minimum of Uout is folded
   into *dt_acc (an FP64 accumulator the caller initialises to 1e30). */
int wamp3_limit_update(wamp3_ctx *ctx, double *const Un[7], double *const U[7],
                       double *const F[48], double *const Uout[7], double alpha, double beta,
                       const double *dt_dev, double *dt_acc, void *stream);

/* Batches of 3D cases (no counterpart in the reference, whose solver is 2D):
   the wamp_batch_* pattern for the 3D step.  Cases of any grid size and EOS
   sharing one scheme and precision advance together; each kernel of a stage
   is one launch over every active case's blocks, and each case's dt, time,
   end time and done flag stay on the device.  Every case's result is bitwise
   the result of wamp3_step on that case alone. */
typedef struct wamp3_batch wamp3_batch;
int wamp3_batch_create(const wamp_scheme *scheme, int32_t precision, int32_t max_cases,
                       wamp3_batch **out);
int wamp3_batch_destroy(wamp3_batch *b);
/* Put a case in a slot (allocates its scratch: 68 FP64 planes of its padded
   grid, 47 plane-equivalents in mixed precision); t_end <= 0: no end time.
   The state U stays the caller's and is advanced in place. */
int wamp3_batch_set_case(wamp3_batch *b, int32_t slot, const wamp_grid3 *grid,
                         const wamp_eos *eos, double *const U[7], double t0, double t_end);
int wamp3_batch_clear_case(wamp3_batch *b, int32_t slot);
/* Upload the case table and work offsets (after set/clear). */
int wamp3_batch_commit(wamp3_batch *b, void *stream);
/* Initial sequence and first dt of the given slots (all occupied if n_slots < 0). */
int wamp3_batch_prepare(wamp3_batch *b, const int32_t *slots, int32_t n_slots, void *stream);
/* Capture `steps` steps in a hipGraph that wamp3_batch_step replays (0: no graph). */
int wamp3_batch_graph(wamp3_batch *b, int32_t steps);
/* Advance every unfinished case; returns the number of steps enqueued (> 0). */
int wamp3_batch_step(wamp3_batch *b, void *stream);
/* t, dt, done of every slot to host arrays of max_cases entries (synchronises). */
int wamp3_batch_status(wamp3_batch *b, double *t, double *dt, int32_t *done, void *stream);

#ifdef __cplusplus
}
#endif
#endif
