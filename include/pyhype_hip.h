/*
 * pyhype_hip.h -- C ABI of the MI355X (gfx950) backend for pyHype's second-order MUSCL
 * residual and explicit Runge-Kutta time stepping.
 *
 * All functions return 0 on success and a negative PHX_E_* code on failure; the message of
 * the last failure on the calling thread is available from phx_last_error().  Device buffers
 * are caller-owned FP64 arrays (typically torch tensors) passed as raw pointers; a plan owns
 * only its own scratch.  `stream` is a hipStream_t passed as void* (NULL = default stream).
 * Calls are stream-ordered; none of them synchronises the host except phx_status and
 * phx_plan_create/destroy.  A plan is not thread-safe; separate plans on separate streams are.
 * On a mesh with both Cartesian and general blocks, each stage phase also runs on a stream the
 * plan owns, forked from and joined back to the caller's stream with events, so the ordering seen
 * from the caller's stream is unchanged (the element-wise stage, and the split stages with one
 * y-face per wave and midpoint faces, instead launch both kinds' items together); two calls on one
 * plan must not overlap in time.
 *
 * State layout: every block is a C-ordered (ny, nx, 4) array of conservative variables
 * (rho, rho*u, rho*v, E), exactly the layout of the reference's ConservativeState.data
 * (states/conservative.py:32-40); blocks are concatenated at their `state_offset`.
 */
#ifndef PYHYPE_HIP_H
#define PYHYPE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHX_ABI_VERSION 3

/* error codes */
#define PHX_OK 0
#define PHX_E_INVALID (-1)   /* bad argument / descriptor (reference: ValueError at construction) */
#define PHX_E_HIP (-2)       /* a HIP runtime call failed */
#define PHX_E_UNREALIZABLE (-3) /* rho <= 0 or E <= 0 (reference: RealizabilityException) */
#define PHX_E_STATE (-4)     /* call out of order for this plan */

/* side indices (utils/utils.py:33-42 Direction.east/west/north/south) */
#define PHX_E 0
#define PHX_W 1
#define PHX_N 2
#define PHX_S 3

/* what lies beyond a block side (blocks/ghost.py:128-143, 187-216; fvm/base.py:305-325) */
#define PHX_SIDE_NEIGHBOR 0  /* bc_type None with a neighbour block: ghost = neighbour cells      */
#define PHX_SIDE_OPEN 1      /* bc_type None, no neighbour: ghost = own boundary cells            */
#define PHX_SIDE_WALL 2      /* "Reflection" / "Slipwall": ghost = reflected own boundary cells   */
#define PHX_SIDE_OUTLET 3    /* "OutletDirichlet": ghost = own cells, face state not reflected    */

/* flux functions (flux/factory.py:30-47) */
#define PHX_FLUX_ROE 0
#define PHX_FLUX_HLLL 1

/* reconstruction variables (SolverConfig.reconstruction_type; fvm/base.py:255-303,
 * blocks/quad_block.py:120-134,225-240) */
#define PHX_RECON_CONSERVATIVE 0
#define PHX_RECON_PRIMITIVE 1

/* limiters (limiters/factory.py:28-39, limiters/limiters.py:25-62) */
#define PHX_LIMITER_VENKATAKRISHNAN 0
#define PHX_LIMITER_BARTH_JESPERSEN 1
#define PHX_LIMITER_VAN_ALBADA 2
#define PHX_LIMITER_VAN_LEER 3

/* geometry kinds of a block's packed geometry array */
#define PHX_GEOM_CARTESIAN 1 /* axis-aligned block, separable 1-D arrays (see below) */
#define PHX_GEOM_GENERAL 0   /* per-face arrays, rotated-frame fluxes (fvm/base.py:366-390,442-486) */

#define PHX_MAX_STAGES 8

/* Gauss points per face (SolverConfig.fvm_num_quadrature_points; mesh/quadratures.py:33-37) */
#define PHX_MAX_QUADRATURE_POINTS 3

/*
 * Packed geometry, all values in FP64, produced on the host by pyhype_amd.geometry from the
 * reference's own formulas (mesh/quad_mesh.py:71-212, mesh/base.py:29-78,
 * mesh/quadratures.py:56-95).
 *
 * PHX_GEOM_CARTESIAN (length 8 + 11*nx + 11*ny + nx*ny):
 *   [0..7]        wall cos/sin per side: cE sE cW sW cN sN cS sS (libm cos/sin of the face angle)
 *   column arrays (11 x nx): xlN xlS ylN ylS LN LS dxqE dxqW dxqN dxqS dxc
 *   row arrays    (11 x ny): xlE xlW ylE ylW LE LW dyqE dyqW dyqN dyqS dyc
 *   A (ny x nx)
 *   where xlF = L_F * n_F.x, ylF = L_F * n_F.y, dxqF = xq_F - xc, dyqF = yq_F - yc,
 *   dxc = dx, dyc = dy (mesh/quad_mesh.py:134-135).
 * PHX_GEOM_GENERAL (length 7*ny*(nx+1) + 7*(ny+1)*nx + 5*ny*nx):
 *   vertical faces ny x (nx+1), 7 arrays: L xl yl xq yq c s  (east orientation, face j between
 *       cells j-1 and j; the west face of a cell uses -xl, -yl)
 *   horizontal faces (ny+1) x nx, 7 arrays: L xl yl xq yq c s (north orientation)
 *   cells ny x nx, 5 arrays: A xc yc dx dy
 * With nq = quadrature_points > 1 either layout is followed by the Gauss-point offsets of every
 * cell, 8*nq arrays of ny x nx: for face F in E, W, N, S, point i < nq (rule order), then
 * component x, y: xq_{F,i} - xc, yq_{F,i} - yc, with the points of
 * QuadraturePointData._transform (mesh/quadratures.py:56-95); the midpoint entries above are
 * then unused by the reconstruction.
 */
typedef struct {
    int32_t nx, ny;
    int32_t case_id;          /* independent case (own dt and time) this block belongs to */
    int32_t geom_kind;        /* PHX_GEOM_* */
    int32_t side_type[4];     /* PHX_SIDE_* per side E, W, N, S */
    int32_t neighbor[4];      /* block index for PHX_SIDE_NEIGHBOR, else -1 */
    int64_t state_offset;     /* offset of block in the state buffer, in doubles, multiple of 4 */
    const double* geom;       /* host pointer to the packed geometry */
    int64_t geom_len;         /* number of doubles in geom */
} phx_block_desc;

typedef struct {
    int32_t abi_version;      /* PHX_ABI_VERSION */
    int32_t nblocks;
    const phx_block_desc* blocks;
    int32_t ncases;
    int32_t flux;             /* PHX_FLUX_* */
    int32_t limiter;          /* PHX_LIMITER_* */
    int32_t nstages;          /* Butcher stages (time_marching/explicit_runge_kutta.py:91-274) */
    const double* butcher_a;  /* nstages x nstages row-major, entry [s][j] for j <= s */
    double gamma;             /* Fluid.gamma() */
    double gm1;               /* gamma - 1      (states/conservative.py:110) */
    double one_over_gm1;      /* 1/(gamma-1.0)  (fluids/base.py:62-64) */
    double g_over_gm1;        /* gamma/(gamma-1.0) (fluids/base.py:58-60) */
    double cfl;               /* SolverConfig.CFL */
    const double* t_final;    /* per case, already scaled by a_inf (solvers/base.py:73) */
    int32_t rows_per_item;    /* rows of one work item; 0 = default */
    int32_t grid_waves;       /* waves of the persistent grid; 0 = one per work item */
    int32_t reconstruction;   /* PHX_RECON_* (SolverConfig.reconstruction_type) */
    int32_t quadrature_points; /* Gauss points per face, 1..PHX_MAX_QUADRATURE_POINTS; > 1 runs
                                  the split stages (phx_set_variant 1, 5 or 6, the default) */
} phx_plan_desc;

typedef struct phx_plan phx_plan;

/* Build a plan: validates descriptors, uploads geometry, allocates scratch.
 * Replaces the construction of Blocks/QuadBlock/fvm objects (blocks/base.py:399-412,515-572;
 * blocks/quad_block.py:244-272) and TimeIntegratorFactory.create (time_marching/factory.py:29-62). */
int phx_plan_create(const phx_plan_desc* desc, phx_plan** out);
int phx_plan_destroy(phx_plan* plan);

/* Total number of doubles of the concatenated state buffer. */
int64_t phx_state_len(const phx_plan* plan);

/* Set simulation time and step counter of every case (solvers/base.py:70-71), clear flags. */
int phx_reset(phx_plan* plan, const double* t0_host, int32_t step0, void* stream);

/* Advance `nsteps` full time steps in place on U: per step dt = Solver.get_dt
 * (solvers/base.py:114-136), then ExplicitRungeKutta.integrate (explicit_runge_kutta.py:47-89)
 * with the ghost fill of Blocks.apply_boundary_condition (blocks/base.py:448-471) after each
 * stage, then the realizability check (solvers/Euler2D.py:144-152).  Replaces the body of the
 * step loop Euler2D._solve (solvers/Euler2D.py:195-210).  use_graph != 0 replays captured
 * hipGraphs (captured on first use): one of 8 steps per full group of 8, then one of a single step
 * per remaining step.  Realizability is reported by phx_status. */
int phx_run(phx_plan* plan, double* U, int32_t nsteps, int32_t use_graph, void* stream);

/* Advance nsteps with a frozen, caller-given time step per step (dt_sched: device pointer,
 * nsteps x ncases); used by the discrete adjoint.  Time is advanced as in phx_run. */
int phx_run_frozen(phx_plan* plan, double* U, int32_t nsteps, const double* dt_sched, void* stream);

/* phx_run_frozen that also records the input state of every stage of every step: stage s of step n
 * goes to stages + (n*nstages + s)*state_len (device, nsteps x nstages x state_len doubles).  The
 * recorded states are the linearisation points of phx_step_vjp.  No counterpart in the reference
 * (the discrete adjoint is new). */
int phx_run_record(phx_plan* plan, double* U, int32_t nsteps, const double* dt_sched, double* stages,
                   void* stream);

/* One stage (0-based) of the plan's Runge-Kutta program on state U and the plan's stage buffers,
 * using the dt of the current step: the body of the stage loop of ExplicitRungeKutta.integrate
 * (time_marching/explicit_runge_kutta.py:63-80) for all blocks at once.  Used to time the stage
 * kernel on its own; phx_run performs whole steps. */
int phx_stage(phx_plan* plan, double* U, int32_t stage, void* stream);

/* dUdt of every block at state U into R (same layout): FiniteVolumeMethod.dUdt
 * (fvm/base.py:108-139) with the ghost cells implied by U (blocks/base.py:448-471). */
int phx_residual(phx_plan* plan, const double* U, double* R, void* stream);

/* Per-case CFL time step of state U (blocks/quad_block.py:423-436, solvers/base.py:114-136,
 * including the clip to t_final - t) into dt_dev (device, ncases).  Does not advance time. */
int phx_dt(phx_plan* plan, const double* U, double* dt_dev, void* stream);

/* Ghost strips of every block (blocks/ghost.py:187-278) for state U, into G (device); block b
 * occupies 2*ny+2*nx cells x 4 starting at doubles offset 4*sum_{c<b}(2ny_c+2nx_c), ordered
 * E (ny), W (ny), N (nx), S (nx).  For tests of the ghost fill. */
int phx_ghosts(phx_plan* plan, const double* U, double* G, void* stream);

/* Host-synchronising status.  bad_step (ncases entries, may be NULL): per case, the first step
 * whose result was unrealizable, or -1 (Euler2D._realizability_check, solvers/Euler2D.py:144-152;
 * a failed case is frozen from then on, as the reference aborts that run); step: steps taken;
 * t_host / dt_host (ncases, may be NULL): per-case time and last dt. */
int phx_status(phx_plan* plan, int32_t* bad_step, int32_t* step, double* t_host, double* dt_host);

/* Reverse-mode product of the residual: lamU = (dR/dU)^T lamR at state U (discrete adjoint of
 * fvm/base.py:108-190 with the active limiter/flux branches; deterministic, no atomics).  Requires
 * mutual neighbour links (PHX_E_INVALID otherwise).  The first adjoint call allocates the plan's
 * adjoint work arrays (about 20 + 4*(nstages-1) + 32*nq doubles per cell, nq Gauss points per
 * face).  No counterpart in the reference. */
int phx_residual_vjp(phx_plan* plan, const double* U, const double* lamR, double* lamU, void* stream);

/* Forward-mode (tangent-linear) product of the residual: dR = (dR/dU) dU at state U, the residual's
 * own arithmetic evaluated on value-and-derivative pairs, so every limiter, Harten and wave-speed
 * branch is the forward's (ties as the reverse pass takes them); the transpose of
 * phx_residual_vjp.  Midpoint faces only (quadrature_points 1; PHX_E_INVALID otherwise).  Uses the
 * adjoint's work arrays (allocated on the first adjoint or tangent call).  No counterpart in the
 * reference. */
int phx_residual_jvp(phx_plan* plan, const double* U, const double* dU, double* dR, void* stream);

/* Forward-mode product of one Runge-Kutta step with a frozen time step: on entry dU holds the
 * derivative of the state before the step, on return that of the state after it
 * (dU_i = dU_0 + dt sum_m a[i-1][m] (dR/dU)(U_m) dU_m).  stages, dt_dev: as for phx_step_vjp.  The
 * first call allocates nstages + 1 state arrays.  No counterpart in the reference. */
int phx_step_jvp(phx_plan* plan, const double* stages, const double* dt_dev, double* dU, void* stream);

/* Size of the static device work array of phx_residual_vjp_noalloc (per GPU). */
#define PHX_NOALLOC_SCRATCH_BYTES (32u * 1024u * 1024u)

/* phx_residual_vjp that allocates nothing: the adjoint work arrays live in one static device array
 * of PHX_NOALLOC_SCRATCH_BYTES per GPU (held by every GPU that loads the library until the process
 * ends; its address is taken once per device at plan creation) instead of arrays the plan
 * allocates on its first adjoint call.  For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  Same kernels, so the
 * result equals phx_residual_vjp's bit for bit.  Needs 16*nq + 20 doubles per cell plus 8*nq per
 * face (nq Gauss points per face): a plan that needs more than the array returns PHX_E_INVALID
 * and phx_last_error() gives the size.  Every call on a GPU uses that GPU's one array: calls that
 * overlap in time on one GPU (other streams, other threads, other plans) race and must be ordered
 * by the caller.  Nothing in the array has to start at zero (each call writes every element it
 * reads).  The plan's device must be current.  No counterpart in the reference. */
int phx_residual_vjp_noalloc(phx_plan* plan, const double* U, const double* lamR, double* lamU, void* stream);

/* Reverse-mode product of one Runge-Kutta step with a frozen time step: on entry lam holds the
 * adjoint of the state after the step, on return the adjoint of the state before it.  stages:
 * the nstages stage-input states of that step as written by phx_run_record (device,
 * nstages x state_len); dt_dev: that step's dt per case (device, ncases).  No counterpart in the
 * reference. */
int phx_step_vjp(phx_plan* plan, const double* stages, const double* dt_dev, double* lam, void* stream);

/* Tunables of the stage launch: rows per work item, persistent grid size in waves (0 = one
 * workgroup per item) and the register budget variant (waves per SIMD, 2, 3 or 4; 0 = 2; the
 * fused kernel uses at most 3); rebuilds the work list.  Used by the autotuner in
 * pyhype_amd/autotune.py.  No counterpart in the reference. */
int phx_set_launch(phx_plan* plan, int32_t rows_per_item, int32_t grid_waves, int32_t waves_per_simd);

/* Stage kernel variant: 0 = fused rolling rows (default; best for large meshes), 2 = fused with
 * workgroups of 4 waves that share their halo rows through LDS (rows_per_item rows per wave),
 * 1 = split into a
 * reconstruction + x-face kernel, a y-face kernel and a residual + update kernel, which evaluates
 * every reconstruction and face flux exactly once (best for small meshes, where short fused items
 * spend most of their work on halo rows); rows_b = rows per work item of the update kernel
 * (0 = keep); 3 = element-wise: one lane per cell and variable for the reconstruction, one lane
 * per face for the fluxes, one lane per cell for the update (no shape parameters); 4 = paired: the
 * fused kernel's work items with two waves each, one per pair of variables, the x- and y-face
 * fluxes split between them; 5 = split with the y-face and update kernels fused (each lane walks
 * rows_b cell rows of its column, evaluating the face below the first row again); 6 = the same with
 * one face per wave (workgroups of rows_b + 1 waves for rows_b <= 7 cell rows, the faces exchanged
 * through LDS); 7 = variant 6 with a reconstruction kernel over the row-major positions of each
 * block (62 owned cells or boundary faces per wave, so rows need not fill whole column strips;
 * one gradient-limiter point per face only); 8 = that reconstruction with a y-face and update
 * kernel over flat cell ranges of each block (workgroups of rows_b + 1 waves evaluate the
 * n + nx y-faces of n consecutive cells into LDS, at least rows_per_item faces per lane; blocks
 * must satisfy nx + 64 <= 2048 faces per range).  All variants
 * give identical results.  Rebuilds the work lists; the split variants allocate about 8*nq + 8
 * doubles per cell, the element-wise variant 16 per cell and 4 per face.  A plan with more than
 * one Gauss point per face runs the split variants only (1, 5 and 6; every other variant returns
 * PHX_E_INVALID).  No counterpart in the reference. */
int phx_set_variant(phx_plan* plan, int32_t variant, int32_t rows_b);

/* Order of the split and element-wise stages' work lists over the 8 XCDs: 0 = one contiguous range
 * of the list per XCD (the default: neighbouring items share halo rows in that XCD's L2); chunk > 0 =
 * chunks of `chunk` consecutive items dealt to the XCDs in turn (every XCD gets a share of every
 * region, which evens out meshes whose cost per cell varies).  The fused stage always deals its items
 * in pairs and ends each XCD's range on half-size items.  Results are identical for every order;
 * rebuilds the work lists.  Used by the autotuner.  No counterpart in the reference. */
int phx_set_xcd_order(phx_plan* plan, int32_t chunk);

/* Number of work items and the persistent grid used for the Cartesian and general kernels. */
int phx_launch_info(const phx_plan* plan, int32_t* items_cart, int32_t* items_gen,
                    int32_t* grid_cart, int32_t* grid_gen);

const char* phx_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* PYHYPE_HIP_H */
