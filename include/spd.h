/*
 * spd.h -- C ABI of the MI355X-native SD-ADER time step (libspd_hip.so).
 *
 * One context describes one simulation (dimension, order, mesh, options) and
 * owns every scratch buffer of the step.  The caller owns the state arrays;
 * all of them are FP64 device pointers in the reference's layouts:
 *
 *   element layout  (V, [Nz,] [Ny,] Nx, [nz,] [ny,] nx)   U_sp, U_cv, W_cv
 *
 * with V = ndim + 2 variables ordered (rho, v_x[, v_y[, v_z]], P) for
 * primitives and (rho, m_x[, m_y[, m_z]], E) for conservatives, and
 * n = p + 1 solution points per dimension.  All calls are asynchronous on the
 * given HIP stream; nothing in spd_step or spd_compute_dt synchronises the
 * host.  Functions return 0 on success or a negative SPD_E* code, and
 * spd_last_error() describes the last failure of the calling thread.  A
 * context is not re-entrant: calls that overlap on one context race.
 *
 * Each entry point cites the reference interface it replaces
 * (paths relative to the reference repository jpalafou/spd).
 */
#ifndef SPD_H
#define SPD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPD_ABI_VERSION 5

enum spd_status {
  SPD_OK = 0,
  SPD_E_INVALID = -1,       /* invalid or unsupported configuration / argument */
  SPD_E_UNSUPPORTED = -2,   /* option the reference has but this build does not */
  SPD_E_HIP = -3,           /* HIP runtime error (allocation, launch) */
  SPD_E_NOMEM = -4
};

enum spd_riemann { SPD_RS_LLF = 0, SPD_RS_HLLC = 1, SPD_RS_LHLLC = 2 };
enum spd_limiter { SPD_LIMITER_MINMOD = 0, SPD_LIMITER_MONCEN = 1 };
/* Boundary kinds (src/sd_boundary.py:32-58, src/fv_simulator.py:224-253):
 * "ic" and "pressure" sides keep the values stored at construction (the
 * element traces of the initial state and the initial condition's ghost
 * layers, spd_set_boundary_values); a pressure side also replaces the boundary
 * element's own trace with the stored one.  A pressure side opposite a
 * periodic side is rejected (the reference's result there depends on the
 * order in which it stores the two sides). */
enum spd_bc {
  SPD_BC_PERIODIC = 0,
  SPD_BC_GRADFREE = 1,
  SPD_BC_REFLECTIVE = 2,
  SPD_BC_IC = 3,
  SPD_BC_PRESSURE = 4
};
enum spd_update { SPD_UPDATE_SD = 0, SPD_UPDATE_FV = 1 };
enum spd_nad { SPD_NAD_RELATIVE = 0, SPD_NAD_DELTA = 1 };
/* Arithmetic of the step.  FP64: the reference's arithmetic (parity form).
 * FP32: predictor, interpolation, element-face and fallback Riemann solvers and
 * MUSCL-Hancock in FP32, with FP64 kept for the state, the conservative update
 * accumulation, the pressure recovery from total energy, trouble detection
 * and the dt reduction (an extension; no reference counterpart). */
enum spd_precision { SPD_PRECISION_FP64 = 0, SPD_PRECISION_FP32 = 1 };

/* Constructor keywords of SDADER_Simulator (src/sdader_simulator.py:15-27,
 * src/sd_simulator.py:20-26, src/fv_simulator.py:11-17, src/simulator.py:10-36)
 * as a plain struct.  Unsupported reference options (viscosity, potential,
 * well-balancing, isothermal, the "eq" side, predictor=False) have no field;
 * the Python facade rejects them.  The slope limiters (minmod, moncen) are
 * `slope_limiter`; the Riemann solvers (llf, hllc, lhllc) are `riemann_sd` /
 * `riemann_fv`. */
typedef struct spd_config {
  int ndim;               /* 1, 2 or 3 */
  int N[3];               /* elements per dimension (x, y, z) */
  int p;                  /* polynomial degree (solution points p+1), 1..5 */
  int m;                  /* ADER order - 1; must equal p in this build */
  double lim[3][2];       /* domain bounds per dimension */
  double gamma;
  double cfl_coeff;
  double min_c2;
  int update;             /* spd_update */
  int FB;                 /* a-posteriori fallback on/off */
  int riemann_sd;         /* spd_riemann of the element faces */
  int riemann_fv;         /* spd_riemann of the fallback faces */
  int bc[3][2];           /* spd_bc per dimension and side (left, right) */
  double tolerance;       /* NAD tolerance */
  int NAD;                /* spd_nad */
  int SED;
  int PAD;
  int blending;
  double min_rho, max_rho, min_P;
  int godunov;
  int precision;          /* spd_precision */
  int slope_limiter;      /* spd_limiter of the MUSCL-Hancock fallback */
} spd_config;

typedef struct spd_ctx spd_ctx;

/* Message of the last failed call on this thread ("" if none). */
const char* spd_last_error(void);

/* ABI version of the library (SPD_ABI_VERSION) and sizeof(spd_config), for
 * bindings to check their struct mirror. */
int spd_abi_version(void);
size_t spd_config_size(void);

/* Host-only (no GPU needed): the operator matrices of order (p, m) in the
 * layout of spd_get_operators, and the ghosted subcell geometry of one
 * dimension with N elements of width h -- face positions (N(p+1)+5), centres
 * (N(p+1)+4), widths (N(p+1)+4), centre spacings (N(p+1)+3), concatenated
 * (src/polynomials.py:5-122, src/sdader_simulator.py:40-92).  Return the
 * number of doubles written or a negative code. */
int spd_host_operators(int p, int m, double* host_out, int len);
int spd_host_geometry(int N, double h, int p, double* host_out, int len);

/* Build the operator matrices (src/polynomials.py:5-122,
 * src/sd_simulator.py:44-52, src/sdader_simulator.py:40-46), the subcell
 * geometry (src/sdader_simulator.py:61-92) and allocate the step's scratch on
 * the current HIP device.  Replaces the constructor of SDADER_Simulator. */
int spd_ctx_create(const spd_config* cfg, spd_ctx** out);
int spd_ctx_destroy(spd_ctx* ctx);

/* Bytes of device scratch the context holds. */
size_t spd_ctx_scratch_bytes(const spd_ctx* ctx);

/* Copy the host operator matrices: sp_to_fp (nf x n), dfp_to_sp (n x nf),
 * sp_to_cv (n x n), cv_to_sp (n x n), invader (T x T), w_tp (T), x_tp (T),
 * row-major, concatenated in that order.  `len` is the capacity in doubles;
 * returns the number of doubles written or a negative code. */
int spd_get_operators(const spd_ctx* ctx, double* host_out, int len);

/* dt = cfl * (h_min/(p+1)) / max(ndim*c_s + sum|v|) over the subcell-mean
 * primitives (src/sd_simulator.py:214-227).  With W_cv != NULL the maximum is
 * taken over that array (element layout); with W_cv == NULL the maximum the
 * last spd_step recorded for its new state is used (the value the reference
 * computes from its post-step W_cv).  Writes dt to dt_dev.  The clamp of dt
 * to reach t_end and the stop when it falls below 1e-14
 * (src/sdader_simulator.py:404-410) are not applied here: the caller's loop
 * applies them (the Python facade's perform_time_evolution), and the device
 * loop applies them on the device (spd_run_begin with clamp = 1). */
int spd_compute_dt(spd_ctx* ctx, const double* W_cv, double* dt_dev, void* stream);

/* spd_compute_dt with W_cv != NULL for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): no context, and no allocation of any kind.  No
 * counterpart in the reference.  The kernels are spd_compute_dt's; their
 * scratch (per-block maxima) is a static device array of 1 MiB + 512 B per
 * GPU, held until the process ends, which limits W_cv to 2^25 = 33,554,432
 * solution points (N_x N_y N_z (p+1)^ndim); larger inputs are rejected with
 * SPD_E_INVALID before anything is launched.  cfg supplies ndim, N, p, lim,
 * gamma, min_c2 and cfl_coeff.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (on different streams or host threads) race. */
int spd_compute_dt_static(const spd_config* cfg, const double* W_cv, double* dt_dev,
                          void* stream);

/* Exchange the maximum wave speed the context holds for its last state (the
 * value spd_compute_dt(ctx, NULL, ...) and the device loop use; spd_step and
 * spd_compute_dt with W_cv != NULL overwrite it) with the device scalar
 * *slot.  Lets a second user of one context keep its own value: the adjoint
 * swaps its own in around its calls so that the simulator's next dt is
 * unchanged.  An extension with no reference counterpart (the reference
 * recomputes dt from dm.W_cv, src/sd_simulator.py:214-227). */
int spd_swap_wave_max(spd_ctx* ctx, double* slot, void* stream);

/* The stored boundary values of "ic" / "pressure" sides, set once before the
 * first step (init_sd_Boundaries / init_fv_Boundaries,
 * src/sdader_simulator.py:415-423, src/fv_simulator.py:213-222): the library
 * takes the element traces of U_sp (the initial state, element layout) at
 * the domain's first and last flux points of every dimension, and copies
 * W_ghosted, the primitive subcell means of the initial condition on the
 * ghosted FV grid (V, Ns_z+4, Ns_y+4, Ns_x+4) (x fastest; two ghost layers per
 * side with the corners, from the reference's element mesh with one ghost
 * element per side; only ghost cells are read).  spd_step and spd_run_steps
 * refuse a context with such a side until this is called; the adjoint does
 * not support them. */
int spd_set_boundary_values(spd_ctx* ctx, const double* U_sp, const double* W_ghosted,
                            void* stream);

/* One step: ADER predictor + SD or FV(/FB) corrector
 * (SDADER_Simulator.perform_update, src/sdader_simulator.py:361-378).
 * U_sp is updated in place.  If `snapshot` is non-zero, U_cv and W_cv receive
 * the post-step subcell means and primitives in element layout (the values
 * perform_update leaves in dm.U_cv / dm.W_cv); otherwise they are not touched
 * and may be NULL.  dt_dev points at the step's dt on the device. */
int spd_step(spd_ctx* ctx, double* U_sp, double* U_cv, double* W_cv, const double* dt_dev,
             int snapshot, void* stream);

/* Device-side time loop (SDADER_Simulator.perform_time_evolution /
 * perform_iterations, src/sdader_simulator.py:394-413).  spd_run_begin sets
 * the device run state: current time and step count, t_end, a step budget
 * (max_steps < 0: unlimited) and whether dt is clamped to reach t_end (with
 * the reference's stop when the clamped dt < 1e-14).  spd_run_steps enqueues
 * `nsteps` iterations of [dt from the last recorded maximum wave speed, clamp,
 * spd_step, time += dt]; once the run is done the remaining iterations' kernels
 * return at once, so a fixed chunk of steps can be captured in a HIP graph
 * and replayed until spd_run_status reports done.  The first dt of a run uses
 * the maximum recorded by spd_compute_dt (call it with the initial W_cv) or by
 * the last step.  spd_run_status synchronises the stream and reads the state. */
int spd_run_begin(spd_ctx* ctx, double time, double t_end, long long n_step, long long max_steps,
                  int clamp, void* stream);
int spd_run_steps(spd_ctx* ctx, double* U_sp, int nsteps, void* stream);
int spd_run_status(spd_ctx* ctx, double* time, long long* n_step, int* done, double* dt,
                   void* stream);

/* Batched steps of several 2-D simulations of one order (an extension with
 * no reference counterpart, for ensembles of small problems, where one launch
 * per kernel and simulation costs more than the work).  spd_batch_run_steps
 * advances every member by nsteps steps of its device-side loop, as
 * spd_run_steps does for one context, with one launch per stage of the step
 * for all members: the grid is the members' grids one after the other, and
 * a block runs the kernel for the member its index falls in.  Members: 2-D,
 * update = FV with the fallback, the same order, precision, Riemann solvers
 * and kinds of sides (sizes, domains, states and parameters may differ); each
 * member's state U_sp (a device array in the element layout) is fixed when
 * the batch is made, and its device loop is begun with spd_run_begin before
 * the steps (spd_run_status reads each member's time and step count).  A
 * member's results equal those of spd_run_steps on its context with the
 * substep form 2 (SPD_LAUNCH_FB_FORM), bit for bit.  The batch holds device
 * launch tables (a few hundred bytes per member and stage); its members must
 * outlive it; calls that overlap on a batch or on one of its members race. */
typedef struct spd_batch spd_batch;
int spd_batch_create(spd_ctx** members, double** U_sp, int count, spd_batch** out);
int spd_batch_run_steps(spd_batch* batch, int nsteps, void* stream);
int spd_batch_destroy(spd_batch* batch);

/* Launch shapes that the kernels take at run time (an extension with no
 * reference counterpart; the autotuner of spd_amd.ensemble sets them):
 * elements per block of the predictor update and of the 2-D high-order
 * face-flux kernel of form 1, and the form of the 2-D FV substep
 * (1: k_ho2d writes the face fluxes, k_apply the candidate, k_detect2d the
 * flags, the fallback corrects cell tiles staged in LDS; 2: one kernel forms
 * the face fluxes, the candidate and the first half of the detection (NAD,
 * PAD), k_sed2d the smooth-extrema half, the fallback marches column strips
 * in registers over the face-flux arrays; 3: the same without the face-flux
 * arrays, the fallback runs on the compacted list of active tiles and forms
 * their face fluxes again; 4, the default: per substep the path of 2 or of
 * 3, chosen on the device from the active tiles of the previous substep;
 * forms 2, 3 and 4 give the same bits, form 1 agrees with them to round-off,
 * so the autotuner leaves the form alone), and the elements per block of the 2-D
 * fused high-order flux + candidate kernel (at most (256 / (p+1)^2), the
 * default; fewer give small grids more blocks; the results do not depend on
 * it).  value 0 restores the default; values above the maximum (256 threads /
 * 48 KiB of LDS per block; 4 forms) are rejected.  spd_get_launch reports the
 * value in effect and the maximum. */
enum spd_launch_key {
  SPD_LAUNCH_PRED_EPB = 0,
  SPD_LAUNCH_HO_EPB = 1,
  SPD_LAUNCH_FB_FORM = 2,
  SPD_LAUNCH_HOC_EPB = 3,
  SPD_LAUNCH_COUNT = 4
};
int spd_set_launch(spd_ctx* ctx, int key, int value);
int spd_get_launch(const spd_ctx* ctx, int key, int* value, int* max_value);

/* Initial condition on the device (SD_Simulator.post_init / quadrature_mean,
 * src/sd_simulator.py:59-90, src/polynomials.py:124-148).  spd_ic_points
 * writes the 2-point Gauss-Legendre nodes of every subcell of the active
 * elements, layout (ndim, E_{d-1}..E_0, l_{d-1}..l_0, q_{d-1}..q_0), from the
 * subcell lower bounds and widths of the reference's element mesh (N_a*(p+1)
 * values per dimension, concatenated x first; device arrays).  The caller
 * evaluates the IC at those points (one variable at a time) and spd_ic_mean
 * reduces the values (E.., l.., q..) to subcell means (E.., l..).
 * spd_sp_to_regular interpolates element-layout values to p+1 equispaced
 * subcell centres (SD_Simulator.regular_mesh, src/sd_simulator.py:92-102). */
int spd_ic_points(spd_ctx* ctx, const double* lows, const double* widths, double* pts,
                  void* stream);
int spd_ic_mean(spd_ctx* ctx, const double* vals, double* W_cv_var, void* stream);
int spd_sp_to_regular(spd_ctx* ctx, const double* in, double* out, void* stream);

/* Element-layout tensor-product transforms (src/transforms.py:27-45 via
 * SD_Simulator.compute_cv_from_sp / compute_sp_from_cv, src/sd_simulator.py:184-188). */
int spd_sp_to_cv(spd_ctx* ctx, const double* in, double* out, void* stream);
int spd_cv_to_sp(spd_ctx* ctx, const double* in, double* out, void* stream);

/* Pointwise conservative -> primitive on `count` points of a (V, count) array
 * (src/hydro.py:48-86). */
int spd_cons_to_prims(spd_ctx* ctx, const double* U, double* W, long long count, void* stream);

/* Pointwise primitive -> conservative on `count` points of a (V, count) array
 * (src/hydro.py:88-125). */
int spd_prims_to_cons(spd_ctx* ctx, const double* W, double* U, long long count, void* stream);

/* Debug tap for flag parity: during the next spd_step, after the trouble
 * detection of FV substep `substep`, copy the troubled-cell flags
 * (Ns_z x Ns_y x Ns_x doubles), the blended coefficients theta on the ghosted
 * grid ((Ns+4)^ndim doubles) and, per dimension, the face coefficients
 * (the reference's dm.troubles, dm.theta and dm.affected_faces_{x,y,z},
 * src/trouble_detection.py:5-87) into the given device buffers.  Pass
 * substep = -1 to disable, or SPD_TAP_ALL to record every substep (the buffers
 * then hold m+1 consecutive blocks).  Any buffer may be NULL. */
#define SPD_TAP_ALL (-2)
int spd_set_debug_tap(spd_ctx* ctx, int substep, double* troubles, double* theta,
                      double* aff_x, double* aff_y, double* aff_z);

/* Reverse-mode derivative (adjoint) of one step, both precision forms (the
 * reverse sweep is FP64; in the FP32-bulk form the tape holds the form's FP32
 * predictor iterates).  No entry of the reference corresponds: the reference
 * has no derivative.  These
 * differentiate perform_update (src/sdader_simulator.py:361-378: the predictor
 * src/sdader_simulator.py:160-200,213-240, the FV corrector :288-349 with
 * src/fv_simulator.py:167-205, src/muscl.py:129-221 and the blend :325-331)
 * with the trouble flags, blending coefficients and dt frozen.
 *
 * spd_adj_step runs one step from U_in into U_out (U_in is not modified) and
 * records its tape in the context: the m Picard iterates, the FV state before
 * each substep and each substep's blending coefficients.  theta_in != NULL
 * replays coefficients recorded earlier (spd_adj_theta_count doubles: m+1
 * substeps of the ghosted grid (Ns+4)^ndim) instead of detecting them;
 * theta_out != NULL receives the coefficients used.  Both need update FV with
 * the fallback.
 *
 * spd_adj_backward takes the cotangent of the step's output in U_bar and
 * replaces it with the cotangent of its input, U_bar^T d(step)/d(U_in), by the
 * reverse sweep over the tape of the last spd_adj_step, which must have been
 * run from the same U_in and dt.  U_bar must be ordinary (coarse-grained)
 * device memory, as hipMalloc returns: the sweep accumulates into it with
 * hardware FP64 atomics.  The tape (m (m+1) V E L values of the bulk
 * type + (m+1)(V Nc + Ngc) doubles) and the adjoint scratch are allocated on
 * first use. */
int spd_adj_step(spd_ctx* ctx, const double* U_in, double* U_out, const double* dt_dev,
                 const double* theta_in, double* theta_out, void* stream);
int spd_adj_backward(spd_ctx* ctx, const double* U_in, const double* dt_dev, double* U_bar,
                     void* stream);
size_t spd_adj_theta_count(const spd_ctx* ctx);

/* Optional HIP-event timing of the step's kernels (no counterpart in the
 * reference, which only times whole runs with a wall clock,
 * src/sdader_simulator.py:383,387).  spd_profile_begin makes the following
 * calls record an event pair around each of at most `max_launches` kernel
 * launches on their own stream; spd_profile_end waits for the last event and
 * returns, per kernel class k < n_classes, the summed duration in ms and the
 * launch count (return value: launches recorded).  Class names:
 * spd_kernel_class_name(k) for k < spd_kernel_class_count(). */
int spd_profile_begin(spd_ctx* ctx, int max_launches);
int spd_profile_end(spd_ctx* ctx, double* ms_out, long long* count_out, int n_classes);
int spd_kernel_class_count(void);
const char* spd_kernel_class_name(int k);

#ifdef __cplusplus
}
#endif
#endif /* SPD_H */
