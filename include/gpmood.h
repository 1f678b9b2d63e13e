/*
 * gpmood.h -- C ABI of the MI355X (gfx950) GP-MOOD 2D Euler step.
 *
 * The library replaces the reference's time-stepping layer (slug-cfd/gp-mood,
 * gp-mood-code/) behind plain C entry points: plain pointers, sizes and ints,
 * no C++ or torch types.  A Fortran program binds them with iso_c_binding
 * (gp-mood_amd/fortran/gpm_capi.f90); Python binds them with ctypes.
 *
 * State layout (host arrays): the reference's U(4, lb:le, nb:ne) with
 * lb = nb = 1-ngc, le = lf+ngc, ne = nf+ngc, ngc = 4 (global_variables.f90:8-26),
 * i.e. a C array double U[nf+8][lf+8][4], variable fastest (rho, rho*u, rho*v, E).
 * Order masks are int32 CellGPO(lb:le, nb:ne) (global_variables.f90:48).
 *
 * Error convention: every entry point returns 0 on success and a negative
 * GPM_E* code on failure; gpm_last_error() gives the message.  A numerically
 * inadmissible state is not an error: it propagates as in the reference and
 * MOOD lowers the order of the cells it reaches (mod_HLLC.f90:166,
 * mod_subroutine_MOOD.f90:43-51).
 *
 * Threading: one context per host thread.  Host-array calls are synchronous;
 * device-resident calls are ordered on the context's HIP stream.
 */
#ifndef GPMOOD_H
#define GPMOOD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Identifiers shared with the reference's constants.f90:16-34. */
#define GPM_FOG 1
#define GPM_GP_MOOD 2
#define GPM_POL_MOOD 200
#define GPM_UNLIM_POL 201
#define GPM_GP_SE 202
#define GPM_NN_GP_MOOD 203
#define GPM_NN_GP_MOOD_CC 204
#define GPM_FE 8
#define GPM_SSP_RK2 9
#define GPM_SSP_RK3 10
#define GPM_SSP_RK4 12
#define GPM_HLLC 100
#define GPM_LLF 101
#define GPM_HLL 102
#define GPM_UNSPLIT 103
#define GPM_NEUMANN 17
#define GPM_PERIODIC 18
#define GPM_DIRICHLET 19
#define GPM_REFLECTIVE 20
#define GPM_KH_BC 403
#define GPM_MACH800_BC 201
#define GPM_DOUBLEMACH800_BC 301
#define GPM_RMI_BC 401
#define GPM_RT_BC 402
#define GPM_DMR_BC 35        /* the reference's BC_type = DMR (the problem id) */
#define GPM_PROBLEM_RT 51    /* problem = RT: gravity source after MOOD (mod_FE.f90:183-186) */

/* Order cascade applied to a cell that fails detection. */
#define GPM_CASCADE_REFERENCE 0 /* 7->3, 5->3, 3->1 (mod_subroutine_MOOD.f90:159-163) */
#define GPM_CASCADE_FULL 1      /* 7->5->3->1 */

#define GPM_OK 0
#define GPM_EINVAL -1   /* invalid configuration or argument */
#define GPM_EHIP -2     /* HIP runtime error */
#define GPM_ENOMEM -3   /* device allocation failed */
#define GPM_EUNSUP -4   /* configuration not supported by this build */
#define GPM_ESTATE -5   /* call out of order (e.g. no state set) */

/* Runtime form of the reference's compile-time `module parameters`
 * (problems/2DRP3_GPMOOD.f90:8-46).  Field names follow that module. */
typedef struct gpm_config {
    int32_t lf, nf;       /* interior cells in x and y (>= 8 each) */
    int32_t mord;         /* Mord: 1, 3, 5 or 7 */
    int32_t ngp;          /* Gauss points per face: 1..5 */
    int32_t method;       /* GPM_GP_MOOD; GPM_NN_GP_MOOD, GPM_NN_GP_MOOD_CC (Mord 3) (copy boundary,
                             no source term; a network from gpm_set_reference_nn, or for
                             NN_GP_MOOD a learned predictor from gpm_set_predictor);
                             GPM_FOG, first-order Godunov (mod_FOG.f90; ngp = 1);
                             GPM_POL_MOOD, polynomial MOOD (mod_pol_MOOD.f90; Mord 3, ngp 2);
                             GPM_GP_SE, unlimited GP (mod_GP_SE.f90), and GPM_UNLIM_POL,
                             unlimited polynomials (mod_Unlim_Pol.f90; Mord 3, ngp 2): one
                             pass, no detection.  FOG, GP_SE and Unlim_POL: FE, SSP_RK2 or
                             SSP_RK3 */
    int32_t integrator;   /* GPM_FE, GPM_SSP_RK2, GPM_SSP_RK3 or GPM_SSP_RK4 (time_stepper.f90:17-21;
                             SSP_RK4 only with a method that has DETECTION) */
    int32_t num_flux;     /* GPM_HLLC (mod_HLLC.f90), GPM_LLF (mod_LLF.f90), GPM_HLL (mod_HLL.f90)
                             or GPM_UNSPLIT (mod_unsplit.f90) */
    int32_t bc_type;      /* copy boundaries GPM_NEUMANN, GPM_PERIODIC, GPM_REFLECTIVE, GPM_KH_BC;
                             formula boundaries GPM_DIRICHLET, GPM_MACH800_BC, GPM_DOUBLEMACH800_BC,
                             GPM_RMI_BC, GPM_DMR_BC, GPM_RT_BC (BC.f90:18-413) */
    int32_t dmp, u2, u2_tol; /* MOOD switches DMP, U2, U2_tol (logical: 0/1) */
    int32_t cascade;      /* GPM_CASCADE_* */
    int32_t problem;      /* the reference's problem id; only GPM_PROBLEM_RT changes the step */
    double cfl;           /* CFL */
    double lx, ly;        /* domain lengths Lx, Ly */
    double ell_o_dx;      /* GP length scale over dx */
} gpm_config;

/* Per-step statistics (the reference's counters in global_variables.f90:62-64
 * and the MOOD timers of mod_FE.f90:52-55,164-173, as counts). */
typedef struct gpm_stats {
    int64_t steps;                 /* steps taken since the last reset */
    int64_t count_detected_cell;   /* decreases of the last step, all stages and passes */
    int64_t detected_total;        /* decreases summed over all steps since reset */
    int32_t passes[3];             /* MOOD passes of the first three RK stages of the last step */
    int32_t max_passes;            /* largest pass count of any stage since reset */
    int64_t second_pass_cells;     /* cells recomputed after the first pass, since reset */
    int64_t a_priori_lowered;      /* cells started below Mord by the learned predictor */
    double t;                      /* simulation time */
    double dt;                     /* dt of the last step */
    /* the reference network (gpm_set_reference_nn); count_detected_cell
       includes its a priori detections, as the reference counts them */
    int64_t count_detected_cell_a_priori; /* a priori detections of the last step */
    int64_t nn_nan_passes;         /* passes of the last step whose check lowered a cell: the
                                      increments of count_steps_NN_produced_NAN (mod_FE.f90:154-157) */
    int64_t nn_nan_passes_total;   /* the same since the last reset */
    int64_t niter;                 /* niter: steps started so far (gpm_set_niter) */
} gpm_stats;

typedef struct gpm_ctx gpm_ctx;

/* Create a context on the current HIP device.  Replaces the set-up part of
 * main.f90:112-125 that the step needs (GP_presim_sphere, GP_init.f90:26-357,
 * runs on the host inside this call with the library's own quad-precision
 * precompute unless gpm_set_gp_tables() is called afterwards). */
int gpm_create(const gpm_config *cfg, gpm_ctx **out);
int gpm_destroy(gpm_ctx *ctx);
const char *gpm_last_error(const gpm_ctx *ctx);

/* GP tables in the reference's layouts (global_variables.f90:37-38,53,61):
 *   zT_sp(Mord, sz_sphere, 4, ngp), GP_d2x_sp(Mord, sz_sphere),
 *   GP_d2y_sp(Mord, sz_sphere), gauss_weight(5,5), ixiy_sp(7,25,2).
 * gpm_set_gp_tables loads tables computed by the caller (e.g. the Fortran
 * GP_presim_sphere), so that a drop-in run uses bit-identical weights.
 * gpm_gp_tables computes them on the host (GP_init.f90:26-357) for a grid of
 * lf x nf cells on an lx x ly domain (dx = lx/lf evaluated in quad precision). */
int gpm_set_gp_tables(gpm_ctx *ctx, const double *zT_sp, const double *gp_d2x_sp,
                      const double *gp_d2y_sp, const double *gauss_weight,
                      const int32_t *ixiy_sp);
int gpm_gp_tables(int32_t mord, int32_t ngp, double ell_o_dx, double lx, double ly,
                  int32_t lf, int32_t nf, double *zT_sp, double *gp_d2x_sp,
                  double *gp_d2y_sp, double *gauss_weight, int32_t *ixiy_sp);

/* Boundary_C (BC.f90:18-413) on a host array U[nf+8][lf+8][4] of the
 * configuration's grid and BC_type, at stage time t_rk (read by the DMR
 * boundary): fills the ghost cells from the interior in place.  Host code
 * only; needs no device and no context.  The kernels apply the same rule. */
int gpm_boundary_c(const gpm_config *cfg, double t_rk, double *U);

/* Host-array state transfer in the reference layout.  gpm_get_state fills the
 * ghost cells as the reference's array holds them: for copy boundaries by
 * Boundary_C (BC.f90:18-135, 383-406); for formula boundaries and problem RT
 * with the values the integrator's whole-array statements leave there (the
 * values the next Setdt reads: RK combinations of the stages' Boundary_C
 * fills and the gravity source), or, before the first step, the ghost cells
 * passed to gpm_set_state.  If cellgpo is non-NULL it receives the order mask
 * of the last RK stage. */
int gpm_set_state(gpm_ctx *ctx, const double *U);
int gpm_get_state(gpm_ctx *ctx, double *U, int32_t *cellgpo);

/* Ghost cells that the next Setdt takes its minimum over (physics.f90:98-99
 * reads lb:le x nb:ne), once, in place of the resident state's own: the
 * ghost cells of U (host array in the layout above; its interior is
 * ignored).  For the reference's restart (main.f90:124-151): InitialCdt and
 * Boundary_C fill U, read() (reader.f90:12-120) overwrites the interior only,
 * and the first Setdt sees the initial condition's ghost cells.  Call after
 * the state is set.  For formula boundaries and problem RT these become the
 * ghost values the library keeps anyway (see gpm_get_state); for copy
 * boundaries they serve the next Setdt only (the stage's Boundary_C refills
 * every ghost cell before it is read again). */
int gpm_set_dt_ghosts(gpm_ctx *ctx, const double *U);

/* Device-resident state transfer: `u_dev` points to nf*lf*4 doubles of device
 * memory holding the interior cells [n][l][k] (no ghosts).  For formula
 * boundaries gpm_set_state_device fills the ghost cells by Boundary_C at the
 * current time. */
int gpm_set_state_device(gpm_ctx *ctx, const double *u_dev);
int gpm_get_state_device(gpm_ctx *ctx, double *u_dev, int8_t *order_dev);

/* Setdt (physics.f90:78-129) on the resident state, ghost cells included:
 * dt = CFL*min(...), clipped to dtfinal. */
int gpm_compute_dt(gpm_ctx *ctx, double dtfinal, double *dt);

/* Drop-in for `subroutine time_stepping(Uin, Uout)` (time_stepper.f90:12-23:
 * Forward_Euler, SSP_RK2_, SSP_RK3_ or SSP_RK4_ by the configured integrator)
 * with the reference's implicit globals made explicit: t and dt are the module
 * variables set by main.f90 and Setdt; the order mask of the last stage is
 * written to cellgpo (may be NULL) and the counters to stats (may be NULL).
 * Uin's ghost cells are refilled as the reference's Boundary_C(Uin) of the
 * first stage does; Uout's ghost cells hold what the reference's Uout holds
 * (see gpm_get_state). */
int gpm_time_stepping(gpm_ctx *ctx, double *Uin, double *Uout, double t, double dt,
                      int32_t *cellgpo, gpm_stats *stats);

/* Device-resident driver loop of main.f90:145-162 for nsteps steps:
 * dtfinal = tmax - t; Setdt; one step of the integrator; t = t + dt.  No host
 * round trip per step (with cooperative MOOD launches, gpm_set_mood_launch).
 * Returns when the steps are done.  If dts is non-NULL it receives the
 * nsteps time steps (host memory).  A MOOD stage may run any number of
 * passes. */
int gpm_run(gpm_ctx *ctx, int32_t nsteps, double tmax, double *dts);

/* gpm_run without a context and without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  One call is gpm_create + gpm_set_gp_tables (the tables
 * in the layouts of gpm_set_gp_tables, host memory) + gpm_set_state_device +
 * gpm_set_time(t) + gpm_run(nsteps, tmax) + gpm_get_state_device: u_dev
 * (device memory, nf*lf*4 doubles, interior cells [n][l][k]) holds the state
 * on entry and the state after nsteps steps on return; t_end (host, may be
 * NULL) receives the final time.  The results equal gpm_run's bit for bit.
 * Work is enqueued on `stream` (a hipStream_t; NULL for the default stream)
 * and the call returns when it is done.
 * Every buffer of the run lives in one static device array of the library
 * (about 80 MiB, held on each GPU the library runs on until the process
 * ends), so: lf * nf <= 1048576 cells (GPM_EINVAL beyond); method GP_MOOD
 * with a copy boundary (Neumann, Periodic, Reflective, KH_BC), no source
 * term, integrator FE, SSP_RK2 or SSP_RK3 (GPM_EUNSUP otherwise).  Calls on
 * one GPU share that array: calls that overlap on one GPU (from several host
 * threads or streams) race.  Errors are reported on stderr (no context). */
int gpm_run_noalloc(const gpm_config *cfg, const double *zT_sp, const double *gp_d2x_sp,
                    const double *gp_d2y_sp, const double *gauss_weight, const int32_t *ixiy_sp,
                    double *u_dev, int32_t nsteps, double t, double tmax, double *t_end,
                    void *stream);

int gpm_get_stats(gpm_ctx *ctx, gpm_stats *stats);
int gpm_reset_stats(gpm_ctx *ctx);
int gpm_set_time(gpm_ctx *ctx, double t);

/* The HIP stream the context enqueues on (a hipStream_t, as void*). */
void *gpm_stream(gpm_ctx *ctx);

/* How the MOOD correction passes of a stage (the MOOD loop of mod_FE.f90:69-174
 * after its first pass) are launched.  cooperative = 1 (the default): one
 * persistent launch per stage runs every pass on the device, with grid
 * barriers between passes.  By default it is a cooperative launch (the
 * hardware's grid barrier).  When the device refuses one (one process at a
 * time holds that resource), or with GPM_MOOD_ORD=1 at gpm_create, it is an
 * ordinary launch whose workgroups first meet at a bounded rendezvous (if
 * they are not all resident together, one workgroup runs every pass alone)
 * and then use grid barriers of their own; a barrier wait that reaches its
 * bound (about a second) fails the call with GPM_EHIP.
 * cooperative = 0: per-pass launches in growing batches, with the stage's
 * pending flag read back between batches (one host synchronisation per
 * batch); the environment variable GPM_MOOD_COOP=0 selects them at
 * gpm_create.  The results are the same bit for bit.
 * With persistent launches, a stage's first passes over long lists (more
 * cells than one trip of the persistent grid) run in two ordinary launches
 * of the whole GPU ahead of the persistent one, while the context's lists
 * are long: each gpm_run or gpm_time_stepping call sets the next call's
 * plan from the longest first list it saw (GPM_MOOD_WIDE=0 at gpm_create
 * switches them off).  The plan changes no result.
 * No counterpart in the reference (its MOOD loop is a host loop). */
int gpm_set_mood_launch(gpm_ctx *ctx, int32_t cooperative);
int gpm_get_mood_launch(gpm_ctx *ctx, int32_t *cooperative);

/* Launch shape and timing of the dominant kernel, for roofline reporting:
 * average duration in ms of the first-pass stage kernel over the launches
 * since the last gpm_reset_stats (0 if timing is off), and launch count.
 * gpm_set_timing: 0 off; 1 HIP events around the stage and MOOD launches of
 * every step; k > 1 around those of every k-th step from now on (a sample:
 * each event pair costs the stream a few microseconds). */
int gpm_set_timing(gpm_ctx *ctx, int32_t enable);
int gpm_kernel_time(gpm_ctx *ctx, double *avg_stage_ms, int64_t *stage_launches,
                    double *avg_mood_ms, int64_t *mood_launches);

/* ------------------------------------------------------------------------
 * Learned a priori order predictor (NN-GP-MOOD).
 *
 * An MLP  in -> 128 -> 128 -> 64 -> C  (ReLU; BF16 weights and activations,
 * FP32 accumulation) on each cell's (2R+1)^2 x 4 conserved stencil, normalised
 * per variable by the stencil max |value|; class c means order 2c+1
 * (C = (Mord+1)/2).  It replaces the reference's 52-input sigmoid network
 * (mod_NN.f90:14-58,168-183); the reference's own role -- CellGPO chosen a
 * priori, MOOD a posteriori check kept -- is unchanged, except that the full
 * PAD + DMP check runs after the prediction (the reference kept PAD only,
 * mod_subroutine_NN_MOOD.f90:13-101).  Parameters are exchanged in PyTorch
 * Linear order (W0 [128][in], b0, W1, b1, W2, b2, W3 [C][64], b3), as the
 * reference's exporter writes them (neural_network/save_NN_for_fortran.py:12-25).
 * Training (forward, backward, AdamW) runs on the device; x / y buffers below
 * are device pointers (float [n][in], int32 [n]).
 * ---------------------------------------------------------------------- */
typedef struct gpm_mlp gpm_mlp;

int gpm_mlp_create(int32_t in_features, int32_t n_classes, int32_t max_batch, gpm_mlp **out);
int gpm_mlp_destroy(gpm_mlp *m);
const char *gpm_mlp_last_error(const gpm_mlp *m);
int64_t gpm_mlp_num_params(const gpm_mlp *m);
int gpm_mlp_set_params(gpm_mlp *m, const float *params);
int gpm_mlp_get_params(gpm_mlp *m, float *params);
int gpm_mlp_get_grads(gpm_mlp *m, float *grads);
/* forward + mean cross-entropy + backward; gradients stay on the device.
 * loss / accuracy may be NULL: then nothing is read back and the call only
 * enqueues work on the network's stream (no synchronisation). */
int gpm_mlp_forward_backward(gpm_mlp *m, const float *x, const int32_t *y, int32_t batch, double *loss,
                             double *accuracy);
/* AdamW on the FP32 master weights (torch.optim.AdamW semantics) */
int gpm_mlp_adamw_step(gpm_mlp *m, double lr, double beta1, double beta2, double eps, double weight_decay);
int gpm_mlp_train_step(gpm_mlp *m, const float *x, const int32_t *y, int32_t batch, double lr,
                       double weight_decay, double *loss, double *accuracy);
/* class logits [n][C] for n feature rows */
int gpm_mlp_logits(gpm_mlp *m, const float *x, int32_t n, float *logits);

/* Attach a predictor to a solver (NULL detaches).  Each RK stage then starts
 * from the predicted orders; gate = 1 sends only cells where the a posteriori
 * check can fail (the shock detector fires, or near vacuum) to the network and
 * starts every other cell at Mord; gate = 0 predicts every cell. */
int gpm_set_predictor(gpm_ctx *ctx, gpm_mlp *m, int32_t gate);
/* Training-data dump (append_to_NN_datatset, mod_append_NN_dataset.f90:11-79):
 * after each RK stage in stage_mask (bit s-1 for stage s), for sampled cells,
 * the features of the stage input and the stage's final order class.  Cells
 * below Mord are always kept, cells at Mord with probability keep_mord;
 * count_dev (uint64) is incremented past capacity when the buffer is full.
 * x_dev = NULL switches the dump off. */
int gpm_set_dump(gpm_ctx *ctx, float *x_dev, int32_t *y_dev, int64_t capacity, uint64_t *count_dev,
                 double keep_mord, uint64_t seed, int32_t gated_only, int32_t stage_mask);
/* features of the listed cells of the resident state (device buffers) */
int gpm_features(gpm_ctx *ctx, const int32_t *cells_dev, int32_t n, float *x_dev);
/* Fixed a priori orders (int8 [nf][lf], odd, 1..Mord; host memory) used as the
 * starting CellGPO of every RK stage in place of Mord -- the hook the
 * reference fills from its network (mod_FE.f90:51-56).  NULL restores Mord.
 * A predictor attached with gpm_set_predictor takes precedence. */
int gpm_set_apriori_orders(gpm_ctx *ctx, const int8_t *orders);
/* a priori orders of the last stage (int8, device) */
int gpm_get_apriori_orders(gpm_ctx *ctx, int8_t *pred_dev);

/* ------------------------------------------------------------------------
 * The reference's own network (methods NN_GP_MOOD and NN_GP_MOOD_CC with a
 * network trained by neural_network/train_function.py and written by
 * save_NN_for_fortran.py).  From step nsteps_with_no_NN + 1 = 11 on
 * (global_variables.f90:97; niter as main.f90:147 counts it), every RK stage
 * predicts each interior cell's order from its R+1 stencil in FP32
 * (compute_CellGPO_with_NN, mod_NN.f90:14-58) and checks only PAD afterwards
 * (NN_DETECTION, mod_subroutine_NN_MOOD.f90:13-101); NN_GP_MOOD_CC also blends
 * the order-3 reconstruction (mod_NN_GP_MOOD_CC.f90:11-85).  Steps 1-10 are
 * plain GP-MOOD.  Mord 3 or 5; L >= 4*szs_sphere(Mord+2) (52 for Mord 3), L <=
 * 128, length <= 64.
 *
 * gpm_set_reference_nn replaces the arrays load_NN fills (global_variables.f90:
 * 74-80) and takes them in their Fortran layouts: weight0(length, L),
 * bias0(length), weight1(length, length), bias1(length), weight2(2, length),
 * bias2(2).  gpm_load_reference_nn is load_NN itself (mod_NN.f90:60-117): it
 * reads "<nn_filename>.txt" (blanks trimmed), one number per record, each
 * weight matrix row by row. */
int gpm_set_reference_nn(gpm_ctx *ctx, int32_t L, int32_t length, const float *weight0, const float *bias0,
                         const float *weight1, const float *bias1, const float *weight2, const float *bias2);
int gpm_load_reference_nn(gpm_ctx *ctx, const char *nn_filename, int32_t L, int32_t length);
/* The reading part of gpm_load_reference_nn alone (host only, no context):
 * fills caller arrays in the Fortran layouts above.  GPM_EINVAL for a missing
 * file, a record that is not a number or fewer records than the network holds
 * (where the reference's read stops with an end-of-file error). */
int gpm_read_reference_nn(const char *nn_filename, int32_t L, int32_t length, float *weight0, float *bias0,
                          float *weight1, float *bias1, float *weight2, float *bias2);
/* Training samples for the reference's trainer (append_to_NN_datatset,
 * mod_append_NN_dataset.f90:11-79, called after the first RK stage of every
 * step, mod_FE.f90:176-181): each interior cell's R+1 stencil of the stage
 * input through format_input (inputs_dev, float [capacity][L], -666 past the
 * stencil) and the stage's final order as the label (labels_dev, float
 * [capacity][2]: (0, 1) for order 3, (1, 0) otherwise).  Constant stencils are
 * skipped; order-3 samples are kept with probability NR0/(NR0+NR1) (0.5 at
 * first) from the running counts nr_dev (int64 NR0, NR1) at the start of the
 * stage, drawn from a counter-based generator of (seed, step, cell): the
 * reference draws Fortran random_number per sample, so the selection is not
 * the reference's, the samples' contents are.  Kept samples take row
 * atomicAdd(*count_dev, 1) while it is below capacity; cells_dev (int64, may
 * be NULL) receives each row's cell id j*lf + i.  The order of rows is not
 * defined.  inputs_dev = NULL switches it off.  Write the file with
 * gpm_io_write_nn_dataset (include/gpmood_io.h). */
int gpm_set_reference_dump(gpm_ctx *ctx, float *inputs_dev, float *labels_dev, int64_t *cells_dev, int64_t capacity,
                           uint64_t *count_dev, int64_t *nr_dev, uint64_t seed, int32_t L);
/* The same samples for the reference's own dataset arrays, as
 * append_to_NN_datatset writes them with write_NN_dataset = .true.
 * (mod_append_NN_dataset.f90:11-79, global_variables.f90:117-120), for the
 * host-array drop-in.  gpm_set_reference_dataset switches them on (L > 0, the
 * input width) with the running counts NR0, NR1 (L = 0 switches them off);
 * every gpm_time_stepping call then keeps its first stage's rows on the
 * device, and gpm_append_reference_dataset, called after the step, appends
 * them to inputs(dataset_size, L) and labels(dataset_size, 2) (real(4),
 * Fortran layouts): in cell order (n outer, l inner) at row *index (1-based),
 * wrapping to 1 with *n_overwrite + 1 after dataset_size rows; *nr0 / *nr1
 * count the rows labelled (1,0) / (0,1) and set the next step's order-3 keep
 * probability NR0/(NR0+NR1).  Which order-3 rows are kept is drawn as in
 * gpm_set_reference_dump (not the reference's random_number); every other
 * row is the reference's.  The next step discards rows not appended.
 * gpm_run refuses to run while the samples are on (GPM_ESTATE). */
int gpm_set_reference_dataset(gpm_ctx *ctx, int32_t L, int32_t nr0, int32_t nr1);
int gpm_append_reference_dataset(gpm_ctx *ctx, float *inputs, float *labels, int32_t dataset_size,
                                 int32_t *index, int32_t *n_overwrite, int32_t *nr0, int32_t *nr1);
/* niter before the next step (main.f90:129,147): every step first advances it.
 * The Fortran time_stepping drop-in sets niter - 1, main having advanced it. */
int gpm_set_niter(gpm_ctx *ctx, int64_t niter);
/* CellGPO_priori(lb:le, nb:ne) as int32 (global_variables.f90:59) after the
 * last step: 1 for a priori detections, 3 for constant stencils, Mord
 * elsewhere and on ghost cells (Mord everywhere after a step without the
 * network). */
int gpm_get_cellgpo_priori(gpm_ctx *ctx, int32_t *priori);
/* The network's FP32 exp on the floats with bit patterns first .. first+n-1,
 * into y_dev (device memory): the expf of the host C library, for tests. */
int gpm_nn_expf_bits(gpm_ctx *ctx, uint32_t first, int64_t n, float *y_dev);
/* the same function compiled for the host, into host memory y */
int gpm_nn_expf_bits_host(uint32_t first, int64_t n, float *y);
/* The step's FP64 division a / b (correctly rounded, as the reference
 * build's IEEE division) on n pairs in device memory, for tests.  No
 * counterpart in the reference.  Synchronous. */
int gpm_ddiv(int64_t n, const double *a_dev, const double *b_dev, double *q_dev);
/* The same quotients through the form the Riemann solvers use for the
 * divisions by one state's density: the divisor's reciprocal computed once
 * and shared (range-guarded, falling back to the full division).  No
 * counterpart in the reference.  Synchronous. */
int gpm_qdiv(int64_t n, const double *a_dev, const double *b_dev, double *q_dev);

/* ------------------------------------------------------------------------
 * Differentiable form.  No counterpart in the reference: the adjoint of K
 * SSP-RK3 steps of GP reconstruction + HLLC + update with the MOOD order mask
 * and dt of every RK stage frozen from a forward run, for sensitivities of an
 * output functional with respect to the initial state.
 *
 * gpm_record_masks: the next max_steps steps taken by gpm_run or
 * gpm_time_stepping write the final CellGPO of each RK stage (int8 orders,
 * interior, [step][stage][nf][lf]) to masks_dev, device memory of
 * max_steps*3*lf*nf bytes.  NULL stops recording.  gpm_recorded_steps reports
 * how many steps were recorded since recording last started.
 *
 * gpm_frozen_run: advances the resident state nsteps steps with the given
 * per-stage orders (no a posteriori detection) and the given dt per step (host
 * array).  With masks recorded from a run, it reproduces that run's states bit
 * for bit.  If stage_states_dev is non-NULL it receives the input of every RK
 * stage, [step][stage][nf][lf][4] doubles.
 *
 * gpm_adjoint: ubar_dev (interior layout, nf*lf*4 doubles, device) holds dJ/dU
 * after nsteps steps on entry and dJ/dU before them on return, for the steps
 * whose stage inputs, masks and dts are given.  Deterministic (no atomics).
 * Supports (Mord, ngp) = (3,2), (5,3), (7,4) with numFlux = HLLC.
 *
 * gpm_hllc_vjp: reverse mode of the HLLC flux (mod_HLLC.f90:11-169) on n state
 * pairs in device memory ([n][4] each): ulbar = (dF/dul)^T fbar, urbar =
 * (dF/dur)^T fbar, for x-faces (dir 0) or y-faces (dir 1).  Synchronous. */
int gpm_record_masks(gpm_ctx *ctx, int8_t *masks_dev, int32_t max_steps);
int gpm_recorded_steps(gpm_ctx *ctx, int32_t *nsteps);
int gpm_frozen_run(gpm_ctx *ctx, const int8_t *masks_dev, const double *dts, int32_t nsteps,
                   double *stage_states_dev);
int gpm_adjoint(gpm_ctx *ctx, const double *stage_states_dev, const int8_t *masks_dev, const double *dts,
                int32_t nsteps, double *ubar_dev);
int gpm_hllc_vjp(int32_t dir, int32_t n, const double *ul, const double *ur, const double *fbar,
                 double *ulbar, double *urbar);

#ifdef __cplusplus
}
#endif

#endif /* GPMOOD_H */
