/*
 * topodda.h -- C ABI of the MI355X DDA solve engine (libtopodda_hip.so).
 *
 * Drop-in boundary for the Aproduct + BiCGSTAB path of Topo-DDA-forWin64.
 * Reference paths are relative to its Topo-DDA-forWin64/ directory.
 *
 *   reference                                          this ABI
 *   AProductCore::AProductCore  AProductCore.cpp:12-249   tdda_op_create
 *     (periodic overload)       AProductCore.cpp:251-536  tdda_op_create_periodic
 *     (AMatrixMethod "FCD")     AProductCore.cpp:625-695  tdda_op_create_ex, tdda_green_tensor_fcd
 *   AProductCore::~AProductCore AProductCore.cpp:538-558  tdda_op_destroy
 *   AProductCore::Aproduct      AProductCore.cpp:697-758  tdda_aproduct
 *   AProductCore::LDR_inter     AProductCore.cpp:588-623  tdda_green_tensor
 *   DDAModel state (P, E, al)   DDAModel.h:14-30          tdda_model_create/_set/_get
 *   DDAModel::Aproductwithalb   DDAModel.cpp:577-586      tdda_model_apply
 *   DDAModel::bicgstab          DDAModel.cpp:131-260      tdda_model_solve
 *   DDAModel ctor (E, al)       DDAModel.cpp:12-69        tdda_incident_field, tdda_polarizability
 *   Get_Alpha                   tools.cpp:617-630         tdda_get_alpha
 *   update_E_in_structure       DDAModel.cpp:567-575      tdda_local_field
 *   CoreStructure::UpdateStr    CoreStructure.cpp:31-104  tdda_design_step (with
 *     + DDAModel::UpdateAlpha   DDAModel.cpp:470-484       tdda_design_create/_get)
 *   FilterOption::update_beta   FilterOption.cpp:24-46    tdda_beta_update
 *   DDAModel::solve_E           DDAModel.cpp:530-565      tdda_field_at_points
 *   DDAModel::output_to_file    DDAModel.cpp:589-697      tdda_write_model_results
 *   CoreStructure::output_to_file CoreStructure.cpp:155-190 tdda_write_structure
 *   ObjPointEDDAModel           ObjDDAModel.cpp:20-97     tdda_objective_point_e (value and
 *   ObjIntegratedEDDAModel      ObjDDAModel.cpp:100-177   tdda_objective_integrated_e  the
 *     + EvoDDAModel::devp       EvoDDAModel.cpp:353-391   adjoint source in closed form)
 *   (kernel.cu A2As/B2Bs/Conv/Conv2B and the cuFFT calls are internal.)
 *   No counterpart (one solve at a time in the reference, T/main.cpp:584-598):
 *   tdda_batch_* (many structures per launch), tdda_server_* (request loop),
 *   tdda_op_autotune / tdda_op_tiles / tdda_op_xpass (tile autotuner), the measurement hooks
 *   tdda_*_time_pass(es).
 *
 * Conventions
 *   - Every function returns TDDA_OK (0) or a negative status; the message of
 *     the last failure on the calling thread is tdda_last_error().
 *   - Complex vectors are interleaved (re, im) doubles; a site vector holds
 *     3N complex values, entry 3*i+c is component c of dipole i
 *     (AProductCore.cpp:705-714).  Batched calls take nrhs such vectors back
 *     to back.
 *   - Handles own their device memory; callers own host buffers.  A handle is
 *     not thread-safe: one host thread at a time per context.  A server runs
 *     its own worker thread and context; its calls are thread-safe.
 *   - No exceptions cross this boundary.
 */
#ifndef TOPODDA_H_
#define TOPODDA_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDDA_OK 0
#define TDDA_ERR_ARG (-1)
#define TDDA_ERR_HIP (-2)
#define TDDA_ERR_OOM (-3)
#define TDDA_ERR_STATE (-4)
#define TDDA_ERR_UNSUPPORTED (-5)
#define TDDA_ERR_IO (-6)

/* precision bits */
#define TDDA_C128 1
#define TDDA_C64 2

/* model vectors */
#define TDDA_VEC_AL 0 /* 1/alpha per component (DDAModel::al)            */
#define TDDA_VEC_E 1  /* right-hand side (DDAModel::E)                     */
#define TDDA_VEC_P 2  /* solution / initial guess (DDAModel::P)            */
#define TDDA_VEC_Y 3  /* output of tdda_model_apply                        */

/* solve status */
#define TDDA_SOLVE_CONVERGED 1
#define TDDA_SOLVE_BREAKDOWN 2 /* r0.rl == 0, DDAModel.cpp:193-216         */
#define TDDA_SOLVE_MAXIT 3     /* "does not converge", DDAModel.cpp:256-259 */

/* solve modes */
#define TDDA_MODE_FP64 0  /* reference algorithm in complex128 on the GPU   */
#define TDDA_MODE_MIXED 1 /* complex64 inner BiCGSTAB, FP64 refinement      */

typedef struct tdda_ctx_s* tdda_ctx;
typedef struct tdda_op_s* tdda_op;
typedef struct tdda_model_s* tdda_model;
typedef struct tdda_batch_s* tdda_batch;
typedef struct tdda_server_s* tdda_server;
typedef struct tdda_design_s* tdda_design;

/* Last error message of the calling thread ("" if none). */
const char* tdda_last_error(void);
int tdda_version(void);

/* Device context: one HIP device and one stream. */
int tdda_ctx_create(int device, tdda_ctx* out);
int tdda_ctx_destroy(tdda_ctx ctx);
int tdda_ctx_sync(tdda_ctx ctx);

/* Operator description.  AProductCore.cpp:12-44 takes these from
 * CoreStructure (N, Nx, Ny, Nz, R, d) and its own arguments (lam, nback).
 * R holds 3N integer lattice coordinates with 0 <= R[3i+a] < N_a; sites must
 * be distinct.  R may be NULL for a full box (N = Nx Ny Nz) in build_a_bulk
 * order (tools.cpp:497-510).  grid = (Mx, My, Mz) chooses the zero-padded FFT grid; 0 lets
 * the planner pick: My, Mz the smallest {2,3,5}-smooth sizes >= 2N_a - 1
 * (My: 75 and 125 skipped for 80 and 128), Mx the length within 30 % above
 * 2Nx - 1 whose x pass costs least on MI355X when that is more than 5 %
 * below the smallest smooth length's (a measured table; the x pass moves
 * the same bytes for any Mx).  Any grid >= 2N_a - 1 gives the same
 * linear convolution as the reference's 2N_a - 1.
 * K = 2 pi nback / lam (AProductCore.cpp:17).                               */
typedef struct {
    int N;
    int Nx, Ny, Nz;
    const int* R;
    double d;
    double lam;
    double nback;
    int Mx, My, Mz;
    int precisions; /* TDDA_C128 | TDDA_C64 */
} tdda_op_desc;

/* AProductCore::AProductCore (AProductCore.cpp:12-249): builds the Green
 * tensor on the device, transforms it and keeps one octant of its spectrum. */
int tdda_op_create(tdda_ctx ctx, const tdda_op_desc* desc, tdda_op* out);
/* Periodic AProductCore (AProductCore.cpp:251-536): the tensor of each grid
 * offset is summed over the images m = -max_m..max_m, n = -max_n..max_n
 * shifted by (m Lm, n Ln, 0) -- physical lengths in the unit of d and lam
 * (e.g. nm), added to the offsets (i d, j d, k d) -- each damped by
 * exp(-(0.01 K r)^4) (A_dic_generator, AProductCore.cpp:573-586).  At most
 * 1e6 images. */
int tdda_op_create_periodic(tdda_ctx ctx, const tdda_op_desc* desc, int max_m, int max_n, double Lm, double Ln,
                            tdda_op* out);
/* Interaction tensor of an operator (A_dic_generator, AProductCore.cpp:560-571). */
#define TDDA_METHOD_LDR 0 /* LDR_inter, AProductCore.cpp:588-623 */
#define TDDA_METHOD_FCD 1 /* FCD_inter, AProductCore.cpp:625-695 (needs K d < pi - 0.1) */
typedef struct {
    int method;         /* TDDA_METHOD_* */
    int max_m, max_n;   /* images of a periodic structure; both < 0: not periodic */
    double Lm, Ln;      /* periods: physical lengths, same unit as d (not multiples of d) */
} tdda_op_ext;
/* AProductCore with the tensor method and, optionally, periodic images
 * (the reference's constructors, AProductCore.cpp:12-249 and 251-536, with
 * their AMatrixMethod argument). */
int tdda_op_create_ex(tdda_ctx ctx, const tdda_op_desc* desc, const tdda_op_ext* ext, tdda_op* out);
/* Tile autotuner (no counterpart in the reference, which leaves tiling to
 * cuFFT): times the fused z + y passes against the separate ones (where the
 * planner fuses), the inverse z pass (P5) with z tiles of 1, 1/2 and 1/4 of
 * the planned width and the forward z pass (P1) with 1 down to 1/8 of it,
 * the y passes with each of their kernels (register-resident where one
 * exists for the length, LDS-staged) and the x pass with each of its kernels
 * at x tiles narrowed 3 columns at a time, each on a scratch model of one
 * right-hand side (reps launches per timing; <= 0: 20), and keeps the
 * fastest per precision.  Affects models created afterwards. */
int tdda_op_autotune(tdda_op op, int reps);
/* tdda_op_autotune on a scratch model of nrhs right-hand sides (1..4096):
 * the tiles for the number of right-hand sides the operator will carry (a
 * structure with several incidences; no counterpart in the reference). */
int tdda_op_autotune_ex(tdda_op op, int reps, int nrhs);
/* Current tiles of one precision: out4 = z pencils per tile (P5), z pencils
 * per tile (P1), y columns per tile, x columns per tile. */
int tdda_op_tiles(tdda_op op, int precision, int* out4);
/* Pass kernels of one precision (no counterpart in the reference): bit 0
 * set when the x pass runs the register-resident kernel (else the
 * LDS-staged one), bit 1 likewise for the y passes, bit 2 when the z and y
 * passes run fused (one workgroup per x slab, the z-transformed slab kept
 * in LDS); < 0 on error.  The environment variables TDDA_XPASS=0 /
 * TDDA_YPASS=0 / TDDA_FUSEZY=0 make the planner keep the LDS-staged or
 * separate kernels for operators created afterwards. */
int tdda_op_xpass(tdda_op op, int precision);
/* AProductCore::~AProductCore (AProductCore.cpp:538-558). */
int tdda_op_destroy(tdda_op op);
/* The checks of tdda_op_create that need no device (sizes, grid, the
 * 32-bit limit Nx * My * Mz * 3 < 2^31 of the pass kernels); grid3 <- the
 * FFT grid the operator would use (may be NULL).  No counterpart in the
 * reference (AProductCore.cpp:41-44 fixes the grid at 2N - 1).            */
int tdda_op_plan(const tdda_op_desc* desc, int* grid3);
/* grid3 <- (Mx, My, Mz); K <- wavenumber.  Either may be NULL. */
int tdda_op_info(tdda_op op, int* grid3, double* K);

/* AProductCore::Aproduct (AProductCore.cpp:697-758): y = A_offdiag b for
 * nrhs site vectors (no diagonal term, AProductCore.h:56).  If al is not
 * NULL, adds al o b per right-hand side (DDAModel::Aproductwithalb,
 * DDAModel.cpp:577-586); al_per_rhs = 0 shares one al across the batch.
 * precision selects the arithmetic (TDDA_C128 or TDDA_C64; inputs and
 * outputs are complex128 either way).  Host buffers, synchronous.           */
int tdda_aproduct(tdda_op op, const double* b, double* y, int nrhs, const double* al, int al_per_rhs,
                  int precision);

/* Allocation-free variant of tdda_aproduct.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  b, y and al are DEVICE pointers, 16-byte aligned,
 * interleaved complex128 (nrhs * 3N values each; al holds nrhs * 3N values
 * if al_per_rhs, else 3N; al NULL: no diagonal term).  The same passes and
 * tiles as tdda_aproduct; the job tables, the pass intermediates and the
 * complex64 staging live in one static device array of
 * TDDA_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  A call
 * that needs more (about (Nx Ny Mz + Nx My Mz) * 48 B per right-hand side
 * in complex128, half in complex64, plus 256 KiB of tables), or nrhs > 64,
 * fails with TDDA_ERR_UNSUPPORTED / TDDA_ERR_ARG.  Runs on the operator's
 * context stream and returns when y is written.  Calls no hipMalloc or
 * hipFree.  Every call on one GPU uses that GPU's array: calls that overlap
 * on one GPU (several threads or contexts) race and must be serialised by
 * the caller.                                                              */
#define TDDA_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int tdda_aproduct_noalloc(tdda_op op, const double* b, double* y, int nrhs, const double* al, int al_per_rhs,
                          int precision);

/* AProductCore::LDR_inter (AProductCore.cpp:588-623) on the host: the six
 * unique components xx xy xz yy yz zz of the interaction tensor at offset
 * (x, y, z) for wavenumber K, as 12 doubles.                                */
int tdda_green_tensor(double K, double x, double y, double z, double* out12);
/* AProductCore::FCD_inter (AProductCore.cpp:625-695) for lattice spacing d,
 * with the reference's Si / Ci table interpolation (SiCi.cpp:27-46). */
int tdda_green_tensor_fcd(double K, double d, double x, double y, double z, double* out12);

/* Host-side DDAModel set-up math (pure host code, no device needed). */
/* Material rule of AProductCore.cpp:19-24: eps /= nback^2, +0.01 if |eps| <= 1.01
 * (nmat complex values, in place).                                         */
int tdda_material_rule(int nmat, double* material, double nback);
/* The periodic constructor's variant (AProductCore.cpp:261-267): +0.01 only
 * if |eps| < 1.01, so a material at exactly |eps| = 1.01 is kept.          */
int tdda_material_rule_periodic(int nmat, double* material, double nback);
/* Get_Alpha (tools.cpp:617-630): LDR polarizability. */
int tdda_get_alpha(double lam, double K, double d, double diel_re, double diel_im, const double* nE3,
                   const double* nK3, double* out2);
/* Incident plane wave E[3i+c] = E0 nE[c] exp(i K d nK.R_i) (DDAModel.cpp:47-51). */
int tdda_incident_field(int N, const int* R, double K, double d, double E0, const double* nK3, const double* nE3,
                        double* E);
/* Permittivity interpolation from the design values diel_old (3N) and the
 * rule-adjusted material list, and al = 1/alpha (DDAModel.cpp:54-63,
 * UpdateAlpha DDAModel.cpp:470-487).  diel may be NULL.                   */
int tdda_polarizability(int n3, const double* diel_old, int nmat, const double* material, double lam, double K,
                        double d, const double* nE3, const double* nK3, double* diel, double* al);
/* Local field E_int = al P / ((2 + eps) / 3) (DDAModel.cpp:567-575). */
int tdda_local_field(int n3, const double* al, const double* P, const double* diel, double* out);

/* DDAModel device state for nrhs right-hand sides sharing one operator:
 * al, E, P per right-hand side, plus solver work space.                    */
int tdda_model_create(tdda_op op, int nrhs, tdda_model* out);
int tdda_model_destroy(tdda_model m);
/* Copy one vector kind (TDDA_VEC_*) to or from the device.  rhs = -1 moves
 * all nrhs vectors (nrhs * 3N complex); otherwise one right-hand side.     */
int tdda_model_set(tdda_model m, int which, int rhs, const double* host);
int tdda_model_get(tdda_model m, int which, int rhs, double* host);
/* Sparse access to one right-hand side: get reads n entries (complex
 * indices into the 3N vector) of a vector; set zeroes the vector and writes n
 * entries.  For adjoint sources concentrated on a few dipoles, as in
 * EvoDDAModel::devp (EvoDDAModel.cpp:353-391) feeding change_E
 * (EvoDDAModel.cpp:596), without moving whole vectors over PCIe.          */
int tdda_model_get_entries(tdda_model m, int which, int rhs, int n, const int* idx, double* vals);
int tdda_model_set_entries(tdda_model m, int which, int rhs, int n, const int* idx, const double* vals);
/* Y = A P + al o P on the device (DDAModel::Aproductwithalb). */
int tdda_model_apply(tdda_model m, int precision);
/* Measurement hook (no reference counterpart): runs reps device-resident
 * applications Y = A P + al o P and returns the mean duration in ms of each
 * of the five passes (z-forward, y-forward, x-fused Green product,
 * y-inverse, z-inverse), timed with HIP events on the model's stream.     */
int tdda_model_time_passes(tdda_model m, int precision, int reps, double* ms5);
/* Measurement hook (no reference counterpart): pass `pass` (0..4, order as
 * above) launched reps times back to back between two HIP events; *ms is
 * the mean per launch.  Overwrites the model's pass intermediates and Y.   */
int tdda_model_time_pass(tdda_model m, int precision, int pass, int reps, double* ms);

typedef struct {
    int max_it;        /* MAX_ITERATION of DDAModel::bicgstab             */
    double tol;        /* MAX_ERROR: stop when ||r|| / ||E|| <= tol       */
    int mode;          /* TDDA_MODE_FP64 or TDDA_MODE_MIXED               */
    int x0_zero;       /* 1: start from P = 0 (skips the A P product)     */
    double inner_tol;  /* mixed: tolerance of an inner solve (loosened to tol / (2 rel) when the residual rel is near tol) */
    int inner_max_it;  /* mixed: iteration cap of each inner solve        */
    int max_outer;     /* mixed: refinement cycles                        */
    int check_every;   /* iterations between host status polls (0: auto)  */
    int true_residual; /* 1: also compute ||E - A P|| / ||E|| in FP64     */
    int precond;       /* TDDA_PRECOND_NONE (the reference's BiCGSTAB) or
                          TDDA_PRECOND_JACOBI: BiCGSTAB on A D^-1 with
                          D = diag(al) (right preconditioning: the same
                          residual E - A P and stop tests, P = D^-1 y at the
                          end); no counterpart in the reference          */
} tdda_solve_opts;

#define TDDA_PRECOND_NONE 0
#define TDDA_PRECOND_JACOBI 1

typedef struct {
    int status;      /* TDDA_SOLVE_*                                        */
    int iterations;  /* ITERATION (FP64 mode) or total inner iterations     */
    double error;    /* recurrence residual ||r|| / ||E|| (DDAModel::Error) */
    double true_residual; /* ||E - A P|| / ||E|| in FP64, if requested      */
    int outer;       /* mixed: refinement cycles used                       */
} tdda_solve_info;

void tdda_solve_opts_default(tdda_solve_opts* o);

/* DDAModel::bicgstab (DDAModel.cpp:131-260) for every right-hand side of
 * the model, on the device.  FP64 mode follows the reference algorithm
 * statement by statement; MIXED runs complex64 BiCGSTAB inner solves inside
 * an FP64 residual-refinement loop whose stop test uses the true residual.
 * info must hold nrhs entries.  P is updated in device memory.            */
int tdda_model_solve(tdda_model m, const tdda_solve_opts* opts, tdda_solve_info* info);

/* Batches (no counterpart in the reference, whose solves run one at a time:
 * T/main.cpp:584-598 loops DDAModel::bicgstab over structures, and
 * EvoDDAModel.cpp:433-441 over models).  A batch gathers the right-hand
 * sides of several models -- different structures, grids, wavelengths and
 * incidences -- and runs every operator pass as one launch over all of
 * them; solver states stay per right-hand side, so each result equals its
 * solve alone.  The batch references the models' device state (models must
 * share one context and stay alive; the batch keeps them referenced).     */
int tdda_batch_create(tdda_model* models, int nmodels, tdda_batch* out);
int tdda_batch_destroy(tdda_batch b);
/* Number of right-hand sides (sum of the models' nrhs, in model order). */
int tdda_batch_size(tdda_batch b);
/* DDAModel::bicgstab for every right-hand side of the batch; info holds
 * tdda_batch_size(b) entries in model order.  P is updated in each model. */
int tdda_batch_solve(tdda_batch b, const tdda_solve_opts* opts, tdda_solve_info* info);
/* Y = A P + al o P for every right-hand side (tdda_model_apply per model). */
int tdda_batch_apply(tdda_batch b, int precision);
/* Measurement hook (no reference counterpart): tdda_model_time_passes over
 * every right-hand side of the batch (each pass timed as all its launches). */
int tdda_batch_time_passes(tdda_batch b, int precision, int reps, double* ms5);
/* Measurement hook (no reference counterpart): tdda_model_time_pass over
 * every right-hand side of the batch (all launches of one pass kind). */
int tdda_batch_time_pass(tdda_batch b, int precision, int pass, int reps, double* ms);

/* Request loop (no counterpart in the reference, which runs one task's
 * solves in sequence: T/main.cpp:584-598, EvoDDAModel.cpp:433-441).  A
 * server owns a worker thread and its own stream on `device`.  Submitted
 * requests -- one structure with nrhs incidences -- wait in a FIFO queue
 * until their device footprint fits the memory budget and nrhs solver
 * slots are free; admitted right-hand sides run together (continuous
 * batching: every tick is one launch group per pass over all running
 * right-hand sides, whatever their structure and phase) and finished ones
 * leave at once.  Each request's al and E are built on the device exactly
 * as the DDAModel constructor builds them (DDAModel.cpp:12-69), or E is a
 * sparse source (adjoint, EvoDDAModel.cpp:596).  Solves start from P = 0
 * and follow tdda_model_solve (solve.mode, tolerances, limits), so a
 * result does not depend on the other requests in flight.                  */
typedef struct {
    double hbm_budget_bytes; /* device memory the server may hold; <= 0: 85% of free memory */
    int max_rhs;             /* right-hand sides in flight (solver slots); <= 0: 4096 */
    tdda_solve_opts solve;   /* x0_zero is implied */
} tdda_server_opts;

typedef struct {
    tdda_op_desc op;          /* structure; R may be NULL for a full box          */
    const double* diel_old;   /* 3N design values t (CoreStructure::diel_old)      */
    int nmat;                 /* material table (complex, after tdda_material_rule) */
    const double* material;
    int nrhs;
    const double* nK;         /* 3 per right-hand side: propagation direction      */
    const double* nE;         /* 3 per right-hand side: polarisation               */
    const double* E0;         /* amplitude per right-hand side, or NULL for 1      */
    int nsrc;                 /* > 0: E = sparse source instead of the plane wave  */
    const int* src_idx;       /* nrhs * nsrc distinct entries of the 3N vector      */
    const double* src_vals;   /* nrhs * nsrc complex values                        */
} tdda_request;

/* The DDAModel constructor on the device (DDAModel.cpp:12-69, with
 * UpdateAlpha's eps interpolation, DDAModel.cpp:54-63, and Get_Alpha,
 * tools.cpp:617-630): al and E of every right-hand side of model m from a
 * request description, as the request loop builds them -- the 3N design
 * values and the incidences cross PCIe instead of nrhs * 3N al and E.  The
 * request's N, Nx, Ny, Nz and nrhs must match m (its op.d, lam and grid are
 * ignored: m's operator supplies K and d); E is the plane wave, or the
 * sparse source when nsrc > 0.  Also sets P = 0.  Synchronous.             */
int tdda_model_setup(tdda_model m, const tdda_request* r);

typedef struct {
    double t_submit, t_admit, t_done; /* seconds, steady clock of the server */
    double footprint_bytes;           /* device memory charged to the request */
} tdda_request_times;

typedef struct {
    long long submitted, completed, failed, ticks;
    int queued, active_requests, active_rhs;
    double used_bytes, peak_bytes, budget_bytes;
    double setup_s, tick_s, release_s; /* worker time: request set-up, ticks, releases */
} tdda_server_stats;

/* Defaults: budget 85% of free memory, 4096 slots, mixed precision to a
 * true residual of 1e-8. */
void tdda_server_opts_default(tdda_server_opts* o);
int tdda_server_create(int device, const tdda_server_opts* o, tdda_server* out);
/* Stops the loop; requests not yet finished fail. */
int tdda_server_destroy(tdda_server s);
/* Copies the request's inputs; returns at once with a ticket. */
int tdda_server_submit(tdda_server s, const tdda_request* r, long long* ticket);
/* 0: queued or running, 1: done, 2: failed (negative: error code). */
int tdda_server_poll(tdda_server s, long long ticket);
/* Blocks until the request is done or failed, or timeout_s passes (< 0:
 * no limit); returns as tdda_server_poll. */
int tdda_server_wait(tdda_server s, long long ticket, double timeout_s);
/* Copies P (nrhs * 3N complex, right-hand side major) and the per-RHS solve
 * information of a finished request and releases its ticket. */
int tdda_server_result(tdda_server s, long long ticket, double* P, tdda_solve_info* info, tdda_request_times* t);
int tdda_server_get_stats(tdda_server s, tdda_server_stats* out);

/* ------------------------------------------------------------------------
 * Device-resident design update of an extruded structure: one free
 * parameter per (x, y) column of a full box (2D extrusion, no symmetry, not
 * periodic), kept in HBM.  Replaces, per optimisation step,
 * CoreStructure::UpdateStr(step, it, max_it) (CoreStructure.cpp:31-104) with
 * the FreeWeight cone filter of SpacePara (SpacePara.cpp:37-48, 1241-1270)
 * and FilterOption::SmoothDensity (FilterOption.cpp:49-60), followed by
 * DDAModel::UpdateAlpha (DDAModel.cpp:470-484) of every right-hand side.
 * ---------------------------------------------------------------------- */
#define TDDA_BETA_EXP 0       /* exp_update, tools.cpp:776-778 */
#define TDDA_BETA_PIECEWISE 1 /* piecewise_update, tools.cpp:780-794 */
#define TDDA_BETA_LINEAR 2    /* linear_update, tools.cpp:796-798 */

/* Design state of model m (its operator must be a full box in build_a_bulk
 * order).  para_origin: Nx * Ny values in [0, 1], column index x * Ny + y
 * (SpacePara's Para order for an extruded box); rfilter: cone-filter radius
 * in lattice units (> 0.1); material: the operator's material table (re, im
 * pairs, after the AProductCore.cpp:19-24 rule), nmat >= 2; inc: nK[3],
 * nE[3] per right-hand side of m.  The design holds a reference to m. */
int tdda_design_create(tdda_model m, const double* para_origin, double rfilter, const double* material, int nmat,
                       const double* inc, tdda_design* out);
/* Any parameterization of SpacePara (SpacePara.h): the parameter of every
 * site (geometryPara), the parameter values (Para: free and fixed), the
 * free parameters (Free) and, with a filter, each free parameter's cone
 * weights in SpacePara's order (FreeWeight, SpacePara.cpp:1232-1380: any
 * radius, the periodic wrap, 4-fold shared parameters, fixed parameters
 * among the weighted ones).  tdda_design_create is this with one free
 * parameter per column of an extruded full box.                          */
typedef struct {
    int npara;                 /* Para values (free and fixed)                     */
    int nfree;                 /* free parameters                                  */
    const int* free_index;     /* nfree: index of each free parameter in Para      */
    const int* site_param;     /* N: parameter of each site of the model's operator */
    const double* para;        /* npara: initial Para (also Para_origin, Para_filtered) */
    int filter;                /* 1: cone filter + SmoothDensity; 0: Para += step  */
    const int* weight_start;   /* nfree + 1 offsets into weight_pos / weight_val   */
    const int* weight_pos;     /* Para index of each weight                        */
    const double* weight_val;  /* the weights                                      */
} tdda_design_desc;
/* Design state of model m for any parameterization (see tdda_design_desc);
 * material / nmat / inc as tdda_design_create. */
int tdda_design_create_ex(tdda_model m, const tdda_design_desc* desc, const double* material, int nmat,
                          const double* inc, tdda_design* out);
/* One CoreStructure::UpdateStr step (CoreStructure.cpp:31-104) for the
 * nfree free parameters (step: nfree host values; NULL: no step): with a
 * filter, Para_origin += step clamped to [0, 1], Para_filtered = weighted
 * mean over the free parameter's weights, Para = SmoothDensity(beta, eta);
 * without, Para += step clamped (beta and eta unused).  Then al of every
 * right-hand side of the model (complex128 and, when present, complex64
 * copies) from diel_old = Para of each site's parameter.  beta follows the
 * caller's schedule (tdda_beta_update). */
int tdda_design_step(tdda_design dz, const double* step, double beta, double eta);
/* Copies Para_origin, Para_filtered and Para (npara values each -- Nx * Ny
 * for tdda_design_create; any may be NULL).  Without a filter only Para
 * changes. */
int tdda_design_get(tdda_design dz, double* para_origin, double* para_filtered, double* para);
void tdda_design_destroy(tdda_design dz);
/* FilterOption::update_beta (FilterOption.cpp:24-46): beta at iteration it
 * of max_it for the TDDA_BETA_* schedule. */
double tdda_beta_update(int kind, double it, double max_it, double beta_min, double beta_max);

/* ------------------------------------------------------------------------
 * Objectives and their adjoint sources on the device.  The value F of an
 * objective for right-hand side `rhs` of model m (its current P), and, when
 * dst is not NULL, the adjoint right-hand side 2 dF/dP in closed form --
 * the quantity EvoDDAModel::devp (EvoDDAModel.cpp:353-391) estimates by
 * one-sided finite differences -- written into the E vector of right-hand
 * side dst_rhs of dst (a model on the same operator, e.g. m itself or the
 * model of the adjoint solve).
 * ---------------------------------------------------------------------- */
/* ObjPointEDDAModel (ObjDDAModel.cpp:20-97): F = |E_ext - sum_i
 * A(r0 - d R_i) P_i| with E_ext = E0 nE exp(i K nK . r0); r0 is a physical
 * position in the same length unit as d (not scaled by d).  Source: -(1/F) sum_a A_ac conj(E_a). */
int tdda_objective_point_e(tdda_model m, int rhs, const double* r0, const double* nK, const double* nE, double E0,
                           double* value, tdda_model dst, int dst_rhs);
/* ObjIntegratedEDDAModel (ObjDDAModel.cpp:100-177): F = log sum over the
 * region's sites of SmoothDensity(diel_old_3i; eta, beta) sum_c
 * |al_c P_c|^pow_num.  box6 = xMin, xMax, yMin, yMax, zMin, zMax tested as
 * the reference writes it, (lo <= v <= hi) with C semantics
 * (ObjDDAModel.cpp:138).  diel_old: 3N host values in [0, 1].  Source:
 * pow_num f |al|^pow_num |P|^(pow_num - 2) conj(P) / E_int (0 where P = 0). */
int tdda_objective_integrated_e(tdda_model m, int rhs, int pow_num, const int* box6, double eta, double beta,
                                const double* diel_old, double* value, tdda_model dst, int dst_rhs);

/* ------------------------------------------------------------------------
 * Fields at points and the reference's text result files.
 * ---------------------------------------------------------------------- */
/* DDAModel::solve_E (DDAModel.cpp:530-565): out (npts complex triples) =
 * E_ext - sum_i A(d pt - d R_i) P_i at the lattice points pts (3 ints each,
 * units of d) with the operator's tensor and without periodic images, as
 * the reference evaluates it.  E_ext is the reference's incident field at
 * the point, including its cosine argument of components y and z
 * (K (nK_x y + nK_y y + nK_z z) and K (nK_x z + nK_y y + nK_z z)). */
int tdda_field_at_points(tdda_model m, int rhs, const int* pts, int npts, const double* nK, const double* nE,
                         double E0, double* out);
/* CoreStructure::output_to_file (CoreStructure.cpp:155-190): full != 0
 * writes Nx, Ny, Nz, N, R (3N), diel_old (3N) and d ("CoreStructure.txt",
 * mode "normal"); full == 0 writes diel_old only (any other mode).  Same
 * bytes as the reference's streams (6 significant digits, Eigen columns). */
int tdda_write_structure(const char* path, int Nx, int Ny, int Nz, int N, const int* R, const double* diel_old,
                         double d, int full);
/* DDAModel::output_to_file (DDAModel.cpp:589-697); P and EResult are nP and
 * nE complex numbers written "re+imj" / "re-imj", one per line. */
#define TDDA_RESULTS_FULL 0 /* P, nK, nE0, nE, RResult (nR ints), EResult: DDAModel.cpp:589-616 */
#define TDDA_RESULTS_E_P 1  /* EResult then P: DDAModel.cpp:618-656 */
#define TDDA_RESULTS_E 2    /* EResult: DDAModel.cpp:658-697 */
int tdda_write_model_results(const char* path, int layout, const double* P, int nP, const double* EResult, int nE,
                             const int* RResult, int nR, const double* nK, const double* nE0);

#ifdef __cplusplus
}
#endif

#endif /* TOPODDA_H_ */
