/* simple_gpu.h — C ABI of the MI355X (gfx950) SIMPLE lid-driven-cavity backend.
 *
 * This is the drop-in boundary for `backend = gpu`. It replaces the reference's C++ class
 * `HipSolver` and `run_simple_hip` (include/simple/solver_hip.hpp:11-40): one entry point per
 * solver phase, the same phase order and semantics as `outer_iteration` (src/solver.cpp:27-93),
 * plus a whole-run entry point (`sg_advance`) that the host's outer loop
 * (src/solver.cpp:106-132, `run_outer_loop`) calls in chunks.
 *
 * Conventions
 *  - Plain C: PODs, pointers and sizes only. No exceptions cross this boundary.
 *  - Every function returns 0 on success, a nonzero SG_E* code on failure; sg_last_error()
 *    returns a thread-local message for the last failure ("GPU error: ..." for HIP errors,
 *    mirroring src/hip/solver_hip.hip:17-23). The C++ host turns a nonzero status into
 *    std::runtime_error, so `error (rank 0): ...` and exit code 1 are unchanged
 *    (src/main.cpp:77-83).
 *  - Host arrays are full fields with halos, nrows x ncols row-major, (i, j) -> i*ncols + j
 *    (include/simple/field2d.hpp:36-65, docs/SPEC.md section 2).
 *  - One host thread per solver; all device work is ordered on the solver's own HIP stream.
 */
#ifndef SIMPLE_GPU_H
#define SIMPLE_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SG_ABI_VERSION 1

/* Status codes. */
enum {
    SG_OK = 0,
    SG_E_ARG = 1,       /* invalid argument (bad params, unknown field id, null pointer) */
    SG_E_HIP = 2,       /* a HIP runtime call failed */
    SG_E_UNSUPPORTED = 3,
    SG_E_NOMEM = 4,
    SG_E_STATE = 5      /* call not valid in the current state */
};

/* Enumerations, same values and spellings as include/simple/params.hpp:9-13. */
enum { SG_UDS = 0, SG_QUICK = 1 };
enum { SG_NORM_ABSOLUTE = 0, SG_NORM_RUNNING_MAX = 1 };
enum { SG_BACKEND_SERIAL = 0, SG_BACKEND_OPENMP = 1, SG_BACKEND_GPU = 2 };
enum { SG_MOM_GS_LEX = 0, SG_MOM_JACOBI = 1, SG_MOM_GS_MULTICOLOUR = 2 };
enum { SG_P_GS_LEX = 0, SG_P_JACOBI = 1, SG_P_GS_REDBLACK = 2, SG_P_CG = 3, SG_P_MG = 4 };
/* Additions with no reference counterpart. */
enum { SG_MG_FP64 = 0, SG_MG_MIXED = 1 };      /* mg_precision */
enum { SG_WARM_NONE = 0, SG_WARM_LEARNED = 1 }; /* pressure_warm_start */
enum { SG_MOMP_FP64 = 0, SG_MOMP_MIXED = 1 };  /* momentum_precision */

/* POD mirror of Params (include/simple/params.hpp:15-58): every numeric key of the config
 * file, the derived sizes, and the additions of this backend. output_dir and the MPI fields
 * stay on the host. */
typedef struct sg_params {
    int NX, NY, H;            /* N_cells_x (rows), N_cells_y (columns), halo width */
    int N_halo_x, N_halo_y;
    double Re;
    int maxit_inner_u, maxit_inner_v, maxit_inner_p;
    double error_tolerance;
    int maxit;
    double relaxation_factor_u, relaxation_factor_v, relaxation_factor_p;
    int convection_scheme;      /* SG_UDS / SG_QUICK */
    int diffusion_order;        /* 2 or 4 */
    int Hybrid_switch;          /* must be 0 */
    int residual_normalisation; /* SG_NORM_* */
    int backend;                /* SG_BACKEND_* */
    int momentum_solver;        /* SG_MOM_* */
    int pressure_solver;        /* SG_P_* */
    double pressure_tolerance;  /* cg: ||r||/||b|| stop */
    int pressure_maxit;         /* cg: iteration cap */
    int mg_cycles, mg_pre, mg_post;
    double mg_omega;
    int write_fields;
    /* derived (src/params.cpp:174-178) */
    int nrows, ncols;
    double dx, dy, nu;
    /* additions */
    int mg_precision;           /* SG_MG_FP64 (reference arithmetic) or SG_MG_MIXED */
    int mg_coarsest;            /* smallest coarse side kept (reference: 4) */
    int mg_dc_maxit;            /* mixed: FP64 defect-correction steps per solve */
    int pressure_warm_start;    /* SG_WARM_* */
    int check_every;            /* device-side stop test: host looks every this many iterations */
    int momentum_precision;     /* SG_MOMP_FP64 (reference arithmetic) or SG_MOMP_MIXED: outer
                                 * iterations sweep an FP32 correction of u and v inside FP64
                                 * defect correction (gs_multicolour, at least one sweep) */
} sg_params;

/* Array identifiers. Host layout of every id is nrows x ncols doubles. */
enum {
    SG_F_U = 0, SG_F_V, SG_F_P,
    SG_F_CU_AE, SG_F_CU_AW, SG_F_CU_AN, SG_F_CU_AS,
    SG_F_CU_AEE, SG_F_CU_AWW, SG_F_CU_ANN, SG_F_CU_ASS, SG_F_CU_AP,
    SG_F_CV_AE, SG_F_CV_AW, SG_F_CV_AN, SG_F_CV_AS,
    SG_F_CV_AEE, SG_F_CV_AWW, SG_F_CV_ANN, SG_F_CV_ASS, SG_F_CV_AP,
    SG_F_AEP, SG_F_AWP, SG_F_ANP, SG_F_ASP, SG_F_APP, SG_F_B, SG_F_PC,
    SG_F_COUNT
};

/* Phase indices of the per-phase timers, same order and names as include/simple/timer.hpp:6-20
 * (assemble_u, bc_u, residual_u, sweep_u, assemble_v, bc_v, residual_v, sweep_v, assemble_p,
 * solve_p, correct). */
#define SG_NPHASES 11

typedef struct sg_solver sg_solver;

/* Last error message of this thread ("" if none). */
const char* sg_last_error(void);
int sg_abi_version(void);

/* Number of visible GPUs (0 if none); name of device 0 into buf. */
int sg_device_count(int* n);
int sg_device_name(char* buf, size_t len);

/* Validates p and fills the derived fields (nrows, ncols, dx, dy, nu), like the validation and
 * derivation block of make_params (src/params.cpp:152-178). */
int sg_params_finalize(sg_params* p);
/* Reference defaults (include/simple/params.hpp:15-58) plus this backend's defaults. */
void sg_params_default(sg_params* p);

/* HipSolver(const Params&, const State&) (include/simple/solver_hip.hpp:13): allocates the
 * device arrays and uploads u, v, p (any may be NULL: zero). Coefficient and p' arrays start 0. */
int sg_create(const sg_params* p, const double* u, const double* v, const double* pf,
              sg_solver** out);
int sg_destroy(sg_solver* s);

/* Upload / download one array (host layout nrows x ncols). download replaces
 * HipSolver::download (include/simple/solver_hip.hpp:34). */
int sg_upload(sg_solver* s, int field_id, const double* host);
int sg_download(sg_solver* s, int field_id, double* host);

/* Per-phase entry points, HipSolver::assemble_u .. correct (include/simple/solver_hip.hpp:19-29).
 * Each is enqueued and completed before returning. Residuals use the reference's summation
 * order (rows j ascending from 0.0, then rows ascending; src/decomp.cpp:176-183). */
int sg_assemble_u(sg_solver* s);
int sg_bc_u(sg_solver* s);
int sg_residual_u(sg_solver* s, double* res);
int sg_sweep_u(sg_solver* s, int nsweeps);
int sg_assemble_v(sg_solver* s);
int sg_bc_v(sg_solver* s);
int sg_residual_v(sg_solver* s, double* res);
int sg_sweep_v(sg_solver* s, int nsweeps);
int sg_assemble_p(sg_solver* s, double* res_p);
int sg_solve_p(sg_solver* s);
int sg_correct(sg_solver* s);

/* HipSolver::outer_iteration (include/simple/solver_hip.hpp:31): one outer iteration in eager
 * mode, residuals into res[3] (u, v, p) and per-phase seconds (hipEvent-timed, SG_NPHASES) added
 * into phase_s (may be NULL). */
int sg_outer_iteration(sg_solver* s, double res[3], double* phase_s);

/* Runs up to max_iters outer iterations as replays of one captured hipGraph, with the stopping
 * rule of run_outer_loop (src/solver.cpp:112-128) evaluated on the device after every
 * iteration: work stops at the converged iteration, never past it. Writes iters_done and, for
 * each iteration done, [res_u, res_v, res_p] into log (3*max_iters doubles, may be NULL);
 * *converged = 1 if the stop rule fired. Iteration numbers continue across calls. */
int sg_advance(sg_solver* s, int max_iters, int* iters_done, double* log, int* converged);

/* Iterations done so far by this solver (sg_outer_iteration and sg_advance). */
int sg_iterations(sg_solver* s, int* iters);

/* Per-phase timing of replayed outer iterations, the measured counterpart of the reference's
 * per-phase timers (src/hip/solver_hip.hip:502-508, written to timing.csv by src/io.cpp:71-90).
 * With on != 0, every sg_advance call runs one of its outer iterations launch by launch with
 * hipEvents at the phase boundaries (the same kernels as the replayed iterations, enqueued behind
 * the first replayed graph; results unchanged) and adds that iteration's device seconds per phase
 * (SG_NPHASES, timer.hpp order) into an accumulator; sg_set_phase_timing clears it. In the replayed iteration the residuals are
 * evaluated inside the first sweeps and summed inside the p' solve, and u's and v's boundary
 * conditions inside the assembly and the first sweep, so those phases carry that work.
 * sg_phase_times reads the accumulated seconds and the number of timed iterations. */
int sg_set_phase_timing(sg_solver* s, int on);
int sg_phase_times(sg_solver* s, double* phase_s, int* timed_iterations);

/* Mean seconds of one launch of the named kernel over the last sg_advance call that ran with
 * timing on (sg_set_kernel_timing), measured with hipEvents on the solver's stream, and the
 * algorithmic bytes of one launch (SG_K_* ids). */
enum {
    SG_K_SWEEP_U = 0,  /* one u multicolour sweep (every colour, one launch) */
    SG_K_SWEEP_V,
    SG_K_COPY,         /* plain device copy of one field array (nrows x ld doubles): calibration */
    SG_K_ASSEMBLE_U,
    SG_K_PSOLVE,       /* one complete p' solve (pressure_solver of the params) */
    SG_K_SWEEP_U_FIRST, /* the first u sweep of an outer iteration (|R| cells fused in) */
    SG_K_ASSEMBLE_V,
    SG_K_SWEEP_V_FIRST, /* the first v sweep of an outer iteration */
    SG_K_SWEEP_U_DC,   /* momentum_precision = mixed: one FP32 u correction sweep (every colour) */
    SG_K_ASSEMBLE_U_DC, /* momentum_precision = mixed: the u assembly with the FP64 residual */
    SG_K_COUNT
};
int sg_kernel_probe(sg_solver* s, int kernel_id, int reps, double* mean_seconds,
                    double* bytes_per_launch, int* launches_per_iteration);

/* Multigrid information: number of levels, and the red-black sweeps of the last coarsest-level
 * solve. */
int sg_mg_levels(sg_solver* s, int* nlevels);
int sg_mg_coarse_sweeps(sg_solver* s, int* sweeps);

/* Launch-shape knobs (no effect on results). SG_TUNE_TAIL_THREADS: threads of the one-workgroup
 * multigrid tail (64..1024, multiple of 64); SG_TUNE_MID_CELLS: levels with at most this many
 * cells run inside that workgroup; SG_TUNE_LDS_CELLS: levels with at most this many cells (that
 * fit) live in LDS there. */
enum {
    SG_TUNE_TAIL_THREADS = 0,
    SG_TUNE_MID_CELLS = 1,
    SG_TUNE_LDS_CELLS = 2,
    SG_TUNE_MOM_TILE = 3, /* momentum sweep tile (rows x columns): 0 32x64, 1 64x64, 2 32x128,
                           * 3 16x128, 4 24x96, 5 48x96, 6 24x64, 7 40x128, 8 32x72, 9 28x80,
                           * 10 36x64, 11 40x56 */
    SG_TUNE_MG_STAGED = 4, /* first multigrid level whose tiles also stage the operator in LDS */
    SG_TUNE_MOM_TILE_FIRST = 5, /* tile of the first sweep (fused residual) of each component;
                                 * unset: SG_TUNE_MOM_TILE */
    SG_TUNE_DC_TILE = 6, /* momentum_precision = mixed: FP32 correction sweep tile (rows x
                          * columns / threads): 0 32x64/256, 1 32x64/512, 2 64x64/512,
                          * 3 32x128/512, 4 48x64/512, 5 16x128/256, 6 24x96/512, 7 40x56/512 */
    SG_TUNE_COUNT
};
int sg_set_tuning(sg_solver* s, int key, long value);
int sg_get_tuning(sg_solver* s, int key, long* value);
/* Autotuner: on the solver's current state, times one momentum sweep per tile shape and the p'
 * solve (hipEvents around graph replays, reps each) for a grid of the multigrid knobs, keeps the
 * fastest of each, restores every array, and writes a one-line-per-candidate report into report
 * (may be NULL); its last line names the multigrid choice. */
int sg_autotune(sg_solver* s, int reps, char* report, size_t len);

/* ---- Batched cavities ----
 * No counterpart in the reference, which runs one cavity per process (src/main.cpp:40-46). A
 * batch holds n cavities (zero initial fields, lid velocity 1) that may differ in N_cells_x,
 * N_cells_y, Re, error_tolerance and the relaxation factors; every other key must agree. Each
 * phase of the outer iteration runs as one persistent launch over all cavities, and one outer
 * iteration of the whole batch is replayed as a captured hipGraph. Every cavity keeps its own
 * stopping rule (src/solver.cpp:112-128) and, once stopped, is skipped; its results are exactly
 * those of the same cavity run alone by sg_advance. */
typedef struct sg_batch sg_batch;
int sg_batch_create(const sg_params* params, int n, sg_batch** out); /* params[n], finalized */
int sg_batch_destroy(sg_batch* b);
int sg_batch_size(sg_batch* b, int* n);
/* Borrowed handle of cavity k (sg_download, sg_iterations, ...); do not destroy it. */
int sg_batch_solver(sg_batch* b, int k, sg_solver** s);
/* Up to max_iters outer iterations of every cavity not yet stopped (the stopping rules are read
 * every check_every iterations of cavity 0). Per cavity (arrays of n, each may be NULL): total
 * outer iterations, status (1 converged, -1 diverged, 0 running) and the last logged residuals
 * (3 per cavity). */
int sg_batch_advance(sg_batch* b, int max_iters, int* iters, int* status, double* last_res);
/* Kernel launches of one batched outer iteration (independent of n). */
int sg_batch_launches(sg_batch* b, int* per_iteration);
int sg_batch_set_tuning(sg_batch* b, int key, long value); /* SG_TUNE_* for every cavity */
int sg_batch_get_tuning(sg_batch* b, int key, long* value);
/* Autotuner: times the batch's own outer iteration for each momentum tile shape and a grid of
 * multigrid tail knobs on the current state, restores u, v, p and the loop state, keeps the
 * fastest; report format as sg_autotune. */
int sg_batch_autotune(sg_batch* b, int reps, char* report, size_t len);

/* ---- Learned warm start of the p' solve ----
 * No counterpart in the reference (it has no learned component). A U-Net (4 levels of 16, 32, 64,
 * 128 channels, 3x3 convolutions, bfloat16 weights and activations, FP32 accumulation) maps the
 * coarse-grid p' system (block sums of b, block means of a_P, log10(Re/1000)) to the block means
 * of the converged p'. Its parameters are one FP32 vector (sg_unet_nparams entries; layout:
 * per convolution W[c_out][c_in][3][3] then b[c_out], then the 1x1 head's 16 weights and bias).
 * Images are NHWC float32 with 8 input channels (3 used), sides multiples of 8. */
typedef struct sg_unet sg_unet;
int sg_unet_nparams(int* n);
int sg_unet_create(sg_unet** out);
int sg_unet_destroy(sg_unet* net);
/* Set / get the FP32 parameters (set also resets the optimiser state). */
int sg_unet_set_params(sg_unet* net, const float* theta, int n);
int sg_unet_get_params(sg_unet* net, float* theta, int n);
/* Forward pass: x [nimg][h][w][8] -> y [nimg][h][w]. */
int sg_unet_forward(sg_unet* net, int nimg, int h, int w, const float* x, float* y);
/* sg_unet_forward without allocation. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). x [nimg][h][w][8] and y [nimg][h][w] are DEVICE float arrays of the
 * network's device (current when it was created, and current at the call); the pass is enqueued
 * on `stream` (a hipStream_t; NULL: the null stream) and not waited for. Every intermediate
 * activation lives in one fixed scratch area per GPU (7 MiB, held until the process ends), so
 * nimg * h * w is limited to 16384 pixels (larger: SG_E_ARG). Calls that overlap in time on one
 * GPU share that area and race: order them on one stream. The network's parameters must not
 * change while a call runs. Same kernels, same results as sg_unet_forward. */
int sg_unet_forward_noalloc(sg_unet* net, int nimg, int h, int w, const float* x, float* y,
                            void* stream);
/* Mean squared error against t [nimg][h][w] and its gradient (n params, may be NULL). */
int sg_unet_loss_grad(sg_unet* net, int nimg, int h, int w, const float* x, const float* t,
                      float* loss, float* grad);
/* One Adam step (beta 0.9, 0.999) on the loss of one batch; the loss before the step. */
int sg_unet_train_step(sg_unet* net, int nimg, int h, int w, const float* x, const float* t,
                       float lr, float* loss);
/* Mean seconds of one forward pass at this shape (hipEvents, reps) and its MFMA flops. */
int sg_unet_mfma_probe(sg_unet* net, int nimg, int h, int w, int reps, double* seconds,
                       double* flops);
/* Use net (not owned) ahead of every p' solve of s when s's params say pressure_warm_start =
 * learned (pressure_solver = mg, mg_precision = mixed), on multigrid level `level` (2^level
 * blocks; N_cells_x and N_cells_y multiples of 8 * 2^level; -1: the coarsest level whose grid
 * is at least 64 x 64, the size the network was trained on). The guess g enters scaled by
 * b.g / g.A.g (the energy minimiser along g: never farther from the solution than zero in the
 * A-norm); a degenerate guess is discarded (p' starts from zero). Do not train net while s runs. */
int sg_attach_warm_start(sg_solver* s, sg_unet* net, int level);
/* The network input of s's current p' system at `level` (x: (NX/2^level)*(NY/2^level)*8 floats)
 * and the scale that turns a prediction into p'. */
int sg_warm_features(sg_solver* s, int level, float* x, double* scale);
/* The level the attached warm start runs on (-1 when none is attached). No counterpart in the
 * reference (the warm start is an addition); reports what level -1 resolved to. */
int sg_warm_level(sg_solver* s, int* level);
/* Guesses kept and discarded so far. */
int sg_warm_stats(sg_solver* s, int* accepted, int* rejected);

#ifdef __cplusplus
}
#endif
#endif /* SIMPLE_GPU_H */
