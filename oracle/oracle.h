/* oracle.h — CPU oracle of the SIMPLE lid-driven-cavity path (TEST INFRASTRUCTURE ONLY).
 *
 * A plain C++17/OpenMP restatement of the reference's serial/OpenMP CPU arithmetic
 * (arfazeli/simple-hpc src/coefficients.cpp, momentum_u.cpp, momentum_v.cpp, pressure.cpp,
 * multigrid.cpp, solver.cpp, decomp.cpp:157-233), written expression for expression in the
 * reference's operation order so that results are bit-identical to it. It is the checker of the
 * GPU backend and the timed CPU baseline; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it. The product (libsimple_gpu.so, the `simple` CLI) never links it.
 *
 * Pins (each function is checked by at least one of):
 *  - the reference's snapshot fixtures (tests/snapshots/{short_16,converged_32x40}), bit-exact;
 *  - brute-force loops of tests/unit/test_inner_solvers.cpp restated in tests/ (LCG seed 12345);
 *  - multigrid/CG invariants of tests/unit/test_multigrid.cpp and test_cg.cpp;
 *  - published outer-iteration counts (docs/RESULTS.md);
 *  - the reference itself, compiled from its sources into oracle/_ref by oracle/ref.mk.
 */
#ifndef SIMPLE_ORACLE_H
#define SIMPLE_ORACLE_H

#include "../include/simple_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_state or_state;

or_state* or_create(const sg_params* p);
void or_destroy(or_state* s);
/* Replace the scalar parameters of a state (same grid); 0 on success. */
int or_set_params(or_state* s, const sg_params* p);
/* Pointer to the host array of a field (nrows x ncols doubles), SG_F_* ids. */
double* or_field(or_state* s, int field_id);
int or_nrows(or_state* s);
int or_ncols(or_state* s);
void or_set_threads(int n);
int or_get_threads(void);

/* Phases (reference functions named in oracle.cpp). */
void or_assemble_u(or_state* s);
void or_bc_u(or_state* s);
double or_residual_u(or_state* s);
void or_sweep_u(or_state* s, int nsweeps);          /* momentum_solver of the params */
void or_assemble_v(or_state* s);
void or_bc_v(or_state* s);
double or_residual_v(or_state* s);
void or_sweep_v(or_state* s, int nsweeps);
double or_assemble_p(or_state* s);
void or_solve_p(or_state* s);                       /* pressure_solver of the params */
void or_correct(or_state* s);
void or_outer_iteration(or_state* s, double res[3]);
/* run_outer_loop: up to maxit iterations with the reference stop rule; log is 4*maxit doubles
 * [iter, res_u, res_v, res_p]; returns iterations done, *converged set. */
int or_run(or_state* s, double* log, int* converged);
/* Fixed work: n outer iterations, wall seconds returned. */
double or_run_fixed(or_state* s, int n, double* last_res);

/* Individual inner solvers (for brute-force parity tests on random systems). */
void or_sweep_u_kind(or_state* s, int kind, int nsweeps); /* SG_MOM_* */
void or_sweep_v_kind(or_state* s, int kind, int nsweeps);
/* momentum_precision = mixed: nsweeps FP32 multicolour sweeps of the correction inside FP64
 * defect correction (the GPU's form, no reference counterpart); or_outer_iteration uses them when
 * the params say so. */
void or_sweep_u_dc(or_state* s, int nsweeps);
void or_sweep_v_dc(or_state* s, int nsweeps);
void or_solve_p_kind(or_state* s, int kind);              /* SG_P_* */
int or_cg_last_iterations(or_state* s);
double or_cg_last_rel(or_state* s);
long long or_cg_total_iterations(or_state* s);

/* Coefficients of one control volume: which = 0 convective, 1 diffusive, 2 their sum;
 * out = AE, AW, AN, AS, AEE, AWW, ANN, ASS, AP. */
void or_coefficients(const sg_params* p, double ue, double uw, double vn, double vs, int i, int j,
                     int which, double out[9]);
int or_in_band(const sg_params* p, int i, int j);

/* Multigrid internals. */
int or_mg_nlevels(or_state* s);
void or_mg_dims(or_state* s, int level, int* nx, int* ny);
void or_mg_build(or_state* s);
/* Copies level l's interior (nx x ny, row-major) of array k (0 AE,1 AW,2 AN,3 AS,4 AP,5 b,6 x,7 r). */
void or_mg_get(or_state* s, int level, int k, double* out);
void or_mg_vcycle(or_state* s);
int or_mg_last_coarse_sweeps(or_state* s);
/* out = A x on the level-0 p' system (interior only; halo entries untouched). */
void or_apply_A_p(or_state* s, const double* x, double* out);

#ifdef __cplusplus
}
#endif
#endif
