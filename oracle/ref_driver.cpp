// ref_driver.cpp — C entry points over the reference's own CPU code, compiled from
// /root/reference/src by oracle/ref.mk into oracle/_ref/libsimple_ref.so. TEST INFRASTRUCTURE
// ONLY: it lets tests check the oracle (oracle.cpp) against the reference itself on the same
// inputs. It is built only where the reference sources are (this build container); the built
// library travels with the tree, and on the GPU box only oracle/_ref/adapter_check links it.
#include "simple/config.hpp"
#include "simple/momentum_u.hpp"
#include "simple/momentum_v.hpp"
#include "simple/multigrid.hpp"
#include "simple/params.hpp"
#include "simple/pressure.hpp"
#include "simple/solver.hpp"

#include <cstring>
#include <memory>
#include <sstream>
#include <string>

namespace {
struct Ref {
    Params p;
    std::unique_ptr<State> s;
    std::string err;
};

Field2D* field(Ref* r, int id) {
    State& s = *r->s;
    Field2D* cu[9] = {&s.cu.AE, &s.cu.AW, &s.cu.AN, &s.cu.AS, &s.cu.AEE,
                      &s.cu.AWW, &s.cu.ANN, &s.cu.ASS, &s.cu.AP};
    Field2D* cv[9] = {&s.cv.AE, &s.cv.AW, &s.cv.AN, &s.cv.AS, &s.cv.AEE,
                      &s.cv.AWW, &s.cv.ANN, &s.cv.ASS, &s.cv.AP};
    Field2D* pw[7] = {&s.pw.AEP, &s.pw.AWP, &s.pw.ANP, &s.pw.ASP, &s.pw.APP, &s.pw.b, &s.pw.pc};
    if (id == 0) return &s.u;
    if (id == 1) return &s.v;
    if (id == 2) return &s.p;
    if (id >= 3 && id < 12) return cu[id - 3];
    if (id >= 12 && id < 21) return cv[id - 12];
    if (id >= 21 && id < 28) return pw[id - 21];
    return nullptr;
}
} // namespace

extern "C" {

// Config text in the reference's key = value format; returns nullptr and fills err on failure.
void* ref_create(const char* cfg, char* err, int errlen) {
    try {
        std::istringstream in(cfg);
        auto* r = new Ref;
        r->p = make_params(parse_config(in));
        r->s = std::make_unique<State>(r->p);
        return r;
    } catch (const std::exception& e) {
        if (err && errlen > 0) {
            std::strncpy(err, e.what(), errlen - 1);
            err[errlen - 1] = 0;
        }
        return nullptr;
    }
}
void ref_destroy(void* h) { delete static_cast<Ref*>(h); }
double* ref_field(void* h, int id) {
    Field2D* f = field(static_cast<Ref*>(h), id);
    return f ? f->data() : nullptr;
}
// Derived parameters: nrows, ncols, dx, dy, nu.
void ref_derived(void* h, int* nrows, int* ncols, double* dx, double* dy, double* nu) {
    const Params& p = static_cast<Ref*>(h)->p;
    *nrows = p.nrows;
    *ncols = p.ncols;
    *dx = p.dx;
    *dy = p.dy;
    *nu = p.nu;
}
void ref_outer_iteration(void* h, double* res) {
    Ref* r = static_cast<Ref*>(h);
    PhaseTimers t;
    Residuals x = outer_iteration(r->p, *r->s, t);
    res[0] = x.u;
    res[1] = x.v;
    res[2] = x.p;
}
// run_simple; log gets 4 * iterations doubles (capacity maxit rows).
int ref_run(void* h, double* log, int* converged) {
    Ref* r = static_cast<Ref*>(h);
    RunResult rr = run_simple(r->p, *r->s);
    for (size_t k = 0; k < rr.log.size(); ++k)
        for (int m = 0; m < 4; ++m)
            log[4 * k + m] = rr.log[k][m];
    *converged = rr.converged ? 1 : 0;
    return rr.iterations;
}
// Individual phases.
void ref_assemble_u(void* h) { Ref* r = static_cast<Ref*>(h); assemble_u(r->p, r->s->u, r->s->v, r->s->cu); }
void ref_bc_u(void* h) { Ref* r = static_cast<Ref*>(h); apply_bc_u(r->p, r->s->u); }
double ref_residual_u(void* h) { Ref* r = static_cast<Ref*>(h); return residual_u(r->p, r->s->cu, r->s->u, r->s->p); }
void ref_sweep_u(void* h, int n) { Ref* r = static_cast<Ref*>(h); sweep_u_batch(r->p, r->s->cu, r->s->u, r->s->p, r->s->scratch, n); }
void ref_assemble_v(void* h) { Ref* r = static_cast<Ref*>(h); assemble_v(r->p, r->s->u, r->s->v, r->s->cv); }
void ref_bc_v(void* h) { Ref* r = static_cast<Ref*>(h); apply_bc_v(r->p, r->s->v); }
double ref_residual_v(void* h) { Ref* r = static_cast<Ref*>(h); return residual_v(r->p, r->s->cv, r->s->v, r->s->p); }
void ref_sweep_v(void* h, int n) { Ref* r = static_cast<Ref*>(h); sweep_v_batch(r->p, r->s->cv, r->s->v, r->s->p, r->s->scratch, n); }
double ref_assemble_p(void* h) { Ref* r = static_cast<Ref*>(h); return assemble_p(r->p, r->s->cu.AP, r->s->cv.AP, r->s->u, r->s->v, r->s->pw); }
void ref_solve_p(void* h) {
    Ref* r = static_cast<Ref*>(h);
    switch (r->p.pressure_solver) {
    case PressureSolver::Jacobi: solve_pc_jacobi(r->p, r->s->pw, r->s->scratch); break;
    case PressureSolver::GsRedBlack: solve_pc_redblack(r->p, r->s->pw); break;
    case PressureSolver::CG: solve_pc_cg(r->p, r->s->pw, r->s->cg); break;
    case PressureSolver::MG: solve_pc_mg(r->p, r->s->pw, r->s->mg); break;
    default: solve_pc_gs_lex(r->p, r->s->pw);
    }
}
void ref_correct(void* h) { Ref* r = static_cast<Ref*>(h); correct(r->p, r->s->pw, r->s->cu.AP, r->s->cv.AP, r->s->u, r->s->v, r->s->p); }
int ref_cg_last_iterations(void* h) { return static_cast<Ref*>(h)->s->cg.last_iterations; }
int ref_mg_nlevels(void* h) { return static_cast<int>(static_cast<Ref*>(h)->s->mg.levels.size()); }
void ref_set_threads(int n);
}

#ifdef _OPENMP
#include <omp.h>
void ref_set_threads(int n) { omp_set_num_threads(n); }
#else
void ref_set_threads(int) {}
#endif
