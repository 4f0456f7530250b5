// solver_sg.cpp — the reference's HipSolver and run_simple_hip (include/simple/solver_hip.hpp:11-40)
// implemented over the MI355X backend's C ABI (include/simple_gpu.h). This file belongs to the
// reference's tree: it replaces src/hip/solver_hip.hip there, compiles against the reference's
// own headers (Params, State, Residuals, PhaseTimers, run_outer_loop) and links
// build/libsimple_gpu.so. It is not part of this repository's build; oracle/ref.mk compiles it
// against the reference's headers to check it (oracle/adapter_check.cpp).
//
// Semantics kept from src/hip/solver_hip.hip:
//  - the constructor uploads every State array the GPU uses (:398-404): u, v, p, the 9 + 9
//    momentum coefficients and the 7 p' arrays;
//  - download(State&) fills every one of them back (:525-536);
//  - a failing call throws std::runtime_error("GPU error: ...") (:17-23), so main's
//    "error (rank 0): ..." path and exit code 1 are unchanged (src/main.cpp:77-83).
#include "simple/solver_hip.hpp"
#include "simple_gpu.h"

#include <cstdio>
#include <stdexcept>
#include <string>

namespace {

void check(int rc) {
    if (rc != SG_OK)
        throw std::runtime_error(sg_last_error()); // "GPU error: ..." for HIP failures
}

sg_params to_sg(const Params& p) {
    sg_params g;
    sg_params_default(&g);
    g.NX = p.NX;
    g.NY = p.NY;
    g.N_halo_x = p.N_halo_x;
    g.N_halo_y = p.N_halo_y;
    g.Re = p.Re;
    g.maxit_inner_u = p.maxit_inner_u;
    g.maxit_inner_v = p.maxit_inner_v;
    g.maxit_inner_p = p.maxit_inner_p;
    g.error_tolerance = p.error_tolerance;
    g.maxit = p.maxit;
    g.relaxation_factor_u = p.relaxation_factor_u;
    g.relaxation_factor_v = p.relaxation_factor_v;
    g.relaxation_factor_p = p.relaxation_factor_p;
    g.convection_scheme = p.convection_scheme == Convection::QUICK ? SG_QUICK : SG_UDS;
    g.diffusion_order = p.diffusion_order;
    g.Hybrid_switch = p.Hybrid_switch ? 1 : 0;
    g.residual_normalisation =
        p.residual_normalisation == ResidualNorm::RunningMax ? SG_NORM_RUNNING_MAX
                                                             : SG_NORM_ABSOLUTE;
    g.backend = SG_BACKEND_GPU;
    g.momentum_solver = static_cast<int>(p.momentum_solver); // same enumerator order
    g.pressure_solver = static_cast<int>(p.pressure_solver); // same enumerator order
    g.pressure_tolerance = p.pressure_tolerance;
    g.pressure_maxit = p.pressure_maxit;
    g.mg_cycles = p.mg_cycles;
    g.mg_pre = p.mg_pre;
    g.mg_post = p.mg_post;
    g.mg_omega = p.mg_omega;
    g.write_fields = p.write_fields ? 1 : 0;
    check(sg_params_finalize(&g)); // H, nrows, ncols, dx, dy, nu as in src/params.cpp:173-178
    return g;
}

// Every State array the GPU holds, by SG_F_* id (the enumerator order of include/simple_gpu.h).
template <class S, class F> auto* state_field(S& st, int id) {
    F* cu[9] = {&st.cu.AE, &st.cu.AW, &st.cu.AN, &st.cu.AS, &st.cu.AEE,
                &st.cu.AWW, &st.cu.ANN, &st.cu.ASS, &st.cu.AP};
    F* cv[9] = {&st.cv.AE, &st.cv.AW, &st.cv.AN, &st.cv.AS, &st.cv.AEE,
                &st.cv.AWW, &st.cv.ANN, &st.cv.ASS, &st.cv.AP};
    F* pw[7] = {&st.pw.AEP, &st.pw.AWP, &st.pw.ANP, &st.pw.ASP, &st.pw.APP, &st.pw.b, &st.pw.pc};
    if (id == SG_F_U)
        return &st.u;
    if (id == SG_F_V)
        return &st.v;
    if (id == SG_F_P)
        return &st.p;
    if (id >= SG_F_CU_AE && id <= SG_F_CU_AP)
        return cu[id - SG_F_CU_AE];
    if (id >= SG_F_CV_AE && id <= SG_F_CV_AP)
        return cv[id - SG_F_CV_AE];
    return pw[id - SG_F_AEP];
}

} // namespace

struct HipSolver::Impl {
    sg_solver* s = nullptr;
};

HipSolver::HipSolver(const Params& p, const State& st) : d_(new Impl) {
    const sg_params g = to_sg(p);
    check(sg_create(&g, st.u.data(), st.v.data(), st.p.data(), &d_->s));
    try {
        for (int id = SG_F_CU_AE; id < SG_F_COUNT; ++id)
            check(sg_upload(d_->s, id, state_field<const State, const Field2D>(st, id)->data()));
    } catch (...) {
        sg_destroy(d_->s);
        throw;
    }
}

HipSolver::~HipSolver() { sg_destroy(d_->s); }

void HipSolver::assemble_u() { check(sg_assemble_u(d_->s)); }
void HipSolver::bc_u() { check(sg_bc_u(d_->s)); }
double HipSolver::residual_u() {
    double r = 0.0;
    check(sg_residual_u(d_->s, &r));
    return r;
}
void HipSolver::sweep_u(int n) { check(sg_sweep_u(d_->s, n)); }
void HipSolver::assemble_v() { check(sg_assemble_v(d_->s)); }
void HipSolver::bc_v() { check(sg_bc_v(d_->s)); }
double HipSolver::residual_v() {
    double r = 0.0;
    check(sg_residual_v(d_->s, &r));
    return r;
}
void HipSolver::sweep_v(int n) { check(sg_sweep_v(d_->s, n)); }
double HipSolver::assemble_p() {
    double r = 0.0;
    check(sg_assemble_p(d_->s, &r));
    return r;
}
void HipSolver::solve_p() { check(sg_solve_p(d_->s)); }
void HipSolver::correct() { check(sg_correct(d_->s)); }

Residuals HipSolver::outer_iteration(PhaseTimers& t) {
    double r[3], ph[SG_NPHASES] = {};
    check(sg_outer_iteration(d_->s, r, ph)); // phases in the order of Phase (timer.hpp)
    for (int k = 0; k < SG_NPHASES; ++k)
        t.seconds[k] += ph[k];
    return {r[0], r[1], r[2]};
}

void HipSolver::download(State& st) const {
    for (int id = 0; id < SG_F_COUNT; ++id)
        check(sg_download(d_->s, id, state_field<State, Field2D>(st, id)->data()));
}

RunResult run_simple_hip(const Params& p, State& s) {
    char name[256];
    if (sg_device_name(name, sizeof name) == SG_OK)
        std::printf("gpu: %s\n", name);
    HipSolver g(p, s);
    // The reference's own loop and stop rule, one eager outer iteration per step.
    RunResult out = run_outer_loop(p, [&](PhaseTimers& t) { return g.outer_iteration(t); });
    g.download(s);
    return out;
}
