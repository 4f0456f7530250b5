// sg_api.hip — the C ABI of include/simple_gpu.h: solver lifetime, uploads/downloads, one entry
// point per solver phase, the eager outer iteration with hipEvent phase timers, and sg_advance,
// which replays one captured hipGraph per outer iteration with the stopping rule evaluated on the
// device (the reference's HipSolver and run_simple_hip, include/simple/solver_hip.hpp:11-40, and
// run_outer_loop, src/solver.cpp:106-132).
#include "sg_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <new>

namespace sg {

namespace {
thread_local std::string g_err;

__device__ __forceinline__ void k_stop_rule(const StopArgs& a, const Blk&) {
    if (threadIdx.x == 0)
        stop_rule(a);
}

__global__ void k_ctl_begin(Ctl* ctl) {
    if (threadIdx.x == 0) {
        ctl->stop = 0;
        ctl->log_pos = 0;
        ctl->converged = 0;
    }
}

// Diffusive coefficients, src/coefficients.cpp:41-63 (host arithmetic, IEEE double, no
// contraction: the same bits as the CPU).
void diffusive(const sg_params& p, C9& d2, C9& d4) {
    const double dx = p.dx, dy = p.dy, nu = p.nu;
    d2.v[0] = d2.v[1] = dy * nu / dx;
    d2.v[2] = d2.v[3] = dx * nu / dy;
    d2.v[4] = d2.v[5] = d2.v[6] = d2.v[7] = 0.0;
    d2.v[8] = -nu * (-2 * dx / dy - 2 * dy / dx);
    d4.v[0] = d4.v[1] = 4.0 / 3 * dy * nu / dx;
    d4.v[2] = d4.v[3] = 4.0 / 3 * dx * nu / dy;
    d4.v[4] = d4.v[5] = -1.0 / 12 * dy * nu / dx;
    d4.v[6] = d4.v[7] = -1.0 / 12 * dx * nu / dy;
    d4.v[8] = -nu * (-2.5 * dx / dy - 2.5 * dy / dx);
}

template <class T> T* dalloc(Solver& s, size_t n) {
    void* q = nullptr;
    SG_HIP(hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1)));
    s.owned.push_back(q);
    SG_HIP(hipMemset(q, 0, sizeof(T) * std::max<size_t>(n, 1)));
    return static_cast<T*>(q);
}

} // namespace

// The arrays' state after replayed outer iterations (their assemblies, solves and corrections
// ran on the device): which form holds the momentum coefficients, the pressure terms stale, the
// lean p' operator, p' in scratch and b with its mean (see Solver).
void mark_iterated(Solver& s) {
    const bool dc = mom_dc_ok(s.p);
    s.face[0] = !dc && face_ok(s, true), s.face[1] = !dc && face_ok(s, false);
    s.dc_stale[0] = s.dc_stale[1] = dc;
    s.pt_stale[0] = s.pt_stale[1] = true;
    s.p_lean = mg_lean_ok(s);
    s.pc_in_scratch = s.p.pressure_solver == SG_P_MG;
    s.b_raw = mg_fuse_b_ok(s);
}

void drop_graphs(Solver& s) {
    for (hipGraphExec_t* g : {&s.graph, &s.graph_n})
        if (*g) {
            SG_HIP(hipGraphExecDestroy(*g));
            *g = nullptr;
        }
}

void ensure_log(Solver& s, int cap) {
    if (cap <= s.log_cap)
        return;
    if (s.dlog) {
        SG_HIP(hipFree(s.dlog));
        s.owned.erase(std::remove(s.owned.begin(), s.owned.end(), static_cast<void*>(s.dlog)),
                      s.owned.end());
    }
    s.log_cap = std::max(cap, 256);
    s.dlog = dalloc<double>(s, 3 * static_cast<size_t>(s.log_cap));
    drop_graphs(s); // the stop-rule node holds the old log pointer
}

// One outer iteration (src/solver.cpp:27-93), enqueued. ev: optional phase events (12).
// The residual sums (they feed only the stop rule) are deferred and done together before it.
void enqueue_iteration(Solver& s, const int* stop, bool eager, hipEvent_t* ev, int fused) {
    const sg_params& p = s.p;
    // per-phase iterations sum each residual in its own phase
    const bool defer = fused < 0 ? !ev : fused != 0;
    auto mark = [&](int k) {
        if (ev)
            SG_HIP(hipEventRecord(ev[k], s.stream));
    };
    mark(0);
    if (mom_dc_ok(p)) {
        // FP32 sweeps of the corrections inside FP64 defect correction: each assembly also
        // evaluates the FP64 residual (its |r| cells give res_u / res_v) and the FP32 pack; u's
        // ghost rows are set by its assembly launch, v's by bc_v after v's assembly (which reads
        // the old ones) and before anything reads them.
        auto res_sum = [&](bool is_u) {
            const int nrow = is_u ? p.NX : p.NX - 1, w = is_u ? p.NY - 1 : p.NY;
            double* slot = s.scal + (is_u ? S_RES_U : S_RES_V);
            if (defer)
                defer_sum(s, is_u ? 1 : 2, s.rcell[is_u ? 0 : 1], nrow, w, slot);
            else
                launch_ordered_sum(s, s.rcell[is_u ? 0 : 1], w, 0, nrow - 1, 0, w - 1, slot,
                                   stop);
        };
        launch_asm_dc(s, true, stop, true);
        mark(1);
        mark(2);
        res_sum(true);
        mark(3);
        launch_dc_sweeps(s, true, p.maxit_inner_u, stop);
        mark(4);
        launch_asm_dc(s, false, stop, false);
        mark(5);
        launch_bc(s, false, stop);
        mark(6);
        res_sum(false);
        mark(7);
        launch_dc_sweeps(s, false, p.maxit_inner_v, stop);
        mark(8);
    } else {
        // Without per-phase events, u's ghost rows are set by the assembly launch itself.
        const bool fuse_bc = defer && s.pku.rows >= 2;
        launch_assemble(s, true, stop, fuse_bc, face_ok(s, true), true);
        mark(1);
        if (!fuse_bc)
            launch_bc(s, true, stop);
        mark(2);
        mark(3);
        launch_mom_sweeps(s, true, p.maxit_inner_u, stop, s.scal + S_RES_U, defer);
        mark(4);
        launch_assemble(s, false, stop, false, face_ok(s, false), true);
        mark(5);
        // Without per-phase events, v's ghost columns are set by its first sweep.
        const bool bcv = defer && bcv_in_sweep(s);
        if (!bcv)
            launch_bc(s, false, stop);
        mark(6);
        mark(7);
        launch_mom_sweeps(s, false, p.maxit_inner_v, stop, s.scal + S_RES_V, defer, bcv);
        mark(8);
    }
    launch_assemble_p(s, s.scal + S_RES_P, stop, defer, mg_lean_ok(s));
    mark(9);
    // The residual sums, all deferred by now, may ride along with the p' solve's first mean
    // removal (multigrid), which sums its right-hand side anyway; the multigrid solve's last mean
    // removal is carried by correct.
    s.ride = defer && !s.rec;
    s.fuse_pc_mean = defer && p.pressure_solver == SG_P_MG;
    s.fuse_b_mean = defer && mg_fuse_b_ok(s);
    launch_solve_p(s, stop, eager);
    s.ride = false;
    s.fuse_pc_mean = s.fuse_b_mean = false;
    const bool rode = s.rode;
    s.rode = false;
    mark(10);
    launch_correct(s, stop);
    mark(11);
    const StopArgs rule{s.ctl, s.scal, s.dlog, s.log_cap, p.residual_normalisation,
                        p.error_tolerance};
    // Captured iterations (every launch skips once the rule has fired): the residual totals and
    // the rule in one launch, unless the totals are done already.
    if (defer && stop == &s.ctl->stop && !rode) {
        flush_sums(s, stop, &rule);
        return;
    }
    if (defer && !rode)
        flush_sums(s, stop);
    emit<StopArgs, k_stop_rule, 64>(s, dim3(1), dim3(64), 0, rule, &s.ctl->stop);
}

namespace {

// Capture `reps` outer iterations into *out. The coefficient form flags are left as they were:
// sg_advance sets them after replaying the graphs.
void capture_graph(Solver& s, int reps, hipGraphExec_t* out) {
    if (*out)
        return;
    struct Restore {
        Solver& s;
        bool f0, f1, p0, p1, d0, d1, lean, pcs, braw;
        ~Restore() {
            s.face[0] = f0, s.face[1] = f1, s.pt_stale[0] = p0, s.pt_stale[1] = p1;
            s.dc_stale[0] = d0, s.dc_stale[1] = d1;
            s.p_lean = lean, s.pc_in_scratch = pcs, s.b_raw = braw;
            s.fuse_pc_mean = s.pc_mean_pending = s.fuse_b_mean = s.b_mean_pending = false;
            s.ap_in_sweep[0] = s.ap_in_sweep[1] = false;
        }
    } restore{s,           s.face[0],   s.face[1], s.pt_stale[0],    s.pt_stale[1],
              s.dc_stale[0], s.dc_stale[1], s.p_lean, s.pc_in_scratch, s.b_raw};
    hipGraph_t g = nullptr;
    SG_HIP(hipStreamBeginCapture(s.stream, hipStreamCaptureModeRelaxed));
    try {
        for (int r = 0; r < reps; ++r)
            enqueue_iteration(s, &s.ctl->stop, false, nullptr, 1);
    } catch (...) {
        s.ride = s.rode = false;
        s.deferred.clear();
        hipGraph_t tmp = nullptr;
        (void)hipStreamEndCapture(s.stream, &tmp);
        if (tmp)
            (void)hipGraphDestroy(tmp);
        throw;
    }
    SG_HIP(hipStreamEndCapture(s.stream, &g));
    hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    SG_HIP(e);
}

void sync(Solver& s) { SG_HIP(hipStreamSynchronize(s.stream)); }

Ctl read_ctl(Solver& s) {
    Ctl c;
    SG_HIP(hipMemcpyAsync(&c, s.ctl, sizeof c, hipMemcpyDeviceToHost, s.stream));
    sync(s);
    return c;
}

template <class F> int guard(F f) {
    try {
        g_err.clear();
        f();
        return SG_OK;
    } catch (const HipError& e) {
        g_err = e.what();
        (void)hipGetLastError(); // the failed call's error is reported here, not by the next call
        return SG_E_HIP;
    } catch (const ArgError& e) {
        g_err = e.what();
        return SG_E_ARG;
    } catch (const std::bad_alloc&) {
        g_err = "out of host memory";
        return SG_E_NOMEM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return SG_E_STATE;
    }
}

// Every entry point reaches its solver here, so p' is back in place (materialize_pc) before any
// of them reads or replaces it.
Solver& S(sg_solver* h) {
    if (!h)
        throw ArgError("null solver");
    Solver& s = *reinterpret_cast<Solver*>(h);
    materialize_pc(s);
    materialize_b(s);
    return s;
}

void validate(const sg_params& p) {
    // The checks of src/params.cpp:152-172, same messages.
    if (p.N_halo_x != p.N_halo_y)
        throw ArgError("N_halo_x must equal N_halo_y");
    if (p.N_halo_x < 2)
        throw ArgError("halo width H must be >= 2");
    if (p.diffusion_order != 2 && p.diffusion_order != 4)
        throw ArgError("diffusion_order must be 2 or 4");
    if (p.NX < 1 || p.NY < 1)
        throw ArgError("N_cells_x and N_cells_y must be positive");
    if (p.Hybrid_switch)
        throw ArgError("Hybrid_switch = True is not supported");
    if (!(p.pressure_tolerance > 0.0))
        throw ArgError("pressure_tolerance must be positive");
    if (p.pressure_maxit < 1)
        throw ArgError("pressure_maxit must be >= 1");
    if (p.mg_cycles < 1)
        throw ArgError("mg_cycles must be >= 1");
    if (p.mg_pre < 0 || p.mg_post < 0)
        throw ArgError("mg_pre and mg_post must be >= 0");
    if (!(p.mg_omega > 0.0 && p.mg_omega <= 2.0))
        throw ArgError("mg_omega must be in (0, 2]");
    // Additions.
    if (p.mg_coarsest < 1)
        throw ArgError("mg_coarsest must be >= 1");
    if (p.mg_precision != SG_MG_FP64 && p.mg_precision != SG_MG_MIXED)
        throw ArgError("mg_precision must be fp64 or mixed");
    if (p.mg_dc_maxit < 1)
        throw ArgError("mg_dc_maxit must be >= 1");
    if (p.check_every < 1)
        throw ArgError("check_every must be >= 1");
    if (p.momentum_precision != SG_MOMP_FP64 && p.momentum_precision != SG_MOMP_MIXED)
        throw ArgError("momentum_precision must be fp64 or mixed");
    if (p.momentum_precision == SG_MOMP_MIXED &&
        (p.momentum_solver != SG_MOM_GS_MULTICOLOUR || p.maxit_inner_u < 1 ||
         p.maxit_inner_v < 1 || p.NX < 2 || p.NY < 2))
        throw ArgError("momentum_precision = mixed needs momentum_solver = gs_multicolour, "
                       "maxit_inner_u and maxit_inner_v >= 1, and at least 2 x 2 cells");
    if (p.pressure_warm_start == SG_WARM_LEARNED &&
        (p.pressure_solver != SG_P_MG || p.mg_precision != SG_MG_MIXED))
        throw ArgError("pressure_warm_start = learned needs pressure_solver = mg and "
                       "mg_precision = mixed");
}

} // namespace

void set_error(const std::string& m) { g_err = m; }

} // namespace sg

using namespace sg;

extern "C" {

const char* sg_last_error(void) { return g_err.c_str(); }
int sg_abi_version(void) { return SG_ABI_VERSION; }

int sg_device_count(int* n) {
    return guard([&] {
        if (!n)
            throw ArgError("null pointer");
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess)
            c = 0;
        *n = c;
    });
}

int sg_device_name(char* buf, size_t len) {
    return guard([&] {
        hipDeviceProp_t prop;
        SG_HIP(hipGetDeviceProperties(&prop, 0));
        if (buf && len) {
            // Some boxes report an empty marketing name: then the architecture leads.
            if (prop.name[0] != '\0')
                std::snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName,
                              prop.multiProcessorCount);
            else
                std::snprintf(buf, len, "%s, %d CUs", prop.gcnArchName,
                              prop.multiProcessorCount);
        }
    });
}

void sg_params_default(sg_params* p) {
    std::memset(p, 0, sizeof *p);
    // include/simple/params.hpp:15-36
    p->NX = p->NY = 80;
    p->H = p->N_halo_x = p->N_halo_y = 2;
    p->Re = 1000.0;
    p->maxit_inner_u = 4;
    p->maxit_inner_v = 4;
    p->maxit_inner_p = 8;
    p->error_tolerance = 1e-4;
    p->maxit = 15000;
    p->relaxation_factor_u = 0.3;
    p->relaxation_factor_v = 0.3;
    p->relaxation_factor_p = 0.2;
    p->convection_scheme = SG_QUICK;
    p->diffusion_order = 4;
    p->Hybrid_switch = 0;
    p->residual_normalisation = SG_NORM_ABSOLUTE;
    p->backend = SG_BACKEND_SERIAL;
    p->momentum_solver = SG_MOM_GS_LEX;
    p->pressure_solver = SG_P_GS_LEX;
    p->pressure_tolerance = 1e-6;
    p->pressure_maxit = 1000;
    p->mg_cycles = 1;
    p->mg_pre = 2;
    p->mg_post = 2;
    p->mg_omega = 1.8;
    p->write_fields = 1;
    // additions
    p->mg_precision = SG_MG_FP64;
    p->mg_coarsest = 4;
    p->mg_dc_maxit = 1;
    p->pressure_warm_start = SG_WARM_NONE;
    p->check_every = 200;
    p->momentum_precision = SG_MOMP_FP64;
}

int sg_params_finalize(sg_params* p) {
    return guard([&] {
        if (!p)
            throw ArgError("null params");
        validate(*p);
        // src/params.cpp:173-178
        p->H = p->N_halo_x;
        p->nrows = p->NX + 2 * p->H + 1;
        p->ncols = p->NY + 2 * p->H + 1;
        p->dx = 1.0 / p->NY;
        p->dy = 1.0 / p->NX;
        p->nu = 1.0 / p->Re;
    });
}

} // extern "C"

namespace sg {

// Launch shapes before any autotuning (results do not depend on them): what the autotuner chose
// on an MI355X for the 1024^2 benchmark workload and at 64^2 to 512^2. Every size kept the same
// multigrid knobs (a 1024-thread tail holding only the levels that fit in LDS, no
// intermediate levels in that workgroup); the sweep tile grows with the grid: 16 x 128 at 64^2,
// 24 x 64 up to 512^2, 40 x 56 (first sweep 28 x 80) at 1024^2.
void default_launch_shapes(Solver& s) {
    s.tail_threads = 1024;
    s.mid_cells = 0;
    s.lds_cells = 4096;
    s.mg_staged = 0;
    const long cells = static_cast<long>(s.p.NX) * s.p.NY;
    s.mom_tile = cells <= 96L * 96 ? 3 : cells < 768L * 768 ? 6 : 11;
    s.mom_tile_first = cells < 768L * 768 ? -1 : 9;
}

Solver* create_solver(const sg_params& p, const double* u, const double* v, const double* pf) {
    validate(p);
    if (p.nrows != p.NX + 2 * p.H + 1 || p.ncols != p.NY + 2 * p.H + 1)
        throw ArgError("params not finalized (call sg_params_finalize)");
    if (p.momentum_solver == SG_MOM_GS_LEX)
        throw ArgError("backend = gpu runs momentum_solver = gs_multicolour or jacobi "
                       "(gs_lex is a serial sweep)");
    if (p.pressure_solver == SG_P_GS_LEX)
        throw ArgError("backend = gpu runs pressure_solver = gs_redblack, jacobi, cg or mg "
                       "(gs_lex is a serial sweep)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        throw HipError("GPU error: no HIP device visible");
    Solver* sp = new Solver;
    Solver& s = *sp;
    try {
        s.p = p;
        SG_HIP(hipGetDevice(&s.device));
        {
            hipDeviceProp_t prop;
            SG_HIP(hipGetDeviceProperties(&prop, s.device));
            s.lds_budget = std::min<size_t>(prop.sharedMemPerBlock, 160 * 1024) - 1024;
        }
        default_launch_shapes(s);
        s.nr = p.nrows;
        s.nc = p.ncols;
        s.ld = (p.ncols + 15) / 16 * 16;
        s.n = static_cast<size_t>(s.nr) * s.ld;
        SG_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
        for (int k = 0; k < SG_F_COUNT; ++k)
            s.f[k] = dalloc<double>(s, s.n);
        s.scratch = dalloc<double>(s, s.n);
        if (p.momentum_solver == SG_MOM_GS_MULTICOLOUR)
            s.scratch2 = dalloc<double>(s, s.n);
        s.cell = dalloc<double>(s, 4 * static_cast<size_t>(std::max(p.NX + 2, 1)) * (p.NY + 2));
        s.rowbuf = dalloc<double>(s, 4 * static_cast<size_t>(s.nr));
        s.scal = dalloc<double>(s, S_COUNT);
        for (auto& c : s.rcell)
            c = dalloc<double>(s, static_cast<size_t>(p.NX + 2) * (p.NY + 2));
        s.ctl = dalloc<Ctl>(s, 1);
        s.never = dalloc<int>(s, 1);
        if (p.pressure_solver == SG_P_CG) {
            for (auto& c : s.cg)
                c = dalloc<double>(s, s.n);
            s.cgctl = dalloc<int>(s, 2);
        }
        s.pku = make_pack(p, true);
        s.pkv = make_pack(p, false);
        s.pku.base = dalloc<double>(s, 10 * s.pku.astride);
        s.pkv.base = dalloc<double>(s, 10 * s.pkv.astride);
        s.fku = make_face_pack(p, true);
        s.fkv = make_face_pack(p, false);
        s.fku.base = dalloc<double>(s, 3 * s.fku.astride);
        s.fkv.base = dalloc<double>(s, 3 * s.fkv.astride);
        if (mom_dc_ok(p)) {
            s.dcf[0] = dalloc<float>(s, 10 * s.pku.astride);
            s.dcf[1] = dalloc<float>(s, 10 * s.pkv.astride);
            for (auto& d : s.dd)
                d = dalloc<float>(s, s.n);
        }
        ensure_log(s, 256);
        diffusive(p, s.diff2, s.diff4);
        if (p.pressure_solver == SG_P_MG)
            mg_setup(s);
        for (auto& e : s.ev)
            SG_HIP(hipEventCreate(&e));
        for (auto& e : s.pev)
            SG_HIP(hipEventCreate(&e));
        SG_HIP(hipEventCreate(&s.kt0));
        SG_HIP(hipEventCreate(&s.kt1));
        const double* init[3] = {u, v, pf};
        for (int k = 0; k < 3; ++k)
            if (init[k])
                SG_HIP(hipMemcpy2D(s.f[k], sizeof(double) * s.ld, init[k],
                                   sizeof(double) * s.nc, sizeof(double) * s.nc, s.nr,
                                   hipMemcpyHostToDevice));
        SG_HIP(hipDeviceSynchronize());
    } catch (...) {
        destroy_solver(sp);
        throw;
    }
    return sp;
}

void destroy_solver(Solver* s) {
    if (s->stream)
        (void)hipStreamSynchronize(s->stream);
    warm_free(*s);
    if (s->graph)
        (void)hipGraphExecDestroy(s->graph);
    if (s->graph_n)
        (void)hipGraphExecDestroy(s->graph_n);
    for (auto e : s->ev)
        if (e)
            (void)hipEventDestroy(e);
    for (auto e : s->pev)
        if (e)
            (void)hipEventDestroy(e);
    if (s->kt0)
        (void)hipEventDestroy(s->kt0);
    if (s->kt1)
        (void)hipEventDestroy(s->kt1);
    for (void* q : s->owned)
        (void)hipFree(q);
    if (s->stream)
        (void)hipStreamDestroy(s->stream);
    delete s;
}

} // namespace sg

extern "C" {

int sg_create(const sg_params* pin, const double* u, const double* v, const double* pf,
              sg_solver** out) {
    return guard([&] {
        if (!pin || !out)
            throw ArgError("null argument");
        *out = reinterpret_cast<sg_solver*>(create_solver(*pin, u, v, pf));
    });
}

int sg_destroy(sg_solver* h) {
    if (h)
        destroy_solver(reinterpret_cast<Solver*>(h));
    return SG_OK;
}

int sg_upload(sg_solver* h, int id, const double* host) {
    return guard([&] {
        Solver& s = S(h);
        if (id < 0 || id >= SG_F_COUNT || !host)
            throw ArgError("bad field id or null pointer");
        if (id >= SG_F_AEP && id <= SG_F_APP)
            ensure_p_full(s);
        SG_HIP(hipMemcpy2DAsync(s.f[id], sizeof(double) * s.ld, host, sizeof(double) * s.nc,
                                sizeof(double) * s.nc, s.nr, hipMemcpyHostToDevice, s.stream));
        sync_packed(s, id, true);
        sync(s);
    });
}

int sg_download(sg_solver* h, int id, double* host) {
    return guard([&] {
        Solver& s = S(h);
        if (id < 0 || id >= SG_F_COUNT || !host)
            throw ArgError("bad field id or null pointer");
        if (id >= SG_F_AEP && id <= SG_F_APP)
            ensure_p_full(s);
        sync_packed(s, id, false);
        SG_HIP(hipMemcpy2DAsync(host, sizeof(double) * s.nc, s.f[id], sizeof(double) * s.ld,
                                sizeof(double) * s.nc, s.nr, hipMemcpyDeviceToHost, s.stream));
        sync(s);
    });
}

static int phase(sg_solver* h, void (*fn)(Solver&)) {
    return guard([&] {
        Solver& s = S(h);
        fn(s);
        sync(s);
    });
}

static int phase_res(sg_solver* h, double* res, int slot, void (*fn)(Solver&)) {
    return guard([&] {
        Solver& s = S(h);
        fn(s);
        if (res) {
            SG_HIP(hipMemcpyAsync(res, s.scal + slot, sizeof(double), hipMemcpyDeviceToHost,
                                  s.stream));
        }
        sync(s);
    });
}

int sg_assemble_u(sg_solver* h) {
    return phase(h, [](Solver& s) { launch_assemble(s, true, s.never); });
}
int sg_bc_u(sg_solver* h) {
    return phase(h, [](Solver& s) { launch_bc(s, true, s.never); });
}
int sg_residual_u(sg_solver* h, double* res) {
    return phase_res(h, res, S_RES_U,
                     [](Solver& s) { launch_mom_residual(s, true, s.scal + S_RES_U, s.never, false); });
}
int sg_sweep_u(sg_solver* h, int n) {
    return guard([&] {
        Solver& s = S(h);
        launch_mom_sweeps(s, true, n, s.never, nullptr, false);
        sync(s);
    });
}
int sg_assemble_v(sg_solver* h) {
    return phase(h, [](Solver& s) { launch_assemble(s, false, s.never); });
}
int sg_bc_v(sg_solver* h) {
    return phase(h, [](Solver& s) { launch_bc(s, false, s.never); });
}
int sg_residual_v(sg_solver* h, double* res) {
    return phase_res(h, res, S_RES_V,
                     [](Solver& s) { launch_mom_residual(s, false, s.scal + S_RES_V, s.never, false); });
}
int sg_sweep_v(sg_solver* h, int n) {
    return guard([&] {
        Solver& s = S(h);
        launch_mom_sweeps(s, false, n, s.never, nullptr, false);
        sync(s);
    });
}
int sg_assemble_p(sg_solver* h, double* res) {
    return phase_res(h, res, S_RES_P,
                     [](Solver& s) { launch_assemble_p(s, s.scal + S_RES_P, s.never, false); });
}
int sg_solve_p(sg_solver* h) {
    return phase(h, [](Solver& s) {
        ensure_p_full(s);
        launch_solve_p(s, s.never, true);
    });
}
int sg_correct(sg_solver* h) {
    return phase(h, [](Solver& s) { launch_correct(s, s.never); });
}

int sg_outer_iteration(sg_solver* h, double res[3], double* phase_s) {
    return guard([&] {
        Solver& s = S(h);
        hipLaunchKernelGGL(k_ctl_begin, dim3(1), dim3(64), 0, s.stream, s.ctl);
        SG_HIP(hipGetLastError());
        enqueue_iteration(s, s.never, true, s.ev);
        if (res)
            SG_HIP(hipMemcpyAsync(res, s.scal + S_RES_U, 3 * sizeof(double),
                                  hipMemcpyDeviceToHost, s.stream));
        sync(s);
        if (phase_s)
            for (int k = 0; k < SG_NPHASES; ++k) {
                float ms = 0.f;
                SG_HIP(hipEventElapsedTime(&ms, s.ev[k], s.ev[k + 1]));
                phase_s[k] += 1e-3 * ms;
            }
        s.iters = read_ctl(s).iter;
    });
}

int sg_advance(sg_solver* h, int max_iters, int* iters_done, double* log, int* converged) {
    return guard([&] {
        Solver& s = S(h);
        if (max_iters < 0)
            throw ArgError("max_iters must be >= 0");
        ensure_log(s, max_iters);
        hipLaunchKernelGGL(k_ctl_begin, dim3(1), dim3(64), 0, s.stream, s.ctl);
        SG_HIP(hipGetLastError());
        SG_HIP(hipGetLastError());
        if (s.p.pressure_solver == SG_P_CG) {
            // CG needs the host between its iterations: eager iterations, stop flag checked.
            for (int k = 0; k < max_iters; ++k) {
                enqueue_iteration(s, &s.ctl->stop, true, nullptr);
                if (read_ctl(s).stop)
                    break;
            }
        } else {
            // Every kernel of an iteration returns at once after the stop rule fires, so replaying
            // GRAPH_ITERS iterations per launch stops at the same iteration.
            // Phase timing: one iteration enqueued launch by launch with events at the phase
            // boundaries (the replayed iteration's launches: fused), behind the first
            // GRAPH_ITERS-iteration graph when there is one, so the host has enqueued it before
            // the device reaches it and the event deltas hold no launch gaps. Every graph is
            // captured first, from the flags at the start of the call, as the eager iteration
            // is enqueued from them.
            const bool timed = s.phase_timing && max_iters > 0;
            const int replayed = max_iters - (timed ? 1 : 0);
            if (replayed >= GRAPH_ITERS)
                capture_graph(s, GRAPH_ITERS, &s.graph_n);
            if (replayed % GRAPH_ITERS)
                capture_graph(s, 1, &s.graph);
            int left = replayed;
            if (left >= GRAPH_ITERS) {
                SG_HIP(hipGraphLaunch(s.graph_n, s.stream));
                left -= GRAPH_ITERS;
            }
            if (timed)
                enqueue_iteration(s, &s.ctl->stop, false, s.pev, 1);
            for (; left >= GRAPH_ITERS; left -= GRAPH_ITERS)
                SG_HIP(hipGraphLaunch(s.graph_n, s.stream));
            for (; left > 0; --left)
                SG_HIP(hipGraphLaunch(s.graph, s.stream));
            if (max_iters > 0)
                mark_iterated(s);
            if (timed) {
                sync(s);
                for (int q = 0; q < SG_NPHASES; ++q) {
                    float ms = 0.f;
                    SG_HIP(hipEventElapsedTime(&ms, s.pev[q], s.pev[q + 1]));
                    s.phase_acc[q] += 1e-3 * ms;
                }
                ++s.phase_n;
            }
        }
        const Ctl c = read_ctl(s);
        const int done = std::min(c.log_pos, max_iters);
        if (log && done > 0) {
            SG_HIP(hipMemcpyAsync(log, s.dlog, sizeof(double) * 3 * done, hipMemcpyDeviceToHost,
                                  s.stream));
            sync(s);
        }
        if (iters_done)
            *iters_done = done;
        if (converged)
            *converged = c.converged ? 1 : (c.diverged ? -1 : 0);
        s.iters = c.iter;
    });
}

int sg_iterations(sg_solver* h, int* iters) {
    return guard([&] {
        Solver& s = S(h);
        *iters = read_ctl(s).iter;
    });
}

int sg_set_phase_timing(sg_solver* h, int on) {
    return guard([&] {
        Solver& s = S(h);
        s.phase_timing = on != 0;
        std::fill(std::begin(s.phase_acc), std::end(s.phase_acc), 0.0);
        s.phase_n = 0;
    });
}

int sg_phase_times(sg_solver* h, double* phase_s, int* timed) {
    return guard([&] {
        Solver& s = S(h);
        if (phase_s)
            for (int k = 0; k < SG_NPHASES; ++k)
                phase_s[k] = s.phase_acc[k];
        if (timed)
            *timed = s.phase_n;
    });
}

int sg_kernel_probe(sg_solver* h, int kid, int reps, double* mean_s, double* bytes,
                    int* per_iter) {
    return guard([&] {
        Solver& s = S(h);
        ensure_p_full(s);
        const sg_params& p = s.p;
        if (reps < 1)
            throw ArgError("reps must be >= 1");
        double cells = 0.0, bpc = 0.0;
        int per = 0;
        std::function<void(int)> one;
        switch (kid) {
        case SG_K_SWEEP_U:
        case SG_K_SWEEP_V: {
            const bool is_u = kid == SG_K_SWEEP_U;
            if (p.momentum_solver != SG_MOM_GS_MULTICOLOUR)
                throw ArgError("sweep probe needs momentum_solver = gs_multicolour");
            cells = is_u ? static_cast<double>(p.NX) * (p.NY - 1)
                         : static_cast<double>(p.NX - 1) * p.NY;
            // one launch is a whole sweep (every colour): the field read and written, and the
            // face pack's ue, vn and pressure term (face form) or the 9 coefficients and the
            // pressure term, 8 B each, per cell
            bpc = s.face[is_u ? 0 : 1] ? 40.0 : 96.0;
            per = is_u ? p.maxit_inner_u : p.maxit_inner_v;
            // field -> scratch: the field itself is left unchanged
            one = [&s, is_u](int) {
                launch_mom_sweep_once(s, is_u, s.F(is_u ? SG_F_U : SG_F_V), s.scratch, false,
                                      s.never);
            };
            break;
        }
        case SG_K_SWEEP_U_FIRST:
        case SG_K_SWEEP_V_FIRST: {
            const bool is_u = kid == SG_K_SWEEP_U_FIRST;
            if (p.momentum_solver != SG_MOM_GS_MULTICOLOUR)
                throw ArgError("sweep probe needs momentum_solver = gs_multicolour");
            cells = is_u ? static_cast<double>(p.NX) * (p.NY - 1)
                         : static_cast<double>(p.NX - 1) * p.NY;
            bpc = s.face[is_u ? 0 : 1] ? 48.0 : 104.0; // as a sweep, plus the |R| cell written
            per = 1;
            one = [&s, is_u](int) {
                launch_mom_sweep_once(s, is_u, s.F(is_u ? SG_F_U : SG_F_V), s.scratch, true,
                                      s.never);
            };
            break;
        }
        case SG_K_SWEEP_U_DC: {
            if (!mom_dc_ok(p))
                throw ArgError("the correction sweep probe needs momentum_precision = mixed");
            cells = static_cast<double>(p.NX) * (p.NY - 1);
            // FP32 arrays: the 10 of the pack read (8 coefficients, a / AP, r0), the correction
            // read and written, 4 B each, per cell
            bpc = 48.0;
            per = p.maxit_inner_u;
            one = [&s](int) { launch_dc_sweep_probe(s, true, s.never); };
            break;
        }
        case SG_K_ASSEMBLE_U_DC: {
            if (!mom_dc_ok(p))
                throw ArgError("the assembly probe needs momentum_precision = mixed");
            // u, v, p read (8 B each); AP written (8 B) over the assembly range; the pack (40 B)
            // and the |r| cell (8 B) written over the sweep range
            cells = static_cast<double>(p.NX) * (p.NY + 1);
            bpc = 32.0 + 48.0 * (p.NY - 1) / (p.NY + 1);
            per = 1;
            one = [&s](int) { launch_asm_dc(s, true, s.never, false); };
            break;
        }
        case SG_K_PSOLVE:
            cells = static_cast<double>(p.NX) * p.NY;
            bpc = 0.0;
            per = 1;
            one = [&s](int) { launch_solve_p(s, s.never, false); };
            break;
        case SG_K_COPY:
            cells = static_cast<double>(s.n);
            bpc = 16.0; // 8 B read + 8 B written per element
            per = 0;
            one = [&s](int) { launch_copy(s, s.F(SG_F_U), s.scratch, s.n, s.never); };
            break;
        case SG_K_ASSEMBLE_U:
        case SG_K_ASSEMBLE_V: {
            // In the form the outer iterations use: u, v and p read, and the face pack and AP
            // (face form) or the 9 coefficients and the pressure term written, 8 B each.
            const bool is_u = kid == SG_K_ASSEMBLE_U;
            const bool face = face_ok(s, is_u);
            cells = is_u ? static_cast<double>(p.NX) * (p.NY + 1)
                         : static_cast<double>(p.NX + 2) * (p.NY + 2);
            bpc = face ? 56.0 : 104.0;
            per = 1;
            one = [&s, is_u, face](int) { launch_assemble(s, is_u, s.never, false, face); };
            break;
        }
        default:
            throw ArgError("unknown kernel id");
        }
        one(0); // warm
        SG_HIP(hipEventRecord(s.kt0, s.stream));
        for (int k = 0; k < reps; ++k)
            one(k);
        SG_HIP(hipEventRecord(s.kt1, s.stream));
        sync(s);
        float ms = 0.f;
        SG_HIP(hipEventElapsedTime(&ms, s.kt0, s.kt1));
        if (mean_s)
            *mean_s = 1e-3 * ms / reps;
        if (bytes)
            *bytes = cells * bpc;
        if (per_iter)
            *per_iter = per;
    });
}

int sg_mg_levels(sg_solver* h, int* n) {
    return guard([&] { *n = static_cast<int>(S(h).mg.size()); });
}

int sg_mg_coarse_sweeps(sg_solver* h, int* n) {
    return guard([&] { *n = read_ctl(S(h)).coarse_sweeps; });
}

static long* knob(Solver& s, int key) {
    static thread_local long tmp;
    switch (key) {
    case SG_TUNE_MID_CELLS: return &s.mid_cells;
    case SG_TUNE_LDS_CELLS: return &s.lds_cells;
    case SG_TUNE_TAIL_THREADS: tmp = s.tail_threads; return &tmp;
    case SG_TUNE_MOM_TILE: tmp = s.mom_tile; return &tmp;
    case SG_TUNE_MOM_TILE_FIRST: tmp = s.mom_tile_first >= 0 ? s.mom_tile_first : s.mom_tile; return &tmp;
    case SG_TUNE_MG_STAGED: tmp = s.mg_staged; return &tmp;
    case SG_TUNE_DC_TILE: tmp = s.dc_tile; return &tmp;
    default: throw ArgError("unknown tuning key");
    }
}

int sg_set_tuning(sg_solver* h, int key, long value) {
    return guard([&] {
        Solver& s = S(h);
        if (key == SG_TUNE_TAIL_THREADS) {
            if (value < 64 || value > 1024 || value % 64)
                throw ArgError("tail threads must be a multiple of 64 in [64, 1024]");
            s.tail_threads = static_cast<int>(value);
        } else if (key == SG_TUNE_MOM_TILE || key == SG_TUNE_MOM_TILE_FIRST) {
            if (value < 0 || value >= N_MOM_TILES)
                throw ArgError("momentum tile shape must be 0.." + std::to_string(N_MOM_TILES - 1));
            (key == SG_TUNE_MOM_TILE ? s.mom_tile : s.mom_tile_first) = static_cast<int>(value);
        } else if (key == SG_TUNE_DC_TILE) {
            if (value < 0 || value >= N_DC_TILES)
                throw ArgError("correction sweep tile shape must be 0.." +
                               std::to_string(N_DC_TILES - 1));
            s.dc_tile = static_cast<int>(value);
        } else if (key == SG_TUNE_MG_STAGED) {
            if (value < 0 || value > 64)
                throw ArgError("staged level must be 0..64");
            s.mg_staged = static_cast<int>(value);
        } else {
            if (value < 0)
                throw ArgError("tuning value must be >= 0");
            *knob(s, key) = value;
        }
        drop_graphs(s); // launch shapes are baked into the captured graphs
    });
}

int sg_get_tuning(sg_solver* h, int key, long* value) {
    return guard([&] { *value = *knob(S(h), key); });
}

// Graph executables of the autotuner's candidates, destroyed together once the stream is idle
// (after the last timing, or when an error unwinds the autotuner).
struct CandidateGraphs {
    hipStream_t stream;
    std::vector<hipGraphExec_t> g;
    ~CandidateGraphs() {
        (void)hipStreamSynchronize(stream);
        for (hipGraphExec_t e : g)
            (void)hipGraphExecDestroy(e);
    }
};

// One p' solve with the current knobs captured into a new executable of `keep`. A failure while
// capturing ends the capture, so the stream never stays in capture mode.
static hipGraphExec_t capture_psolve(Solver& s, CandidateGraphs& keep) {
    hipGraph_t g = nullptr;
    SG_HIP(hipStreamBeginCapture(s.stream, hipStreamCaptureModeRelaxed));
    try {
        launch_solve_p(s, s.never, false);
    } catch (...) {
        hipGraph_t tmp = nullptr;
        (void)hipStreamEndCapture(s.stream, &tmp);
        if (tmp)
            (void)hipGraphDestroy(tmp);
        throw;
    }
    SG_HIP(hipStreamEndCapture(s.stream, &g));
    hipGraphExec_t ge = nullptr;
    const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    SG_HIP(e);
    keep.g.push_back(ge);
    return ge;
}

// Device time of one p' solve with the current knobs: the solve captured once and replayed.
static double time_psolve(Solver& s, int reps, CandidateGraphs& keep) {
    const hipGraphExec_t ge = capture_psolve(s, keep);
    SG_HIP(hipGraphLaunch(ge, s.stream)); // warm
    SG_HIP(hipEventRecord(s.kt0, s.stream));
    for (int k = 0; k < reps; ++k)
        SG_HIP(hipGraphLaunch(ge, s.stream));
    SG_HIP(hipEventRecord(s.kt1, s.stream));
    SG_HIP(hipStreamSynchronize(s.stream));
    float ms = 0.f;
    SG_HIP(hipEventElapsedTime(&ms, s.kt0, s.kt1));
    return 1e-3 * ms / reps;
}

// b and p' (every p' solve rewrites both) kept in device copies while the autotuner times solves,
// and the multigrid knobs as they were: the destructor puts all of them back, also when an error
// unwinds the autotuner, and frees the copies.
struct PsolveState {
    Solver& s;
    double* kb = nullptr;
    double* kp = nullptr;
    int threads, staged;
    long mid, lds;
    bool keep_knobs = false; // set once the autotuner has chosen its knobs
    explicit PsolveState(Solver& sv)
        : s(sv), threads(sv.tail_threads), staged(sv.mg_staged), mid(sv.mid_cells),
          lds(sv.lds_cells) {
        SG_HIP(hipMalloc(reinterpret_cast<void**>(&kb), sizeof(double) * s.n));
        SG_HIP(hipMalloc(reinterpret_cast<void**>(&kp), sizeof(double) * s.n));
        SG_HIP(hipMemcpyAsync(kb, s.F(SG_F_B), sizeof(double) * s.n, hipMemcpyDeviceToDevice,
                              s.stream));
        SG_HIP(hipMemcpyAsync(kp, s.F(SG_F_PC), sizeof(double) * s.n, hipMemcpyDeviceToDevice,
                              s.stream));
    }
    ~PsolveState() {
        (void)hipStreamSynchronize(s.stream);
        (void)hipMemcpyAsync(s.F(SG_F_B), kb, sizeof(double) * s.n, hipMemcpyDeviceToDevice,
                             s.stream);
        (void)hipMemcpyAsync(s.F(SG_F_PC), kp, sizeof(double) * s.n, hipMemcpyDeviceToDevice,
                             s.stream);
        (void)hipStreamSynchronize(s.stream);
        (void)hipFree(kb);
        (void)hipFree(kp);
        if (!keep_knobs) {
            s.tail_threads = threads, s.mg_staged = staged;
            s.mid_cells = mid, s.lds_cells = lds;
        }
    }
};

int sg_autotune(sg_solver* h, int reps, char* report, size_t len) {
    return guard([&] {
        Solver& s = S(h);
        ensure_p_full(s); // the candidates include solves that read every p' array
        if (reps < 1)
            throw ArgError("reps must be >= 1");
        std::string rep;
        if (s.p.momentum_solver == SG_MOM_GS_MULTICOLOUR && !mom_dc_ok(s.p)) {
            // Momentum sweep tile: one u sweep (field -> scratch) per candidate shape. (With
            // momentum_precision = mixed, outer iterations run the correction sweeps instead.)
            double best_t = 1e30;
            int best_k = s.mom_tile;
            for (int k = 0; k < N_MOM_TILES; ++k) {
                s.mom_tile = k;
                launch_mom_sweep_once(s, true, s.F(SG_F_U), s.scratch, false, s.never);
                SG_HIP(hipEventRecord(s.kt0, s.stream));
                for (int r = 0; r < reps; ++r)
                    launch_mom_sweep_once(s, true, s.F(SG_F_U), s.scratch, false, s.never);
                SG_HIP(hipEventRecord(s.kt1, s.stream));
                sync(s);
                float ms = 0.f;
                SG_HIP(hipEventElapsedTime(&ms, s.kt0, s.kt1));
                const double t = 1e-3 * ms / reps;
                char line[160];
                std::snprintf(line, sizeof line, "momentum tile %-6s : %9.2f us per sweep\n",
                              tile_name(k).c_str(), 1e6 * t);
                rep += line;
                if (t < best_t)
                    best_t = t, best_k = k;
            }
            s.mom_tile = best_k;
            char line[160];
            std::snprintf(line, sizeof line, "best momentum tile: %s (%d)\n",
                          tile_name(best_k).c_str(), best_k);
            rep += line;
            // The first sweep (fused residual, more LDS) on its own.
            best_t = 1e30;
            int best_f = best_k;
            for (int k = 0; k < N_MOM_TILES; ++k) {
                s.mom_tile_first = k;
                launch_mom_sweep_once(s, true, s.F(SG_F_U), s.scratch, true, s.never);
                SG_HIP(hipEventRecord(s.kt0, s.stream));
                for (int r = 0; r < reps; ++r)
                    launch_mom_sweep_once(s, true, s.F(SG_F_U), s.scratch, true, s.never);
                SG_HIP(hipEventRecord(s.kt1, s.stream));
                sync(s);
                float ms = 0.f;
                SG_HIP(hipEventElapsedTime(&ms, s.kt0, s.kt1));
                const double t = 1e-3 * ms / reps;
                std::snprintf(line, sizeof line,
                              "first-sweep tile %-6s : %9.2f us per sweep\n",
                              tile_name(k).c_str(), 1e6 * t);
                rep += line;
                if (t < best_t)
                    best_t = t, best_f = k;
            }
            s.mom_tile_first = best_f;
            std::snprintf(line, sizeof line, "best first-sweep tile: %s (%d)\n",
                          tile_name(best_f).c_str(), best_f);
            rep += line;
        }
        if (mom_dc_ok(s.p)) {
            // FP32 correction sweep tile: one middle u sweep per candidate shape.
            double best_t = 1e30;
            int best_k = s.dc_tile;
            for (int k = 0; k < N_DC_TILES; ++k) {
                s.dc_tile = k;
                launch_dc_sweep_probe(s, true, s.never);
                SG_HIP(hipEventRecord(s.kt0, s.stream));
                for (int r = 0; r < reps; ++r)
                    launch_dc_sweep_probe(s, true, s.never);
                SG_HIP(hipEventRecord(s.kt1, s.stream));
                sync(s);
                float ms = 0.f;
                SG_HIP(hipEventElapsedTime(&ms, s.kt0, s.kt1));
                const double t = 1e-3 * ms / reps;
                char line[160];
                std::snprintf(line, sizeof line, "correction sweep tile %-10s : %9.2f us\n",
                              dc_tile_name(k).c_str(), 1e6 * t);
                rep += line;
                if (t < best_t)
                    best_t = t, best_k = k;
            }
            s.dc_tile = best_k;
            char line[160];
            std::snprintf(line, sizeof line, "best correction sweep tile: %s (%d)\n",
                          dc_tile_name(best_k).c_str(), best_k);
            rep += line;
        }
        if (s.p.pressure_solver != SG_P_MG) {
            drop_graphs(s); // the captured iterations hold the previous tile shapes
            if (report && len) {
                std::strncpy(report, rep.c_str(), len - 1);
                report[len - 1] = 0;
            }
            return;
        }
        // b and p' are restored (and the candidate graphs destroyed) when these go out of
        // scope, on success or error; a p' solve does not change the solver's other arrays.
        PsolveState keep_state(s);
        CandidateGraphs graphs{s.stream, {}};
        double best = 1e30;
        int bt = s.tail_threads;
        long bm = s.mid_cells, bl = s.lds_cells;
        for (int threads : {64, 256, 1024})
            for (long mid : {0L, 4096L, 16384L, 65536L})
                for (long lds : {256L, 1024L, 4096L}) {
                    s.tail_threads = threads;
                    s.mid_cells = mid;
                    s.lds_cells = lds;
                    const double t = time_psolve(s, reps, graphs);
                    char line[160];
                    std::snprintf(line, sizeof line,
                                  "tail_threads %4d mid_cells %6ld lds_cells %5ld : %9.2f us\n",
                                  threads, mid, lds, 1e6 * t);
                    rep += line;
                    if (t < best) {
                        best = t;
                        bt = threads, bm = mid, bl = lds;
                    }
                }
        s.tail_threads = bt;
        s.mid_cells = bm;
        s.lds_cells = bl;
        char line[160];
        std::snprintf(line, sizeof line,
                      "best: tail_threads %d mid_cells %ld lds_cells %ld : %.2f us\n", bt, bm, bl,
                      1e6 * best);
        rep += line;
        // Staged tiles from level 0, 1, 2 or none.
        int bs = s.mg_staged;
        for (int st : {0, 1, 2, 64}) {
            s.mg_staged = st;
            const double t = time_psolve(s, reps, graphs);
            std::snprintf(line, sizeof line, "staged from level %2d : %9.2f us\n", st, 1e6 * t);
            rep += line;
            if (t < best) {
                best = t;
                bs = st;
            }
        }
        s.mg_staged = bs;
        keep_state.keep_knobs = true;
        std::snprintf(line, sizeof line, "best: staged from level %d : %.2f us\n", bs, 1e6 * best);
        rep += line;
        drop_graphs(s);
        if (report && len) {
            std::strncpy(report, rep.c_str(), len - 1);
            report[len - 1] = 0;
        }
    });
}

} // extern "C"
