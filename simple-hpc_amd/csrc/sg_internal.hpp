// sg_internal.hpp — device state and kernel launchers of the gfx950 SIMPLE backend.
//
// Device layout: every level-0 field is an nrows x ld FP64 array (ld = ncols rounded up to 16
// doubles, so each row starts on a 128-byte line), row-major, (i, j) -> i*ld + j, the same (i, j)
// as the host arrays. Fields stay resident in HBM for the whole run; the host sees them only
// through sg_upload / sg_download.
//
// Every kernel that runs inside the outer iteration takes `const int* stop` and returns at once
// when *stop != 0: the device-side stopping rule sets it, so graph replays queued past the
// converged iteration do no work.
#pragma once
#include "../../include/simple_gpu.h"
#include "sg_launch.hpp"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace sg {

void set_error(const std::string& msg);

// SG_HIP, HipError and ArgError are defined in sg_launch.hpp.

// Nine coefficients in the order AE, AW, AN, AS, AEE, AWW, ANN, ASS, AP.
struct C9 {
    double v[9];
};

// A multigrid level on the device. Level 0 aliases the p' arrays (stride ld, offset of interior
// cell (H, H+1)); coarse levels own (nx+2) x (ny+2) arrays with a zero halo.
struct MgLevel {
    int nx = 0, ny = 0, stride = 0;
    double *AE = nullptr, *AW = nullptr, *AN = nullptr, *AS = nullptr, *AP = nullptr;
    double *b = nullptr, *x = nullptr, *r = nullptr;
    double* xt = nullptr; // pre-smoothed iterate (tiled V-cycle), same layout as x, zero halo
    float *fAE = nullptr, *fAW = nullptr, *fAN = nullptr, *fAS = nullptr, *fAP = nullptr;
    float *fb = nullptr, *fx = nullptr, *fr = nullptr; // FP32 copies (mixed precision)
    float* fxt = nullptr;
};

// Colour-packed layout of one momentum component's coefficients (sg_momentum.hip): 10 arrays
// (AE, AW, AN, AS, AEE, AWW, ANN, ASS, AP, pressure term), each ncol x rows x kp doubles.
struct Pack {
    int ncol;           // colours: 5, or 2 for UDS with 2nd-order diffusion
    int ai0, ai1, aj0, aj1; // assembly range (inclusive)
    int rows, kp;       // rows of the range, cells per colour per row (padded to 16)
    size_t astride;     // doubles per array
    double* base;
};

// Momentum sweep tile shapes (rows x columns of cells), indexed by the momentum tile tuning knobs.
// The autotuner times them all: which one wins depends on how many workgroups a shape makes (one
// resident wave over the 256 CUs without a partial second one) against its halo overhead.
struct TileShape {
    int ti, tj;
};
constexpr TileShape MOM_TILES[] = {{32, 64}, {64, 64}, {32, 128}, {16, 128}, {24, 96}, {48, 96},
                                   {24, 64}, {40, 128}, {32, 72}, {28, 80}, {36, 64}, {40, 56}};
constexpr int N_MOM_TILES = sizeof(MOM_TILES) / sizeof(MOM_TILES[0]);
inline std::string tile_name(int k) {
    return std::to_string(MOM_TILES[k].ti) + "x" + std::to_string(MOM_TILES[k].tj);
}

// Tile shapes of the FP32 correction sweeps (momentum_precision = mixed): rows x columns of cells
// and threads, indexed by SG_TUNE_DC_TILE.
struct DcTile {
    int ti, tj, nt;
};
constexpr DcTile DC_TILES[] = {{32, 64, 256}, {32, 64, 512}, {64, 64, 512}, {32, 128, 512},
                               {48, 64, 512}, {16, 128, 256}, {24, 96, 512}, {40, 56, 512}};
constexpr int N_DC_TILES = sizeof(DC_TILES) / sizeof(DC_TILES[0]);
inline std::string dc_tile_name(int k) {
    return std::to_string(DC_TILES[k].ti) + "x" + std::to_string(DC_TILES[k].tj) + "/" +
           std::to_string(DC_TILES[k].nt);
}

// Device control block of the outer loop.
struct Ctl {
    int stop;       // nonzero: skip all work (set by the stop rule)
    int iter;       // outer iterations completed
    int converged;  // stop rule fired
    int diverged;   // a residual was not finite
    int log_pos;    // entries written into the log buffer in the current sg_advance call
    int coarse_sweeps;
    int warm_ok, warm_rejected; // learned warm start: guesses kept / discarded
    double mu, mv, mp; // running maxima (residual_normalisation = running_max)
};

// Device scalar slots.
enum {
    S_RES_U = 0, S_RES_V, S_RES_P, S_SUM_A, S_SUM_B, S_SUM_C, S_SUM_D, S_BSUM,
    S_CG = 8, // CG's recurrence scalars (8 slots)
    S_COUNT = 32
};

struct Warm;

// An ordered sum of a compact nrow x w cell buffer into *slot, enqueued later (flush_sums).
struct DeferredSum {
    const double* cell; // null: empty range (the sum is 0.0)
    int nrow, w, lane;
    double* slot;
};

struct Solver {
    sg_params p{};
    int nr = 0, nc = 0, ld = 0;
    size_t n = 0; // nr * ld
    hipStream_t stream = nullptr;
    int device = 0;
    double* f[SG_F_COUNT] = {};
    double* scratch = nullptr;         // Jacobi previous iterate, multicolour sweep ping-pong
    double* scratch2 = nullptr;        // multicolour: third buffer of odd sweep counts
    double* cell = nullptr;            // per-cell values for ordered sums (4 planes of n)
    double* rowbuf = nullptr;          // row sums, one nr-slice per ordered-sum lane (4 lanes)
    double* rcell[3] = {};             // |R| cells of res_u, res_v, res_p (deferred sums)
    double* scal = nullptr;            // S_COUNT device scalars
    double* cg[4] = {};                // CG r, z, d, q
    int* cgctl = nullptr;              // CG: done flag (the CG launches' stop flag), iterations
    Ctl* ctl = nullptr;                // device control block
    int* never = nullptr;              // device int that stays 0 (eager per-phase calls)
    double* dlog = nullptr;            // device log, 3 doubles per iteration (log_cap entries)
    int log_cap = 0;
    std::vector<MgLevel> mg;
    double* mg_r0 = nullptr;           // level-0 residual
    // Multigrid tail (tunable; sg_tune): levels with at most mid_cells cells run in one workgroup
    // of tail_threads threads, those with at most lds_cells cells (and fitting lds_budget) in LDS.
    // Defaults for every launch-shape knob: default_launch_shapes.
    long mid_cells = 0, lds_cells = 4096;
    int tail_threads = 1024;
    bool mg_tiled = true;              // multigrid levels above the tail: down/up tile kernels
    int mg_staged = 0;                 // first level whose tiles stage the operator in LDS too
    size_t lds_budget = 65536 - 1024;  // dynamic LDS of the multigrid tail (set from the device)
    std::vector<void*> owned;          // allocations to free
    Pack pku{}, pkv{};                 // packed momentum coefficients
    Pack fku{}, fkv{};                 // face packs: ue, vn, pressure term (face form)
    // momentum_precision = mixed: FP32 packs of u and v (8 neighbour coefficients, a / AP, r0;
    // geometry of pku / pkv) and, per component, the two FP32 correction arrays its sweeps
    // alternate between (u: 0, 1; v: 2, 3). Each component's pair is zero outside its own sweep
    // range, which the sweeps read as the fixed boundary values' zero correction.
    float* dcf[2] = {};
    float* dd[4] = {};
    bool dc_stale[2] = {};             // u, v: the last outer iteration kept no FP64 coefficients
    bool face[2] = {};                 // u, v: the face pack holds the current coefficients (the
                                       // coefficient pack is stale until ensure_packed)
    bool pt_stale[2] = {};             // u, v: p changed since the pressure terms were evaluated
    bool ap_in_sweep[2] = {};          // u, v: row-major AP of the sweep range not yet written:
                                       // the next first (residual) face-form sweep writes it
    bool p_lean = false;               // AWP, ASP, APP not written by the last assemble_p (stale
                                       // until ensure_p_full)
    // p' mean removal carried by correct (outer iterations with the multigrid p' solve): fuse_pc_mean
    // arms it for the next solve, which then only sums p' into S_SUM_D (pc_mean_pending); correct
    // reads p' - mean and writes it to scratch (pc_in_scratch), from where materialize_pc copies it
    // back before any other use of the solver.
    bool fuse_pc_mean = false, pc_mean_pending = false, pc_in_scratch = false;
    // b's mean removal carried by the V-cycle (the same outer iterations): fuse_b_mean arms it for
    // the next multigrid solve, which then only sums b into S_BSUM and zeroes p' outside the
    // interior; b_mean_pending: level 0's kernels read b - mean and start from p' = 0 unread;
    // b_raw: b still holds its mean, until materialize_b subtracts it.
    bool fuse_b_mean = false, b_mean_pending = false, b_raw = false;
    int mom_tile = 0;                  // momentum sweep tile shape, an index of MOM_TILES (tunable)
    int mom_tile_first = -1;           // the first sweep's tile shape (-1: mom_tile)
    int dc_tile = 6;                   // FP32 correction sweep tile, an index of DC_TILES (24x96:
                                       // the fastest at 1024^2)
    C9 diff2{}, diff4{};               // diffusive coefficients: order 2 (and band), order 4
    hipGraphExec_t graph = nullptr;    // one captured outer iteration
    hipGraphExec_t graph_n = nullptr;  // GRAPH_ITERS of them (fewer launch gaps)
    // Phase timing of replayed iterations (sg_set_phase_timing): in every sg_advance call one
    // outer iteration is enqueued launch by launch with events at its phase boundaries, between
    // the graph replays; its per-phase device seconds accumulate in phase_acc.
    bool phase_timing = false;
    hipEvent_t pev[SG_NPHASES + 1] = {};
    double phase_acc[SG_NPHASES] = {};
    int phase_n = 0;
    int iters = 0;                     // host copy of completed iterations
    long long cg_total = 0;
    int cg_last = 0;
    hipEvent_t ev[SG_NPHASES + 1] = {};
    // kernel timing
    hipEvent_t kt0 = nullptr, kt1 = nullptr;
    // Recording (batched execution): launches are appended here instead of enqueued; tag is the
    // structural label of the launches being enqueued (multigrid level and cycle, else 0).
    std::vector<Rec>* rec = nullptr;
    int tag = 0;
    int vcount = 0; // V-cycles enqueued in the current p' solve
    Warm* warm = nullptr; // learned warm start of the p' solve (sg_unet.hip)
    std::vector<DeferredSum> deferred; // residual sums waiting for flush_sums
    // ride: the next ordered mean removal also takes the deferred residual sums (their rows in its
    // row-sum launch, their totals in its total launch); rode: it did, so the iteration's end has
    // only the stopping rule left.
    bool ride = false, rode = false;

    const double* F(int id) const { return f[id]; }
    double* F(int id) { return f[id]; }
};

// Launch body B over `grid` for this cavity on its stream, or record it.
template <class A, void (*B)(const A&, const Blk&), int NT>
void emit(Solver& s, dim3 grid, dim3 block, size_t lds, const A& a, const int* stop) {
    if (grid.x == 0 || grid.y == 0 || grid.z == 0)
        return;
    if (s.rec) {
        Rec r;
        r.vt = &Op<A, B, NT>::vt;
        r.args.resize(sizeof(A));
        std::memcpy(r.args.data(), &a, sizeof(A));
        r.grid = grid;
        r.block = block;
        r.lds = lds;
        r.stop = stop;
        r.tag = s.tag;
        for (const Rec& q : *s.rec)
            r.occ += (q.vt == r.vt && q.tag == r.tag) ? 1 : 0;
        s.rec->push_back(std::move(r));
        return;
    }
    hipLaunchKernelGGL((k_one<A, B, NT>), grid, block, lds, s.stream, a, stop);
    SG_HIP(hipGetLastError());
}

// ---- launchers (each enqueues on s.stream; `stop` may be s.never) ----
// momentum (sg_momentum.hip)
// with_bc (u only): the same launch also sets u's ghost rows (launch_bc).
// face: the face form (face pack and AP only; the sweeps then recompute the coefficients).
// ap_in_sweep (face form, the outer iteration): AP of the sweep range is left to the first sweep,
// which evaluates every coefficient of those cells for the residual anyway (ap_in_sweep flag).
void launch_assemble(Solver& s, bool is_u, const int* stop, bool with_bc = false,
                     bool face = false, bool ap_in_sweep = false);
// Whether outer iterations use the face form for this component (multicolour sweeps, at least one).
bool face_ok(const Solver& s, bool is_u);
// Bring the coefficient pack up to date if the face pack holds the current coefficients.
void ensure_packed(Solver& s, bool is_u);
// Re-evaluate the component's pressure terms from the current p if it changed since (correct,
// an upload of p), in the form that holds the coefficients.
void refresh_pterm(Solver& s, bool is_u);
Pack make_face_pack(const sg_params& p, bool is_u);
void launch_bc(Solver& s, bool is_u, const int* stop);
// momentum_precision = mixed (FP32 sweeps inside FP64 defect correction) applies to these params
bool mom_dc_ok(const sg_params& p);
// Throws if the component's FP64 coefficients were not kept (after mixed-momentum iterations).
void need_coefs(const Solver& s, bool is_u);
// The assembly with the FP64 residual r0 and the FP32 pack (and u's ghost rows with with_bc).
void launch_asm_dc(Solver& s, bool is_u, const int* stop, bool with_bc);
// nsweeps FP32 multicolour sweeps of the correction; the last writes x0 + d into the field.
void launch_dc_sweeps(Solver& s, bool is_u, int nsweeps, const int* stop);
// One middle sweep of the correction (neither first nor last), correction array 0 -> 1: probes.
void launch_dc_sweep_probe(Solver& s, bool is_u, const int* stop);
void launch_mom_residual(Solver& s, bool is_u, double* slot, const int* stop, bool defer);
// bcv (v, multicolour, at least one sweep): v's ghost columns are set by the first sweep, which
// loads them as the boundary condition gives them, instead of a launch_bc before the sweeps.
bool bcv_in_sweep(const Solver& s);
void launch_mom_sweeps(Solver& s, bool is_u, int nsweeps, const int* stop, double* res_slot,
                       bool defer, bool bcv = false);
void launch_mom_sweep_once(Solver& s, bool is_u, const double* src, double* dst, bool first,
                           const int* stop, bool bcv = false);
int ncolours(const sg_params& p);
void launch_copy(Solver& s, const double* a, double* b, size_t n, const int* stop);
Pack make_pack(const sg_params& p, bool is_u);
void sync_packed(Solver& s, int field_id, bool to_packed);
// pressure (sg_pressure.hip)
// lean: write AE, AN, b only (and |b|): AW, AS and AP follow from them (the operator's symmetry)
// and are derived on demand (ensure_p_full).
void launch_assemble_p(Solver& s, double* slot, const int* stop, bool defer, bool lean = false);
// Derive AWP, ASP and APP from AEP and ANP if the last assemble_p was lean.
void ensure_p_full(Solver& s);
// Whether the multigrid p' solve reads only AE and AN of level 0 (every level from 0 to the tail
// runs staged tiles, and the tail starts at level 2 or below; no learned warm start).
bool mg_lean_ok(Solver& s);
void launch_solve_p(Solver& s, const int* stop, bool eager);
void launch_correct(Solver& s, const int* stop);
// The multigrid solve's final p' mean removal, or only its sum when fuse_pc_mean is armed.
void finish_pc_mean(Solver& s, const int* stop);
// Copy the mean-free p' back from scratch if correct left it there.
void materialize_pc(Solver& s);
// Subtract b's mean from b if the multigrid solve left it in (b_raw).
void materialize_b(Solver& s);
// b's sums for a fused mean removal (rows, then totals into S_BSUM with the deferred residual
// totals, and p' zeroed outside the interior).
void launch_b_mean_fused(Solver& s, const int* stop);
// Whether b's mean removal can be carried by the V-cycle: multigrid staged from level 0, level 0
// above the tail, no learned warm start.
bool mg_fuse_b_ok(Solver& s);
void launch_fill(Solver& s, double* a, size_t n, double v, const int* stop);
// ordered sums (sg_reduce.hip): rows [i0, i1], columns [j0, j1] of a pitched array, row sums j
// ascending from 0.0 into rowbuf, then rows ascending from 0.0 into *slot.
// lane: which row-sum slice (0 for immediate sums; 1..3 for the deferred residual sums).
void launch_ordered_sum(Solver& s, const double* a, int pitch, int i0, int i1, int j0, int j1,
                        double* slot, const int* stop, int lane = 0);
// The ordered sum of x over rows [i0, i1], columns [j0, j1] into *slot, then x -= sum / n over
// that range.
void launch_ordered_mean_sub(Solver& s, double* x, int pitch, int i0, int i1, int j0, int j1,
                             double n, double* slot, const int* stop, double* zero = nullptr,
                             size_t nzero = 0);
// Record an ordered sum of a compact nrow x w cell buffer (row-sum lane 1..3) for flush_sums.
void defer_sum(Solver& s, int lane, const double* cell, int nrow, int w, double* slot);
// Stopping rule of run_outer_loop (src/solver.cpp:112-128), one thread: log the residuals,
// update the running maxima, test max(...) <= tol && iter > 1 with std::max's comparison order.
// A residual that is not finite stops the loop (diverged), which the reference does not detect.
struct StopArgs {
    Ctl* ctl;
    const double* scal;
    double* log;
    int log_cap, norm;
    double tol;
};

__device__ __forceinline__ void stop_rule(const StopArgs& a) {
    Ctl* ctl = a.ctl;
    const double* scal = a.scal;
    if (ctl->stop)
        return;
    const int iter = ++ctl->iter;
    const double ru = scal[S_RES_U], rv = scal[S_RES_V], rp = scal[S_RES_P];
    if (ctl->log_pos < a.log_cap) {
        double* e = a.log + 3 * ctl->log_pos;
        e[0] = ru;
        e[1] = rv;
        e[2] = rp;
    }
    ++ctl->log_pos;
    if (!isfinite(ru) || !isfinite(rv) || !isfinite(rp)) {
        ctl->diverged = 1;
        ctl->stop = 1;
        return;
    }
    double tu = ru, tv = rv, tp = rp;
    if (a.norm == SG_NORM_RUNNING_MAX) {
        ctl->mu = ctl->mu < ru ? ru : ctl->mu;
        ctl->mv = ctl->mv < rv ? rv : ctl->mv;
        ctl->mp = ctl->mp < rp ? rp : ctl->mp;
        tu = ctl->mu > 0.0 ? ru / ctl->mu : 0.0;
        tv = ctl->mv > 0.0 ? rv / ctl->mv : 0.0;
        tp = ctl->mp > 0.0 ? rp / ctl->mp : 0.0;
    }
    double m = tu;
    if (m < tv)
        m = tv;
    if (m < tp)
        m = tp;
    if (m <= a.tol && iter > 1) {
        ctl->converged = 1;
        ctl->stop = 1;
    }
}

// Outer iterations per replay of Solver::graph_n, and dropping both captured graphs (their launch
// shapes and pointers are baked in).
constexpr int GRAPH_ITERS = 8;
void drop_graphs(Solver& s);
// The arrays' state flags after outer iterations ran as graph replays.
void mark_iterated(Solver& s);

// Enqueue every deferred sum: one launch for all their rows, one for their totals (with rule:
// the totals and then the stopping rule in that second launch).
void flush_sums(Solver& s, const int* stop, const StopArgs* rule = nullptr);
// multigrid (sg_mg.hip)
void mg_setup(Solver& s);
void launch_mg_solve(Solver& s, const int* stop, bool has_guess);
void launch_rb_staged(Solver& s, const int* stop);
// The red-black solve of a grid this small runs in one workgroup (k_pc_rb_one).
bool rb_one_fits(const sg_params& p);
// With zero, the same launches also fill zero[0..s.n) with 0.0.
void launch_remove_mean0(Solver& s, double* a, const int* stop, double* zero = nullptr);
// learned warm start (sg_unet.hip): the network's guess of p' and its decision (kept only if it
// lowers the energy); the caller applies it to the level-0 x
struct GuessArgs;
GuessArgs launch_warm_guess(Solver& s, const int* stop, bool bfused);
void default_launch_shapes(Solver& s);
void warm_free(Solver& s);
// outer loop control (sg_api.hip)

// One outer iteration (src/solver.cpp:27-93), enqueued or recorded. ev: optional phase events.
// ev: optional events (SG_NPHASES + 1) recorded at the phase boundaries. fused (default: no ev):
// the replayed iteration's launch structure (residual sums deferred, boundary conditions in the
// assemblies and sweeps), else every phase on its own as the eager per-phase iteration runs it.
void enqueue_iteration(Solver& s, const int* stop, bool eager, hipEvent_t* ev, int fused = -1);
void ensure_log(Solver& s, int cap);
Solver* create_solver(const sg_params& p, const double* u, const double* v, const double* pf);
void destroy_solver(Solver* s);

inline unsigned cdiv(long a, long b) { return static_cast<unsigned>((a + b - 1) / b); }

// Device helpers.
__device__ __forceinline__ bool stopped(const int* stop) { return *stop != 0; }

} // namespace sg
