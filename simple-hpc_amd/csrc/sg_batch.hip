// sg_batch.hip — many cavities of uneven size and Reynolds number advanced together: one
// persistent launch per phase for all cavities, one captured hipGraph per outer iteration of the
// whole batch, and an autotuner that times the batch's own iteration.
//
// Each cavity is an ordinary Solver (its own fields, coefficients, multigrid hierarchy, control
// block and stopping rule). To build the batch's iteration, every cavity's outer iteration is
// enqueued in recording mode (emit() appends instead of launching). A record carries its kernel
// instantiation, its argument struct and its grid, plus a label: (kernel, structural tag,
// occurrence). The tag names the multigrid level and V-cycle, so cavities with different numbers of
// levels still put "level 1 of V-cycle 0, going down" in the same launch. The labels of all
// cavities form a DAG (each cavity's records in order); a topological order of it gives the
// batched launches, each a k_many over the cavities that carry that label. A cavity whose stopping
// rule has fired is skipped by every launch, so the batch runs until its slowest cavity converges.
#include "sg_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <queue>
#include <set>
#include <tuple>

namespace sg {

namespace {

__global__ void k_batch_begin(Ctl* ctl) {
    if (threadIdx.x == 0) {
        ctl->stop = 0;
        ctl->log_pos = 0;
    }
}

using Label = std::tuple<const void*, int, int>;

Label label_of(const Rec& r) { return Label{r.vt, r.tag, r.occ}; }

struct Batch {
    std::vector<Solver*> cav;
    hipStream_t stream = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    std::vector<std::vector<Rec>> recs;
    std::vector<Group> plan;
    std::vector<void*> tables; // device tables of the plan
    hipGraphExec_t graph = nullptr;
    int cus = 1;
    int launches = 0;          // kernel launches per batched iteration
    // per-cavity host state
    std::vector<int> done;     // 1 converged, -1 diverged, 0 running
    std::vector<int> iters;
    std::vector<double> last;  // last logged residuals, 3 per cavity
    // snapshot for the autotuner (u, v, p and the control block of every cavity)
    std::vector<double*> snap;
};

void free_plan(Batch& b) {
    if (b.graph) {
        (void)hipGraphExecDestroy(b.graph);
        b.graph = nullptr;
    }
    for (void* q : b.tables)
        (void)hipFree(q);
    b.tables.clear();
    b.plan.clear();
    b.recs.clear();
}

// Record one outer iteration of every cavity and merge the records into batched launches.
void build_plan(Batch& b) {
    free_plan(b);
    const int n = static_cast<int>(b.cav.size());
    b.recs.assign(n, {});
    for (int c = 0; c < n; ++c) {
        Solver& s = *b.cav[c];
        // Recording runs nothing: the coefficient form and pressure-term flags stay.
        const bool f0 = s.face[0], f1 = s.face[1], p0 = s.pt_stale[0], p1 = s.pt_stale[1];
        const bool d0 = s.dc_stale[0], d1 = s.dc_stale[1];
        const bool lean = s.p_lean, pcs = s.pc_in_scratch, braw = s.b_raw;
        auto restore = [&] {
            s.face[0] = f0, s.face[1] = f1, s.pt_stale[0] = p0, s.pt_stale[1] = p1;
            s.dc_stale[0] = d0, s.dc_stale[1] = d1;
            s.p_lean = lean, s.pc_in_scratch = pcs, s.b_raw = braw;
            s.fuse_pc_mean = s.pc_mean_pending = s.fuse_b_mean = s.b_mean_pending = false;
            s.ap_in_sweep[0] = s.ap_in_sweep[1] = false;
        };
        s.rec = &b.recs[c];
        try {
            enqueue_iteration(s, &s.ctl->stop, false, nullptr);
        } catch (...) {
            s.rec = nullptr;
            restore();
            throw;
        }
        s.rec = nullptr;
        s.tag = 0;
        restore();
    }
    // Label DAG.
    std::map<Label, int> id;
    std::vector<Label> labels;
    for (int c = 0; c < n; ++c)
        for (const Rec& r : b.recs[c])
            if (id.emplace(label_of(r), static_cast<int>(labels.size())).second)
                labels.push_back(label_of(r));
    const int L = static_cast<int>(labels.size());
    std::vector<std::set<int>> succ(L);
    std::vector<int> indeg(L, 0), first_pos(L, 1 << 30);
    std::vector<std::vector<std::pair<int, const Rec*>>> members(L);
    for (int c = 0; c < n; ++c) {
        const auto& rs = b.recs[c];
        for (size_t k = 0; k < rs.size(); ++k) {
            const int v = id[label_of(rs[k])];
            members[v].emplace_back(c, &rs[k]);
            first_pos[v] = std::min(first_pos[v], static_cast<int>(k));
            if (k > 0) {
                const int u = id[label_of(rs[k - 1])];
                if (succ[u].insert(v).second)
                    ++indeg[v];
            }
        }
    }
    using Item = std::pair<int, int>; // (first position, label)
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> ready;
    for (int v = 0; v < L; ++v)
        if (indeg[v] == 0)
            ready.emplace(first_pos[v], v);
    std::vector<int> order;
    while (!ready.empty()) {
        const int v = ready.top().second;
        ready.pop();
        order.push_back(v);
        for (int w : succ[v])
            if (--indeg[w] == 0)
                ready.emplace(first_pos[w], w);
    }
    if (static_cast<int>(order.size()) != L)
        throw ArgError("batch: the cavities' launch sequences cannot be merged");
    b.launches = 0;
    for (int v : order) {
        Group g;
        g.vt = static_cast<const OpVT*>(std::get<0>(labels[v]));
        for (const auto& m : members[v])
            g.m.push_back(m.second);
        {
            for (const Rec* r : g.m)
                if (r->block.x != g.m[0]->block.x || r->block.y != g.m[0]->block.y ||
                    r->block.z != g.m[0]->block.z)
                    throw ArgError("batch: cavities disagree on a launch's workgroup shape");
            g.vt->prepare(g, b.tables, b.cus);
            ++b.launches;
        }
        b.plan.push_back(std::move(g));
    }
    // Capture the batched iteration.
    hipGraph_t graph = nullptr;
    SG_HIP(hipStreamBeginCapture(b.stream, hipStreamCaptureModeRelaxed));
    hipError_t err = hipSuccess;
    for (const Group& g : b.plan) {
        g.vt->launch(g, b.stream);
        err = err ? err : hipGetLastError();
    }
    const hipError_t e2 = hipStreamEndCapture(b.stream, &graph);
    SG_HIP(err);
    SG_HIP(e2);
    const hipError_t e3 = hipGraphInstantiate(&b.graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    SG_HIP(e3);
}

// Every batch entry point reaches its batch here: each cavity's p' goes back in place first
// (materialize_pc), before snapshots, tuning or replays touch the cavities' arrays.
Batch& B(sg_batch* h) {
    if (!h)
        throw ArgError("null batch");
    Batch& b = *reinterpret_cast<Batch*>(h);
    for (Solver* c : b.cav) {
        materialize_pc(*c);
        materialize_b(*c);
    }
    return b;
}

// Replay k batched iterations (every cavity not yet stopped advances by one per replay).
void replay(Batch& b, int k) {
    if (!b.graph)
        build_plan(b);
    for (int i = 0; i < k; ++i)
        SG_HIP(hipGraphLaunch(b.graph, b.stream));
}

void read_state(Batch& b) {
    const int n = static_cast<int>(b.cav.size());
    std::vector<Ctl> ctl(n);
    for (int c = 0; c < n; ++c)
        SG_HIP(hipMemcpyAsync(&ctl[c], b.cav[c]->ctl, sizeof(Ctl), hipMemcpyDeviceToHost,
                              b.stream));
    SG_HIP(hipStreamSynchronize(b.stream));
    for (int c = 0; c < n; ++c) {
        Solver& s = *b.cav[c];
        b.iters[c] = ctl[c].iter;
        s.iters = ctl[c].iter;
        if (ctl[c].diverged)
            b.done[c] = -1;
        else if (ctl[c].converged)
            b.done[c] = 1;
        const int pos = std::min(ctl[c].log_pos, s.log_cap);
        if (pos > 0)
            SG_HIP(hipMemcpy(&b.last[3 * c], s.dlog + 3 * (pos - 1), 3 * sizeof(double),
                             hipMemcpyDeviceToHost));
    }
}

// Re-arm the cavities still running (stop = 0, log position 0).
void begin_chunk(Batch& b) {
    for (size_t c = 0; c < b.cav.size(); ++c) {
        if (b.done[c] != 0)
            continue;
        hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(64), 0, b.stream, b.cav[c]->ctl);
        SG_HIP(hipGetLastError());
    }
}

// Snapshot / restore of the iteration state (u, v, p, control block) of every cavity.
void snapshot(Batch& b, bool save) {
    const int n = static_cast<int>(b.cav.size());
    if (save && b.snap.empty())
        for (int c = 0; c < n; ++c) {
            const Solver& s = *b.cav[c];
            void* q = nullptr;
            SG_HIP(hipMalloc(&q, 3 * s.n * sizeof(double) + sizeof(Ctl)));
            b.snap.push_back(static_cast<double*>(q));
        }
    for (int c = 0; c < n; ++c) {
        Solver& s = *b.cav[c];
        double* q = b.snap[c];
        for (int k = 0; k < 3; ++k) {
            double* f = s.f[k == 0 ? SG_F_U : k == 1 ? SG_F_V : SG_F_P];
            SG_HIP(hipMemcpyAsync(save ? q + k * s.n : f, save ? f : q + k * s.n,
                                  s.n * sizeof(double), hipMemcpyDeviceToDevice, b.stream));
        }
        void* qc = q + 3 * s.n;
        SG_HIP(hipMemcpyAsync(save ? qc : static_cast<void*>(s.ctl),
                              save ? static_cast<const void*>(s.ctl) : qc, sizeof(Ctl),
                              hipMemcpyDeviceToDevice, b.stream));
    }
    SG_HIP(hipStreamSynchronize(b.stream));
}

// Mean time of one batched iteration over reps replays (state restored afterwards).
double time_iteration(Batch& b, int reps) {
    build_plan(b);
    snapshot(b, false);
    begin_chunk(b);
    replay(b, 1); // warm-up
    SG_HIP(hipEventRecord(b.t0, b.stream));
    replay(b, reps);
    SG_HIP(hipEventRecord(b.t1, b.stream));
    SG_HIP(hipEventSynchronize(b.t1));
    float ms = 0.0f;
    SG_HIP(hipEventElapsedTime(&ms, b.t0, b.t1));
    snapshot(b, false);
    return 1e-3 * ms / reps;
}

long* knob_of(Solver& s, int key) {
    switch (key) {
    case SG_TUNE_TAIL_THREADS:
        return nullptr;
    case SG_TUNE_MID_CELLS:
        return &s.mid_cells;
    case SG_TUNE_LDS_CELLS:
        return &s.lds_cells;
    default:
        return nullptr;
    }
}

void set_knob(Batch& b, int key, long v) {
    for (Solver* s : b.cav) {
        switch (key) {
        case SG_TUNE_TAIL_THREADS:
            if (v < 64 || v > 1024 || v % 64)
                throw ArgError("tail_threads must be a multiple of 64 in [64, 1024]");
            s->tail_threads = static_cast<int>(v);
            break;
        case SG_TUNE_MOM_TILE:
        case SG_TUNE_MOM_TILE_FIRST:
            if (v < 0 || v >= N_MOM_TILES)
                throw ArgError("momentum tile must be 0.." + std::to_string(N_MOM_TILES - 1));
            (key == SG_TUNE_MOM_TILE ? s->mom_tile : s->mom_tile_first) = static_cast<int>(v);
            break;
        case SG_TUNE_MG_STAGED:
            if (v < 0 || v > 64)
                throw ArgError("staged level must be 0..64");
            s->mg_staged = static_cast<int>(v);
            break;
        case SG_TUNE_MID_CELLS:
        case SG_TUNE_LDS_CELLS:
            if (v < 0)
                throw ArgError("cell counts must be >= 0");
            *knob_of(*s, key) = v;
            break;
        case SG_TUNE_DC_TILE:
            if (v < 0 || v >= N_DC_TILES)
                throw ArgError("correction sweep tile must be 0.." + std::to_string(N_DC_TILES - 1));
            s->dc_tile = static_cast<int>(v);
            break;
        default:
            throw ArgError("unknown tuning key");
        }
    }
    free_plan(b);
}

long get_knob(Batch& b, int key) {
    Solver& s = *b.cav[0];
    switch (key) {
    case SG_TUNE_TAIL_THREADS:
        return s.tail_threads;
    case SG_TUNE_MOM_TILE:
        return s.mom_tile;
    case SG_TUNE_MG_STAGED:
        return s.mg_staged;
    case SG_TUNE_MOM_TILE_FIRST:
        return s.mom_tile_first >= 0 ? s.mom_tile_first : s.mom_tile;
    case SG_TUNE_MID_CELLS:
        return s.mid_cells;
    case SG_TUNE_LDS_CELLS:
        return s.lds_cells;
    case SG_TUNE_DC_TILE:
        return s.dc_tile;
    default:
        throw ArgError("unknown tuning key");
    }
}

} // namespace

} // namespace sg

using namespace sg;

namespace {
template <class F> int bguard(F f) {
    try {
        f();
        return SG_OK;
    } catch (const HipError& e) {
        set_error(e.what());
        (void)hipGetLastError(); // reported here, not by the next call
        return SG_E_HIP;
    } catch (const ArgError& e) {
        set_error(e.what());
        return SG_E_ARG;
    } catch (const std::bad_alloc&) {
        set_error("out of host or device memory");
        return SG_E_NOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SG_E_STATE;
    }
}

void destroy_batch(Batch* b) {
    if (!b)
        return;
    if (b->stream)
        (void)hipStreamSynchronize(b->stream);
    free_plan(*b);
    for (double* q : b->snap)
        (void)hipFree(q);
    for (Solver* s : b->cav)
        destroy_solver(s);
    for (hipEvent_t e : {b->t0, b->t1})
        if (e)
            (void)hipEventDestroy(e);
    if (b->stream)
        (void)hipStreamDestroy(b->stream);
    delete b;
}
} // namespace

extern "C" {

int sg_batch_create(const sg_params* ps, int n, sg_batch** out) {
    Batch* bp = nullptr;
    const int rc = bguard([&] {
        if (!ps || !out || n < 1)
            throw ArgError("batch: null argument or empty batch");
        const sg_params& p0 = ps[0];
        for (int c = 0; c < n; ++c) {
            const sg_params& p = ps[c];
            // Everything but the size and the Reynolds number is shared, so that every cavity
            // enqueues the same sequence of phases.
            const bool same =
                p.N_halo_x == p0.N_halo_x && p.maxit_inner_u == p0.maxit_inner_u &&
                p.maxit_inner_v == p0.maxit_inner_v && p.maxit_inner_p == p0.maxit_inner_p &&
                p.convection_scheme == p0.convection_scheme &&
                p.diffusion_order == p0.diffusion_order &&
                p.momentum_solver == p0.momentum_solver &&
                p.pressure_solver == p0.pressure_solver && p.mg_cycles == p0.mg_cycles &&
                p.mg_pre == p0.mg_pre && p.mg_post == p0.mg_post &&
                p.mg_precision == p0.mg_precision && p.mg_dc_maxit == p0.mg_dc_maxit &&
                p.momentum_precision == p0.momentum_precision &&
                p.residual_normalisation == p0.residual_normalisation;
            if (!same)
                throw ArgError("batch: cavities may differ only in N_cells_x, N_cells_y, Re, "
                               "tolerance and relaxation factors");
            if (p.pressure_solver == SG_P_CG)
                throw ArgError("batch: pressure_solver = cg needs the host between iterations");
        }
        bp = new Batch;
        Batch& b = *bp;
        for (int c = 0; c < n; ++c) {
            b.cav.push_back(create_solver(ps[c], nullptr, nullptr, nullptr));
            ensure_log(*b.cav.back(), std::max(256, ps[c].check_every));
        }
        // The batched launches share one set of launch shapes: the largest cavity's defaults.
        const Solver* big = b.cav[0];
        for (const Solver* s : b.cav)
            if (static_cast<long>(s->p.NX) * s->p.NY > static_cast<long>(big->p.NX) * big->p.NY)
                big = s;
        for (Solver* s : b.cav) {
            s->tail_threads = big->tail_threads, s->mid_cells = big->mid_cells;
            s->lds_cells = big->lds_cells, s->mg_staged = big->mg_staged;
            s->mom_tile = big->mom_tile, s->mom_tile_first = big->mom_tile_first;
            s->dc_tile = big->dc_tile;
        }
        int dev = 0;
        SG_HIP(hipGetDevice(&dev));
        SG_HIP(hipDeviceGetAttribute(&b.cus, hipDeviceAttributeMultiprocessorCount, dev));
        SG_HIP(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
        SG_HIP(hipEventCreate(&b.t0));
        SG_HIP(hipEventCreate(&b.t1));
        b.done.assign(n, 0);
        b.iters.assign(n, 0);
        b.last.assign(3 * static_cast<size_t>(n), 0.0);
    });
    if (rc != SG_OK) {
        destroy_batch(bp);
        return rc;
    }
    *out = reinterpret_cast<sg_batch*>(bp);
    return SG_OK;
}

int sg_batch_destroy(sg_batch* h) {
    destroy_batch(reinterpret_cast<Batch*>(h));
    return SG_OK;
}

int sg_batch_size(sg_batch* h, int* n) {
    return bguard([&] {
        if (!n)
            throw ArgError("null pointer");
        *n = static_cast<int>(B(h).cav.size());
    });
}

int sg_batch_solver(sg_batch* h, int k, sg_solver** s) {
    return bguard([&] {
        Batch& b = B(h);
        if (!s || k < 0 || k >= static_cast<int>(b.cav.size()))
            throw ArgError("batch: bad cavity index");
        *s = reinterpret_cast<sg_solver*>(b.cav[k]);
    });
}

int sg_batch_advance(sg_batch* h, int max_iters, int* iters, int* status, double* last_res) {
    return bguard([&] {
        Batch& b = B(h);
        if (max_iters < 0)
            throw ArgError("max_iters must be >= 0");
        const int n = static_cast<int>(b.cav.size());
        const int chunk = std::max(1, b.cav[0]->p.check_every);
        if (!b.graph)
            build_plan(b);
        int left = max_iters;
        while (left > 0) {
            if (std::all_of(b.done.begin(), b.done.end(), [](int d) { return d != 0; }))
                break;
            const int k = std::min(left, chunk);
            // Cavities still running assemble in this chunk, in their iterations' form.
            for (int c = 0; c < n; ++c)
                if (!b.done[c])
                    mark_iterated(*b.cav[c]);
            begin_chunk(b);
            replay(b, k);
            read_state(b);
            left -= k;
        }
        for (int c = 0; c < n; ++c) {
            if (iters)
                iters[c] = b.iters[c];
            if (status)
                status[c] = b.done[c];
            if (last_res)
                std::memcpy(last_res + 3 * c, &b.last[3 * c], 3 * sizeof(double));
        }
    });
}

int sg_batch_launches(sg_batch* h, int* per_iteration) {
    return bguard([&] {
        Batch& b = B(h);
        if (!per_iteration)
            throw ArgError("null pointer");
        if (!b.graph)
            build_plan(b);
        *per_iteration = b.launches;
    });
}

int sg_batch_set_tuning(sg_batch* h, int key, long value) {
    return bguard([&] { set_knob(B(h), key, value); });
}

int sg_batch_get_tuning(sg_batch* h, int key, long* value) {
    return bguard([&] {
        if (!value)
            throw ArgError("null pointer");
        *value = get_knob(B(h), key);
    });
}

int sg_batch_autotune(sg_batch* h, int reps, char* report, size_t len) {
    return bguard([&] {
        Batch& b = B(h);
        reps = std::max(reps, 1);
        std::string out;
        char line[256];
        snapshot(b, true);
        int bt = 0;
        double best = 1e30;
        for (int t = 0; t < N_MOM_TILES; ++t) {
            set_knob(b, SG_TUNE_MOM_TILE, t);
            const double us = 1e6 * time_iteration(b, reps);
            std::snprintf(line, sizeof line, "momentum tile %-7s: %10.2f us per batch iteration\n",
                          tile_name(t).c_str(), us);
            out += line;
            if (us < best) {
                best = us;
                bt = t;
            }
        }
        set_knob(b, SG_TUNE_MOM_TILE, bt);
        std::snprintf(line, sizeof line, "best momentum tile: %s (%d)\n", tile_name(bt).c_str(),
                      bt);
        out += line;
        int bf = bt;
        for (int t = 0; t < N_MOM_TILES; ++t) {
            set_knob(b, SG_TUNE_MOM_TILE_FIRST, t);
            const double us = 1e6 * time_iteration(b, reps);
            std::snprintf(line, sizeof line,
                          "first-sweep tile %-7s: %10.2f us per batch iteration\n",
                          tile_name(t).c_str(), us);
            out += line;
            if (us < best) {
                best = us;
                bf = t;
            }
        }
        set_knob(b, SG_TUNE_MOM_TILE_FIRST, bf);
        std::snprintf(line, sizeof line, "best first-sweep tile: %s (%d)\n",
                      tile_name(bf).c_str(), bf);
        out += line;
        long bth = get_knob(b, SG_TUNE_TAIL_THREADS), bmid = get_knob(b, SG_TUNE_MID_CELLS),
             blds = get_knob(b, SG_TUNE_LDS_CELLS);
        best = 1e30;
        for (long th : {64L, 256L, 1024L})
            for (long mid : {0L, 16384L})
                for (long lds : {1024L, 4096L}) {
                    set_knob(b, SG_TUNE_TAIL_THREADS, th);
                    set_knob(b, SG_TUNE_MID_CELLS, mid);
                    set_knob(b, SG_TUNE_LDS_CELLS, lds);
                    const double us = 1e6 * time_iteration(b, reps);
                    std::snprintf(line, sizeof line,
                                  "tail_threads %4ld mid_cells %6ld lds_cells %5ld : %10.2f us\n",
                                  th, mid, lds, us);
                    out += line;
                    if (us < best) {
                        best = us;
                        bth = th;
                        bmid = mid;
                        blds = lds;
                    }
                }
        set_knob(b, SG_TUNE_TAIL_THREADS, bth);
        set_knob(b, SG_TUNE_MID_CELLS, bmid);
        set_knob(b, SG_TUNE_LDS_CELLS, blds);
        std::snprintf(line, sizeof line,
                      "best: tail_threads %ld mid_cells %ld lds_cells %ld : %.2f us\n", bth, bmid,
                      blds, best);
        out += line;
        long bst = get_knob(b, SG_TUNE_MG_STAGED);
        for (long st : {0L, 1L, 64L}) {
            set_knob(b, SG_TUNE_MG_STAGED, st);
            const double us = 1e6 * time_iteration(b, reps);
            std::snprintf(line, sizeof line, "staged from level %2ld : %10.2f us\n", st, us);
            out += line;
            if (us < best) {
                best = us;
                bst = st;
            }
        }
        set_knob(b, SG_TUNE_MG_STAGED, bst);
        std::snprintf(line, sizeof line, "best: staged from level %ld : %.2f us\n", bst, best);
        out += line;
        snapshot(b, false);
        build_plan(b);
        if (report && len) {
            std::strncpy(report, out.c_str(), len - 1);
            report[len - 1] = '\0';
        }
    });
}

} // extern "C"
