// The outer loop of run_outer_loop (src/solver.cpp:106-132) over the C ABI. The first iteration
// runs phase by phase with hipEvent timers (the reference's per-phase timing); later iterations
// run as graph replays in chunks of check_every, with the stopping rule evaluated on the device
// after every iteration, so the fields stop exactly at the converged iteration.
#include "host.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace sh {

namespace {
struct Handle {
    sg_solver* s = nullptr;
    sg_unet* net = nullptr;
    ~Handle() {
        if (s)
            sg_destroy(s);
        if (net)
            sg_unet_destroy(net);
    }
};

// Weights file of the warm-start network: the line "SGUNET1 <n>" then n little-endian float32.
std::vector<float> read_weights(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f)
        throw std::runtime_error("cannot open warm_start_weights file " + path);
    char head[64] = {0};
    int n = 0;
    const bool ok = std::fgets(head, sizeof head, f) && std::sscanf(head, "SGUNET1 %d", &n) == 1 &&
                    n > 0;
    std::vector<float> w(ok ? static_cast<size_t>(n) : 0);
    const bool full = ok && std::fread(w.data(), sizeof(float), w.size(), f) == w.size();
    std::fclose(f);
    if (!full)
        throw std::runtime_error(path + ": not a warm-start weights file (SGUNET1)");
    return w;
}
} // namespace

RunResult run_simple_gpu(const Params& P, std::vector<double>& u, std::vector<double>& v,
                         std::vector<double>& pf) {
    const sg_params& p = P.g;
    RunResult out;
    char name[256] = {0};
    if (sg_device_name(name, sizeof name) == 0)
        out.device = name;
    Handle h;
    check(sg_create(&p, u.data(), v.data(), pf.data(), &h.s));
    if (p.pressure_warm_start == SG_WARM_LEARNED) {
        if (P.warm_start_weights.empty())
            throw std::runtime_error("pressure_warm_start = learned needs warm_start_weights");
        const std::vector<float> w = read_weights(P.warm_start_weights);
        check(sg_unet_create(&h.net));
        check(sg_unet_set_params(h.net, w.data(), static_cast<int>(w.size())));
        check(sg_attach_warm_start(h.s, h.net, P.warm_start_level));
    }
    const auto t0 = std::chrono::steady_clock::now();
    std::array<double, SG_NPHASES> ph1{};
    double eager_s = 0.0;
    auto log_row = [&](int iter, const double* r) {
        out.log.push_back({static_cast<double>(iter), r[0], r[1], r[2]});
        if (iter == 1 || iter % 100 == 0)
            std::printf("iter %d  res_u %.6e  res_v %.6e  res_p %.6e\n", iter, r[0], r[1], r[2]);
        if (!std::isfinite(r[0]) || !std::isfinite(r[1]) || !std::isfinite(r[2]))
            out.diverged = true;
    };
    int done = 0;
    if (p.maxit >= 1) {
        // Iteration 1, phase by phase. The stop rule needs iter > 1, so it cannot stop here.
        double r[3];
        const auto e0 = std::chrono::steady_clock::now();
        check(sg_outer_iteration(h.s, r, ph1.data()));
        eager_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - e0).count();
        done = 1;
        log_row(1, r);
    }
    // Launch shapes timed on the running state (a fraction of a second; every array is restored
    // and results do not depend on the shapes), for runs long enough to repay it.
    if (P.autotune && p.maxit >= 1000 && static_cast<long>(p.NX) * p.NY >= 64L * 64) {
        const auto a0 = std::chrono::steady_clock::now();
        check(sg_autotune(h.s, 5, nullptr, 0));
        out.autotune_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - a0).count();
    }
    // Phase timing of the replayed iterations: the first of every chunk is timed on the device.
    check(sg_set_phase_timing(h.s, 1));
    const int chunk = p.check_every > 0 ? p.check_every : 200;
    std::vector<double> buf(3 * static_cast<size_t>(chunk));
    while (done < p.maxit && !out.diverged) {
        const int n = std::min(chunk, p.maxit - done);
        int k = 0, conv = 0;
        check(sg_advance(h.s, n, &k, buf.data(), &conv));
        for (int m = 0; m < k; ++m)
            log_row(done + m + 1, &buf[3 * static_cast<size_t>(m)]);
        done += k;
        if (conv == 1) {
            out.converged = true;
            break;
        }
        if (conv < 0) {
            out.diverged = true;
            break;
        }
        if (k < n)
            break;
    }
    out.iterations = done;
    if (h.net) {
        check(sg_warm_stats(h.s, &out.warm_kept, &out.warm_discarded));
        check(sg_warm_level(h.s, &out.warm_level));
    }
    check(sg_download(h.s, SG_F_U, u.data()));
    check(sg_download(h.s, SG_F_V, v.data()));
    check(sg_download(h.s, SG_F_P, pf.data()));
    out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // Phase seconds: iteration 1 measured per phase (eager, hipEvents); the replayed iterations
    // from the device-timed first iteration of every chunk (event-record nodes in its graph),
    // their mean per phase times the replayed iterations. What the phases do not cover (gaps
    // between graph launches, the host's reads of the stop flag) stays in timing.csv's
    // "unaccounted" row.
    std::array<double, SG_NPHASES> acc{};
    int timed = 0;
    check(sg_phase_times(h.s, acc.data(), &timed));
    for (int k = 0; k < SG_NPHASES; ++k)
        out.phase_s[k] = ph1[k] + (timed > 0 ? acc[k] / timed * (done - 1) : 0.0);
    out.phase_sampled = timed;
    (void)eager_s;
    if (out.diverged)
        std::printf("stopped: non-finite residual at iteration %d\n", done);
    return out;
}

} // namespace sh
