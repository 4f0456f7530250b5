// Output files in the reference's formats (src/io.cpp:10-151): full fields with halos as %.18e
// CSV (round-trips exactly), cell-centred velocities, the residual log, phase timing and
// run_info.json.
#include "host.hpp"

#include <chrono>
#include <cstdio>
#include <filesystem>

#ifndef SH_GIT_HASH
#define SH_GIT_HASH "unknown"
#endif
#ifndef SH_COMPILER
#define SH_COMPILER "unknown"
#endif

namespace sh {

namespace {
struct File {
    std::FILE* f;
    File(const std::string& path, const char* what) : f(std::fopen(path.c_str(), "w")) {
        if (!f)
            throw std::runtime_error(std::string("cannot write ") + what);
    }
    ~File() { std::fclose(f); }
};

std::string json_escape(const std::string& s) {
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\')
            o += '\\';
        o += c;
    }
    return o;
}
} // namespace

void write_csv(const std::string& path, const double* f, int rows, int cols) {
    std::FILE* fp = std::fopen(path.c_str(), "w");
    if (!fp)
        throw std::runtime_error("cannot open for writing: " + path);
    for (int i = 0; i < rows; ++i) {
        for (int j = 0; j < cols; ++j)
            std::fprintf(fp, j ? ",%.18e" : "%.18e", f[static_cast<size_t>(i) * cols + j]);
        std::fputc('\n', fp);
    }
    std::fclose(fp);
}

void write_npy(const std::string& path, const double* f, int rows, int cols) {
    std::FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp)
        throw std::runtime_error("cannot open for writing: " + path);
    // Magic, version 1.0, header length, then the header dict padded with spaces to a multiple of
    // 64 bytes in all and ended by a newline.
    std::string hdr = "{'descr': '<f8', 'fortran_order': False, 'shape': (" +
                      std::to_string(rows) + ", " + std::to_string(cols) + "), }";
    const size_t total = (10 + hdr.size() + 1 + 63) / 64 * 64;
    hdr.append(total - 10 - hdr.size() - 1, ' ');
    hdr += '\n';
    const unsigned char pre[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0,
                                   static_cast<unsigned char>(hdr.size() & 0xff),
                                   static_cast<unsigned char>(hdr.size() >> 8)};
    const size_t n = static_cast<size_t>(rows) * cols;
    const bool ok = std::fwrite(pre, 1, 10, fp) == 10 &&
                    std::fwrite(hdr.data(), 1, hdr.size(), fp) == hdr.size() &&
                    std::fwrite(f, sizeof(double), n, fp) == n; // x86-64: little-endian
    std::fclose(fp);
    if (!ok)
        throw std::runtime_error("cannot write " + path);
}

void write_outputs(const Params& P, const std::vector<double>& u, const std::vector<double>& v,
                   const std::vector<double>& pf, const RunResult& r, const std::string& cfg_path) {
    namespace fs = std::filesystem;
    const sg_params& p = P.g;
    std::array<double, kPhases> ph = r.phase_s;
    const fs::path d(P.output_dir);
    const auto t0 = std::chrono::steady_clock::now();
    fs::create_directories(d);
    if (p.write_fields) {
        const int nr = p.nrows, nc = p.ncols, H = p.H;
        const bool npy = P.field_format == "npy";
        auto field = [&](const char* name, const double* a, int rows, int cols) {
            if (npy)
                write_npy((d / (std::string(name) + ".npy")).string(), a, rows, cols);
            else
                write_csv((d / (std::string(name) + ".csv")).string(), a, rows, cols);
        };
        field("u", u.data(), nr, nc);
        field("v", v.data(), nr, nc);
        field("p", pf.data(), nr, nc);
        // Cell-centred velocities (src/io.cpp:22-31, docs/SPEC.md section 6).
        std::vector<double> ux(static_cast<size_t>(p.NX) * p.NY), uy(ux.size());
        for (int rr = 0; rr < p.NX; ++rr)
            for (int c = 0; c < p.NY; ++c) {
                const size_t q = static_cast<size_t>(rr) * p.NY + c;
                ux[q] = (u[static_cast<size_t>(H + rr) * nc + H + c] +
                         u[static_cast<size_t>(H + rr) * nc + H + c + 1]) / 2;
                uy[q] = (v[static_cast<size_t>(H + rr) * nc + H + 1 + c] +
                         v[static_cast<size_t>(H + rr + 1) * nc + H + 1 + c]) / 2;
            }
        field("ux_centre", ux.data(), p.NX, p.NY);
        field("uy_centre", uy.data(), p.NX, p.NY);
    }
    {
        File al((d / "alog.csv").string(), "alog.csv");
        std::fputs("Iteration,Residual_u,Residual_v,Residual_p\n", al.f);
        for (const auto& row : r.log)
            std::fprintf(al.f, "%d,%.18e,%.18e,%.18e\n", static_cast<int>(row[0]), row[1], row[2],
                         row[3]);
    }
    ph[kPhases - 1] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    {
        double total = 0.0;
        for (double x : ph)
            total += x;
        const double timed = total - ph[kPhases - 1];
        File tf((d / "timing.csv").string(), "timing.csv");
        std::fputs("phase,seconds,fraction\n", tf.f);
        const double denom = total > 0.0 ? total : 1.0;
        for (int k = 0; k < kPhases; ++k)
            std::fprintf(tf.f, "%s,%.9f,%.6f\n", phase_name(k), ph[k], ph[k] / denom);
        std::fprintf(tf.f, "total,%.9f,1.000000\n", total);
        std::fprintf(tf.f, "unaccounted,%.9f,%.6f\n", r.seconds - timed,
                     r.seconds > 0.0 ? (r.seconds - timed) / r.seconds : 0.0);
    }
    File js((d / "run_info.json").string(), "run_info.json");
    std::FILE* fp = js.f;
    std::fprintf(fp, "{\n  \"config\": \"%s\",\n  \"git_hash\": \"%s\",\n",
                 json_escape(cfg_path).c_str(), SH_GIT_HASH);
    std::fprintf(fp, "  \"compiler\": \"%s\",\n", SH_COMPILER);
    std::fprintf(fp, "  \"backend\": \"%s\",\n  \"threads\": %d,\n", backend_name(p.backend), 1);
    std::fprintf(fp, "  \"device\": \"%s\",\n", json_escape(r.device).c_str());
    std::fprintf(fp, "  \"momentum_solver\": \"%s\",\n  \"pressure_solver\": \"%s\",\n",
                 momentum_name(p.momentum_solver), pressure_name(p.pressure_solver));
    std::fprintf(fp, "  \"N_cells_x\": %d,\n  \"N_cells_y\": %d,\n  \"N_halo_x\": %d,\n", p.NX,
                 p.NY, p.N_halo_x);
    std::fprintf(fp, "  \"N_halo_y\": %d,\n  \"Re\": %.17g,\n", p.N_halo_y, p.Re);
    std::fprintf(fp, "  \"maxit_inner_u\": %d,\n  \"maxit_inner_v\": %d,\n", p.maxit_inner_u,
                 p.maxit_inner_v);
    std::fprintf(fp, "  \"maxit_inner_p\": %d,\n  \"error_tolerance\": %.17g,\n", p.maxit_inner_p,
                 p.error_tolerance);
    std::fprintf(fp, "  \"maxit\": %d,\n", p.maxit);
    std::fprintf(fp, "  \"relaxation_factor_u\": %.17g,\n  \"relaxation_factor_v\": %.17g,\n",
                 p.relaxation_factor_u, p.relaxation_factor_v);
    std::fprintf(fp, "  \"relaxation_factor_p\": %.17g,\n", p.relaxation_factor_p);
    std::fprintf(fp, "  \"convection_scheme\": \"%s\",\n  \"diffusion_order\": %d,\n",
                 scheme_name(p.convection_scheme), p.diffusion_order);
    std::fprintf(fp, "  \"Hybrid_switch\": %s,\n", p.Hybrid_switch ? "true" : "false");
    std::fprintf(fp, "  \"residual_normalisation\": \"%s\",\n",
                 p.residual_normalisation == SG_NORM_ABSOLUTE ? "absolute" : "running_max");
    std::fprintf(fp, "  \"output_dir\": \"%s\",\n", json_escape(P.output_dir).c_str());
    std::fprintf(fp, "  \"pressure_tolerance\": %.17g,\n  \"pressure_maxit\": %d,\n",
                 p.pressure_tolerance, p.pressure_maxit);
    std::fprintf(fp, "  \"mg_cycles\": %d,\n  \"mg_pre\": %d,\n  \"mg_post\": %d,\n", p.mg_cycles,
                 p.mg_pre, p.mg_post);
    std::fprintf(fp, "  \"mg_omega\": %.17g,\n", p.mg_omega);
    std::fprintf(fp, "  \"mg_precision\": \"%s\",\n  \"mg_coarsest\": %d,\n",
                 p.mg_precision == SG_MG_MIXED ? "mixed" : "fp64", p.mg_coarsest);
    std::fprintf(fp, "  \"momentum_precision\": \"%s\",\n",
                 p.momentum_precision == SG_MOMP_MIXED ? "mixed" : "fp64");
    std::fprintf(fp, "  \"pressure_warm_start\": \"%s\",\n",
                 p.pressure_warm_start == SG_WARM_LEARNED ? "learned" : "none");
    std::fprintf(fp, "  \"cg_iterations_total\": %lld,\n  \"cg_iterations_mean\": %.6f,\n",
                 r.cg_iterations,
                 r.iterations > 0 ? static_cast<double>(r.cg_iterations) / r.iterations : 0.0);
    std::fprintf(fp, "  \"ranks\": %d,\n", 1);
    std::fprintf(fp, "  \"halo_seconds_mean\": %.6f,\n  \"halo_seconds_max\": %.6f,\n", 0.0, 0.0);
    std::fprintf(fp, "  \"collective_seconds_mean\": %.6f,\n  \"collective_seconds_max\": %.6f,\n",
                 0.0, 0.0);
    std::fprintf(fp, "  \"diverged\": %s,\n", r.diverged ? "true" : "false");
    std::fprintf(fp, "  \"phase_timed_iterations\": %d,\n", r.phase_sampled);
    std::fprintf(fp, "  \"autotune_seconds\": %.6f,\n", r.autotune_s);
    if (p.pressure_warm_start == SG_WARM_LEARNED)
        std::fprintf(fp, "  \"warm_start\": {\"weights\": \"%s\", \"level\": %d, \"kept\": %d, "
                         "\"discarded\": %d},\n",
                     json_escape(P.warm_start_weights).c_str(), r.warm_level, r.warm_kept,
                     r.warm_discarded);
    std::fprintf(fp,
                 "  \"iterations\": %d,\n  \"converged\": %s,\n  \"seconds\": %.6f,\n"
                 "  \"seconds_per_iteration\": %.9f\n}\n",
                 r.iterations, r.converged ? "true" : "false", r.seconds,
                 r.iterations > 0 ? r.seconds / r.iterations : 0.0);
}

} // namespace sh
