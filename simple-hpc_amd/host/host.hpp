// host.hpp — C++17 host of the MI355X SIMPLE backend: config file, parameters, the outer loop
// and the output files, with the reference's keys, formats and error behaviour
// (src/config.cpp, src/params.cpp, src/solver.cpp:106-132, src/io.cpp, src/main.cpp). The
// numerics run behind the C ABI of include/simple_gpu.h.
#pragma once
#include "../../include/simple_gpu.h"

#include <array>
#include <iosfwd>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace sh {

using ConfigMap = std::map<std::string, std::string>;

// `key = value` lines, '#' comments, surrounding whitespace trimmed; malformed lines and duplicate
// keys throw with the line number (src/config.cpp:16-39).
ConfigMap parse_config(std::istream& in);
ConfigMap load_config(const std::string& path);

struct Params {
    sg_params g{};                      // numerics, passed through the C ABI
    std::string output_dir = "results/run";
    std::string mpi_grid = "auto";      // accepted for config compatibility; one GPU runs 1x1
    bool halo_overlap = true;
    std::string warm_start_weights;     // pressure_warm_start = learned: network weights file
    int warm_start_level = -1; // multigrid level the network runs on (2^level blocks; -1: automatic)
    std::string field_format = "csv";   // field files: csv (the reference's), or npy (binary)
    bool autotune = true;               // time the launch shapes on the running state after
                                        // iteration 1 (results do not depend on them)
};

// Known keys only, typed parsing, validation and derived sizes (src/params.cpp:50-180), plus this
// backend's keys.
Params make_params(const ConfigMap& cfg);
void print_params(std::ostream& os, const Params& p);
const char* scheme_name(int v);
const char* backend_name(int v);
const char* momentum_name(int v);
const char* pressure_name(int v);

// Per-phase timers, same names as include/simple/timer.hpp:6-20.
constexpr int kPhases = 12; // 11 solver phases + output
const char* phase_name(int k);

struct RunResult {
    int iterations = 0;
    bool converged = false;
    bool diverged = false;              // a residual became NaN/Inf (the loop stops there)
    std::vector<std::array<double, 4>> log;
    double seconds = 0.0;
    std::array<double, kPhases> phase_s{};
    long long cg_iterations = 0;
    std::string device;
    int warm_kept = 0, warm_discarded = 0; // learned warm start: guesses used / discarded
    int warm_level = -1;                   // the level it ran on
    int phase_sampled = 0;              // replayed iterations timed per phase on the device
    double autotune_s = -1.0;           // seconds the autotuner took (-1: not run)
};

// C ABI status -> std::runtime_error with sg_last_error() (the "GPU error: ..." path of
// src/hip/solver_hip.hip:17-23).
void check(int rc);

// run_simple_hip (include/simple/solver_hip.hpp:40): uploads u, v, p, runs the outer loop with the
// stopping rule of src/solver.cpp:106-132 (graph replays in chunks with the stop rule evaluated
// on the device), downloads u, v, p. u, v, pf are nrows x ncols host arrays.
RunResult run_simple_gpu(const Params& p, std::vector<double>& u, std::vector<double>& v,
                         std::vector<double>& pf);

// Output files of src/io.cpp:43-151: u/v/p.csv, ux/uy_centre.csv (%.18e; .npy with
// field_format = npy), alog.csv, timing.csv, run_info.json.
void write_csv(const std::string& path, const double* f, int rows, int cols);
// The same array as a NumPy .npy file (format 1.0, little-endian float64, C order): 8 bytes per
// value instead of ~25 characters, and exact. The reference's run loader (postprocess/io.py:11-19)
// reads NAME.npy when NAME.csv is absent.
void write_npy(const std::string& path, const double* f, int rows, int cols);
void write_outputs(const Params& p, const std::vector<double>& u, const std::vector<double>& v,
                   const std::vector<double>& pf, const RunResult& r, const std::string& cfg_path);

} // namespace sh
