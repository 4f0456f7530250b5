// C entry points of the host library (libsimple_host.so) for the Python bindings.
#include "host.hpp"

#include <cstring>
#include <sstream>

extern "C" {

// Config text -> sg_params (and output_dir) with the host's parser; 0 on success, else 1 with the
// message that make_params would throw.
int sh_params_from_config(const char* text, sg_params* out, char* output_dir, int outlen,
                          char* err, int errlen) {
    try {
        std::istringstream in(text ? text : "");
        const sh::Params p = sh::make_params(sh::parse_config(in));
        *out = p.g;
        if (output_dir && outlen > 0) {
            std::strncpy(output_dir, p.output_dir.c_str(), outlen - 1);
            output_dir[outlen - 1] = 0;
        }
        return 0;
    } catch (const std::exception& e) {
        if (err && errlen > 0) {
            std::strncpy(err, e.what(), errlen - 1);
            err[errlen - 1] = 0;
        }
        return 1;
    }
}

// One array as a .npy file (the field_format = npy writer); 0 on success.
int sh_write_npy(const char* path, const double* f, int rows, int cols) {
    try {
        sh::write_npy(path, f, rows, cols);
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

// Writes the output files of a finished run (u, v, pf: nrows x ncols; log: n rows of
// [iter, res_u, res_v, res_p]) to output_dir with the host's writers; 0 on success.
int sh_write_outputs(const sg_params* p, const char* output_dir, const double* u, const double* v,
                     const double* pf, const double* log, int n, int converged, double seconds,
                     const char* cfg_path, char* err, int errlen) {
    try {
        sh::Params P;
        P.g = *p;
        P.output_dir = output_dir;
        const size_t cells = static_cast<size_t>(p->nrows) * p->ncols;
        std::vector<double> U(u, u + cells), V(v, v + cells), PF(pf, pf + cells);
        sh::RunResult r;
        r.iterations = n;
        r.converged = converged != 0;
        r.seconds = seconds;
        for (int k = 0; k < n; ++k)
            r.log.push_back({log[4 * k], log[4 * k + 1], log[4 * k + 2], log[4 * k + 3]});
        sh::write_outputs(P, U, V, PF, r, cfg_path ? cfg_path : "");
        return 0;
    } catch (const std::exception& e) {
        if (err && errlen > 0) {
            std::strncpy(err, e.what(), errlen - 1);
            err[errlen - 1] = 0;
        }
        return 1;
    }
}
}
