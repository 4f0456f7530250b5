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
}
