// Parameters from a config map: the reference's known-key set, typed parsing and messages
// (src/params.cpp:50-180), plus this backend's keys. Validation and the derived sizes are done
// by sg_params_finalize so that the C ABI and the host agree.
#include "host.hpp"

#include <cstddef>
#include <initializer_list>
#include <iomanip>
#include <ostream>
#include <set>
#include <utility>

namespace sh {

namespace {
int as_int(const std::string& k, const std::string& v) {
    try {
        size_t used = 0;
        const int r = std::stoi(v, &used);
        if (used == v.size())
            return r;
    } catch (const std::exception&) {
    }
    throw std::runtime_error("cannot parse integer for '" + k + "': " + v);
}

double as_double(const std::string& k, const std::string& v) {
    try {
        size_t used = 0;
        const double r = std::stod(v, &used);
        if (used == v.size())
            return r;
    } catch (const std::exception&) {
    }
    throw std::runtime_error("cannot parse number for '" + k + "': " + v);
}

bool as_bool(const std::string& k, const std::string& v) {
    if (v == "True")
        return true;
    if (v == "False")
        return false;
    throw std::runtime_error("invalid boolean for '" + k + "' (use True/False): " + v);
}

int as_enum(const std::string& k, const std::string& v,
            std::initializer_list<std::pair<const char*, int>> names) {
    for (const auto& n : names)
        if (v == n.first)
            return n.second;
    throw std::runtime_error("unknown value for '" + k + "': " + v);
}

const std::set<std::string>& known_keys() {
    // The reference's 30 keys (src/params.cpp:51-59) and this backend's additions.
    static const std::set<std::string> k = {
        "N_cells_x", "N_cells_y", "N_halo_x", "N_halo_y", "Re", "maxit_inner_u", "maxit_inner_v",
        "maxit_inner_p", "error_tolerance", "maxit", "relaxation_factor_u", "relaxation_factor_v",
        "relaxation_factor_p", "convection_scheme", "diffusion_order", "Hybrid_switch", "backend",
        "momentum_solver", "pressure_solver", "output_dir", "residual_normalisation",
        "pressure_tolerance", "pressure_maxit", "mg_cycles", "mg_pre", "mg_post", "mg_omega",
        "mpi_grid", "halo_overlap", "write_fields",
        // additions
        "mg_precision", "mg_coarsest", "mg_dc_maxit", "pressure_warm_start",
        "warm_start_weights", "warm_start_level", "check_every", "field_format", "autotune",
        "momentum_precision"};
    return k;
}
} // namespace

const char* scheme_name(int v) { return v == SG_UDS ? "UDS" : "QUICK"; }
const char* backend_name(int v) {
    return v == SG_BACKEND_SERIAL ? "serial" : v == SG_BACKEND_OPENMP ? "openmp" : "gpu";
}
const char* momentum_name(int v) {
    return v == SG_MOM_GS_LEX ? "gs_lex" : v == SG_MOM_JACOBI ? "jacobi" : "gs_multicolour";
}
const char* pressure_name(int v) {
    switch (v) {
    case SG_P_GS_LEX: return "gs_lex";
    case SG_P_JACOBI: return "jacobi";
    case SG_P_GS_REDBLACK: return "gs_redblack";
    case SG_P_CG: return "cg";
    default: return "mg";
    }
}

Params make_params(const ConfigMap& cfg) {
    for (const auto& kv : cfg)
        if (!known_keys().count(kv.first))
            throw std::runtime_error("unknown config key: " + kv.first);
    Params P;
    sg_params& p = P.g;
    sg_params_default(&p);
    for (const auto& kv : cfg) {
        const std::string &k = kv.first, &v = kv.second;
        if (k == "N_cells_x") p.NX = as_int(k, v);
        else if (k == "N_cells_y") p.NY = as_int(k, v);
        else if (k == "N_halo_x") p.N_halo_x = as_int(k, v);
        else if (k == "N_halo_y") p.N_halo_y = as_int(k, v);
        else if (k == "Re") p.Re = as_double(k, v);
        else if (k == "maxit_inner_u") p.maxit_inner_u = as_int(k, v);
        else if (k == "maxit_inner_v") p.maxit_inner_v = as_int(k, v);
        else if (k == "maxit_inner_p") p.maxit_inner_p = as_int(k, v);
        else if (k == "error_tolerance") p.error_tolerance = as_double(k, v);
        else if (k == "maxit") p.maxit = as_int(k, v);
        else if (k == "relaxation_factor_u") p.relaxation_factor_u = as_double(k, v);
        else if (k == "relaxation_factor_v") p.relaxation_factor_v = as_double(k, v);
        else if (k == "relaxation_factor_p") p.relaxation_factor_p = as_double(k, v);
        else if (k == "convection_scheme")
            p.convection_scheme = as_enum(k, v, {{"UDS", SG_UDS}, {"QUICK", SG_QUICK}});
        else if (k == "diffusion_order") p.diffusion_order = as_int(k, v);
        else if (k == "Hybrid_switch") p.Hybrid_switch = as_bool(k, v) ? 1 : 0;
        else if (k == "backend")
            p.backend = as_enum(k, v, {{"serial", SG_BACKEND_SERIAL},
                                       {"openmp", SG_BACKEND_OPENMP},
                                       {"gpu", SG_BACKEND_GPU}});
        else if (k == "momentum_solver")
            p.momentum_solver = as_enum(k, v, {{"gs_lex", SG_MOM_GS_LEX},
                                               {"jacobi", SG_MOM_JACOBI},
                                               {"gs_multicolour", SG_MOM_GS_MULTICOLOUR}});
        else if (k == "pressure_solver")
            p.pressure_solver = as_enum(k, v, {{"gs_lex", SG_P_GS_LEX},
                                               {"jacobi", SG_P_JACOBI},
                                               {"gs_redblack", SG_P_GS_REDBLACK},
                                               {"cg", SG_P_CG},
                                               {"mg", SG_P_MG}});
        else if (k == "residual_normalisation")
            p.residual_normalisation = as_enum(
                k, v, {{"absolute", SG_NORM_ABSOLUTE}, {"running_max", SG_NORM_RUNNING_MAX}});
        else if (k == "output_dir") P.output_dir = v;
        else if (k == "pressure_tolerance") p.pressure_tolerance = as_double(k, v);
        else if (k == "pressure_maxit") p.pressure_maxit = as_int(k, v);
        else if (k == "mg_cycles") p.mg_cycles = as_int(k, v);
        else if (k == "mg_pre") p.mg_pre = as_int(k, v);
        else if (k == "mg_post") p.mg_post = as_int(k, v);
        else if (k == "mg_omega") p.mg_omega = as_double(k, v);
        else if (k == "halo_overlap") P.halo_overlap = as_bool(k, v);
        else if (k == "write_fields") p.write_fields = as_bool(k, v) ? 1 : 0;
        else if (k == "mpi_grid") {
            if (v != "auto") { // RxC, e.g. 2x2 (src/params.cpp:135-143)
                const size_t x = v.find('x');
                if (x == std::string::npos)
                    throw std::runtime_error("mpi_grid must be auto or RxC (e.g. 2x2): " + v);
                const int r = as_int(k, v.substr(0, x)), c = as_int(k, v.substr(x + 1));
                if (r < 1 || c < 1)
                    throw std::runtime_error("mpi_grid sizes must be positive: " + v);
            }
            P.mpi_grid = v;
        }
        else if (k == "mg_precision")
            p.mg_precision = as_enum(k, v, {{"fp64", SG_MG_FP64}, {"mixed", SG_MG_MIXED}});
        else if (k == "momentum_precision")
            p.momentum_precision =
                as_enum(k, v, {{"fp64", SG_MOMP_FP64}, {"mixed", SG_MOMP_MIXED}});
        else if (k == "mg_coarsest") p.mg_coarsest = as_int(k, v);
        else if (k == "mg_dc_maxit") p.mg_dc_maxit = as_int(k, v);
        else if (k == "pressure_warm_start")
            p.pressure_warm_start =
                as_enum(k, v, {{"none", SG_WARM_NONE}, {"learned", SG_WARM_LEARNED}});
        else if (k == "warm_start_weights") P.warm_start_weights = v;
        else if (k == "warm_start_level") P.warm_start_level = as_int(k, v);
        else if (k == "check_every") p.check_every = as_int(k, v);
        else if (k == "autotune") P.autotune = as_bool(k, v);
        else if (k == "field_format") {
            if (v != "csv" && v != "npy")
                throw std::runtime_error("field_format must be csv or npy: " + v);
            P.field_format = v;
        }
    }
    check(sg_params_finalize(&p));
    return P;
}

void print_params(std::ostream& os, const Params& P) {
    const sg_params& p = P.g;
    os << std::setprecision(10);
    os << "NX = " << p.NX << "\nNY = " << p.NY << "\nH = " << p.H << "\n"
       << "nrows = " << p.nrows << "\nncols = " << p.ncols << "\n"
       << "dx = " << p.dx << "\ndy = " << p.dy << "\nRe = " << p.Re << "\nnu = " << p.nu << "\n"
       << "maxit = " << p.maxit << "\nmaxit_inner_u = " << p.maxit_inner_u
       << "\nmaxit_inner_v = " << p.maxit_inner_v << "\nmaxit_inner_p = " << p.maxit_inner_p
       << "\nerror_tolerance = " << p.error_tolerance
       << "\nrelaxation_factor_u = " << p.relaxation_factor_u
       << "\nrelaxation_factor_v = " << p.relaxation_factor_v
       << "\nrelaxation_factor_p = " << p.relaxation_factor_p
       << "\nconvection_scheme = " << scheme_name(p.convection_scheme)
       << "\ndiffusion_order = " << p.diffusion_order
       << "\nHybrid_switch = " << (p.Hybrid_switch ? "True" : "False")
       << "\nbackend = " << backend_name(p.backend)
       << "\nmomentum_solver = " << momentum_name(p.momentum_solver)
       << "\npressure_solver = " << pressure_name(p.pressure_solver)
       << "\nresidual_normalisation = "
       << (p.residual_normalisation == SG_NORM_ABSOLUTE ? "absolute" : "running_max")
       << "\noutput_dir = " << P.output_dir << "\npressure_tolerance = " << p.pressure_tolerance
       << "\npressure_maxit = " << p.pressure_maxit << "\nmg_cycles = " << p.mg_cycles
       << "\nmg_pre = " << p.mg_pre << "\nmg_post = " << p.mg_post << "\nmg_omega = " << p.mg_omega
       << "\nhalo_overlap = " << (P.halo_overlap ? "True" : "False")
       << "\nwrite_fields = " << (p.write_fields ? "True" : "False")
       << "\nfield_format = " << P.field_format
       << "\nmg_precision = " << (p.mg_precision == SG_MG_MIXED ? "mixed" : "fp64")
       << "\nmomentum_precision = " << (p.momentum_precision == SG_MOMP_MIXED ? "mixed" : "fp64")
       << "\nmg_coarsest = " << p.mg_coarsest
       << "\npressure_warm_start = " << (p.pressure_warm_start == SG_WARM_LEARNED ? "learned" : "none")
       << "\nautotune = " << (P.autotune ? "True" : "False") << "\n";
}

const char* phase_name(int k) {
    static const char* n[kPhases] = {"assemble_u", "bc_u",       "residual_u", "sweep_u",
                                     "assemble_v", "bc_v",       "residual_v", "sweep_v",
                                     "assemble_p", "solve_p",    "correct",    "output"};
    return (k >= 0 && k < kPhases) ? n[k] : "?";
}

void check(int rc) {
    if (rc != 0)
        throw std::runtime_error(sg_last_error());
}

} // namespace sh
