// `simple <config>`: the reference's command line (src/main.cpp:18-89) for backend = gpu.
// Errors print `error (rank 0): <message>` and return 1.
#include "host.hpp"

#include <iostream>

int main(int argc, char** argv) {
    if (argc != 2) {
        std::cerr << "usage: simple <config>\n";
        return 1;
    }
    try {
        const sh::Params p = sh::make_params(sh::load_config(argv[1]));
        if (p.g.backend != SG_BACKEND_GPU)
            throw std::runtime_error(std::string("backend = ") + sh::backend_name(p.g.backend) +
                                     ": this build runs backend = gpu only");
        if (p.mpi_grid != "auto" && p.mpi_grid != "1x1")
            throw std::runtime_error("mpi_grid = " + p.mpi_grid + ": this build runs on one GPU");
        sh::print_params(std::cout, p);
        const size_t n = static_cast<size_t>(p.g.nrows) * p.g.ncols;
        std::vector<double> u(n, 0.0), v(n, 0.0), pf(n, 0.0);
        sh::RunResult r = sh::run_simple_gpu(p, u, v, pf);
        sh::write_outputs(p, u, v, pf, r, argv[1]);
        std::cout << (r.converged ? "converged" : "not converged") << " after " << r.iterations
                  << " iterations in " << r.seconds << " s\n";
    } catch (const std::exception& e) {
        std::cerr << "error (rank 0): " << e.what() << "\n";
        return 1;
    }
    return 0;
}
