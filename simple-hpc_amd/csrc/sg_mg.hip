// sg_mg.hip — geometric multigrid for p' on gfx950 (src/multigrid.cpp): Galerkin 2x2
// aggregation, red-black smoother, residual, restriction, prolongation with over-correction,
// and the coarsest-level solve in one workgroup. FP64 arithmetic is the reference's expression
// for expression and every sum keeps its fixed order, so p' is bit-identical to the CPU.
#include "sg_internal.hpp"

#include <algorithm>

namespace sg {

namespace {

constexpr int BX = 64, BY = 4;

struct Lv {
    int nx, ny, stride;
    double *AE, *AW, *AN, *AS, *AP, *b, *x, *r;
};

Lv lv(const MgLevel& L) {
    return {L.nx, L.ny, L.stride, L.AE, L.AW, L.AN, L.AS, L.AP, L.b, L.x, L.r};
}

// mg_build (src/multigrid.cpp:232-261): coarse links are the sums of the two fine links crossing
// each coarse face; AP = ae + aw + an + as.
__global__ void __launch_bounds__(BX* BY) k_mg_build(Lv F, Lv C, const int* stop) {
    if (stopped(stop))
        return;
    const int J = blockIdx.x * BX + threadIdx.x, I = blockIdx.y * BY + threadIdx.y;
    if (I >= C.nx || J >= C.ny)
        return;
    const long fs = F.stride;
    const size_t f = 2 * I * fs + 2 * J, c = static_cast<size_t>(I) * C.stride + J;
    const double ae = F.AE[f + 1] + F.AE[f + fs + 1];
    const double aw = F.AW[f] + F.AW[f + fs];
    const double an = F.AN[f] + F.AN[f + 1];
    const double as = F.AS[f + fs] + F.AS[f + fs + 1];
    C.AE[c] = ae;
    C.AW[c] = aw;
    C.AN[c] = an;
    C.AS[c] = as;
    C.AP[c] = ae + aw + an + as;
}

// Two-statement smoother update (src/multigrid.cpp:62-64).
__device__ __forceinline__ void smooth_cell(const Lv& L, size_t q) {
    const long s = L.stride;
    double* x = L.x;
    double t = L.b[q] / L.AP[q];
    t += (L.AE[q] * x[q + 1] + L.AW[q] * x[q - 1] + L.AS[q] * x[q + s] + L.AN[q] * x[q - s]) /
         L.AP[q];
    x[q] = t;
}

__device__ __forceinline__ double resid_cell(const Lv& L, size_t q) {
    const long s = L.stride;
    const double* x = L.x;
    return L.b[q] - (L.AP[q] * x[q] -
                     (L.AE[q] * x[q + 1] + L.AW[q] * x[q - 1] + L.AS[q] * x[q + s] +
                      L.AN[q] * x[q - s]));
}

// One colour of the red-black smoother, colour (I + J) & 1 (src/multigrid.cpp:46-70).
__global__ void __launch_bounds__(BX* BY) k_mg_smooth(Lv L, int col, const int* stop) {
    if (stopped(stop))
        return;
    const int I = blockIdx.y * BY + threadIdx.y;
    const int J = ((I + col) & 1) + 2 * (blockIdx.x * BX + threadIdx.x);
    if (I >= L.nx || J >= L.ny)
        return;
    smooth_cell(L, static_cast<size_t>(I) * L.stride + J);
}

// r = b - A x (src/multigrid.cpp:73-87).
__global__ void __launch_bounds__(BX* BY) k_mg_resid(Lv L, const int* stop) {
    if (stopped(stop))
        return;
    const int J = blockIdx.x * BX + threadIdx.x, I = blockIdx.y * BY + threadIdx.y;
    if (I >= L.nx || J >= L.ny)
        return;
    const size_t q = static_cast<size_t>(I) * L.stride + J;
    L.r[q] = resid_cell(L, q);
}

// Coarse b = sum of the 4 fine residuals, coarse x = 0 (src/multigrid.cpp:90-103).
__global__ void __launch_bounds__(BX* BY) k_mg_restrict(Lv F, Lv C, const int* stop) {
    if (stopped(stop))
        return;
    const int J = blockIdx.x * BX + threadIdx.x, I = blockIdx.y * BY + threadIdx.y;
    if (I >= C.nx || J >= C.ny)
        return;
    const long fs = F.stride;
    const size_t f = 2 * I * fs + 2 * J, c = static_cast<size_t>(I) * C.stride + J;
    C.b[c] = F.r[f] + F.r[f + 1] + F.r[f + fs] + F.r[f + fs + 1];
    C.x[c] = 0.0;
}

// Fine x += omega * coarse x (src/multigrid.cpp:106-116).
__global__ void __launch_bounds__(BX* BY) k_mg_prolong(Lv C, Lv F, double omega, const int* stop) {
    if (stopped(stop))
        return;
    const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
    if (i >= F.nx || j >= F.ny)
        return;
    F.x[static_cast<size_t>(i) * F.stride + j] +=
        omega * C.x[static_cast<size_t>(i / 2) * C.stride + j / 2];
}

// coarse_solve (src/multigrid.cpp:120-140) in one workgroup: mean-free b, red-black sweeps until
// ||r|| < 1e-10 ||b|| or 200 sweeps, mean-free x. det_sum order: each row by one thread (J
// ascending from 0.0), then the rows by thread 0 (I ascending from 0.0).
constexpr int CT = 1024;

__device__ double wg_det_sum(const Lv& L, double* rows, int which) {
    __shared__ double total;
    const int t = threadIdx.x;
    for (int I = t; I < L.nx; I += CT) {
        double s = 0.0;
        const size_t r0 = static_cast<size_t>(I) * L.stride;
        for (int J = 0; J < L.ny; ++J) {
            const size_t q = r0 + J;
            const double v = which == 0 ? L.b[q] : which == 1 ? L.b[q] * L.b[q]
                                         : which == 2 ? L.r[q] * L.r[q] : L.x[q];
            s += v;
        }
        rows[I] = s;
    }
    __syncthreads();
    if (t == 0) {
        double s = 0.0;
        for (int I = 0; I < L.nx; ++I)
            s += rows[I];
        total = s;
    }
    __syncthreads();
    const double v = total;
    __syncthreads();
    return v;
}

__device__ void wg_sub(const Lv& L, double* x, double m) {
    const int n = L.nx * L.ny;
    for (int k = threadIdx.x; k < n; k += CT) {
        const int I = k / L.ny, J = k - I * L.ny;
        x[static_cast<size_t>(I) * L.stride + J] -= m;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(CT) k_mg_coarse(Lv L, double* rows, Ctl* ctl, const int* stop) {
    if (stopped(stop))
        return;
    const double ncell = static_cast<double>(L.nx) * L.ny;
    wg_sub(L, L.b, wg_det_sum(L, rows, 0) / ncell);
    const double bb = wg_det_sum(L, rows, 1);
    int sweeps = 0;
    if (bb > 0.0) {
        const double bn = sqrt(bb);
        const int half = (L.ny + 1) / 2;
        while (sweeps < 200) {
            for (int c = 0; c < 2; ++c) {
                for (int k = threadIdx.x; k < L.nx * half; k += CT) {
                    const int I = k / half;
                    const int J = ((I + c) & 1) + 2 * (k - I * half);
                    if (J < L.ny)
                        smooth_cell(L, static_cast<size_t>(I) * L.stride + J);
                }
                __syncthreads();
            }
            ++sweeps;
            for (int k = threadIdx.x; k < L.nx * L.ny; k += CT) {
                const int I = k / L.ny, J = k - I * L.ny;
                const size_t q = static_cast<size_t>(I) * L.stride + J;
                L.r[q] = resid_cell(L, q);
            }
            __syncthreads();
            const double rr = wg_det_sum(L, rows, 2);
            if (sqrt(rr) < 1e-10 * bn)
                break;
        }
    }
    wg_sub(L, L.x, wg_det_sum(L, rows, 3) / ncell);
    if (threadIdx.x == 0)
        ctl->coarse_sweeps = sweeps;
}

dim3 lgrid(int nx, int ny) { return dim3(cdiv(ny, BX), cdiv(nx, BY)); }

} // namespace

// Levels (src/multigrid.cpp:143-185): halve while both sides are even and the halves stay at
// least mg_coarsest (4 in the reference). Level 0 aliases the p' arrays.
void mg_setup(Solver& s) {
    const sg_params& p = s.p;
    const int coarsest = p.mg_coarsest > 0 ? p.mg_coarsest : 4;
    std::vector<std::pair<int, int>> dims{{p.NX, p.NY}};
    int nx = p.NX, ny = p.NY;
    while (nx % 2 == 0 && ny % 2 == 0 && nx / 2 >= coarsest && ny / 2 >= coarsest) {
        nx /= 2;
        ny /= 2;
        dims.emplace_back(nx, ny);
    }
    s.mg.assign(dims.size(), MgLevel{});
    SG_HIP(hipMalloc(reinterpret_cast<void**>(&s.mg_r0), sizeof(double) * s.n));
    s.owned.push_back(s.mg_r0);
    SG_HIP(hipMemset(s.mg_r0, 0, sizeof(double) * s.n));
    const size_t base0 = static_cast<size_t>(p.H) * s.ld + p.H + 1;
    for (size_t l = 0; l < dims.size(); ++l) {
        MgLevel& L = s.mg[l];
        L.nx = dims[l].first;
        L.ny = dims[l].second;
        if (l == 0) {
            L.stride = s.ld;
            L.AE = s.F(SG_F_AEP) + base0;
            L.AW = s.F(SG_F_AWP) + base0;
            L.AN = s.F(SG_F_ANP) + base0;
            L.AS = s.F(SG_F_ASP) + base0;
            L.AP = s.F(SG_F_APP) + base0;
            L.b = s.F(SG_F_B) + base0;
            L.x = s.F(SG_F_PC) + base0;
            L.r = s.mg_r0 + base0;
            continue;
        }
        L.stride = L.ny + 2;
        const size_t n = static_cast<size_t>(L.nx + 2) * (L.ny + 2);
        double* blk = nullptr;
        SG_HIP(hipMalloc(reinterpret_cast<void**>(&blk), sizeof(double) * n * 8));
        SG_HIP(hipMemset(blk, 0, sizeof(double) * n * 8));
        s.owned.push_back(blk);
        double** ptr[8] = {&L.AE, &L.AW, &L.AN, &L.AS, &L.AP, &L.b, &L.x, &L.r};
        for (int k = 0; k < 8; ++k)
            *ptr[k] = blk + k * n + L.stride + 1;
    }
}

// solve_pc_mg (src/multigrid.cpp:347-357).
void launch_mg_solve(Solver& s, const int* stop) {
    const sg_params& p = s.p;
    const int nl = static_cast<int>(s.mg.size());
    for (int l = 1; l < nl; ++l) {
        const Lv F = lv(s.mg[l - 1]), C = lv(s.mg[l]);
        hipLaunchKernelGGL(k_mg_build, lgrid(C.nx, C.ny), dim3(BX, BY), 0, s.stream, F, C, stop);
        SG_HIP(hipGetLastError());
    }
    launch_fill(s, s.F(SG_F_PC), s.n, 0.0, stop);
    launch_remove_mean0(s, s.F(SG_F_B), stop);
    for (int cyc = 0; cyc < p.mg_cycles; ++cyc) {
        // mg_vcycle (src/multigrid.cpp:299-323)
        for (int l = 0; l + 1 < nl; ++l) {
            const Lv F = lv(s.mg[l]), C = lv(s.mg[l + 1]);
            for (int k = 0; k < p.mg_pre; ++k)
                for (int c = 0; c < 2; ++c) {
                    hipLaunchKernelGGL(k_mg_smooth, lgrid(F.nx, cdiv(F.ny, 2)), dim3(BX, BY), 0,
                                       s.stream, F, c, stop);
                    SG_HIP(hipGetLastError());
                }
            hipLaunchKernelGGL(k_mg_resid, lgrid(F.nx, F.ny), dim3(BX, BY), 0, s.stream, F, stop);
            SG_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_mg_restrict, lgrid(C.nx, C.ny), dim3(BX, BY), 0, s.stream, F, C,
                               stop);
            SG_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(CT), 0, s.stream, lv(s.mg[nl - 1]),
                           s.rowbuf, s.ctl, stop);
        SG_HIP(hipGetLastError());
        for (int l = nl - 2; l >= 0; --l) {
            const Lv F = lv(s.mg[l]), C = lv(s.mg[l + 1]);
            hipLaunchKernelGGL(k_mg_prolong, lgrid(F.nx, F.ny), dim3(BX, BY), 0, s.stream, C, F,
                               p.mg_omega, stop);
            SG_HIP(hipGetLastError());
            for (int k = 0; k < p.mg_post; ++k)
                for (int c = 0; c < 2; ++c) {
                    hipLaunchKernelGGL(k_mg_smooth, lgrid(F.nx, cdiv(F.ny, 2)), dim3(BX, BY), 0,
                                       s.stream, F, c, stop);
                    SG_HIP(hipGetLastError());
                }
        }
    }
    launch_remove_mean0(s, s.F(SG_F_PC), stop);
}

} // namespace sg
