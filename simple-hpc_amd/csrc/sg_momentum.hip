// sg_momentum.hip — u- and v-momentum on gfx950: coefficient assembly, ghost cells, residual,
// and the multicolour Gauss-Seidel / Jacobi sweeps.
//
// Arithmetic is the reference's expression for expression (compiled with -ffp-contract=off, IEEE
// division), so every cell value is bit-identical to src/momentum_u.cpp / momentum_v.cpp. The
// residual keeps the CPU summation order (ordered row sums, sg_reduce.hip).
#include "sg_internal.hpp"

namespace sg {

namespace {

constexpr int BX = 64, BY = 4;

// std::max(0.0, a) of src/coefficients.cpp:18.
__device__ __forceinline__ double pos(double a) { return 0.0 < a ? a : 0.0; }

struct AsmArgs {
    const double* u;
    const double* v;
    double* c[9];
    int ld, H, NX, NY;
    int i0, i1, j0, j1;
    double dx, dy;
    int uds;
    int order;
    C9 d2, d4;
};

// One control volume's coefficients: convection (src/coefficients.cpp:15-39) plus diffusion
// (:41-63, precomputed on the host as d2 / d4), summed element-wise (:5-8).
template <bool IS_U>
__global__ void __launch_bounds__(BX* BY) k_assemble(AsmArgs a, const int* stop) {
    if (stopped(stop))
        return;
    const int j = a.j0 + blockIdx.x * BX + threadIdx.x;
    const int i = a.i0 + blockIdx.y * BY + threadIdx.y;
    if (i > a.i1 || j > a.j1)
        return;
    const int ld = a.ld;
    const size_t q = static_cast<size_t>(i) * ld + j;
    double ue, uw, vn, vs;
    if (IS_U) { // src/momentum_u.cpp:13-16
        ue = (a.u[q] + a.u[q + 1]) / 2;
        uw = (a.u[q] + a.u[q - 1]) / 2;
        vn = (a.v[q] + a.v[q + 1]) / 2;
        vs = (a.v[q + ld] + a.v[q + ld + 1]) / 2;
    } else { // src/momentum_v.cpp:14-17
        ue = (a.u[q - ld] + a.u[q]) / 2;
        uw = (a.u[q - 1] + a.u[q - ld - 1]) / 2;
        vn = (a.v[q] + a.v[q - ld]) / 2;
        vs = (a.v[q] + a.v[q + ld]) / 2;
    }
    const bool band = j == a.H + 1 || j == a.NY + a.H || i == a.NX + a.H - 1 || i == a.H;
    const double dx = a.dx, dy = a.dy;
    double e, w, n, s, ee, ww, nn, ss;
    if (a.uds || band) {
        e = dy * pos(-ue);
        w = dy * pos(uw);
        n = dx * pos(-vn);
        s = dx * pos(vs);
        ee = ww = nn = ss = 0.0;
    } else {
        e = -dy * (-0.75 * pos(-ue) + 0.375 * pos(ue) - 0.125 * pos(-uw));
        w = -dy * (-0.125 * pos(ue) + 0.375 * pos(-uw) - 0.75 * pos(uw));
        n = -dx * (-0.75 * pos(-vn) + 0.375 * pos(vn) - 0.125 * pos(-vs));
        s = -dx * (-0.125 * pos(vn) + 0.375 * pos(-vn) - 0.75 * pos(vs));
        ee = -0.125 * dy * pos(-ue);
        ww = -0.125 * dy * pos(uw);
        nn = -0.125 * dx * pos(-vn);
        ss = -0.125 * dx * pos(vs);
    }
    const double ap = s + w + n + e + ss + ww + nn + ee;
    const C9& d = (a.order == 2 || band) ? a.d2 : a.d4;
    a.c[0][q] = e + d.v[0];
    a.c[1][q] = w + d.v[1];
    a.c[2][q] = n + d.v[2];
    a.c[3][q] = s + d.v[3];
    a.c[4][q] = ee + d.v[4];
    a.c[5][q] = ww + d.v[5];
    a.c[6][q] = nn + d.v[6];
    a.c[7][q] = ss + d.v[7];
    a.c[8][q] = ap + d.v[8];
}

// Ghost cells: u lid 2 - u and south -u (src/momentum_u.cpp:24-34); v west/east -v
// (src/momentum_v.cpp:24-34).
__global__ void k_bc_u(double* u, int ld, int H, int NX, int NY, const int* stop) {
    if (stopped(stop))
        return;
    const int j = H + 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (j > NY + H)
        return;
    u[static_cast<size_t>(H - 1) * ld + j] = 2 - u[static_cast<size_t>(H) * ld + j];
    u[static_cast<size_t>(NX + H) * ld + j] = -u[static_cast<size_t>(NX + H - 1) * ld + j];
}

__global__ void k_bc_v(double* v, int ld, int H, int NX, int NY, const int* stop) {
    if (stopped(stop))
        return;
    const int i = H + blockIdx.x * blockDim.x + threadIdx.x;
    if (i > NX + H - 1)
        return;
    const size_t r = static_cast<size_t>(i) * ld;
    v[r + H] = -v[r + H + 1];
    v[r + NY + H + 1] = -v[r + NY + H];
}

struct MomArgs {
    double* x;
    const double* c[9];
    const double* pf;
    int ld;
    int i0, i1, j0, j1; // inclusive cell range
    double alpha, one_minus_alpha, h;
    long poff;          // offset of the second pressure cell (+1 east for u, -ld north for v)
};

// (E + W + N + S) + (EE + WW + NN + SS) (src/momentum_u.cpp:36-41).
__device__ __forceinline__ double nb9(const MomArgs& m, const double* x, size_t q) {
    const long ld = m.ld;
    return (m.c[0][q] * x[q + 1] + m.c[1][q] * x[q - 1] + m.c[2][q] * x[q - ld] +
            m.c[3][q] * x[q + ld]) +
           (m.c[4][q] * x[q + 2] + m.c[5][q] * x[q - 2] + m.c[6][q] * x[q - 2 * ld] +
            m.c[7][q] * x[q + 2 * ld]);
}
// 5-point sum of the 2-colour case (src/momentum_u.cpp:127-132).
__device__ __forceinline__ double nb5(const MomArgs& m, const double* x, size_t q) {
    const long ld = m.ld;
    return m.c[0][q] * x[q + 1] + m.c[1][q] * x[q - 1] + m.c[2][q] * x[q - ld] +
           m.c[3][q] * x[q + ld];
}

// |R| per cell, R = (nb + pterm) - AP x (src/momentum_u.cpp:48-50), into a compact buffer.
__global__ void __launch_bounds__(BX* BY) k_mom_residual(MomArgs m, double* cell, const int* stop) {
    if (stopped(stop))
        return;
    const int j = m.j0 + blockIdx.x * BX + threadIdx.x;
    const int i = m.i0 + blockIdx.y * BY + threadIdx.y;
    if (i > m.i1 || j > m.j1)
        return;
    const size_t q = static_cast<size_t>(i) * m.ld + j;
    const double R = nb9(m, m.x, q) + (m.pf[q] - m.pf[q + m.poff]) * m.h - m.c[8][q] * m.x[q];
    cell[static_cast<size_t>(i - m.i0) * (m.j1 - m.j0 + 1) + (j - m.j0)] = fabs(R);
}

// One colour pass of the multicolour sweep (src/momentum_u.cpp:130-138): thread (k, row) updates
// the k-th cell of colour `col` in its row. Cells of one colour are never stencil neighbours, so
// the in-place update is race free and equals the CPU's.
template <int NCOL>
__global__ void __launch_bounds__(BX* BY) k_mom_pass(MomArgs m, int col, const int* stop) {
    if (stopped(stop))
        return;
    const int i = m.i0 + blockIdx.y * BY + threadIdx.y;
    if (i > m.i1)
        return;
    int j = m.j0;
    if (NCOL == 2)
        j += ((i + j) & 1) != col;
    else
        while ((i + 3 * j) % 5 != col)
            ++j;
    j += NCOL * (blockIdx.x * BX + threadIdx.x);
    if (j > m.j1)
        return;
    const size_t q = static_cast<size_t>(i) * m.ld + j;
    const double nb = NCOL == 2 ? nb5(m, m.x, q) : nb9(m, m.x, q);
    const double pterm = (m.pf[q] - m.pf[q + m.poff]) * m.h;
    m.x[q] = m.alpha * (nb + pterm) / m.c[8][q] + m.one_minus_alpha * m.x[q];
}

// One Jacobi pass, dst from src (src/momentum_u.cpp:105-113).
__global__ void __launch_bounds__(BX* BY)
    k_mom_jacobi(MomArgs m, const double* src, double* dst, const int* stop) {
    if (stopped(stop))
        return;
    const int j = m.j0 + blockIdx.x * BX + threadIdx.x;
    const int i = m.i0 + blockIdx.y * BY + threadIdx.y;
    if (i > m.i1 || j > m.j1)
        return;
    const size_t q = static_cast<size_t>(i) * m.ld + j;
    const double pterm = (m.pf[q] - m.pf[q + m.poff]) * m.h;
    dst[q] = m.alpha * (nb9(m, src, q) + pterm) / m.c[8][q] + m.one_minus_alpha * src[q];
}

__global__ void k_copy(const double* a, double* b, size_t n, const int* stop) {
    if (stopped(stop))
        return;
    for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; k < n;
         k += static_cast<size_t>(gridDim.x) * blockDim.x)
        b[k] = a[k];
}

MomArgs mom_args(Solver& s, bool is_u) {
    const sg_params& p = s.p;
    MomArgs m{};
    const int id0 = is_u ? SG_F_CU_AE : SG_F_CV_AE;
    m.x = s.F(is_u ? SG_F_U : SG_F_V);
    for (int k = 0; k < 9; ++k)
        m.c[k] = s.F(id0 + k);
    m.pf = s.F(SG_F_P);
    m.ld = s.ld;
    const int H = p.H;
    if (is_u) { // rows [H, NX+H-1], columns [H+1, NY+H-1], pterm (p(i,j) - p(i,j+1)) dy
        m.i0 = H, m.i1 = p.NX + H - 1, m.j0 = H + 1, m.j1 = p.NY + H - 1;
        m.alpha = p.relaxation_factor_u;
        m.h = p.dy;
        m.poff = 1;
    } else { // rows [H+1, NX+H-1], columns [H+1, NY+H], pterm (p(i,j) - p(i-1,j)) dx
        m.i0 = H + 1, m.i1 = p.NX + H - 1, m.j0 = H + 1, m.j1 = p.NY + H;
        m.alpha = p.relaxation_factor_v;
        m.h = p.dx;
        m.poff = -static_cast<long>(s.ld);
    }
    m.one_minus_alpha = 1 - m.alpha;
    return m;
}

int ncolours(const sg_params& p) {
    return (p.convection_scheme == SG_UDS && p.diffusion_order == 2) ? 2 : 5;
}

} // namespace

void launch_assemble(Solver& s, bool is_u, const int* stop) {
    const sg_params& p = s.p;
    AsmArgs a{};
    a.u = s.F(SG_F_U);
    a.v = s.F(SG_F_V);
    const int id0 = is_u ? SG_F_CU_AE : SG_F_CV_AE;
    for (int k = 0; k < 9; ++k)
        a.c[k] = s.F(id0 + k);
    a.ld = s.ld, a.H = p.H, a.NX = p.NX, a.NY = p.NY;
    a.i0 = p.H;
    a.j0 = p.H;
    a.i1 = is_u ? p.NX + p.H - 1 : p.NX + p.H + 1; // src/momentum_u.cpp:11, momentum_v.cpp:12
    a.j1 = is_u ? p.NY + p.H : p.NY + p.H + 1;
    a.dx = p.dx, a.dy = p.dy;
    a.uds = p.convection_scheme == SG_UDS;
    a.order = p.diffusion_order;
    a.d2 = s.diff2;
    a.d4 = s.diff4;
    const dim3 grid(cdiv(a.j1 - a.j0 + 1, BX), cdiv(a.i1 - a.i0 + 1, BY));
    if (is_u)
        hipLaunchKernelGGL(k_assemble<true>, grid, dim3(BX, BY), 0, s.stream, a, stop);
    else
        hipLaunchKernelGGL(k_assemble<false>, grid, dim3(BX, BY), 0, s.stream, a, stop);
    SG_HIP(hipGetLastError());
}

void launch_bc(Solver& s, bool is_u, const int* stop) {
    const sg_params& p = s.p;
    if (is_u)
        hipLaunchKernelGGL(k_bc_u, dim3(cdiv(p.NY, 256)), dim3(256), 0, s.stream, s.F(SG_F_U),
                           s.ld, p.H, p.NX, p.NY, stop);
    else
        hipLaunchKernelGGL(k_bc_v, dim3(cdiv(p.NX, 256)), dim3(256), 0, s.stream, s.F(SG_F_V),
                           s.ld, p.H, p.NX, p.NY, stop);
    SG_HIP(hipGetLastError());
}

void launch_mom_residual(Solver& s, bool is_u, double* slot, const int* stop) {
    const MomArgs m = mom_args(s, is_u);
    const int nrow = m.i1 - m.i0 + 1, w = m.j1 - m.j0 + 1;
    if (nrow <= 0 || w <= 0) {
        launch_ordered_sum(s, s.cell, 1, 0, -1, 0, -1, slot, stop); // empty range: 0.0
        return;
    }
    hipLaunchKernelGGL(k_mom_residual, dim3(cdiv(w, BX), cdiv(nrow, BY)), dim3(BX, BY), 0,
                       s.stream, m, s.cell, stop);
    SG_HIP(hipGetLastError());
    launch_ordered_sum(s, s.cell, w, 0, nrow - 1, 0, w - 1, slot, stop);
}

void launch_mom_colour_pass(Solver& s, bool is_u, int col, const int* stop) {
    const MomArgs m = mom_args(s, is_u);
    const int ncol = ncolours(s.p);
    const int nrow = m.i1 - m.i0 + 1, w = m.j1 - m.j0 + 1;
    if (nrow <= 0 || w <= 0)
        return;
    const dim3 grid(cdiv(cdiv(w, ncol), BX), cdiv(nrow, BY));
    if (ncol == 2)
        hipLaunchKernelGGL(k_mom_pass<2>, grid, dim3(BX, BY), 0, s.stream, m, col, stop);
    else
        hipLaunchKernelGGL(k_mom_pass<5>, grid, dim3(BX, BY), 0, s.stream, m, col, stop);
    SG_HIP(hipGetLastError());
}

// sweep_u_batch / sweep_v_batch (src/momentum_u.cpp:164-171) for the two parallel solvers.
void launch_mom_sweeps(Solver& s, bool is_u, int nsweeps, const int* stop) {
    const sg_params& p = s.p;
    if (nsweeps <= 0)
        return;
    if (p.momentum_solver == SG_MOM_GS_MULTICOLOUR) {
        const int ncol = ncolours(p);
        for (int k = 0; k < nsweeps; ++k)
            for (int c = 0; c < ncol; ++c)
                launch_mom_colour_pass(s, is_u, c, stop);
        return;
    }
    // Jacobi batch: full copy, then alternate so that the last pass writes the field.
    const MomArgs m = mom_args(s, is_u);
    const int nrow = m.i1 - m.i0 + 1, w = m.j1 - m.j0 + 1;
    hipLaunchKernelGGL(k_copy, dim3(1024), dim3(256), 0, s.stream, m.x, s.scratch, s.n, stop);
    SG_HIP(hipGetLastError());
    if (nrow <= 0 || w <= 0)
        return;
    for (int k = 0; k < nsweeps; ++k) {
        const bool to_x = (nsweeps - k) % 2 == 1;
        const double* src = to_x ? s.scratch : m.x;
        double* dst = to_x ? m.x : s.scratch;
        hipLaunchKernelGGL(k_mom_jacobi, dim3(cdiv(w, BX), cdiv(nrow, BY)), dim3(BX, BY), 0,
                           s.stream, m, src, dst, stop);
        SG_HIP(hipGetLastError());
    }
}

} // namespace sg
