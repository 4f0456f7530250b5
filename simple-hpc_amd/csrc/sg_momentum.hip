// sg_momentum.hip — u- and v-momentum on gfx950: coefficient assembly, ghost cells, residual,
// and the multicolour Gauss-Seidel / Jacobi sweeps.
//
// Layout. A colour pass of the multicolour sweep updates every ncol-th cell of a row (colour
// (i + 3j) mod 5, or (i + j) mod 2), so reading the coefficients from row-major arrays would
// stream whole cache lines for a fifth of their bytes. The assembly therefore writes the eight
// neighbour coefficients, AP and the pressure term of every cell colour-packed: array a, colour
// c, row i, k-th cell of that colour in the row at
//     pk.base + a * astride + (c * rows + (i - ai0)) * kp + k,
// so a pass reads 10 contiguous streams. AP is also kept row-major (assemble_p and correct read
// it at neighbouring cells). The row-major copies of the other coefficients are refreshed only
// when the host downloads them.
//
// The field itself (u or v) is also colour-packed while it is swept, over the whole array
// (rows [0, nrows), columns [0, ncols)): the prep kernel packs it (and evaluates the residual from
// the row-major copy in the same pass), the colour passes read and write only packed data (the
// neighbours of a colour-c cell sit at the same or an adjacent packed index of another colour's
// rows), and an unpack kernel restores the row-major field after the last pass.
//
// Arithmetic is the reference's expression for expression (compiled with -ffp-contract=off, IEEE
// division), so every value is bit-identical to src/momentum_u.cpp / momentum_v.cpp; the
// pressure term (p(i,j) - p(i,j+1)) * dy is evaluated once, at assembly, with the same operands.
#include "sg_internal.hpp"

namespace sg {

namespace {

constexpr int BX = 64, BY = 4;

__device__ __forceinline__ double pos(double a) { return 0.0 < a ? a : 0.0; } // std::max(0.0, a)

__device__ __forceinline__ int colour_of(int ncol, int i, int j) {
    return ncol == 2 ? (i + j) & 1 : (i + 3 * j) % 5;
}

// First column >= aj0 of colour c in row i.
__device__ __forceinline__ int jfirst(const Pack& k, int c, int i) {
    const int c0 = colour_of(k.ncol, i, k.aj0);
    const int t = k.ncol == 2 ? ((c - c0) & 1) : (2 * (c - c0) + 10) % 5;
    return k.aj0 + t;
}

__device__ __forceinline__ size_t pidx(const Pack& k, int c, int i, int kk) {
    return static_cast<size_t>(c * k.rows + (i - k.ai0)) * k.kp + kk;
}

struct AsmArgs {
    const double* u;
    const double* v;
    const double* pf;
    double* APn; // row-major AP
    Pack pk;
    int ld, H, NX, NY;
    double dx, dy;
    double h;    // face length of the pressure term (dy for u, dx for v)
    long poff;   // second pressure cell (+1 for u, -ld for v)
    int uds, order;
    C9 d2, d4;
};

// One control volume: convection (src/coefficients.cpp:15-39) plus diffusion (:41-63,
// precomputed on the host as d2 / d4), summed element-wise (:5-8); face velocities of
// src/momentum_u.cpp:13-16 or src/momentum_v.cpp:14-17.
template <bool IS_U>
__global__ void __launch_bounds__(BX* BY) k_assemble(AsmArgs a, const int* stop) {
    if (stopped(stop))
        return;
    const Pack& pk = a.pk;
    const int j = pk.aj0 + blockIdx.x * BX + threadIdx.x;
    const int i = pk.ai0 + blockIdx.y * BY + threadIdx.y;
    if (i > pk.ai1 || j > pk.aj1)
        return;
    const long ld = a.ld;
    const size_t q = static_cast<size_t>(i) * ld + j;
    double ue, uw, vn, vs;
    if (IS_U) {
        ue = (a.u[q] + a.u[q + 1]) / 2;
        uw = (a.u[q] + a.u[q - 1]) / 2;
        vn = (a.v[q] + a.v[q + 1]) / 2;
        vs = (a.v[q + ld] + a.v[q + ld + 1]) / 2;
    } else {
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
    const int c = colour_of(pk.ncol, i, j);
    const size_t o = pidx(pk, c, i, (j - jfirst(pk, c, i)) / pk.ncol);
    double* b = pk.base + o;
    const size_t as = pk.astride;
    b[0 * as] = e + d.v[0];
    b[1 * as] = w + d.v[1];
    b[2 * as] = n + d.v[2];
    b[3 * as] = s + d.v[3];
    b[4 * as] = ee + d.v[4];
    b[5 * as] = ww + d.v[5];
    b[6 * as] = nn + d.v[6];
    b[7 * as] = ss + d.v[7];
    const double app = ap + d.v[8];
    b[8 * as] = app;
    a.APn[q] = app;
    b[9 * as] = (a.pf[q] - a.pf[q + a.poff]) * a.h;
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
    Pack pk;
    int ld;
    int i0, i1, j0, j1; // sweep / residual range
    double alpha, one_minus_alpha;
};

// (E + W + N + S) + (EE + WW + NN + SS) (src/momentum_u.cpp:36-41) with packed coefficients.
__device__ __forceinline__ double nb9(const double* c, size_t as, const double* x, size_t q,
                                      long ld) {
    return (c[0] * x[q + 1] + c[as] * x[q - 1] + c[2 * as] * x[q - ld] + c[3 * as] * x[q + ld]) +
           (c[4 * as] * x[q + 2] + c[5 * as] * x[q - 2] + c[6 * as] * x[q - 2 * ld] +
            c[7 * as] * x[q + 2 * ld]);
}
// Cell (i, j) of colour c handled by this thread in the sweep range, or false; o = its
// coefficient-pack index.
__device__ __forceinline__ bool packed_cell(const MomArgs& m, int c, int& i, int& j, size_t& o) {
    i = m.i0 + blockIdx.y * BY + threadIdx.y;
    if (i > m.i1)
        return false;
    const int jf = jfirst(m.pk, c, i);
    const int k = (jf < m.j0 ? 1 : 0) + blockIdx.x * BX + threadIdx.x;
    j = jf + m.pk.ncol * k;
    if (j > m.j1)
        return false;
    o = pidx(m.pk, c, i, k);
    return true;
}

// Packed index of any cell (i, j) in the whole-array field pack fp.
__device__ __forceinline__ size_t fidx(const Pack& fp, int i, int j) {
    const int c = colour_of(fp.ncol, i, j);
    return pidx(fp, c, i, (j - jfirst(fp, c, i)) / fp.ncol);
}

// Prep, one thread per cell of the whole array: pack the field (row-major -> colour-packed) and,
// inside the residual range, |R| with R = (nb + pterm) - AP x (src/momentum_u.cpp:48-50) from the
// row-major field, into a compact buffer for the ordered sum.
__global__ void __launch_bounds__(BX* BY)
    k_mom_prep(MomArgs m, Pack fp, double* up, double* cell, int nr, int nc, int want_res,
               const int* stop) {
    if (stopped(stop))
        return;
    const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
    if (i >= nr || j >= nc)
        return;
    const long ld = m.ld;
    const size_t q = static_cast<size_t>(i) * ld + j;
    up[fidx(fp, i, j)] = m.x[q];
    if (!want_res || i < m.i0 || i > m.i1 || j < m.j0 || j > m.j1)
        return;
    const Pack& pk = m.pk;
    const int c = colour_of(pk.ncol, i, j);
    const double* co = pk.base + pidx(pk, c, i, (j - jfirst(pk, c, i)) / pk.ncol);
    const double R = nb9(co, pk.astride, m.x, q, ld) + co[9 * pk.astride] -
                     co[8 * pk.astride] * m.x[q];
    cell[static_cast<size_t>(i - m.i0) * (m.j1 - m.j0 + 1) + (j - m.j0)] = fabs(R);
}

// Packed field -> row-major field over the whole array.
__global__ void __launch_bounds__(BX* BY)
    k_field_unpack(Pack fp, const double* up, double* x, int ld, int nr, int nc, const int* stop) {
    if (stopped(stop))
        return;
    const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
    if (i < nr && j < nc)
        x[static_cast<size_t>(i) * ld + j] = up[fidx(fp, i, j)];
}

// One colour pass of the multicolour sweep (src/momentum_u.cpp:130-138) on the packed field.
// Cells of one colour are never stencil neighbours, so the in-place update is race free and
// equals the CPU's. Neighbour sums keep the order of nb_sum / near_sum.
template <int NCOL>
__global__ void __launch_bounds__(BX* BY)
    k_mom_pass(MomArgs m, Pack fp, double* up, int col, const int* stop) {
    if (stopped(stop))
        return;
    int i, j;
    size_t o;
    if (!packed_cell(m, col, i, j, o))
        return;
    const size_t as = m.pk.astride;
    const double* c = m.pk.base + o;
    const size_t self = fidx(fp, i, j);
    double nb = c[0] * up[fidx(fp, i, j + 1)] + c[as] * up[fidx(fp, i, j - 1)] +
                c[2 * as] * up[fidx(fp, i - 1, j)] + c[3 * as] * up[fidx(fp, i + 1, j)];
    if (NCOL != 2)
        nb = nb + (c[4 * as] * up[fidx(fp, i, j + 2)] + c[5 * as] * up[fidx(fp, i, j - 2)] +
                   c[6 * as] * up[fidx(fp, i - 2, j)] + c[7 * as] * up[fidx(fp, i + 2, j)]);
    up[self] = m.alpha * (nb + c[9 * as]) / c[8 * as] + m.one_minus_alpha * up[self];
}

// One Jacobi pass, dst from src (src/momentum_u.cpp:105-113); colour = blockIdx.z.
__global__ void __launch_bounds__(BX* BY)
    k_mom_jacobi(MomArgs m, const double* src, double* dst, const int* stop) {
    if (stopped(stop))
        return;
    int i, j;
    size_t o;
    if (!packed_cell(m, blockIdx.z, i, j, o))
        return;
    const long ld = m.ld;
    const size_t q = static_cast<size_t>(i) * ld + j, as = m.pk.astride;
    const double* c = m.pk.base + o;
    dst[q] = m.alpha * (nb9(c, as, src, q, ld) + c[9 * as]) / c[8 * as] +
             m.one_minus_alpha * src[q];
}

__global__ void k_copy(const double* a, double* b, size_t n, const int* stop) {
    if (stopped(stop))
        return;
    for (size_t k = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; k < n;
         k += static_cast<size_t>(gridDim.x) * blockDim.x)
        b[k] = a[k];
}

// Row-major <-> packed copies of coefficient array a (0..8) over the assembly range.
__global__ void __launch_bounds__(BX* BY) k_pack(Pack pk, double* nat, int ld, int a, int to_packed) {
    const int j = pk.aj0 + blockIdx.x * BX + threadIdx.x;
    const int i = pk.ai0 + blockIdx.y * BY + threadIdx.y;
    if (i > pk.ai1 || j > pk.aj1)
        return;
    const int c = colour_of(pk.ncol, i, j);
    double* p = pk.base + a * pk.astride + pidx(pk, c, i, (j - jfirst(pk, c, i)) / pk.ncol);
    double* n = nat + static_cast<size_t>(i) * ld + j;
    if (to_packed)
        *p = *n;
    else
        *n = *p;
}

// Pressure term of every cell of the assembly range from the current p.
__global__ void __launch_bounds__(BX* BY) k_pterm(Pack pk, const double* pf, int ld, long poff, double h) {
    const int j = pk.aj0 + blockIdx.x * BX + threadIdx.x;
    const int i = pk.ai0 + blockIdx.y * BY + threadIdx.y;
    if (i > pk.ai1 || j > pk.aj1)
        return;
    const int c = colour_of(pk.ncol, i, j);
    const size_t q = static_cast<size_t>(i) * ld + j;
    pk.base[9 * pk.astride + pidx(pk, c, i, (j - jfirst(pk, c, i)) / pk.ncol)] =
        (pf[q] - pf[q + poff]) * h;
}

MomArgs mom_args(Solver& s, bool is_u) {
    const sg_params& p = s.p;
    MomArgs m{};
    m.x = s.F(is_u ? SG_F_U : SG_F_V);
    m.pk = is_u ? s.pku : s.pkv;
    m.ld = s.ld;
    const int H = p.H;
    if (is_u) { // rows [H, NX+H-1], columns [H+1, NY+H-1]
        m.i0 = H, m.i1 = p.NX + H - 1, m.j0 = H + 1, m.j1 = p.NY + H - 1;
        m.alpha = p.relaxation_factor_u;
    } else { // rows [H+1, NX+H-1], columns [H+1, NY+H]
        m.i0 = H + 1, m.i1 = p.NX + H - 1, m.j0 = H + 1, m.j1 = p.NY + H;
        m.alpha = p.relaxation_factor_v;
    }
    m.one_minus_alpha = 1 - m.alpha;
    return m;
}

dim3 pass_grid(const MomArgs& m, int nz) {
    return dim3(cdiv(m.pk.kp, BX), cdiv(m.i1 - m.i0 + 1, BY), nz);
}

} // namespace

int ncolours(const sg_params& p) {
    return (p.convection_scheme == SG_UDS && p.diffusion_order == 2) ? 2 : 5;
}

// Packed-layout geometry over the assembly range (src/momentum_u.cpp:11-12, momentum_v.cpp:12-13).
Pack make_pack(const sg_params& p, bool is_u) {
    Pack k{};
    k.ncol = ncolours(p);
    k.ai0 = p.H;
    k.aj0 = p.H;
    k.ai1 = is_u ? p.NX + p.H - 1 : p.NX + p.H + 1;
    k.aj1 = is_u ? p.NY + p.H : p.NY + p.H + 1;
    k.rows = k.ai1 - k.ai0 + 1;
    const int w = k.aj1 - k.aj0 + 1;
    k.kp = ((w + k.ncol - 1) / k.ncol + 15) / 16 * 16;
    k.astride = static_cast<size_t>(k.ncol) * k.rows * k.kp;
    return k;
}

void launch_assemble(Solver& s, bool is_u, const int* stop) {
    const sg_params& p = s.p;
    AsmArgs a{};
    a.u = s.F(SG_F_U);
    a.v = s.F(SG_F_V);
    a.pf = s.F(SG_F_P);
    a.APn = s.F(is_u ? SG_F_CU_AP : SG_F_CV_AP);
    a.pk = is_u ? s.pku : s.pkv;
    a.ld = s.ld, a.H = p.H, a.NX = p.NX, a.NY = p.NY;
    a.dx = p.dx, a.dy = p.dy;
    a.h = is_u ? p.dy : p.dx;
    a.poff = is_u ? 1 : -static_cast<long>(s.ld);
    a.uds = p.convection_scheme == SG_UDS;
    a.order = p.diffusion_order;
    a.d2 = s.diff2;
    a.d4 = s.diff4;
    const dim3 grid(cdiv(a.pk.aj1 - a.pk.aj0 + 1, BX), cdiv(a.pk.rows, BY));
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

// Pack the field and evaluate the residual (residual_u / residual_v, src/momentum_u.cpp:43-59)
// in one pass, then the ordered sum into *slot.
// side: the sum runs on the side stream (joined before the stop rule).
void launch_mom_residual(Solver& s, bool is_u, double* slot, const int* stop, bool side) {
    const MomArgs m = mom_args(s, is_u);
    const int nrow = m.i1 - m.i0 + 1, w = m.j1 - m.j0 + 1;
    double* cell = s.rcell[is_u ? 0 : 1];
    hipLaunchKernelGGL(k_mom_prep, dim3(cdiv(s.nc, BX), cdiv(s.nr, BY)), dim3(BX, BY), 0,
                       s.stream, m, s.fpk, s.fbuf, cell, s.nr, s.nc, 1, stop);
    SG_HIP(hipGetLastError());
    if (side)
        side_sum(s, is_u ? 1 : 2, cell, nrow, w, slot, stop);
    else if (nrow <= 0 || w <= 0)
        launch_ordered_sum(s, cell, 1, 0, -1, 0, -1, slot, stop); // empty range: 0.0
    else
        launch_ordered_sum(s, cell, w, 0, nrow - 1, 0, w - 1, slot, stop);
}

void launch_mom_pack(Solver& s, bool is_u, const int* stop) {
    const MomArgs m = mom_args(s, is_u);
    hipLaunchKernelGGL(k_mom_prep, dim3(cdiv(s.nc, BX), cdiv(s.nr, BY)), dim3(BX, BY), 0,
                       s.stream, m, s.fpk, s.fbuf, s.rcell[0], s.nr, s.nc, 0, stop);
    SG_HIP(hipGetLastError());
}

void launch_mom_colour_pass(Solver& s, bool is_u, int col, const int* stop) {
    const MomArgs m = mom_args(s, is_u);
    if (m.i1 < m.i0 || m.j1 < m.j0)
        return;
    if (m.pk.ncol == 2)
        hipLaunchKernelGGL(k_mom_pass<2>, pass_grid(m, 1), dim3(BX, BY), 0, s.stream, m, s.fpk,
                           s.fbuf, col, stop);
    else
        hipLaunchKernelGGL(k_mom_pass<5>, pass_grid(m, 1), dim3(BX, BY), 0, s.stream, m, s.fpk,
                           s.fbuf, col, stop);
    SG_HIP(hipGetLastError());
}

// sweep_u_batch / sweep_v_batch (src/momentum_u.cpp:164-171) for the two parallel solvers.
// Multicolour: the packed field must hold the current field (packed = the residual's prep ran
// since the last change, else it is packed here); it is unpacked after the last pass.
void launch_mom_sweeps(Solver& s, bool is_u, int nsweeps, const int* stop, bool packed) {
    const sg_params& p = s.p;
    if (nsweeps <= 0)
        return;
    const MomArgs m = mom_args(s, is_u);
    if (p.momentum_solver == SG_MOM_GS_MULTICOLOUR) {
        if (!packed)
            launch_mom_pack(s, is_u, stop);
        const int ncol = ncolours(p);
        for (int k = 0; k < nsweeps; ++k)
            for (int c = 0; c < ncol; ++c)
                launch_mom_colour_pass(s, is_u, c, stop);
        hipLaunchKernelGGL(k_field_unpack, dim3(cdiv(s.nc, BX), cdiv(s.nr, BY)), dim3(BX, BY), 0,
                           s.stream, s.fpk, s.fbuf, m.x, s.ld, s.nr, s.nc, stop);
        SG_HIP(hipGetLastError());
        return;
    }
    // Jacobi batch: full copy, then alternate so that the last pass writes the field.
    hipLaunchKernelGGL(k_copy, dim3(1024), dim3(256), 0, s.stream, m.x, s.scratch, s.n, stop);
    SG_HIP(hipGetLastError());
    if (m.i1 < m.i0 || m.j1 < m.j0)
        return;
    for (int k = 0; k < nsweeps; ++k) {
        const bool to_x = (nsweeps - k) % 2 == 1;
        const double* src = to_x ? s.scratch : m.x;
        double* dst = to_x ? m.x : s.scratch;
        hipLaunchKernelGGL(k_mom_jacobi, pass_grid(m, m.pk.ncol), dim3(BX, BY), 0, s.stream, m,
                           src, dst, stop);
        SG_HIP(hipGetLastError());
    }
}

// Whole-array field pack geometry (all rows and columns).
Pack make_field_pack(const sg_params& p) {
    Pack k{};
    k.ncol = ncolours(p);
    k.ai0 = 0;
    k.aj0 = 0;
    k.ai1 = p.nrows - 1;
    k.aj1 = p.ncols - 1;
    k.rows = p.nrows;
    k.kp = ((p.ncols + k.ncol - 1) / k.ncol + 15) / 16 * 16;
    k.astride = static_cast<size_t>(k.ncol) * k.rows * k.kp;
    return k;
}

// Row-major array id <-> packed layout (coefficient ids only); p uploads refresh the pressure
// terms of both components.
void sync_packed(Solver& s, int id, bool to_packed) {
    const bool cu = id >= SG_F_CU_AE && id <= SG_F_CU_AP;
    const bool cv = id >= SG_F_CV_AE && id <= SG_F_CV_AP;
    if (cu || cv) {
        const Pack& pk = cu ? s.pku : s.pkv;
        const int a = id - (cu ? SG_F_CU_AE : SG_F_CV_AE);
        hipLaunchKernelGGL(k_pack, dim3(cdiv(pk.aj1 - pk.aj0 + 1, BX), cdiv(pk.rows, BY)),
                           dim3(BX, BY), 0, s.stream, pk, s.F(id), s.ld, a, to_packed ? 1 : 0);
        SG_HIP(hipGetLastError());
    } else if (id == SG_F_P && to_packed) {
        for (int k = 0; k < 2; ++k) {
            const Pack& pk = k == 0 ? s.pku : s.pkv;
            hipLaunchKernelGGL(k_pterm, dim3(cdiv(pk.aj1 - pk.aj0 + 1, BX), cdiv(pk.rows, BY)),
                               dim3(BX, BY), 0, s.stream, pk, s.F(SG_F_P), s.ld,
                               k == 0 ? 1L : -static_cast<long>(s.ld), k == 0 ? s.p.dy : s.p.dx);
            SG_HIP(hipGetLastError());
        }
    }
}

} // namespace sg
