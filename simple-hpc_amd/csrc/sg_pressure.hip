// sg_pressure.hip — pressure correction on gfx950: p' system assembly, the red-black / damped
// Jacobi / CG solvers for p', and the u, v, p corrections (src/pressure.cpp). Multigrid is in
// sg_mg.hip. Arithmetic is the reference's expression for expression.
#include "sg_internal.hpp"

#include <cmath>

namespace sg {

namespace {

// Workgroups of 256 cells x 4 rows (assemble_p and correct read each cell's row and the one
// above or below: 256 x 4 measured 3.4 us per outer iteration faster at 1024^2 than 64 x 16).
constexpr int BX = 256, BY = 4;

struct PArgs {
    double *AE, *AW, *AN, *AS, *AP, *b, *pc;
    int ld, H, NX, NY;
};

PArgs p_args(Solver& s) {
    return {s.F(SG_F_AEP), s.F(SG_F_AWP), s.F(SG_F_ANP), s.F(SG_F_ASP), s.F(SG_F_APP),
            s.F(SG_F_B),   s.F(SG_F_PC),  s.ld,          s.p.H,         s.p.NX,
            s.p.NY};
}

// assemble_p (src/pressure.cpp:14-54) for interior cell (i, j): links dy^2/APU, dx^2/APV, the
// wall links zeroed, APP = AE + AW + AN + AS, mass source b, |b| for res_p.
struct ApArgs {
    PArgs a;
    const double *APU, *APV, *u, *v;
    double dx, dy;
    double* cell;
    int nr, nc; // array rows and columns (the wall links extend over the halo)
    int lean;   // AE, AN, b only
};

__device__ __forceinline__ void k_assemble_p(const ApArgs& g, const Blk& bk) {
    const PArgs& a = g.a;
    const double *APU = g.APU, *APV = g.APV, *u = g.u, *v = g.v;
    const double dx = g.dx, dy = g.dy;
    double* cell = g.cell;
    const int H = a.H;
    const int j = H + 1 + bk.x * BX + threadIdx.x;
    const int i = H + bk.y * BY + threadIdx.y;
    const long ld = a.ld;
    // The wall-link zeroing of src/pressure.cpp:30-41, which also covers halo rows and columns,
    // by the first row of workgroups (the cells below write the same zeros inside the range).
    if (bk.y == 0) {
        for (int k = bk.x * (BX * BY) + threadIdx.y * BX + threadIdx.x; k < max(g.nr, g.nc);
             k += bk.nx * (BX * BY)) {
            if (k < g.nr) {
                if (!g.lean)
                    a.AW[k * ld + H + 1] = 0.0;
                a.AE[k * ld + a.NY + H] = 0.0;
            }
            if (k < g.nc) {
                if (!g.lean)
                    a.AS[(a.NX + H - 1) * ld + k] = 0.0;
                a.AN[H * ld + k] = 0.0;
            }
        }
    }
    if (i > a.NX + H - 1 || j > a.NY + H)
        return;
    const size_t q = static_cast<size_t>(i) * ld + j;
    double ae = dy * dy / APU[q];
    double aw = dy * dy / APU[q - 1];
    double an = dx * dx / APV[q];
    double as = dx * dx / APV[q + ld];
    if (j == H + 1) aw = 0.0;
    if (j == a.NY + H) ae = 0.0;
    if (i == a.NX + H - 1) as = 0.0;
    if (i == H) an = 0.0;
    a.AE[q] = ae;
    a.AN[q] = an;
    if (!g.lean) {
        a.AW[q] = aw;
        a.AS[q] = as;
        a.AP[q] = ae + aw + an + as;
    }
    const double bq = -dx * (-v[q + ld] + v[q]) - dy * (-u[q - 1] + u[q]);
    a.b[q] = bq;
    cell[static_cast<size_t>(i - H) * a.NY + (j - H - 1)] = fabs(bq);
}

// pc_update (src/pressure.cpp:70-75): p' = b/APP; p' += (E + W + S + N)/APP, neighbours from src.
__device__ __forceinline__ void pc_update(const PArgs& a, double* dst, const double* src, size_t q) {
    const long ld = a.ld;
    double t = a.b[q] / a.AP[q];
    t += (a.AE[q] * src[q + 1] + a.AW[q] * src[q - 1] + a.AS[q] * src[q + ld] +
          a.AN[q] * src[q - ld]) /
         a.AP[q];
    dst[q] = t;
}

// Red-black solve (solve_pc_redblack, src/pressure.cpp:114-134): colour c is the cells with
// (i + j) mod 2 = c, i and j the padded array's indices. A grid up to 2048 cells (the 32 x 32
// cavity's size) runs the whole solve in one workgroup, where launches per colour pass would be
// pure latency: p' of the interior and its ring of zeros in LDS, (NX + 2) x (NY + 2) doubles; the
// operator read from global memory (it stays in the caches); maxit_inner_p sweeps of colour 0
// then colour 1 with a barrier after every colour pass; the interior written back at the end (p'
// outside it is zero, as the fill before this launch left it). Every update is pc_update's
// expression on the same neighbour values in the same colour order as the reference's loops, so
// the result is theirs bit for bit. Larger grids run in staged tiles (launch_rb_staged).
constexpr int RB1_T = 1024;
constexpr size_t RB1_MAX_LDS = 65536; // bytes: within the default dynamic-LDS limit
struct Rb1Args {
    PArgs a;
    int sweeps;
};

__device__ __forceinline__ void k_pc_rb_one(const Rb1Args& g, const Blk&) {
    extern __shared__ __align__(16) double sp[];
    const PArgs& a = g.a;
    const int NX = a.NX, NY = a.NY, H = a.H, W = NY + 2;
    const long ld = a.ld;
    const int n = (NX + 2) * W, half = (NY + 1) / 2;
    for (int e = threadIdx.x; e < n; e += RB1_T)
        sp[e] = 0.0;
    __syncthreads();
    for (int k = 0; k < g.sweeps; ++k)
        for (int c = 0; c < 2; ++c) {
            for (int e = threadIdx.x; e < NX * half; e += RB1_T) {
                const int r = e / half, i = H + r;
                const int j = H + 1 + ((i + H + 1 + c) & 1) + 2 * (e - r * half);
                if (j > NY + H)
                    continue;
                const size_t q = static_cast<size_t>(i) * ld + j;
                double* x = sp + (r + 1) * W + (j - H);
                double t = a.b[q] / a.AP[q];
                t += (a.AE[q] * x[1] + a.AW[q] * x[-1] + a.AS[q] * x[W] + a.AN[q] * x[-W]) /
                     a.AP[q];
                *x = t;
            }
            __syncthreads();
        }
    for (int e = threadIdx.x; e < NX * NY; e += RB1_T) {
        const int r = e / NY, cj = e - r * NY;
        a.pc[static_cast<size_t>(H + r) * ld + H + 1 + cj] = sp[(r + 1) * W + cj + 1];
    }
}

// Damped Jacobi pass, weight 0.8 (src/pressure.cpp:98-104).
struct PJacArgs {
    PArgs a;
    const double* src;
    double* dst;
};

__device__ __forceinline__ void k_pc_jacobi(const PJacArgs& g, const Blk& bk) {
    const PArgs& a = g.a;
    const double* src = g.src;
    double* dst = g.dst;
    const int H = a.H;
    const int j = H + 1 + bk.x * BX + threadIdx.x;
    const int i = H + bk.y * BY + threadIdx.y;
    if (i > a.NX + H - 1 || j > a.NY + H)
        return;
    const size_t q = static_cast<size_t>(i) * a.ld + j;
    const double w = 0.8;
    pc_update(a, dst, src, q);
    dst[q] = w * dst[q] + (1 - w) * src[q];
}

// correct (src/pressure.cpp:136-157).
// AW, AS and AP from AE and AN after a lean assemble_p: AW(i, j) = AE(i, j - 1) and AS(i, j) =
// AN(i + 1, j) (the same divisions, src/pressure.cpp:21-24), the wall links zeroed as there
// (:28-39), AP = AE + AW + AN + AS in that order (:45): the arrays a full assemble_p writes.
__global__ void __launch_bounds__(BX* BY) k_p_derive(PArgs a, int nr, int nc) {
    const int H = a.H;
    const long ld = a.ld;
    if (blockIdx.y == 0)
        for (int k = blockIdx.x * (BX * BY) + threadIdx.y * BX + threadIdx.x; k < max(nr, nc);
             k += gridDim.x * (BX * BY)) {
            if (k < nr)
                a.AW[k * ld + H + 1] = 0.0;
            if (k < nc)
                a.AS[(a.NX + H - 1) * ld + k] = 0.0;
        }
    const int j = H + 1 + blockIdx.x * BX + threadIdx.x;
    const int i = H + blockIdx.y * BY + threadIdx.y;
    if (i > a.NX + H - 1 || j > a.NY + H)
        return;
    const size_t q = static_cast<size_t>(i) * ld + j;
    const double ae = a.AE[q], an = a.AN[q];
    const double aw = j == H + 1 ? 0.0 : a.AE[q - 1];
    const double as = i == a.NX + H - 1 ? 0.0 : a.AN[q + ld];
    a.AW[q] = aw;
    a.AS[q] = as;
    a.AP[q] = ae + aw + an + as;
}

struct CorrArgs {
    PArgs a;
    const double *APU, *APV;
    double *u, *v, *pf;
    double dx, dy, ap;
    const double* mslot; // non-null: p' still holds its mean, total at *mslot over n cells
    double n;
    double* pcout;       // with mslot: p' - mean of every interior cell, written here
};

__device__ __forceinline__ void k_correct(const CorrArgs& g, const Blk& bk) {
    const PArgs& a = g.a;
    const double *APU = g.APU, *APV = g.APV;
    double *u = g.u, *v = g.v, *pf = g.pf;
    const double dx = g.dx, dy = g.dy, ap = g.ap;
    const int H = a.H;
    const int j = H + 1 + bk.x * BX + threadIdx.x;
    const int i = H + bk.y * BY + threadIdx.y;
    if (i > a.NX + H - 1 || j > a.NY + H)
        return;
    const long ld = a.ld;
    const size_t q = static_cast<size_t>(i) * ld + j;
    const double* pc = a.pc;
    // With mslot, every p' read is p' - mean, the value the mean removal (k_total_sub) would
    // have stored (the same subtraction), so the corrections have the same bits.
    const bool sub = g.mslot != nullptr;
    const double m = sub ? *g.mslot / g.n : 0.0;
    auto pm = [&](size_t k) { return sub ? pc[k] - m : pc[k]; };
    const double pq = pm(q);
    if (j <= a.NY + H - 1)
        u[q] += dy * (pq - pm(q + 1)) / APU[q];
    if (i >= H + 1)
        v[q] += dx * (pq - pm(q - ld)) / APV[q];
    pf[q] += ap * pq;
    if (sub)
        g.pcout[q] = pq;
}

struct FillArgs {
    double* a;
    size_t n;
    double v;
};
constexpr int FT = 256;

__device__ __forceinline__ void k_fill(const FillArgs& g, const Blk& bk) {
    for (size_t k = bk.x * static_cast<size_t>(FT) + threadIdx.x; k < g.n;
         k += static_cast<size_t>(bk.nx) * FT)
        g.a[k] = g.v;
}

// ---- CG (src/pressure.cpp:208-317). Everything on the device: the per-cell work, the ordered
// sums, and the scalar recurrence (alpha, beta, the stop test) in one-thread kernels with the
// reference's expressions. Every CG launch takes the solve's done flag as its stop flag, so the
// host enqueues iterations in chunks and reads the flag once per chunk.

struct CgArgs {
    PArgs a;
    double *r, *z, *d, *q, *cell;
    size_t plane; // stride between the 4 value planes of cell
    const double* sc; // the recurrence's scalars (CG_*)
};

// Scalars of the recurrence, in the solver's scalar slots from S_CG on.
enum { CG_RR = 0, CG_RZ, CG_MZ, CG_NB, CG_ALPHA, CG_BETA, CG_N };

__device__ __forceinline__ size_t cidx(const PArgs& a, int i, int j) {
    return static_cast<size_t>(i - a.H) * a.NY + (j - a.H - 1);
}

// r = b - mb, z = r / APP; values r*r, z, r*z, r.
__global__ void __launch_bounds__(BX* BY) k_cg_init(CgArgs c, const double* sb, double n) {
    const PArgs& a = c.a;
    const int H = a.H;
    const int j = H + 1 + blockIdx.x * BX + threadIdx.x;
    const int i = H + blockIdx.y * BY + threadIdx.y;
    if (i > a.NX + H - 1 || j > a.NY + H)
        return;
    const size_t q = static_cast<size_t>(i) * a.ld + j, k = cidx(a, i, j);
    const double mb = *sb / n;
    const double ri = a.b[q] - mb;
    c.r[q] = ri;
    const double zi = ri / a.AP[q];
    c.z[q] = zi;
    c.cell[k] = ri * ri;
    c.cell[k + c.plane] = zi;
    c.cell[k + 2 * c.plane] = ri * zi;
    c.cell[k + 3 * c.plane] = ri;
}

// The recurrence, one thread each (src/pressure.cpp:254-302). sums: the 4 ordered sums
// (r.r, sum z, r.z, sum r) or dq in sums[0]; ctl: [done, iterations].
struct CgStep {
    double* sc;
    const double* sums;
    int* ctl;
    double n, tol;
    int maxit;
};
// After the initial sums: stop at once if r.r = 0, else |b| and the mean-free r.z.
__global__ void k_cg_start(CgStep g) {
    if (threadIdx.x != 0)
        return;
    const double rr = g.sums[0], sz = g.sums[1], sr = g.sums[3];
    g.ctl[1] = 0;
    if (rr == 0.0) {
        g.ctl[0] = 1;
        return;
    }
    g.ctl[0] = 0;
    const double mz = sz / g.n;
    g.sc[CG_RR] = rr;
    g.sc[CG_NB] = sqrt(rr);
    g.sc[CG_MZ] = mz;
    g.sc[CG_RZ] = g.sums[2] - mz * sr;
}
// After d.q: alpha = r.z / d.q, or stop if d.q is not positive.
__global__ void k_cg_alpha(CgStep g) {
    if (threadIdx.x != 0 || g.ctl[0])
        return;
    const double dq = g.sums[0];
    if (!(dq > 0.0)) {
        g.ctl[0] = 1;
        return;
    }
    g.sc[CG_ALPHA] = g.sc[CG_RZ] / dq;
}
// After the update's sums: one more iteration; stop at |r| / |b| <= tol or the iteration cap,
// else the new mean-free r.z and beta.
__global__ void k_cg_next(CgStep g) {
    if (threadIdx.x != 0 || g.ctl[0])
        return;
    const double rr = g.sums[0], sz = g.sums[1], sr = g.sums[3];
    double rzn = g.sums[2];
    const int k = ++g.ctl[1];
    g.sc[CG_RR] = rr;
    const double rel = sqrt(rr) / g.sc[CG_NB];
    if (rel <= g.tol || k >= g.maxit) {
        g.ctl[0] = 1;
        return;
    }
    const double mz = sz / g.n;
    rzn -= mz * sr;
    g.sc[CG_MZ] = mz;
    g.sc[CG_BETA] = rzn / g.sc[CG_RZ];
    g.sc[CG_RZ] = rzn;
}

__device__ __forceinline__ bool cg_cell(const PArgs& a, int& i, int& j, const Blk& bk) {
    const int H = a.H;
    j = H + 1 + bk.x * BX + threadIdx.x;
    i = H + bk.y * BY + threadIdx.y;
    return i <= a.NX + H - 1 && j <= a.NY + H;
}

// d = (z - mz) + beta d (first: d = z - mz).
struct CgDirArgs {
    CgArgs c;
    int first;
};
__device__ __forceinline__ void k_cg_dir(const CgDirArgs& g, const Blk& bk) {
    const CgArgs& c = g.c;
    int i, j;
    if (!cg_cell(c.a, i, j, bk))
        return;
    const size_t q = static_cast<size_t>(i) * c.a.ld + j;
    const double mz = c.sc[CG_MZ];
    c.d[q] = g.first ? c.z[q] - mz : (c.z[q] - mz) + c.sc[CG_BETA] * c.d[q];
}

// q = A d (apply_A_p, src/pressure.cpp:165-173); value d*q.
__device__ __forceinline__ void k_cg_matvec(const CgArgs& c, const Blk& bk) {
    const PArgs& a = c.a;
    int i, j;
    if (!cg_cell(a, i, j, bk))
        return;
    const long ld = a.ld;
    const size_t q = static_cast<size_t>(i) * ld + j;
    const double* d = c.d;
    const double qi = a.AP[q] * d[q] - (a.AE[q] * d[q + 1] + a.AW[q] * d[q - 1] +
                                        a.AS[q] * d[q + ld] + a.AN[q] * d[q - ld]);
    c.q[q] = qi;
    c.cell[cidx(a, i, j)] = d[q] * qi;
}

// x += alpha d; r -= alpha q; z = r / APP; values r*r, z, r*z, r.
__device__ __forceinline__ void k_cg_update(const CgArgs& c, const Blk& bk) {
    const PArgs& a = c.a;
    int i, j;
    if (!cg_cell(a, i, j, bk))
        return;
    const size_t q = static_cast<size_t>(i) * a.ld + j, k = cidx(a, i, j);
    const double alpha = c.sc[CG_ALPHA];
    a.pc[q] += alpha * c.d[q];
    const double ri = c.r[q] - alpha * c.q[q];
    c.r[q] = ri;
    const double zi = ri / a.AP[q];
    c.z[q] = zi;
    c.cell[k] = ri * ri;
    c.cell[k + c.plane] = zi;
    c.cell[k + 2 * c.plane] = ri * zi;
    c.cell[k + 3 * c.plane] = ri;
}

dim3 pgrid(const sg_params& p) { return dim3(cdiv(p.NY, BX), cdiv(p.NX, BY)); }

// CG iterations enqueued per host check of the done flag.
constexpr int CG_CHUNK = 32;

void solve_cg(Solver& s) {
    const sg_params& p = s.p;
    const PArgs a = p_args(s);
    const size_t plane = static_cast<size_t>(p.NX) * p.NY;
    double* slot = s.scal + S_SUM_A;
    double* sc = s.scal + S_CG;
    CgArgs c{a, s.cg[0], s.cg[1], s.cg[2], s.cg[3], s.cell, plane, sc};
    const double n = static_cast<double>(p.NX) * p.NY;
    const int* never = s.never;
    const int* done = s.cgctl;
    const CgStep step{sc, slot, s.cgctl, n, p.pressure_tolerance, p.pressure_maxit};
    auto sum_plane = [&](int m, double* dst, const int* stop) {
        launch_ordered_sum(s, s.cell + m * plane, p.NY, 0, p.NX - 1, 0, p.NY - 1, dst, stop);
    };
    const dim3 g = pgrid(p), b(BX, BY);
    launch_fill(s, a.pc, s.n, 0.0, never);
    // sb = sum b, r = b - mb, z = r / APP, and the four sums.
    launch_ordered_sum(s, a.b, s.ld, p.H, p.NX + p.H - 1, p.H + 1, p.NY + p.H, slot, never);
    hipLaunchKernelGGL(k_cg_init, g, b, 0, s.stream, c, slot, n);
    SG_HIP(hipGetLastError());
    for (int m = 0; m < 4; ++m)
        sum_plane(m, slot + m, never);
    hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(64), 0, s.stream, step);
    SG_HIP(hipGetLastError());
    emit<CgDirArgs, k_cg_dir, BX * BY>(s, g, b, 0, CgDirArgs{c, 1}, done);
    int ctl[2] = {0, 0};
    for (int k = 0; k < p.pressure_maxit; k += CG_CHUNK) {
        for (int m = 0; m < CG_CHUNK && k + m < p.pressure_maxit; ++m) {
            emit<CgArgs, k_cg_matvec, BX * BY>(s, g, b, 0, c, done);
            sum_plane(0, slot, done);
            hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(64), 0, s.stream, step);
            SG_HIP(hipGetLastError());
            emit<CgArgs, k_cg_update, BX * BY>(s, g, b, 0, c, done);
            for (int q = 0; q < 4; ++q)
                sum_plane(q, slot + q, done);
            hipLaunchKernelGGL(k_cg_next, dim3(1), dim3(64), 0, s.stream, step);
            SG_HIP(hipGetLastError());
            emit<CgDirArgs, k_cg_dir, BX * BY>(s, g, b, 0, CgDirArgs{c, 0}, done);
        }
        SG_HIP(hipMemcpyAsync(ctl, s.cgctl, sizeof ctl, hipMemcpyDeviceToHost, s.stream));
        SG_HIP(hipStreamSynchronize(s.stream));
        if (ctl[0])
            break;
    }
    launch_ordered_mean_sub(s, a.pc, s.ld, p.H, p.NX + p.H - 1, p.H + 1, p.NY + p.H, n, slot,
                            never);
    s.cg_last = ctl[1];
}

} // namespace

void launch_fill(Solver& s, double* a, size_t n, double v, const int* stop) {
    emit<FillArgs, k_fill, FT>(s, dim3(std::min<size_t>(2048, (n + FT - 1) / FT + 1)), dim3(FT), 0,
                               FillArgs{a, n, v}, stop);
}

void launch_assemble_p(Solver& s, double* slot, const int* stop, bool defer, bool lean) {
    const sg_params& p = s.p;
    const PArgs a = p_args(s);
    emit<ApArgs, k_assemble_p, BX * BY>(s, pgrid(p), dim3(BX, BY), 0,
                                        ApArgs{a, s.F(SG_F_CU_AP), s.F(SG_F_CV_AP), s.F(SG_F_U),
                                               s.F(SG_F_V), p.dx, p.dy, s.rcell[2], s.nr, s.nc,
                                               lean ? 1 : 0},
                                        stop);
    s.p_lean = lean;
    s.b_raw = false; // a new b, as the reference's assemble_p leaves it
    if (defer)
        defer_sum(s, 3, s.rcell[2], p.NX, p.NY, slot);
    else
        launch_ordered_sum(s, s.rcell[2], p.NY, 0, p.NX - 1, 0, p.NY - 1, slot, stop);
}

void ensure_p_full(Solver& s) {
    if (!s.p_lean)
        return;
    s.p_lean = false;
    hipLaunchKernelGGL(k_p_derive, pgrid(s.p), dim3(BX, BY), 0, s.stream, p_args(s), s.nr, s.nc);
    SG_HIP(hipGetLastError());
}

void launch_correct(Solver& s, const int* stop) {
    const sg_params& p = s.p;
    s.pt_stale[0] = s.pt_stale[1] = true; // p changes
    const bool fused = s.pc_mean_pending;
    s.pc_mean_pending = false;
    const CorrArgs c{p_args(s),
                     s.F(SG_F_CU_AP),
                     s.F(SG_F_CV_AP),
                     s.F(SG_F_U),
                     s.F(SG_F_V),
                     s.F(SG_F_P),
                     p.dx,
                     p.dy,
                     p.relaxation_factor_p,
                     fused ? s.scal + S_SUM_D : nullptr,
                     static_cast<double>(p.NX) * p.NY,
                     fused ? s.scratch : nullptr};
    emit<CorrArgs, k_correct, BX * BY>(s, pgrid(p), dim3(BX, BY), 0, c, stop);
    if (fused)
        s.pc_in_scratch = true;
}

void finish_pc_mean(Solver& s, const int* stop) {
    if (!s.fuse_pc_mean) {
        launch_remove_mean0(s, s.F(SG_F_PC), stop);
        return;
    }
    const sg_params& p = s.p;
    launch_ordered_sum(s, s.F(SG_F_PC), s.ld, p.H, p.NX + p.H - 1, p.H + 1, p.NY + p.H,
                       s.scal + S_SUM_D, stop, 0);
    s.pc_mean_pending = true;
}

void materialize_pc(Solver& s) {
    if (!s.pc_in_scratch)
        return;
    s.pc_in_scratch = false;
    const sg_params& p = s.p;
    const size_t o = static_cast<size_t>(p.H) * s.ld + p.H + 1;
    SG_HIP(hipMemcpy2DAsync(s.F(SG_F_PC) + o, sizeof(double) * s.ld, s.scratch + o,
                            sizeof(double) * s.ld, sizeof(double) * p.NY, p.NX,
                            hipMemcpyDeviceToDevice, s.stream));
}

void launch_remove_mean0(Solver& s, double* x, const int* stop, double* zero) {
    const sg_params& p = s.p;
    launch_ordered_mean_sub(s, x, s.ld, p.H, p.NX + p.H - 1, p.H + 1, p.NY + p.H,
                            static_cast<double>(p.NX) * p.NY, s.scal + S_SUM_D, stop, zero,
                            zero ? s.n : 0);
}

// One workgroup pays off while a colour pass is at most one update per thread: at 32^2 (config 1)
// an outer iteration took 0.131 -> 0.109 ms, at 80^2 0.142 -> 0.153 ms (one CU doing 3 updates
// per thread per pass loses to 16 launches spread over the chip).
bool rb_one_fits(const sg_params& p) {
    return static_cast<long>(p.NX) * p.NY <= 2L * RB1_T &&
           sizeof(double) * (p.NX + 2) * static_cast<size_t>(p.NY + 2) <= RB1_MAX_LDS;
}

// solve_p for every GPU pressure solver. CG needs the host between iterations (eager only).
void launch_solve_p(Solver& s, const int* stop, bool eager) {
    const sg_params& p = s.p;
    const PArgs a = p_args(s);
    switch (p.pressure_solver) {
    case SG_P_GS_REDBLACK:
        launch_fill(s, a.pc, s.n, 0.0, stop);
        if (rb_one_fits(p)) {
            const size_t lds = sizeof(double) * (p.NX + 2) * (p.NY + 2);
            emit<Rb1Args, k_pc_rb_one, RB1_T>(s, dim3(1), dim3(RB1_T), lds,
                                              Rb1Args{a, p.maxit_inner_p}, stop);
            return;
        }
        launch_rb_staged(s, stop);
        return;
    case SG_P_JACOBI: {
        const int n = p.maxit_inner_p;
        launch_fill(s, a.pc, s.n, 0.0, stop);
        launch_fill(s, s.scratch, s.n, 0.0, stop);
        for (int k = 0; k < n; ++k) {
            const bool to_pc = (n - k) % 2 == 1;
            const double* src = to_pc ? s.scratch : a.pc;
            double* dst = to_pc ? a.pc : s.scratch;
            emit<PJacArgs, k_pc_jacobi, BX * BY>(s, pgrid(p), dim3(BX, BY), 0,
                                                 PJacArgs{a, src, dst}, stop);
        }
        return;
    }
    case SG_P_CG:
        if (!eager)
            throw ArgError("pressure_solver = cg cannot be captured in a graph");
        solve_cg(s);
        s.cg_total += s.cg_last;
        return;
    case SG_P_MG:
        launch_mg_solve(s, stop, false);
        return;
    default:
        throw ArgError("backend = gpu does not run pressure_solver = gs_lex (a serial sweep)");
    }
}

} // namespace sg
