// oracle.cpp — CPU oracle of the SIMPLE lid-driven-cavity path. TEST INFRASTRUCTURE ONLY: the
// checker of the GPU backend and the timed CPU baseline, never the product (see oracle.h).
//
// Every expression keeps the reference's operation order (docs/SPEC.md section 7; built with
// -ffp-contract=off), so results are bit-identical to arfazeli/simple-hpc's CPU path. Loops that
// are independent per cell run under OpenMP; every sum uses the reference's fixed order: each row
// summed with j ascending from 0.0, then the row sums in ascending row order
// (src/decomp.cpp:176-183, src/multigrid.cpp:18-34), so the bits never depend on the thread count.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using Vec = std::vector<double>;

// std::max(0.0, a) of src/coefficients.cpp:18: (0.0 < a) ? a : 0.0 (so -0.0 and NaN give +0.0).
inline double pos(double a) { return 0.0 < a ? a : 0.0; }

struct C9 {
    double e, w, n, s, ee, ww, nn, ss, p;
};

struct Level {
    int nx = 0, ny = 0, stride = 0;
    double *AE = nullptr, *AW = nullptr, *AN = nullptr, *AS = nullptr, *AP = nullptr;
    double *b = nullptr, *x = nullptr, *r = nullptr;
    Vec store[8]; // coarse levels: padded (nx+2) x (ny+2) arrays, zero halo
};

} // namespace

struct or_state {
    sg_params p{};
    int nr = 0, nc = 0;
    Vec f[SG_F_COUNT];
    Vec scratch;           // Jacobi previous iterate (u, v, p')
    Vec cr, cz, cd, cq;    // CG work (src/pressure.cpp:163)
    Vec r0;                // multigrid level-0 residual
    std::vector<Level> mg; // multigrid hierarchy
    int cg_last = 0;
    double cg_rel = 0.0;
    long long cg_total = 0;
    int coarse_sweeps = 0;

    double& at(int id, int i, int j) { return f[id][static_cast<size_t>(i) * nc + j]; }
    double* a(int id) { return f[id].data(); }
};

namespace {

// ---- ordered sums (CellSums on one rank, src/decomp.cpp:157-233; det_sum,
// src/multigrid.cpp:18-34): row sums j ascending from 0.0, then rows ascending from 0.0.
template <class F> double ordered_sum(int i0, int i1, int j0, int j1, F val) {
    const int nrow = i1 - i0 + 1;
    Vec rows(std::max(nrow, 0), 0.0);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nrow; ++k) {
        const int i = i0 + k;
        double s = 0.0;
        for (int j = j0; j <= j1; ++j)
            s += val(i, j);
        rows[k] = s;
    }
    double s = 0.0;
    for (int k = 0; k < nrow; ++k)
        s += rows[k];
    return s;
}

// K sums at once in the same order (cg_sums, src/pressure.cpp:178-206). val(i, j, v) fills v[K]
// and may update per-cell state.
template <int K, class F> void ordered_sums(int i0, int i1, int j0, int j1, F val, double* out) {
    const int nrow = i1 - i0 + 1;
    Vec rows(static_cast<size_t>(std::max(nrow, 0)) * K, 0.0);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nrow; ++k) {
        const int i = i0 + k;
        double s[K] = {}, v[K];
        for (int j = j0; j <= j1; ++j) {
            val(i, j, v);
            for (int m = 0; m < K; ++m)
                s[m] += v[m];
        }
        for (int m = 0; m < K; ++m)
            rows[static_cast<size_t>(k) * K + m] = s[m];
    }
    for (int m = 0; m < K; ++m) {
        double s = 0.0;
        for (int k = 0; k < nrow; ++k)
            s += rows[static_cast<size_t>(k) * K + m];
        out[m] = s;
    }
}

// ---- coefficients (src/coefficients.cpp:5-63) ----

// Boundary band, src/coefficients.cpp:10-13.
inline bool band(const sg_params& p, int i, int j) {
    return j == p.N_halo_y + 1 || j == p.NY + p.N_halo_y || i == p.NX + p.N_halo_x - 1 ||
           i == p.N_halo_x;
}

// Convective part, src/coefficients.cpp:15-39 (AP summed S, W, N, E, SS, WW, NN, EE as in the code).
C9 convective(const sg_params& p, double ue, double uw, double vn, double vs, int i, int j) {
    const double dx = p.dx, dy = p.dy;
    C9 c;
    if (p.convection_scheme == SG_UDS || band(p, i, j)) {
        c.e = dy * pos(-ue);
        c.w = dy * pos(uw);
        c.n = dx * pos(-vn);
        c.s = dx * pos(vs);
        c.ee = c.ww = c.nn = c.ss = 0.0;
    } else {
        c.e = -dy * (-0.75 * pos(-ue) + 0.375 * pos(ue) - 0.125 * pos(-uw));
        c.w = -dy * (-0.125 * pos(ue) + 0.375 * pos(-uw) - 0.75 * pos(uw));
        c.n = -dx * (-0.75 * pos(-vn) + 0.375 * pos(vn) - 0.125 * pos(-vs));
        c.s = -dx * (-0.125 * pos(vn) + 0.375 * pos(-vn) - 0.75 * pos(vs));
        c.ee = -0.125 * dy * pos(-ue);
        c.ww = -0.125 * dy * pos(uw);
        c.nn = -0.125 * dx * pos(-vn);
        c.ss = -0.125 * dx * pos(vs);
    }
    c.p = c.s + c.w + c.n + c.e + c.ss + c.ww + c.nn + c.ee;
    return c;
}

// Diffusive part, src/coefficients.cpp:41-63 (literal forms kept: 4.0 / 3 * dy * nu / dx ...).
C9 diffusive(const sg_params& p, int i, int j) {
    const double dx = p.dx, dy = p.dy, nu = p.nu;
    C9 c;
    if (p.diffusion_order == 2 || band(p, i, j)) {
        c.e = c.w = dy * nu / dx;
        c.n = c.s = dx * nu / dy;
        c.ee = c.ww = c.nn = c.ss = 0.0;
        c.p = -nu * (-2 * dx / dy - 2 * dy / dx);
    } else {
        c.e = c.w = 4.0 / 3 * dy * nu / dx;
        c.n = c.s = 4.0 / 3 * dx * nu / dy;
        c.ee = c.ww = -1.0 / 12 * dy * nu / dx;
        c.nn = c.ss = -1.0 / 12 * dx * nu / dy;
        c.p = -nu * (-2.5 * dx / dy - 2.5 * dy / dx);
    }
    return c;
}

// Element-wise sum, src/coefficients.cpp:5-8, stored into the 9 arrays starting at id0
// (order AE, AW, AN, AS, AEE, AWW, ANN, ASS, AP).
inline void store9(or_state* s, int id0, int i, int j, const C9& a, const C9& b) {
    const size_t q = static_cast<size_t>(i) * s->nc + j;
    s->f[id0 + 0][q] = a.e + b.e;
    s->f[id0 + 1][q] = a.w + b.w;
    s->f[id0 + 2][q] = a.n + b.n;
    s->f[id0 + 3][q] = a.s + b.s;
    s->f[id0 + 4][q] = a.ee + b.ee;
    s->f[id0 + 5][q] = a.ww + b.ww;
    s->f[id0 + 6][q] = a.nn + b.nn;
    s->f[id0 + 7][q] = a.ss + b.ss;
    s->f[id0 + 8][q] = a.p + b.p;
}

// Neighbour sums of the momentum equations: (E + W + N + S) + (EE + WW + NN + SS)
// (src/momentum_u.cpp:36-41), and the 5-point form of the 2-colour case (:127-132).
inline double nb9(const double* const* c, const double* x, size_t q, int nc) {
    return (c[0][q] * x[q + 1] + c[1][q] * x[q - 1] + c[2][q] * x[q - nc] + c[3][q] * x[q + nc]) +
           (c[4][q] * x[q + 2] + c[5][q] * x[q - 2] + c[6][q] * x[q - 2 * nc] +
            c[7][q] * x[q + 2 * nc]);
}
inline double nb5(const double* const* c, const double* x, size_t q, int nc) {
    return c[0][q] * x[q + 1] + c[1][q] * x[q - 1] + c[2][q] * x[q - nc] + c[3][q] * x[q + nc];
}

// Colours of the multicolour momentum sweep, include/simple/momentum_u.hpp:15-17 and
// src/momentum_u.cpp:70-72.
inline int ncolours(const sg_params& p) {
    return (p.convection_scheme == SG_UDS && p.diffusion_order == 2) ? 2 : 5;
}
inline int colour_of(int ncol, int i, int j) { return ncol == 2 ? (i + j) % 2 : (i + 3 * j) % 5; }

// Generic momentum component. u: coefficient block id0 = SG_F_CU_AE, rows [H, NX+H-1],
// columns [H+1, NY+H-1], pressure term (p(i,j) - p(i,j+1)) * dy. v: id0 = SG_F_CV_AE, rows
// [H+1, NX+H-1], columns [H+1, NY+H], pressure term (p(i,j) - p(i-1,j)) * dx.
struct Mom {
    int field, id0;
    int i0, i1, j0, j1;
    double alpha, h;   // relaxation, face length multiplying the pressure difference
    int poff;          // offset of the second pressure cell: +1 (u, east) or -nc (v, north)
};

Mom mom_u(const or_state* s) {
    const sg_params& p = s->p;
    return {SG_F_U, SG_F_CU_AE, p.H, p.NX + p.H - 1, p.H + 1, p.NY + p.H - 1,
            p.relaxation_factor_u, p.dy, 1};
}
Mom mom_v(const or_state* s) {
    const sg_params& p = s->p;
    return {SG_F_V, SG_F_CV_AE, p.H + 1, p.NX + p.H - 1, p.H + 1, p.NY + p.H,
            p.relaxation_factor_v, p.dx, -s->nc};
}

// residual_u / residual_v (src/momentum_u.cpp:43-59, src/momentum_v.cpp:36-55):
// sum |(nb + pterm) - AP x| in CellSums order.
double mom_residual(or_state* s, const Mom& m) {
    const int nc = s->nc;
    const double* c[9];
    for (int k = 0; k < 9; ++k)
        c[k] = s->a(m.id0 + k);
    const double* x = s->a(m.field);
    const double* pf = s->a(SG_F_P);
    return ordered_sum(m.i0, m.i1, m.j0, m.j1, [&](int i, int j) {
        const size_t q = static_cast<size_t>(i) * nc + j;
        const double R = nb9(c, x, q, nc) + (pf[q] - pf[q + m.poff]) * m.h - c[8][q] * x[q];
        return std::fabs(R);
    });
}

// One relaxed Gauss-Seidel / Jacobi update: a * (nb + pterm) / AP + (1 - a) * x_old
// (src/momentum_u.cpp:66-67).
inline double relax(double a, double nb, double pterm, double ap, double xold) {
    return a * (nb + pterm) / ap + (1 - a) * xold;
}

// gs_lex: rows descending, columns ascending, in place (src/momentum_u.cpp:61-68, momentum_v.cpp:57-70).
void mom_gs_lex(or_state* s, const Mom& m, int nsweeps) {
    const int nc = s->nc;
    const double* c[9];
    for (int k = 0; k < 9; ++k)
        c[k] = s->a(m.id0 + k);
    double* x = s->a(m.field);
    const double* pf = s->a(SG_F_P);
    for (int k = 0; k < nsweeps; ++k)
        for (int i = m.i1; i >= m.i0; --i)
            for (int j = m.j0; j <= m.j1; ++j) {
                const size_t q = static_cast<size_t>(i) * nc + j;
                x[q] = relax(m.alpha, nb9(c, x, q, nc), (pf[q] - pf[q + m.poff]) * m.h, c[8][q], x[q]);
            }
}

// Jacobi batch: one full copy into the scratch buffer, then buffers alternate so that the last
// pass writes the field (src/momentum_u.cpp:89-116, momentum_v.cpp:85-114).
void mom_jacobi(or_state* s, const Mom& m, int nsweeps) {
    if (nsweeps <= 0)
        return;
    const int nc = s->nc;
    const double* c[9];
    for (int k = 0; k < 9; ++k)
        c[k] = s->a(m.id0 + k);
    double* x = s->a(m.field);
    double* old = s->scratch.data();
    const double* pf = s->a(SG_F_P);
    std::memcpy(old, x, sizeof(double) * s->f[m.field].size());
    for (int k = 0; k < nsweeps; ++k) {
        const bool to_x = (nsweeps - k) % 2 == 1;
        const double* src = to_x ? old : x;
        double* dst = to_x ? x : old;
#pragma omp parallel for schedule(static)
        for (int i = m.i1; i >= m.i0; --i)
            for (int j = m.j0; j <= m.j1; ++j) {
                const size_t q = static_cast<size_t>(i) * nc + j;
                dst[q] = relax(m.alpha, nb9(c, src, q, nc), (pf[q] - pf[q + m.poff]) * m.h, c[8][q],
                               src[q]);
            }
    }
}

// Multicolour Gauss-Seidel: per sweep, colours 0..ncol-1, each a parallel in-place pass over its
// cells; the 2-colour case (UDS, order 2) uses the near sum only (src/momentum_u.cpp:123-143).
void mom_multicolour(or_state* s, const Mom& m, int nsweeps) {
    const int nc = s->nc, ncol = ncolours(s->p);
    const double* c[9];
    for (int k = 0; k < 9; ++k)
        c[k] = s->a(m.id0 + k);
    double* x = s->a(m.field);
    const double* pf = s->a(SG_F_P);
    for (int sw = 0; sw < nsweeps; ++sw)
        for (int col = 0; col < ncol; ++col) {
#pragma omp parallel for schedule(static)
            for (int i = m.i1; i >= m.i0; --i) {
                int j = m.j0;
                while (colour_of(ncol, i, j) != col)
                    ++j;
                for (; j <= m.j1; j += ncol) {
                    const size_t q = static_cast<size_t>(i) * nc + j;
                    const double nb = ncol == 2 ? nb5(c, x, q, nc) : nb9(c, x, q, nc);
                    x[q] = relax(m.alpha, nb, (pf[q] - pf[q + m.poff]) * m.h, c[8][q], x[q]);
                }
            }
        }
}

// momentum_precision = mixed: FP32 sweeps of the correction d inside FP64 defect correction, as
// the GPU runs them (no reference counterpart). r0 = nb(x0) + pterm - AP x0 in FP64 (mom_residual's
// expression), then per sweep and colour d = (a / AP) (r0 + nb(d)) + (1 - a) d in float with the
// coefficients, a / AP and r0 rounded to float once and d = 0 outside the sweep range, then
// x = x0 + d. Same operations in the same order as the GPU kernel, so the same bits
// (-ffp-contract=off on both sides).
void mom_dc(or_state* s, const Mom& m, int nsweeps) {
    if (nsweeps <= 0)
        return;
    const int nc = s->nc, ncol = ncolours(s->p);
    const double* c[9];
    for (int k = 0; k < 9; ++k)
        c[k] = s->a(m.id0 + k);
    double* x = s->a(m.field);
    const double* pf = s->a(SG_F_P);
    const size_t n = s->f[m.field].size();
    std::vector<float> d(n, 0.0f), r0(n, 0.0f), w(n, 0.0f);
    std::vector<float> cf[8];
    for (auto& v : cf)
        v.assign(n, 0.0f);
#pragma omp parallel for schedule(static)
    for (int i = m.i0; i <= m.i1; ++i)
        for (int j = m.j0; j <= m.j1; ++j) {
            const size_t q = static_cast<size_t>(i) * nc + j;
            r0[q] = static_cast<float>(nb9(c, x, q, nc) + (pf[q] - pf[q + m.poff]) * m.h -
                                       c[8][q] * x[q]);
            w[q] = static_cast<float>(m.alpha / c[8][q]);
            for (int k = 0; k < 8; ++k)
                cf[k][q] = static_cast<float>(c[k][q]);
        }
    const float oma = static_cast<float>(1 - m.alpha);
    const long L = nc;
    for (int sw = 0; sw < nsweeps; ++sw)
        for (int col = 0; col < ncol; ++col) {
#pragma omp parallel for schedule(static)
            for (int i = m.i0; i <= m.i1; ++i) {
                int j = m.j0;
                while (colour_of(ncol, i, j) != col)
                    ++j;
                for (; j <= m.j1; j += ncol) {
                    const size_t q = static_cast<size_t>(i) * nc + j;
                    float nb = cf[0][q] * d[q + 1] + cf[1][q] * d[q - 1] + cf[2][q] * d[q - L] +
                               cf[3][q] * d[q + L];
                    if (ncol != 2)
                        nb += cf[4][q] * d[q + 2] + cf[5][q] * d[q - 2] + cf[6][q] * d[q - 2 * L] +
                              cf[7][q] * d[q + 2 * L];
                    d[q] = w[q] * (r0[q] + nb) + oma * d[q];
                }
            }
        }
#pragma omp parallel for schedule(static)
    for (int i = m.i0; i <= m.i1; ++i)
        for (int j = m.j0; j <= m.j1; ++j) {
            const size_t q = static_cast<size_t>(i) * nc + j;
            x[q] = x[q] + static_cast<double>(d[q]);
        }
}

void mom_sweep(or_state* s, const Mom& m, int kind, int nsweeps) {
    if (kind == SG_MOM_JACOBI)
        mom_jacobi(s, m, nsweeps);
    else if (kind == SG_MOM_GS_MULTICOLOUR)
        mom_multicolour(s, m, nsweeps);
    else
        mom_gs_lex(s, m, nsweeps);
}

// ---- pressure correction (src/pressure.cpp) ----

// pc_update: two-statement update of SPEC 4.3 step 4 with neighbours E, W, S, N from src
// (src/pressure.cpp:70-75).
inline void pc_update(or_state* s, double* dst, const double* src, size_t q) {
    const int nc = s->nc;
    const double* AE = s->a(SG_F_AEP);
    const double* AW = s->a(SG_F_AWP);
    const double* AN = s->a(SG_F_ANP);
    const double* AS = s->a(SG_F_ASP);
    const double* AP = s->a(SG_F_APP);
    const double* b = s->a(SG_F_B);
    dst[q] = b[q] / AP[q];
    dst[q] += (AE[q] * src[q + 1] + AW[q] * src[q - 1] + AS[q] * src[q + nc] + AN[q] * src[q - nc]) /
              AP[q];
}

void fill0(Vec& v) { std::fill(v.begin(), v.end(), 0.0); }

// solve_pc_gs_lex, src/pressure.cpp:56-67.
void pc_gs_lex(or_state* s) {
    const sg_params& p = s->p;
    fill0(s->f[SG_F_PC]);
    double* x = s->a(SG_F_PC);
    for (int k = 0; k < p.maxit_inner_p; ++k)
        for (int i = p.NX + p.H - 1; i >= p.H; --i)
            for (int j = p.H + 1; j <= p.NY + p.H; ++j)
                pc_update(s, x, x, static_cast<size_t>(i) * s->nc + j);
}

// solve_pc_jacobi, damped by 0.8 (src/pressure.cpp:77-112, include/simple/pressure.hpp:22).
void pc_jacobi(or_state* s) {
    const sg_params& p = s->p;
    const double w = 0.8;
    const int n = p.maxit_inner_p;
    fill0(s->f[SG_F_PC]);
    double* x = s->a(SG_F_PC);
    double* old = s->scratch.data();
    std::memcpy(old, x, sizeof(double) * s->f[SG_F_PC].size());
    for (int k = 0; k < n; ++k) {
        const bool to_pc = (n - k) % 2 == 1;
        const double* src = to_pc ? old : x;
        double* dst = to_pc ? x : old;
#pragma omp parallel for schedule(static)
        for (int i = p.NX + p.H - 1; i >= p.H; --i)
            for (int j = p.H + 1; j <= p.NY + p.H; ++j) {
                const size_t q = static_cast<size_t>(i) * s->nc + j;
                pc_update(s, dst, src, q);
                dst[q] = w * dst[q] + (1 - w) * src[q];
            }
    }
}

// solve_pc_redblack: colour (i + j) & 1, colour 0 first (src/pressure.cpp:114-134).
void pc_redblack(or_state* s) {
    const sg_params& p = s->p;
    fill0(s->f[SG_F_PC]);
    double* x = s->a(SG_F_PC);
    for (int k = 0; k < p.maxit_inner_p; ++k)
        for (int c = 0; c < 2; ++c) {
#pragma omp parallel for schedule(static)
            for (int i = p.NX + p.H - 1; i >= p.H; --i)
                for (int j = p.H + 1 + ((i + p.H + 1 + c) & 1); j <= p.NY + p.H; j += 2)
                    pc_update(s, x, x, static_cast<size_t>(i) * s->nc + j);
        }
}

// apply_A_p, src/pressure.cpp:165-173.
inline double apply_A(or_state* s, const double* x, size_t q) {
    const int nc = s->nc;
    return s->f[SG_F_APP][q] * x[q] -
           (s->f[SG_F_AEP][q] * x[q + 1] + s->f[SG_F_AWP][q] * x[q - 1] +
            s->f[SG_F_ASP][q] * x[q + nc] + s->f[SG_F_ANP][q] * x[q - nc]);
}

// Jacobi-preconditioned CG on the mean-projected system (src/pressure.cpp:208-317).
void pc_cg(or_state* s) {
    const sg_params& p = s->p;
    const int H = p.H, nc = s->nc;
    const int i0 = H, i1 = p.NX + H - 1, j0 = H + 1, j1 = p.NY + H;
    const double n = static_cast<double>(p.NX) * p.NY;
    double *r = s->cr.data(), *z = s->cz.data(), *d = s->cd.data(), *q = s->cq.data();
    double* x = s->a(SG_F_PC);
    const double* b = s->a(SG_F_B);
    const double* APP = s->a(SG_F_APP);
    s->cg_last = 0;
    s->cg_rel = 0.0;
    fill0(s->f[SG_F_PC]);
    auto Q = [nc](int i, int j) { return static_cast<size_t>(i) * nc + j; };

    const double sb = ordered_sum(i0, i1, j0, j1, [&](int i, int j) { return b[Q(i, j)]; });
    const double mb = sb / n;
    double s4[4];
    ordered_sums<4>(i0, i1, j0, j1,
                    [&](int i, int j, double* v) {
                        const size_t c = Q(i, j);
                        const double ri = b[c] - mb;
                        r[c] = ri;
                        const double zi = ri / APP[c];
                        z[c] = zi;
                        v[0] = ri * ri;
                        v[1] = zi;
                        v[2] = ri * zi;
                        v[3] = ri;
                    },
                    s4);
    double rr = s4[0], sz = s4[1], rz = s4[2], sr = s4[3];
    if (rr == 0.0)
        return;
    const double nb = std::sqrt(rr);
    double mz = sz / n;
    rz -= mz * sr;
#pragma omp parallel for schedule(static)
    for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j)
            d[Q(i, j)] = z[Q(i, j)] - mz;
    int k = 0;
    double rel = 1.0;
    while (k < p.pressure_maxit) {
        double dq = 0.0;
        ordered_sums<1>(i0, i1, j0, j1,
                        [&](int i, int j, double* v) {
                            const size_t c = Q(i, j);
                            const double qi = apply_A(s, d, c);
                            q[c] = qi;
                            v[0] = d[c] * qi;
                        },
                        &dq);
        if (!(dq > 0.0))
            break;
        const double alpha = rz / dq;
        ordered_sums<4>(i0, i1, j0, j1,
                        [&](int i, int j, double* v) {
                            const size_t c = Q(i, j);
                            x[c] += alpha * d[c];
                            const double ri = r[c] - alpha * q[c];
                            r[c] = ri;
                            const double zi = ri / APP[c];
                            z[c] = zi;
                            v[0] = ri * ri;
                            v[1] = zi;
                            v[2] = ri * zi;
                            v[3] = ri;
                        },
                        s4);
        rr = s4[0];
        sz = s4[1];
        double rzn = s4[2];
        sr = s4[3];
        ++k;
        rel = std::sqrt(rr) / nb;
        if (rel <= p.pressure_tolerance)
            break;
        mz = sz / n;
        rzn -= mz * sr;
        const double beta = rzn / rz;
        rz = rzn;
#pragma omp parallel for schedule(static)
        for (int i = i0; i <= i1; ++i)
            for (int j = j0; j <= j1; ++j)
                d[Q(i, j)] = (z[Q(i, j)] - mz) + beta * d[Q(i, j)];
    }
    const double sx = ordered_sum(i0, i1, j0, j1, [&](int i, int j) { return x[Q(i, j)]; });
    const double mx = sx / n;
#pragma omp parallel for schedule(static)
    for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j)
            x[Q(i, j)] -= mx;
    s->cg_last = k;
    s->cg_rel = rel;
}

// ---- multigrid (src/multigrid.cpp) ----

// Level dimensions: halve while both sides are even and the halves stay >= the coarsest side
// (4 in the reference, src/multigrid.cpp:143-151).
std::vector<std::pair<int, int>> level_dims(int nx, int ny, int coarsest) {
    std::vector<std::pair<int, int>> d{{nx, ny}};
    while (nx % 2 == 0 && ny % 2 == 0 && nx / 2 >= coarsest && ny / 2 >= coarsest) {
        nx /= 2;
        ny /= 2;
        d.emplace_back(nx, ny);
    }
    return d;
}

// Level storage (MGWork, src/multigrid.cpp:153-185; mg_view :187-220): level 0 aliases the p'
// arrays at interior cell (H, H+1) with stride ncols, and owns only r; coarse levels are padded
// (nx+2) x (ny+2) zero-halo arrays.
void mg_setup(or_state* s) {
    const sg_params& p = s->p;
    const int coarsest = p.mg_coarsest > 0 ? p.mg_coarsest : 4;
    const auto dims = level_dims(p.NX, p.NY, coarsest);
    s->mg.clear();
    s->mg.resize(dims.size());
    s->r0.assign(static_cast<size_t>(s->nr) * s->nc, 0.0);
    for (size_t l = 0; l < dims.size(); ++l) {
        Level& L = s->mg[l];
        L.nx = dims[l].first;
        L.ny = dims[l].second;
        if (l == 0) {
            L.stride = s->nc;
            const size_t base = static_cast<size_t>(p.H) * s->nc + p.H + 1;
            L.AE = s->a(SG_F_AEP) + base;
            L.AW = s->a(SG_F_AWP) + base;
            L.AN = s->a(SG_F_ANP) + base;
            L.AS = s->a(SG_F_ASP) + base;
            L.AP = s->a(SG_F_APP) + base;
            L.b = s->a(SG_F_B) + base;
            L.x = s->a(SG_F_PC) + base;
            L.r = s->r0.data() + base;
        } else {
            L.stride = L.ny + 2;
            const size_t n = static_cast<size_t>(L.nx + 2) * (L.ny + 2);
            for (auto& v : L.store)
                v.assign(n, 0.0);
            const size_t base = L.stride + 1;
            double** ptr[8] = {&L.AE, &L.AW, &L.AN, &L.AS, &L.AP, &L.b, &L.x, &L.r};
            for (int k = 0; k < 8; ++k)
                *ptr[k] = L.store[k].data() + base;
        }
    }
}

// det_sum on a level (src/multigrid.cpp:18-34).
template <class F> double det_sum(const Level& L, F f) {
    Vec rows(L.nx);
    for (int I = 0; I < L.nx; ++I) {
        double s = 0.0;
        for (int J = 0; J < L.ny; ++J)
            s += f(I * L.stride + J);
        rows[I] = s;
    }
    double s = 0.0;
    for (int I = 0; I < L.nx; ++I)
        s += rows[I];
    return s;
}

// Red-black smoother, colours (I + J) & 1 in level indices, colour 0 first, two-statement update
// with neighbours E, W, S, N (src/multigrid.cpp:46-70).
void smooth(const Level& L, int sweeps) {
    const int s = L.stride;
    double* x = L.x;
    for (int k = 0; k < sweeps; ++k)
        for (int c = 0; c < 2; ++c) {
#pragma omp parallel for schedule(static) if (L.nx * L.ny >= 4096)
            for (int I = L.nx - 1; I >= 0; --I)
                for (int J = (I + c) & 1; J < L.ny; J += 2) {
                    const int q = I * s + J;
                    x[q] = L.b[q] / L.AP[q];
                    x[q] += (L.AE[q] * x[q + 1] + L.AW[q] * x[q - 1] + L.AS[q] * x[q + s] +
                             L.AN[q] * x[q - s]) /
                            L.AP[q];
                }
        }
}

// r = b - (AP x - (E + W + S + N)), src/multigrid.cpp:73-87.
void residual(const Level& L) {
    const int s = L.stride;
    const double* x = L.x;
#pragma omp parallel for schedule(static) if (L.nx * L.ny >= 4096)
    for (int I = L.nx - 1; I >= 0; --I)
        for (int J = 0; J < L.ny; ++J) {
            const int q = I * s + J;
            L.r[q] = L.b[q] - (L.AP[q] * x[q] - (L.AE[q] * x[q + 1] + L.AW[q] * x[q - 1] +
                                                 L.AS[q] * x[q + s] + L.AN[q] * x[q - s]));
        }
}

// Coarse b = NW + NE + SW + SE fine residuals, coarse x = 0 (src/multigrid.cpp:90-103).
void restrict_to(const Level& F, const Level& C) {
    const int fs = F.stride, cs = C.stride;
#pragma omp parallel for schedule(static) if (C.nx * C.ny >= 4096)
    for (int I = C.nx - 1; I >= 0; --I)
        for (int J = 0; J < C.ny; ++J) {
            const int f = 2 * I * fs + 2 * J;
            C.b[I * cs + J] = F.r[f] + F.r[f + 1] + F.r[f + fs] + F.r[f + fs + 1];
            C.x[I * cs + J] = 0.0;
        }
}

// Fine x += omega * coarse x on the 4 fine cells (src/multigrid.cpp:106-116).
void prolong(const Level& C, const Level& F, double omega) {
    const int fs = F.stride, cs = C.stride;
#pragma omp parallel for schedule(static) if (F.nx * F.ny >= 4096)
    for (int i = F.nx - 1; i >= 0; --i)
        for (int j = 0; j < F.ny; ++j)
            F.x[i * fs + j] += omega * C.x[(i / 2) * cs + j / 2];
}

// mg_remove_mean, src/multigrid.cpp:263-276.
void remove_mean(const Level& L, double* x) {
    const double m = det_sum(L, [x](int q) { return x[q]; }) / (static_cast<double>(L.nx) * L.ny);
    for (int I = L.nx - 1; I >= 0; --I)
        for (int J = 0; J < L.ny; ++J)
            x[I * L.stride + J] -= m;
}

// coarse_solve, src/multigrid.cpp:120-140: mean-free b, red-black sweeps until
// ||r|| < 1e-10 ||b|| or 200 sweeps, mean-free x.
int coarse_solve(const Level& L) {
    remove_mean(L, L.b);
    const double* b = L.b;
    const double* r = L.r;
    const double bb = det_sum(L, [b](int q) { return b[q] * b[q]; });
    int sweeps = 0;
    if (bb > 0.0) {
        const double bn = std::sqrt(bb);
        while (sweeps < 200) {
            smooth(L, 1);
            ++sweeps;
            residual(L);
            const double rr = det_sum(L, [r](int q) { return r[q] * r[q]; });
            if (std::sqrt(rr) < 1e-10 * bn)
                break;
        }
    }
    remove_mean(L, L.x);
    return sweeps;
}

// mg_build: Galerkin 2x2-aggregation links, AP = ae + aw + an + as (src/multigrid.cpp:232-261).
void mg_build(or_state* s) {
    for (size_t l = 1; l < s->mg.size(); ++l) {
        const Level& F = s->mg[l - 1];
        const Level& C = s->mg[l];
        const int fs = F.stride, cs = C.stride;
#pragma omp parallel for schedule(static) if (C.nx * C.ny >= 4096)
        for (int I = C.nx - 1; I >= 0; --I)
            for (int J = 0; J < C.ny; ++J) {
                const int f = 2 * I * fs + 2 * J, c = I * cs + J;
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
    }
}

// mg_vcycle, src/multigrid.cpp:299-323.
void mg_vcycle(or_state* s) {
    const sg_params& p = s->p;
    const int nl = static_cast<int>(s->mg.size());
    for (int l = 0; l + 1 < nl; ++l) {
        smooth(s->mg[l], p.mg_pre);
        residual(s->mg[l]);
        restrict_to(s->mg[l], s->mg[l + 1]);
    }
    s->coarse_sweeps = coarse_solve(s->mg[nl - 1]);
    for (int l = nl - 2; l >= 0; --l) {
        prolong(s->mg[l + 1], s->mg[l], p.mg_omega);
        smooth(s->mg[l], p.mg_post);
    }
}

// remove_mean0: interior mean of a level-0 field in CellSums order (src/multigrid.cpp:327-345).
void remove_mean0(or_state* s, int id) {
    const sg_params& p = s->p;
    const int H = p.H, nc = s->nc;
    double* f = s->a(id);
    const double m = ordered_sum(H, p.NX + H - 1, H + 1, p.NY + H,
                                 [&](int i, int j) { return f[static_cast<size_t>(i) * nc + j]; }) /
                     (static_cast<double>(p.NX) * p.NY);
#pragma omp parallel for schedule(static)
    for (int i = p.NX + H - 1; i >= H; --i)
        for (int j = H + 1; j <= p.NY + H; ++j)
            f[static_cast<size_t>(i) * nc + j] -= m;
}

// solve_pc_mg, src/multigrid.cpp:347-357.
void pc_mg(or_state* s) {
    mg_build(s);
    fill0(s->f[SG_F_PC]);
    remove_mean0(s, SG_F_B);
    for (int c = 0; c < s->p.mg_cycles; ++c)
        mg_vcycle(s);
    remove_mean0(s, SG_F_PC);
}

void solve_p_kind(or_state* s, int kind) {
    switch (kind) {
    case SG_P_JACOBI: pc_jacobi(s); break;
    case SG_P_GS_REDBLACK: pc_redblack(s); break;
    case SG_P_CG:
        pc_cg(s);
        s->cg_total += s->cg_last;
        break;
    case SG_P_MG: pc_mg(s); break;
    default: pc_gs_lex(s);
    }
}

} // namespace

extern "C" {

or_state* or_create(const sg_params* p) {
    or_state* s = new or_state;
    s->p = *p;
    s->nr = p->nrows;
    s->nc = p->ncols;
    const size_t n = static_cast<size_t>(s->nr) * s->nc;
    for (auto& v : s->f)
        v.assign(n, 0.0);
    s->scratch.assign(n, 0.0);
    s->cr.assign(n, 0.0);
    s->cz.assign(n, 0.0);
    s->cd.assign(n, 0.0);
    s->cq.assign(n, 0.0);
    mg_setup(s);
    return s;
}

void or_destroy(or_state* s) { delete s; }
// Replace the scalar parameters (sizes must not change): tests vary tolerances and counts.
int or_set_params(or_state* s, const sg_params* p) {
    if (p->NX != s->p.NX || p->NY != s->p.NY || p->H != s->p.H || p->mg_coarsest != s->p.mg_coarsest)
        return 1;
    s->p = *p;
    return 0;
}
double* or_field(or_state* s, int id) { return (id >= 0 && id < SG_F_COUNT) ? s->a(id) : nullptr; }
int or_nrows(or_state* s) { return s->nr; }
int or_ncols(or_state* s) { return s->nc; }

void or_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int or_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// assemble_u, src/momentum_u.cpp:8-22: rows [H, NX+H-1], columns [H, NY+H].
void or_assemble_u(or_state* s) {
    const sg_params& p = s->p;
    const int H = p.H, nc = s->nc;
    const double* u = s->a(SG_F_U);
    const double* v = s->a(SG_F_V);
#pragma omp parallel for schedule(static)
    for (int i = p.NX + H - 1; i >= H; --i)
        for (int j = H; j <= p.NY + H; ++j) {
            const size_t q = static_cast<size_t>(i) * nc + j;
            const double ue = (u[q] + u[q + 1]) / 2;
            const double uw = (u[q] + u[q - 1]) / 2;
            const double vn = (v[q] + v[q + 1]) / 2;
            const double vs = (v[q + nc] + v[q + nc + 1]) / 2;
            store9(s, SG_F_CU_AE, i, j, convective(p, ue, uw, vn, vs, i, j), diffusive(p, i, j));
        }
}

// apply_bc_u, src/momentum_u.cpp:24-34: lid ghost 2 - u, south ghost -u.
void or_bc_u(or_state* s) {
    const sg_params& p = s->p;
    const int H = p.H;
    for (int j = H + 1; j <= p.NY + H; ++j)
        s->at(SG_F_U, H - 1, j) = 2 - s->at(SG_F_U, H, j);
    for (int j = H + 1; j <= p.NY + H; ++j)
        s->at(SG_F_U, p.NX + H, j) = -s->at(SG_F_U, p.NX + H - 1, j);
}

double or_residual_u(or_state* s) { return mom_residual(s, mom_u(s)); }
void or_sweep_u(or_state* s, int n) { mom_sweep(s, mom_u(s), s->p.momentum_solver, n); }
void or_sweep_u_dc(or_state* s, int n) { mom_dc(s, mom_u(s), n); }
void or_sweep_u_kind(or_state* s, int kind, int n) { mom_sweep(s, mom_u(s), kind, n); }

// assemble_v, src/momentum_v.cpp:9-22: rows [H, NX+H+1], columns [H, NY+H+1].
void or_assemble_v(or_state* s) {
    const sg_params& p = s->p;
    const int H = p.H, nc = s->nc;
    const double* u = s->a(SG_F_U);
    const double* v = s->a(SG_F_V);
#pragma omp parallel for schedule(static)
    for (int i = p.NX + H + 1; i >= H; --i)
        for (int j = H; j <= p.NY + H + 1; ++j) {
            const size_t q = static_cast<size_t>(i) * nc + j;
            const double ue = (u[q - nc] + u[q]) / 2;
            const double uw = (u[q - 1] + u[q - nc - 1]) / 2;
            const double vn = (v[q] + v[q - nc]) / 2;
            const double vs = (v[q] + v[q + nc]) / 2;
            store9(s, SG_F_CV_AE, i, j, convective(p, ue, uw, vn, vs, i, j), diffusive(p, i, j));
        }
}

// apply_bc_v, src/momentum_v.cpp:24-34: west and east ghosts -v.
void or_bc_v(or_state* s) {
    const sg_params& p = s->p;
    const int H = p.H;
    for (int i = H; i <= p.NX + H - 1; ++i)
        s->at(SG_F_V, i, H) = -s->at(SG_F_V, i, H + 1);
    for (int i = H; i <= p.NX + H - 1; ++i)
        s->at(SG_F_V, i, p.NY + H + 1) = -s->at(SG_F_V, i, p.NY + H);
}

double or_residual_v(or_state* s) { return mom_residual(s, mom_v(s)); }
void or_sweep_v(or_state* s, int n) { mom_sweep(s, mom_v(s), s->p.momentum_solver, n); }
void or_sweep_v_dc(or_state* s, int n) { mom_dc(s, mom_v(s), n); }
void or_sweep_v_kind(or_state* s, int kind, int n) { mom_sweep(s, mom_v(s), kind, n); }

// assemble_p, src/pressure.cpp:14-54: links, zeroed wall links, APP, mass source b, res_p.
double or_assemble_p(or_state* s) {
    const sg_params& p = s->p;
    const int H = p.H, nc = s->nc;
    const double dx = p.dx, dy = p.dy;
    const double* APU = s->a(SG_F_CU_AP);
    const double* APV = s->a(SG_F_CV_AP);
    const double* u = s->a(SG_F_U);
    const double* v = s->a(SG_F_V);
    double *AE = s->a(SG_F_AEP), *AW = s->a(SG_F_AWP), *AN = s->a(SG_F_ANP), *AS = s->a(SG_F_ASP);
    double *AP = s->a(SG_F_APP), *b = s->a(SG_F_B);
#pragma omp parallel for schedule(static)
    for (int i = H; i <= p.NX + H - 1; ++i)
        for (int j = H + 1; j <= p.NY + H; ++j) {
            const size_t q = static_cast<size_t>(i) * nc + j;
            AE[q] = dy * dy / APU[q];
            AW[q] = dy * dy / APU[q - 1];
            AN[q] = dx * dx / APV[q];
            AS[q] = dx * dx / APV[q + nc];
        }
    for (int i = 0; i < s->nr; ++i) {
        AW[static_cast<size_t>(i) * nc + H + 1] = 0.0;
        AE[static_cast<size_t>(i) * nc + p.NY + H] = 0.0;
    }
    for (int j = 0; j < nc; ++j) {
        AS[static_cast<size_t>(p.NX + H - 1) * nc + j] = 0.0;
        AN[static_cast<size_t>(H) * nc + j] = 0.0;
    }
    return ordered_sum(H, p.NX + H - 1, H + 1, p.NY + H, [&](int i, int j) {
        const size_t q = static_cast<size_t>(i) * nc + j;
        AP[q] = AE[q] + AW[q] + AN[q] + AS[q];
        b[q] = -dx * (-v[q + nc] + v[q]) - dy * (-u[q - 1] + u[q]);
        return std::fabs(b[q]);
    });
}

void or_solve_p(or_state* s) { solve_p_kind(s, s->p.pressure_solver); }
void or_solve_p_kind(or_state* s, int kind) { solve_p_kind(s, kind); }
int or_cg_last_iterations(or_state* s) { return s->cg_last; }
double or_cg_last_rel(or_state* s) { return s->cg_rel; }
long long or_cg_total_iterations(or_state* s) { return s->cg_total; }

// correct, src/pressure.cpp:136-157: velocity corrections with no relaxation, p += alpha_p p'.
void or_correct(or_state* s) {
    const sg_params& p = s->p;
    const int H = p.H, nc = s->nc;
    const double dy = p.dy, dx = p.dx, ap = p.relaxation_factor_p;
    const double* pc = s->a(SG_F_PC);
    const double* APU = s->a(SG_F_CU_AP);
    const double* APV = s->a(SG_F_CV_AP);
    double *u = s->a(SG_F_U), *v = s->a(SG_F_V), *pf = s->a(SG_F_P);
#pragma omp parallel for schedule(static)
    for (int i = H; i <= p.NX + H - 1; ++i)
        for (int j = H + 1; j <= p.NY + H - 1; ++j) {
            const size_t q = static_cast<size_t>(i) * nc + j;
            u[q] += dy * (pc[q] - pc[q + 1]) / APU[q];
        }
#pragma omp parallel for schedule(static)
    for (int i = H + 1; i <= p.NX + H - 1; ++i)
        for (int j = H + 1; j <= p.NY + H; ++j) {
            const size_t q = static_cast<size_t>(i) * nc + j;
            v[q] += dx * (pc[q] - pc[q - nc]) / APV[q];
        }
#pragma omp parallel for schedule(static)
    for (int i = H; i <= p.NX + H - 1; ++i)
        for (int j = H + 1; j <= p.NY + H; ++j) {
            const size_t q = static_cast<size_t>(i) * nc + j;
            pf[q] += ap * pc[q];
        }
}

// outer_iteration, src/solver.cpp:27-93.
void or_outer_iteration(or_state* s, double res[3]) {
    const bool dc = s->p.momentum_precision == SG_MOMP_MIXED;
    or_assemble_u(s);
    or_bc_u(s);
    res[0] = or_residual_u(s);
    if (dc)
        or_sweep_u_dc(s, s->p.maxit_inner_u);
    else
        or_sweep_u(s, s->p.maxit_inner_u);
    or_assemble_v(s);
    or_bc_v(s);
    res[1] = or_residual_v(s);
    if (dc)
        or_sweep_v_dc(s, s->p.maxit_inner_v);
    else
        or_sweep_v(s, s->p.maxit_inner_v);
    res[2] = or_assemble_p(s);
    or_solve_p(s);
    or_correct(s);
}

// run_outer_loop, src/solver.cpp:106-132: stop when max of the tested residuals <= tolerance and
// iter > 1; std::max of an initializer list keeps the first of equal/unordered values.
int or_run(or_state* s, double* log, int* converged) {
    const sg_params& p = s->p;
    double mu = 0.0, mv = 0.0, mp = 0.0;
    *converged = 0;
    int iters = 0;
    for (int iter = 1; iter <= p.maxit; ++iter) {
        double r[3];
        or_outer_iteration(s, r);
        if (log) {
            log[4 * (iter - 1) + 0] = iter;
            log[4 * (iter - 1) + 1] = r[0];
            log[4 * (iter - 1) + 2] = r[1];
            log[4 * (iter - 1) + 3] = r[2];
        }
        iters = iter;
        double tu = r[0], tv = r[1], tp = r[2];
        if (p.residual_normalisation == SG_NORM_RUNNING_MAX) {
            mu = std::max(mu, r[0]);
            mv = std::max(mv, r[1]);
            mp = std::max(mp, r[2]);
            tu = mu > 0.0 ? r[0] / mu : 0.0;
            tv = mv > 0.0 ? r[1] / mv : 0.0;
            tp = mp > 0.0 ? r[2] / mp : 0.0;
        }
        if (std::max({tu, tv, tp}) <= p.error_tolerance && iter > 1) {
            *converged = 1;
            break;
        }
    }
    return iters;
}

double or_run_fixed(or_state* s, int n, double* last) {
    const auto t0 = std::chrono::steady_clock::now();
    double r[3] = {0, 0, 0};
    for (int k = 0; k < n; ++k)
        or_outer_iteration(s, r);
    if (last)
        for (int k = 0; k < 3; ++k)
            last[k] = r[k];
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Per-cell coefficients of src/coefficients.cpp: which = 0 convective, 1 diffusive, 2 sum.
void or_coefficients(const sg_params* p, double ue, double uw, double vn, double vs, int i, int j,
                     int which, double out[9]) {
    const C9 c = convective(*p, ue, uw, vn, vs, i, j), d = diffusive(*p, i, j);
    const double cv[9] = {c.e, c.w, c.n, c.s, c.ee, c.ww, c.nn, c.ss, c.p};
    const double dv[9] = {d.e, d.w, d.n, d.s, d.ee, d.ww, d.nn, d.ss, d.p};
    for (int k = 0; k < 9; ++k)
        out[k] = which == 0 ? cv[k] : which == 1 ? dv[k] : cv[k] + dv[k];
}
int or_in_band(const sg_params* p, int i, int j) { return band(*p, i, j) ? 1 : 0; }

int or_mg_nlevels(or_state* s) { return static_cast<int>(s->mg.size()); }
void or_mg_dims(or_state* s, int l, int* nx, int* ny) {
    *nx = s->mg[l].nx;
    *ny = s->mg[l].ny;
}
void or_mg_build(or_state* s) { mg_build(s); }
void or_mg_get(or_state* s, int l, int k, double* out) {
    const Level& L = s->mg[l];
    const double* src[8] = {L.AE, L.AW, L.AN, L.AS, L.AP, L.b, L.x, L.r};
    for (int I = 0; I < L.nx; ++I)
        for (int J = 0; J < L.ny; ++J)
            out[I * L.ny + J] = src[k][I * L.stride + J];
}
void or_mg_vcycle(or_state* s) { mg_vcycle(s); }
int or_mg_last_coarse_sweeps(or_state* s) { return s->coarse_sweeps; }
void or_apply_A_p(or_state* s, const double* x, double* out) {
    const sg_params& p = s->p;
#pragma omp parallel for schedule(static)
    for (int i = p.H; i <= p.NX + p.H - 1; ++i)
        for (int j = p.H + 1; j <= p.NY + p.H; ++j) {
            const size_t q = static_cast<size_t>(i) * s->nc + j;
            out[q] = apply_A(s, x, q);
        }
}

} // extern "C"
