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
// Sweeps. One launch per multicolour sweep (all colours): a workgroup loads a tile of the field
// plus a halo of 2*ncol cells into LDS and runs the colour passes there, pass k over the tile
// grown by 2*(ncol-1-k) cells (the cells whose new values later passes of the tile still read),
// so it reproduces the colour-by-colour sweep exactly; the tile's cells are then written to a
// second array (tiles read their halo from the first), and sweeps alternate between the field and
// a scratch copy. The first sweep of an outer iteration also evaluates the residual from the
// loaded field before any update.
//
// Arithmetic is the reference's expression for expression (compiled with -ffp-contract=off, IEEE
// division), so every value is bit-identical to src/momentum_u.cpp / momentum_v.cpp; the
// pressure term (p(i,j) - p(i,j+1)) * dy is evaluated once, at assembly, with the same operands.
#include "sg_internal.hpp"

#include <utility>

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
    double* ubc; // u only: also set u's ghost rows (bc_u) in this launch, or null
    int uds, order;
    C9 d2, d4;
};

// One control volume: convection (src/coefficients.cpp:15-39) plus diffusion (:41-63,
// precomputed on the host as d2 / d4), summed element-wise (:5-8); face velocities of
// src/momentum_u.cpp:13-16 or src/momentum_v.cpp:14-17. Written at packed entry b (and AP
// row-major at q).
__device__ __forceinline__ void assemble_cell(const AsmArgs& a, int i, int j, size_t q, double* b,
                                              double ue, double uw, double vn, double vs,
                                              double pt) {
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
    const size_t as = a.pk.astride;
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
    b[9 * as] = pt;
}

// Colour-major: thread (x, y) takes the kk-th cell of every colour in row i, so each store to a
// packed array is one contiguous run per wave. All loads of the thread's cells come first. AY
// rows per workgroup: the rows above and below that a cell's faces read are mostly its
// neighbours' own rows, already in L1.
constexpr int AY = 16;
template <bool IS_U> __device__ __forceinline__ void k_assemble(const AsmArgs& a, const Blk& bk) {
    const Pack& pk = a.pk;
    const int kk = bk.x * BX + threadIdx.x;
    const int i = pk.ai0 + bk.y * AY + threadIdx.y;
    if (i > pk.ai1)
        return;
    const long ld = a.ld;
    constexpr int MC = 5;
    double fue[MC], fuw[MC], fvn[MC], fvs[MC], fpt[MC];
    int jj[MC];
#pragma unroll
    for (int c = 0; c < MC; ++c) {
        const int j = jfirst(pk, c < pk.ncol ? c : 0, i) + pk.ncol * kk;
        jj[c] = (c < pk.ncol && j <= pk.aj1) ? j : -1;
        if (jj[c] < 0)
            continue;
        const size_t q = static_cast<size_t>(i) * ld + j;
        if (IS_U) {
            fue[c] = (a.u[q] + a.u[q + 1]) / 2;
            fuw[c] = (a.u[q] + a.u[q - 1]) / 2;
            fvn[c] = (a.v[q] + a.v[q + 1]) / 2;
            fvs[c] = (a.v[q + ld] + a.v[q + ld + 1]) / 2;
        } else {
            fue[c] = (a.u[q - ld] + a.u[q]) / 2;
            fuw[c] = (a.u[q - 1] + a.u[q - ld - 1]) / 2;
            fvn[c] = (a.v[q] + a.v[q - ld]) / 2;
            fvs[c] = (a.v[q] + a.v[q + ld]) / 2;
        }
        fpt[c] = (a.pf[q] - a.pf[q + a.poff]) * a.h;
    }
#pragma unroll
    for (int c = 0; c < MC; ++c) {
        if (jj[c] < 0)
            continue;
        const size_t q = static_cast<size_t>(i) * ld + jj[c];
        assemble_cell(a, i, jj[c], q, pk.base + pidx(pk, c, i, kk), fue[c], fuw[c], fvn[c],
                      fvs[c], fpt[c]);
    }
    // Ghost rows of u (src/momentum_u.cpp:24-34), which the assembly above does not read: the
    // threads of the first row take the lid, those of the second the south wall, for the
    // columns of their cells.
    if (IS_U && a.ubc && i - pk.ai0 < 2) {
        double* u = a.ubc;
        const int H = a.H;
#pragma unroll
        for (int c = 0; c < MC; ++c) {
            const int j = jj[c];
            if (j < H + 1 || j > a.NY + H)
                continue;
            if (i == pk.ai0)
                u[static_cast<size_t>(H - 1) * ld + j] = 2 - u[static_cast<size_t>(H) * ld + j];
            else
                u[static_cast<size_t>(a.NX + H) * ld + j] =
                    -u[static_cast<size_t>(a.NX + H - 1) * ld + j];
        }
    }
}

// Ghost cells: u lid 2 - u and south -u (src/momentum_u.cpp:24-34); v west/east -v
// (src/momentum_v.cpp:24-34).
struct BcArgs {
    double* x;
    int ld, H, NX, NY;
};
constexpr int BC_T = 256;

__device__ __forceinline__ void k_bc_u(const BcArgs& a, const Blk& bk) {
    double* u = a.x;
    const int ld = a.ld, H = a.H, NX = a.NX, NY = a.NY;
    const int j = H + 1 + bk.x * BC_T + threadIdx.x;
    if (j > NY + H)
        return;
    u[static_cast<size_t>(H - 1) * ld + j] = 2 - u[static_cast<size_t>(H) * ld + j];
    u[static_cast<size_t>(NX + H) * ld + j] = -u[static_cast<size_t>(NX + H - 1) * ld + j];
}

__device__ __forceinline__ void k_bc_v(const BcArgs& a, const Blk& bk) {
    double* v = a.x;
    const int ld = a.ld, H = a.H, NX = a.NX, NY = a.NY;
    const int i = H + bk.x * BC_T + threadIdx.x;
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
__device__ __forceinline__ bool packed_cell(const MomArgs& m, const Blk& bk, int c, int& i,
                                            int& j, size_t& o) {
    i = m.i0 + bk.y * BY + threadIdx.y;
    if (i > m.i1)
        return false;
    const int jf = jfirst(m.pk, c, i);
    const int k = (jf < m.j0 ? 1 : 0) + bk.x * BX + threadIdx.x;
    j = jf + m.pk.ncol * k;
    if (j > m.j1)
        return false;
    o = pidx(m.pk, c, i, k);
    return true;
}

// |R| per cell in the residual range, R = (nb + pterm) - AP x (src/momentum_u.cpp:48-50), into a
// compact buffer (one thread per cell; the per-phase entry point and sweepless iterations).
struct ResArgs {
    MomArgs m;
    double* cell;
};

__device__ __forceinline__ void k_mom_residual(const ResArgs& ra, const Blk& bk) {
    const MomArgs& m = ra.m;
    double* cell = ra.cell;
    const int j = m.j0 + bk.x * BX + threadIdx.x, i = m.i0 + bk.y * BY + threadIdx.y;
    if (i > m.i1 || j > m.j1)
        return;
    const long ld = m.ld;
    const size_t q = static_cast<size_t>(i) * ld + j;
    const Pack& pk = m.pk;
    const int c = colour_of(pk.ncol, i, j);
    const double* co = pk.base + pidx(pk, c, i, (j - jfirst(pk, c, i)) / pk.ncol);
    const double R = nb9(co, pk.astride, m.x, q, ld) + co[9 * pk.astride] -
                     co[8 * pk.astride] * m.x[q];
    cell[static_cast<size_t>(i - m.i0) * (m.j1 - m.j0 + 1) + (j - m.j0)] = fabs(R);
}

constexpr int ST = 1024; // sweep-tile workgroup size

// One multicolour sweep (src/momentum_u.cpp:123-143) on a TI x TJ tile, in LDS: src -> dst.
// first: also |R| of the tile's cells from the loaded field, into cell.
struct SweepArgs {
    MomArgs m;
    const double* src;
    double* dst;
    double* cell;
    int nr, nc, first;
};

// FIRST: a second LDS copy keeps the loaded (pre-sweep) values of the tile and the ring of R
// cells its stencils reach, and each core cell's residual |R| = |(nb + pterm) - AP x|
// (src/momentum_u.cpp:48-50) is evaluated from that copy during the cell's own colour pass, with
// the coefficients the update loads anyway. (Only that ring, not the whole halo: the two copies
// then fit two workgroups per CU.)
template <int NCOL, int TI, int TJ, bool FIRST, int NT = ST>
__device__ __forceinline__ void k_mom_sweep(const SweepArgs& sa, const Blk& bk) {
    const MomArgs& m = sa.m;
    const double* src = sa.src;
    double* dst = sa.dst;
    double* cell = sa.cell;
    const int nr = sa.nr, nc = sa.nc;
    constexpr int R = NCOL == 2 ? 1 : 2; // stencil reach
    constexpr int HL = R * NCOL;          // halo read by the first pass
    constexpr int LW = TJ + 2 * HL, LH = TI + 2 * HL;
    constexpr int W0 = TJ + 2 * R, H0 = TI + 2 * R; // pre-sweep copy: tile and stencil ring
    __shared__ double su[LH * LW];
    __shared__ double su0[FIRST ? H0 * W0 : 1];
    __shared__ double res[FIRST ? TI * TJ : 1]; // |R| of the core, written out row by row
    int tx, ty;
    xcd_tile(bk, tx, ty);
    const int ti0 = ty * TI, tj0 = tx * TJ;
    const int gi0 = ti0 - HL, gj0 = tj0 - HL; // array cell of su[0]
    const long ld = m.ld;
    // Batched loads: every load of a chunk is in flight before the first LDS store.
    constexpr int NL = (LH * LW + NT - 1) / NT;
    {
        double v[NL];
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int e = k * NT + threadIdx.x, a = e / LW, b = e - a * LW, i = gi0 + a, j = gj0 + b;
            v[k] = (e < LH * LW && i >= 0 && i < nr && j >= 0 && j < nc)
                       ? src[static_cast<size_t>(i) * ld + j]
                       : 0.0;
        }
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int e = k * NT + threadIdx.x;
            if (e < LH * LW) {
                su[e] = v[k];
                if (FIRST) {
                    const int a = e / LW - (HL - R), b = e % LW - (HL - R);
                    if (a >= 0 && a < H0 && b >= 0 && b < W0)
                        su0[a * W0 + b] = v[k];
                }
            }
        }
    }
    __syncthreads();
    const Pack& pk = m.pk;
    const size_t as = pk.astride;
    const int rw = m.j1 - m.j0 + 1;
#define SU(i, j) su[((i) - gi0) * LW + ((j) - gj0)]
#define S0(i, j) su0[((i) - (ti0 - R)) * W0 + ((j) - (tj0 - R))]
    for (int k = 0; k < NCOL; ++k) {
        const int d = R * (NCOL - 1 - k);
        const int ia = max(m.i0, ti0 - d), ib = min(m.i1, ti0 + TI - 1 + d);
        const int ja = max(m.j0, tj0 - d), jb = min(m.j1, tj0 + TJ - 1 + d);
        if (ia <= ib && ja <= jb) {
            const int cpr = (jb - ja + NCOL) / NCOL;
            const int n = (ib - ia + 1) * cpr;
            for (int e = threadIdx.x; e < n; e += NT) {
                const int r = e / cpr, t = e - r * cpr, i = ia + r;
                const int c0 = colour_of(NCOL, i, ja);
                const int j = ja + (NCOL == 2 ? ((k - c0) & 1) : (2 * (k - c0) + 10) % 5) + NCOL * t;
                if (j > jb)
                    continue;
                const double* co = pk.base + pidx(pk, k, i, (j - jfirst(pk, k, i)) / NCOL);
                const double ce = co[0], cw = co[as], cn = co[2 * as], cs = co[3 * as];
                const double cp = co[8 * as], ct = co[9 * as];
                double cee = 0.0, cww = 0.0, cnn = 0.0, css = 0.0;
                const bool core = FIRST && i >= ti0 && i < ti0 + TI && j >= tj0 && j < tj0 + TJ;
                if (NCOL != 2 || core) {
                    cee = co[4 * as];
                    cww = co[5 * as];
                    cnn = co[6 * as];
                    css = co[7 * as];
                }
                if (core) {
                    const double nb0 = (ce * S0(i, j + 1) + cw * S0(i, j - 1) + cn * S0(i - 1, j) +
                                        cs * S0(i + 1, j)) +
                                       (cee * S0(i, j + 2) + cww * S0(i, j - 2) +
                                        cnn * S0(i - 2, j) + css * S0(i + 2, j));
                    res[(i - ti0) * TJ + (j - tj0)] = fabs(nb0 + ct - cp * S0(i, j));
                }
                double nb = ce * SU(i, j + 1) + cw * SU(i, j - 1) + cn * SU(i - 1, j) +
                            cs * SU(i + 1, j);
                if (NCOL != 2)
                    nb = nb + (cee * SU(i, j + 2) + cww * SU(i, j - 2) + cnn * SU(i - 2, j) +
                               css * SU(i + 2, j));
                SU(i, j) = m.alpha * (nb + ct) / cp + m.one_minus_alpha * SU(i, j);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < (TI * TJ + NT - 1) / NT; ++k) {
        const int e = k * NT + threadIdx.x, a = e / TJ, b = e - a * TJ, i = ti0 + a, j = tj0 + b;
        if (e < TI * TJ && i < nr && j < nc)
            dst[static_cast<size_t>(i) * ld + j] = SU(i, j);
        if (FIRST && e < TI * TJ && i >= m.i0 && i <= m.i1 && j >= m.j0 && j <= m.j1)
            cell[static_cast<size_t>(i - m.i0) * rw + (j - m.j0)] = res[e];
    }
#undef S0
#undef SU
}

// One Jacobi pass, dst from src (src/momentum_u.cpp:105-113); colour = block index z.
struct JacArgs {
    MomArgs m;
    const double* src;
    double* dst;
};

__device__ __forceinline__ void k_mom_jacobi(const JacArgs& ja, const Blk& bk) {
    const MomArgs& m = ja.m;
    const double* src = ja.src;
    double* dst = ja.dst;
    int i, j;
    size_t o;
    if (!packed_cell(m, bk, bk.z, i, j, o))
        return;
    const long ld = m.ld;
    const size_t q = static_cast<size_t>(i) * ld + j, as = m.pk.astride;
    const double* c = m.pk.base + o;
    dst[q] = m.alpha * (nb9(c, as, src, q, ld) + c[9 * as]) / c[8 * as] +
             m.one_minus_alpha * src[q];
}

struct CopyArgs {
    const double* a;
    double* b;
    size_t n;
};
constexpr int CP_T = 256;

__device__ __forceinline__ void k_copy(const CopyArgs& c, const Blk& bk) {
    for (size_t k = bk.x * static_cast<size_t>(CP_T) + threadIdx.x; k < c.n;
         k += static_cast<size_t>(bk.nx) * CP_T)
        c.b[k] = c.a[k];
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

void launch_copy(Solver& s, const double* a, double* b, size_t n, const int* stop) {
    emit<CopyArgs, k_copy, CP_T>(s, dim3(1024), dim3(CP_T), 0, CopyArgs{a, b, n}, stop);
}

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

void launch_assemble(Solver& s, bool is_u, const int* stop, bool with_bc) {
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
    a.ubc = is_u && with_bc ? s.F(SG_F_U) : nullptr;
    a.uds = p.convection_scheme == SG_UDS;
    a.order = p.diffusion_order;
    a.d2 = s.diff2;
    a.d4 = s.diff4;
    const dim3 grid(cdiv(a.pk.kp, BX), cdiv(a.pk.rows, AY));
    if (is_u)
        emit<AsmArgs, k_assemble<true>, BX * AY>(s, grid, dim3(BX, AY), 0, a, stop);
    else
        emit<AsmArgs, k_assemble<false>, BX * AY>(s, grid, dim3(BX, AY), 0, a, stop);
}

void launch_bc(Solver& s, bool is_u, const int* stop) {
    const sg_params& p = s.p;
    if (is_u)
        emit<BcArgs, k_bc_u, BC_T>(s, dim3(cdiv(p.NY, BC_T)), dim3(BC_T), 0,
                                   BcArgs{s.F(SG_F_U), s.ld, p.H, p.NX, p.NY}, stop);
    else
        emit<BcArgs, k_bc_v, BC_T>(s, dim3(cdiv(p.NX, BC_T)), dim3(BC_T), 0,
                                   BcArgs{s.F(SG_F_V), s.ld, p.H, p.NX, p.NY}, stop);
}

// residual_u / residual_v (src/momentum_u.cpp:43-59): |R| per cell, then the ordered sum into
// *slot (deferred to flush_sums if defer).
void launch_mom_residual(Solver& s, bool is_u, double* slot, const int* stop, bool defer) {
    const MomArgs m = mom_args(s, is_u);
    const int nrow = m.i1 - m.i0 + 1, w = m.j1 - m.j0 + 1;
    double* cell = s.rcell[is_u ? 0 : 1];
    if (nrow > 0 && w > 0)
        emit<ResArgs, k_mom_residual, BX * BY>(s, dim3(cdiv(w, BX), cdiv(nrow, BY)), dim3(BX, BY),
                                               0, ResArgs{m, cell}, stop);
    if (defer)
        defer_sum(s, is_u ? 1 : 2, cell, nrow, w, slot);
    else if (nrow <= 0 || w <= 0)
        launch_ordered_sum(s, cell, 1, 0, -1, 0, -1, slot, stop); // empty range: 0.0
    else
        launch_ordered_sum(s, cell, w, 0, nrow - 1, 0, w - 1, slot, stop);
}

namespace {
// Threads of a regular sweep's workgroup (the first sweep, with its extra LDS, uses ST).
constexpr int SW_NT = 512;

template <int NCOL, bool FIRST, int K>
void sweep_shape(Solver& s, const SweepArgs& a, const int* stop) {
    constexpr int NT = FIRST ? ST : SW_NT, TI = MOM_TILES[K].ti, TJ = MOM_TILES[K].tj;
    emit<SweepArgs, k_mom_sweep<NCOL, TI, TJ, FIRST, NT>, NT>(
        s, dim3(cdiv(s.nc, TJ), cdiv(s.nr, TI)), dim3(NT), 0, a, stop);
}

template <int NCOL, bool FIRST, int... K>
void sweep_tile_f(Solver& s, const SweepArgs& a, const int* stop,
                  std::integer_sequence<int, K...>) {
    static_assert(sizeof...(K) == N_MOM_TILES);
    const int t = FIRST && s.mom_tile_first >= 0 ? s.mom_tile_first : s.mom_tile;
    ((t == K ? sweep_shape<NCOL, FIRST, K>(s, a, stop) : void()), ...);
}

template <int NCOL>
void sweep_tile(Solver& s, const MomArgs& m, const double* src, double* dst, double* cell,
                int first, const int* stop) {
    const SweepArgs a{m, src, dst, cell, s.nr, s.nc, first};
    constexpr auto shapes = std::make_integer_sequence<int, N_MOM_TILES>{};
    if (first)
        sweep_tile_f<NCOL, true>(s, a, stop, shapes);
    else
        sweep_tile_f<NCOL, false>(s, a, stop, shapes);
}
} // namespace

// One multicolour sweep of the field into dst (tile shape s.mom_tile); first: also the residual
// cells.
void launch_mom_sweep_once(Solver& s, bool is_u, const double* src, double* dst, bool first,
                           const int* stop) {
    const MomArgs m = mom_args(s, is_u);
    double* cell = s.rcell[is_u ? 0 : 1];
    if (m.pk.ncol == 2)
        sweep_tile<2>(s, m, src, dst, cell, first ? 1 : 0, stop);
    else
        sweep_tile<5>(s, m, src, dst, cell, first ? 1 : 0, stop);
}

// sweep_u_batch / sweep_v_batch (src/momentum_u.cpp:164-171) for the two parallel solvers. With
// res_slot (multicolour only), the residual is evaluated inside the first sweep and summed (on
// flush_sums if defer) into *res_slot.
void launch_mom_sweeps(Solver& s, bool is_u, int nsweeps, const int* stop, double* res_slot,
                       bool defer) {
    const sg_params& p = s.p;
    const MomArgs m = mom_args(s, is_u);
    const bool fuse = res_slot && p.momentum_solver == SG_MOM_GS_MULTICOLOUR && nsweeps > 0;
    if (res_slot && !fuse)
        launch_mom_residual(s, is_u, res_slot, stop, defer);
    if (nsweeps <= 0)
        return;
    if (p.momentum_solver == SG_MOM_GS_MULTICOLOUR) {
        // Alternate field -> scratch -> field ...; an odd count ends with a copy back.
        for (int k = 0; k < nsweeps; ++k) {
            const bool even = (k % 2) == 0;
            launch_mom_sweep_once(s, is_u, even ? m.x : s.scratch, even ? s.scratch : m.x,
                                  fuse && k == 0, stop);
        }
        if (nsweeps % 2)
            launch_copy(s, s.scratch, m.x, s.n, stop);
        if (fuse) {
            const int nrow = m.i1 - m.i0 + 1, w = m.j1 - m.j0 + 1;
            if (defer)
                defer_sum(s, is_u ? 1 : 2, s.rcell[is_u ? 0 : 1], nrow, w, res_slot);
            else
                launch_ordered_sum(s, s.rcell[is_u ? 0 : 1], w, 0, nrow - 1, 0, w - 1, res_slot,
                                   stop);
        }
        return;
    }
    // Jacobi batch: full copy, then alternate so that the last pass writes the field.
    launch_copy(s, m.x, s.scratch, s.n, stop);
    if (m.i1 < m.i0 || m.j1 < m.j0)
        return;
    for (int k = 0; k < nsweeps; ++k) {
        const bool to_x = (nsweeps - k) % 2 == 1;
        const double* src = to_x ? s.scratch : m.x;
        double* dst = to_x ? m.x : s.scratch;
        emit<JacArgs, k_mom_jacobi, BX * BY>(s, pass_grid(m, m.pk.ncol), dim3(BX, BY), 0,
                                             JacArgs{m, src, dst}, stop);
    }
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
