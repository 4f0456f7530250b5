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
// plus a halo into LDS and runs the colour passes there, pass k over the tile grown by the cells
// whose new values later passes of the tile still read (SweepReach, below), so it reproduces the
// colour-by-colour sweep exactly; the tile's cells are then written to a second array (tiles read
// their halo from the first), and sweeps alternate between the field and a scratch copy. The
// first sweep of an outer iteration also evaluates the residual from the loaded field before any
// update.
//
// Arithmetic is the reference's expression for expression (compiled with -ffp-contract=off, IEEE
// division), so every value is bit-identical to src/momentum_u.cpp / momentum_v.cpp; the
// pressure term (p(i,j) - p(i,j+1)) * dy is evaluated once, at assembly, with the same operands.
#include "sg_internal.hpp"

#include <stdexcept>
#include <string>
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

// Reach of a temporally blocked multicolour sweep: how far beyond its tile (north = smaller i,
// south, west = smaller j, east) colour pass k must update cells, and which cells the tile loads.
// Pass k (colour k) must cover the tile and every cell of colour k that a later pass m > k reads
// inside its own region: a cell x of colour m reads x + o for the stencil offsets o (E, W, N, S
// and, with 5 colours, EE, WW, NN, SS) whose colour is k, i.e. (o_i + 3 o_j) mod 5 = k - m (mod 2
// of o_i + o_j with 2 colours). Then every cell of every pass region sees the values the
// colour-by-colour sweep gives it: a neighbour of an earlier colour c lies in pass c's region by
// construction, a neighbour of a later colour still holds its loaded, pre-sweep value, and no
// stencil offset has colour 0. The loaded cells are the pass regions grown by the stencil reach.
// With 5 colours the regions come out lopsided (pass 0: 4 north, 2 south, 8 west, 2 east), not
// the 2 (4 - k) cells on every side of the symmetric cone: 1.17 instead of 1.39 updates per tile
// cell at 40 x 56, and 23 % fewer cells in LDS.
struct Ext {
    int n, s, w, e;
};
template <int NCOL> struct SweepReach {
    static constexpr int R = NCOL == 2 ? 1 : 2; // stencil reach
    Ext pass[NCOL];
    Ext load;
    constexpr SweepReach() : pass{}, load{0, 0, 0, 0} {
        constexpr int NS = NCOL == 2 ? 4 : 8;
        constexpr int oi[8] = {0, 0, -1, 1, 0, 0, -2, 2}, oj[8] = {1, -1, 0, 0, 2, -2, 0, 0};
        for (int k = NCOL - 1; k >= 0; --k) {
            Ext x{0, 0, 0, 0};
            for (int m = k + 1; m < NCOL; ++m)
                for (int t = 0; t < NS; ++t) {
                    const int c = NCOL == 2 ? ((oi[t] + oj[t]) % 2 + 2) % 2
                                            : ((oi[t] + 3 * oj[t]) % 5 + 5) % 5;
                    if (c != ((k - m) % NCOL + NCOL) % NCOL)
                        continue;
                    const Ext& p = pass[m];
                    x.n = x.n > p.n - oi[t] ? x.n : p.n - oi[t];
                    x.s = x.s > p.s + oi[t] ? x.s : p.s + oi[t];
                    x.w = x.w > p.w - oj[t] ? x.w : p.w - oj[t];
                    x.e = x.e > p.e + oj[t] ? x.e : p.e + oj[t];
                }
            pass[k] = x;
            load.n = load.n > x.n + R ? load.n : x.n + R;
            load.s = load.s > x.s + R ? load.s : x.s + R;
            load.w = load.w > x.w + R ? load.w : x.w + R;
            load.e = load.e > x.e + R ? load.e : x.e + R;
        }
    }
    // Cells of colour k per thread of an NT-thread workgroup on a TI x TJ tile, over all passes.
    template <int TI, int TJ, int NT> constexpr int per_thread() const {
        int mu = 0;
        for (int k = 0; k < NCOL; ++k) {
            const int c = (TI + pass[k].n + pass[k].s) *
                          ((TJ + pass[k].w + pass[k].e + NCOL) / NCOL);
            mu = mu > c ? mu : c;
        }
        return (mu + NT - 1) / NT;
    }
};
template <int NCOL> constexpr SweepReach<NCOL> REACH{};
static_assert(REACH<5>.pass[0].n == 4 && REACH<5>.pass[0].s == 2 && REACH<5>.pass[0].w == 8 &&
                  REACH<5>.pass[0].e == 2 && REACH<5>.pass[3].w == 2 && REACH<5>.pass[4].w == 0,
              "5-colour pass regions");
static_assert(REACH<5>.load.n == 6 && REACH<5>.load.s == 4 && REACH<5>.load.w == 10 &&
                  REACH<5>.load.e == 4,
              "5-colour load halo");
static_assert(REACH<2>.pass[0].n == 1 && REACH<2>.pass[0].e == 1 && REACH<2>.load.w == 2,
              "2-colour regions");

// What a control volume's coefficients depend on besides its face velocities.
struct MomConv {
    int H, NX, NY, uds, order;
    double dx, dy;
    C9 d2, d4; // diffusive coefficients (src/coefficients.cpp:41-63, evaluated on the host)
};

// The coefficients of control volume (i, j) from its face velocities, in the order AE, AW, AN,
// AS, AEE, AWW, ANN, ASS, AP: convection (src/coefficients.cpp:15-39) plus diffusion, summed
// element-wise (:5-8). Every coefficient any kernel uses comes from here, so the assembled
// arrays and the coefficients a sweep recomputes from stored face velocities are the same bits.
// The two schemes of src/coefficients.cpp:15-39 and the sum with diffusion, for one cell. Here
// max(0, a) is one v_max_f64 (fmax), where pos's select returns +0 for a = -0: the two differ
// only in the sign of a zero, and no coefficient can see it. A nonzero product or sum does not
// depend on a zero operand's sign, so a convective coefficient differs at most in the sign of a
// zero result; AE, AW, AN, AS and AP then add a nonzero diffusive coefficient, and AEE..ASS add a
// negative one or +0 (and -0 + +0 = +0 + +0 = +0).
__device__ __forceinline__ double posm(double a) { return fmax(a, 0.0); }
__device__ __forceinline__ void conv_uds(const MomConv& a, double ue, double uw, double vn,
                                         double vs, double* c) {
    c[0] = a.dy * posm(-ue), c[1] = a.dy * posm(uw), c[2] = a.dx * posm(-vn);
    c[3] = a.dx * posm(vs);
    c[4] = c[5] = c[6] = c[7] = 0.0;
}
__device__ __forceinline__ void conv_quick(const MomConv& a, double ue, double uw, double vn,
                                           double vs, double* c) {
    const double dx = a.dx, dy = a.dy;
    const double pue = posm(ue), nue = posm(-ue), puw = posm(uw), nuw = posm(-uw);
    const double pvn = posm(vn), nvn = posm(-vn), pvs = posm(vs), nvs = posm(-vs);
    c[0] = -dy * (-0.75 * nue + 0.375 * pue - 0.125 * nuw);
    c[1] = -dy * (-0.125 * pue + 0.375 * nuw - 0.75 * puw);
    c[2] = -dx * (-0.75 * nvn + 0.375 * pvn - 0.125 * nvs);
    c[3] = -dx * (-0.125 * pvn + 0.375 * nvn - 0.75 * pvs);
    c[4] = -0.125 * dy * nue;
    c[5] = -0.125 * dy * puw;
    c[6] = -0.125 * dx * nvn;
    c[7] = -0.125 * dx * pvs;
}
__device__ __forceinline__ C9 conv_plus_diff(const double* c, const C9& d) {
    // ap = s + w + n + e + ss + ww + nn + ee (src/coefficients.cpp:36-37)
    const double ap = c[3] + c[1] + c[2] + c[0] + c[7] + c[5] + c[6] + c[4];
    C9 r;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        r.v[k] = c[k] + d.v[k];
    r.v[8] = ap + d.v[8];
    return r;
}

// mom_coefs with the scheme chosen per wave where it can be: UDS runs take the UDS expressions
// only, and a QUICK wave with no cell in the boundary band the QUICK ones only (the same
// expressions as mom_coefs, so the same bits, without evaluating both schemes per lane).
__device__ __forceinline__ C9 mom_coefs_w(const MomConv& a, int i, int j, double ue, double uw,
                                          double vn, double vs) {
    const bool band = j == a.H + 1 || j == a.NY + a.H || i == a.NX + a.H - 1 || i == a.H;
    double c[8];
    if (a.uds) {
        conv_uds(a, ue, uw, vn, vs, c);
        return conv_plus_diff(c, (a.order == 2 || band) ? a.d2 : a.d4);
    }
    if (!__any(band)) {
        conv_quick(a, ue, uw, vn, vs, c);
        return conv_plus_diff(c, a.order == 2 ? a.d2 : a.d4);
    }
    if (band)
        conv_uds(a, ue, uw, vn, vs, c);
    else
        conv_quick(a, ue, uw, vn, vs, c);
    return conv_plus_diff(c, (a.order == 2 || band) ? a.d2 : a.d4);
}

__device__ __forceinline__ C9 mom_coefs(const MomConv& a, int i, int j, double ue, double uw,
                                        double vn, double vs) {
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
    C9 r;
    r.v[0] = e + d.v[0];
    r.v[1] = w + d.v[1];
    r.v[2] = n + d.v[2];
    r.v[3] = s + d.v[3];
    r.v[4] = ee + d.v[4];
    r.v[5] = ww + d.v[5];
    r.v[6] = nn + d.v[6];
    r.v[7] = ss + d.v[7];
    r.v[8] = ap + d.v[8];
    return r;
}

// Face form. The face velocities of src/momentum_u.cpp:13-16 and src/momentum_v.cpp:14-17 are
// shared bit for bit by neighbouring control volumes: the west face of (i, j) is the east face of
// (i, j - 1) and the south face of (i, j) the north face of (i + 1, j) (the same two velocities
// added, in the other order, which gives the same sum). So a component's coefficients are fixed by
// three colour-packed arrays, ue, vn and the pressure term, over the assembly range grown by a
// column to the west and a row to the south (F_UE, F_VN, F_PT of the face pack), and a sweep
// recomputes them with mom_coefs instead of streaming the nine assembled arrays.
enum { F_UE = 0, F_VN = 1, F_PT = 2, F_NARR = 3 };

// Face velocities (ue, vn) of the face-pack cell (i, j) for u or v.
template <bool IS_U>
__device__ __forceinline__ double face_ue(const double* u, size_t q, long ld) {
    return IS_U ? (u[q] + u[q + 1]) / 2 : (u[q - ld] + u[q]) / 2;
}
template <bool IS_U>
__device__ __forceinline__ double face_vn(const double* v, size_t q, long ld) {
    return IS_U ? (v[q] + v[q + 1]) / 2 : (v[q] + v[q - ld]) / 2;
}

// The face values of cell (i, j) of colour c from the face pack fk: ue, uw, vn, vs and the
// pressure term.
// Packed index of cell (i, j) of colour c: the cells of one colour in a row are NCOL columns
// apart starting less than NCOL columns after aj0, so the k-th of them is (j - aj0) / NCOL. (An
// array of a face pack has fewer than 2^31 entries: ncol x rows x kp.)
template <int NCOL> __device__ __forceinline__ int fidx(const Pack& k, int c, int i, int j) {
    return (c * k.rows + (i - k.ai0)) * k.kp + (j - k.aj0) / NCOL;
}
// The west and south indices follow from the cell's own: the same row (one colour plane over),
// one entry back when the cell is its colour's first of the row's NCOL-column group; the next
// row of the next colour's plane.
template <int NCOL>
__device__ __forceinline__ void face_load(const Pack& fk, int c, int i, int j, double* f) {
    const double* fue = fk.base + F_UE * fk.astride;
    const double* fvn = fk.base + F_VN * fk.astride;
    const double* fpt = fk.base + F_PT * fk.astride;
    const int cw = NCOL == 2 ? (c + 1) & 1 : (c + 2) % 5; // colour of (i, j - 1)
    const int cs = NCOL == 2 ? (c + 1) & 1 : (c + 1) % 5; // colour of (i + 1, j)
    const int dj = j - fk.aj0, kk = dj / NCOL;
    const int plane = fk.rows * fk.kp;
    const int o = (c * fk.rows + (i - fk.ai0)) * fk.kp + kk;
    f[0] = fue[o];
    f[1] = fue[o + (cw - c) * plane - (dj == kk * NCOL ? 1 : 0)];
    f[2] = fvn[o];
    f[3] = fvn[o + (cs - c) * plane + fk.kp];
    f[4] = fpt[o];
}

// The coefficients of cell (i, j) of colour c from the face pack fk.
template <int NCOL>
__device__ __forceinline__ C9 face_coefs(const Pack& fk, const MomConv& cv, int c, int i, int j,
                                         double& pt) {
    double f[5];
    face_load<NCOL>(fk, c, i, j, f);
    pt = f[4];
    return mom_coefs(cv, i, j, f[0], f[1], f[2], f[3]);
}

struct AsmArgs {
    const double* u;
    const double* v;
    const double* pf;
    double* APn; // row-major AP
    Pack pk;     // the coefficient pack, or the face pack (face form)
    int ld;
    double h;    // face length of the pressure term (dy for u, dx for v)
    long poff;   // second pressure cell (+1 for u, -ld for v)
    double* ubc; // u only: also set u's ghost rows (bc_u) in this launch, or null
    MomConv cv;
    // Face form: AP of cells in rows ap_i0..ap_i1, columns ap_j0..ap_j1 is left to the first sweep
    // (an empty range: AP of every cell here).
    int ap_i0, ap_i1, ap_j0, ap_j1;
};

// One control volume, written at packed entry b (and AP row-major at q).
__device__ __forceinline__ void assemble_cell(const AsmArgs& a, int i, int j, size_t q, double* b,
                                              double ue, double uw, double vn, double vs,
                                              double pt) {
    const C9 c = mom_coefs(a.cv, i, j, ue, uw, vn, vs);
    const size_t as = a.pk.astride;
#pragma unroll
    for (int k = 0; k < 9; ++k)
        b[k * as] = c.v[k];
    a.APn[q] = c.v[8];
    b[9 * as] = pt;
}

// Colour-major: thread (x, y) takes the kk-th cell of every colour in row i, so each store to a
// packed array is one contiguous run per wave. All loads of the thread's cells come first. AY
// rows per workgroup: the rows above and below that a cell's faces read are mostly its
// neighbours' own rows, already in L1.
constexpr int AY = 16;

// Ghost rows of u (src/momentum_u.cpp:24-34), which the assembly does not read, set by the
// threads of the first two rows of an assembly launch for the columns of their cells: row 0
// the lid, row 1 the south wall.
__device__ __forceinline__ void ghost_rows_u(const AsmArgs& a, int row, const int* jj) {
    double* u = a.ubc;
    const long ld = a.ld;
    const int H = a.cv.H;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const int j = jj[c];
        if (j < H + 1 || j > a.cv.NY + H)
            continue;
        if (row == 0)
            u[static_cast<size_t>(H - 1) * ld + j] = 2 - u[static_cast<size_t>(H) * ld + j];
        else
            u[static_cast<size_t>(a.cv.NX + H) * ld + j] =
                -u[static_cast<size_t>(a.cv.NX + H - 1) * ld + j];
    }
}

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
    if (IS_U && a.ubc && i - pk.ai0 < 2)
        ghost_rows_u(a, i - pk.ai0, jj);
}

// Face form of the assembly: ue, vn and the pressure term of every face-pack cell (the extra
// column needs only ue and the extra row only vn: they are the west faces of the first column and
// the south faces of the last row), and AP, row-major (assemble_p and correct read it), of the
// assembly-range cells outside [ap_i0, ap_i1] x [ap_j0, ap_j1] (the first sweep stores the rest).
// With ubc, u's ghost rows as in k_assemble. One thread per cell in row-major order, FRX
// consecutive columns by FRY rows per workgroup (the workgroups handed to the XCDs in contiguous
// bands, xcd_tile): the u, v and p reads and the AP writes are coalesced, and each cell's face
// values go to its colour's plane of the face pack. (One thread per five cells of one colour each
// read and wrote five columns apart: 18.3 / 20.0 us for u / v at 1024^2 then.)
constexpr int FRX = 64, FRY = 4;
template <bool IS_U, int NCOL>
__device__ __forceinline__ void k_assemble_face_rm(const AsmArgs& a, const Blk& bk) {
    const Pack& fk = a.pk;
    int tx, ty;
    xcd_tile(bk, tx, ty);
    const int j = fk.aj0 + tx * FRX + threadIdx.x;
    const int i = fk.ai0 + ty * FRY + threadIdx.y;
    if (i > fk.ai1 || j > fk.aj1)
        return;
    const long ld = a.ld;
    const size_t fs = fk.astride;
    const bool xrow = i == fk.ai1, xcol = j == fk.aj0; // the extra row and column
    const size_t q = static_cast<size_t>(i) * ld + j;
    double* b = fk.base + fidx<NCOL>(fk, colour_of(NCOL, i, j), i, j);
    double ue = 0.0, vn = 0.0;
    if (!xcol) {
        vn = face_vn<IS_U>(a.v, q, ld);
        b[F_VN * fs] = vn;
    }
    if (!xrow) {
        ue = face_ue<IS_U>(a.u, q, ld);
        b[F_UE * fs] = ue;
    }
    if (!xrow && !xcol) {
        b[F_PT * fs] = (a.pf[q] - a.pf[q + a.poff]) * a.h;
        if (i < a.ap_i0 || i > a.ap_i1 || j < a.ap_j0 || j > a.ap_j1) {
            double uw, vs;
            if (IS_U) {
                uw = (a.u[q] + a.u[q - 1]) / 2;
                vs = (a.v[q + ld] + a.v[q + ld + 1]) / 2;
            } else {
                uw = (a.u[q - 1] + a.u[q - ld - 1]) / 2;
                vs = (a.v[q] + a.v[q + ld]) / 2;
            }
            a.APn[q] = mom_coefs(a.cv, i, j, ue, uw, vn, vs).v[8];
        }
    }
    if (IS_U && a.ubc && i - fk.ai0 < 2 && j >= a.cv.H + 1 && j <= a.cv.NY + a.cv.H) {
        double* u = a.ubc;
        const int H = a.cv.H;
        if (i == fk.ai0)
            u[static_cast<size_t>(H - 1) * ld + j] = 2 - u[static_cast<size_t>(H) * ld + j];
        else
            u[static_cast<size_t>(a.cv.NX + H) * ld + j] =
                -u[static_cast<size_t>(a.cv.NX + H - 1) * ld + j];
    }
}

// ---- FP32 momentum smoothing inside FP64 defect correction (momentum_precision = mixed) ----
//
// A multicolour sweep is linear in the field: with x = x0 + d, the update
// x = a (nb(x) + pterm) / AP + (1 - a) x of src/momentum_u.cpp:123-143 becomes
// d = a (r0 + nb(d)) / AP + (1 - a) d, r0 = nb(x0) + pterm - AP x0 the residual of x0
// (src/momentum_u.cpp:43-59), with d = 0 to start. So the outer iteration's assembly also evaluates
// r0 in FP64, exactly as the reference's residual does (its |r0| is res_u), the sweeps run on the
// correction d in FP32, and the last sweep writes x0 + d in FP64. The fixed point of the outer
// loop is still that of the FP64 residual: d = 0 where r0 = 0.
//
// The FP32 pack holds, colour-packed like the coefficient pack (array a, colour c, row i, k-th
// cell of the colour), the 8 neighbour coefficients (AE, AW, AN, AS, AEE, AWW, ANN, ASS), a / AP
// and r0 of every cell of the sweep range: DC_NARR arrays of floats.
constexpr int DC_W = 8, DC_R = 9, DC_NARR = 10;

struct AsmDcArgs {
    AsmArgs a;       // the assembly (row-major AP; ubc: u's ghost rows set in the same launch)
    float* dc;       // FP32 pack over the geometry of a.pk (sweep-range cells only)
    double alpha;    // relaxation factor of the component
    double* cell;    // |r0| of the sweep range (the residual's cells), row-major
    int i0, i1, j0, j1; // sweep range
};

// The field after the boundary conditions of this phase: u's ghost rows (src/momentum_u.cpp:24-34)
// and v's ghost columns (src/momentum_v.cpp:24-34) as bc_u / bc_v store them, from the inner
// values, so the residual sees what the reference's residual sees after bc.
template <bool IS_U>
__device__ __forceinline__ double bc_val(const AsmArgs& a, int i, int j) {
    const long ld = a.ld;
    const int H = a.cv.H, NX = a.cv.NX, NY = a.cv.NY;
    const double* x = IS_U ? a.u : a.v;
    if (IS_U) {
        if (j >= H + 1 && j <= NY + H) {
            if (i == H - 1)
                return 2 - x[static_cast<size_t>(H) * ld + j];
            if (i == NX + H)
                return -x[static_cast<size_t>(NX + H - 1) * ld + j];
        }
    } else if (i >= H && i <= NX + H - 1) {
        if (j == H)
            return -x[static_cast<size_t>(i) * ld + H + 1];
        if (j == NY + H + 1)
            return -x[static_cast<size_t>(i) * ld + NY + H];
    }
    return x[static_cast<size_t>(i) * ld + j];
}

// Workgroups of k_asm_dc: FBX cells of each colour (one thread per five cells, one of each colour)
// by FAY rows.
constexpr int FBX = 32, FAY = 16;

// The assembly of k_assemble (coefficients from the faces, AP row-major; u's ghost rows) plus, for
// every cell of the sweep range, r0 = ((E + W + N + S) + (EE + WW + NN + SS)) + pterm - AP x0
// (src/momentum_u.cpp:48-50, the same operands and order) into the |r0| cells, and the FP32 pack.
template <bool IS_U> __device__ __forceinline__ void k_asm_dc(const AsmDcArgs& d, const Blk& bk) {
    const AsmArgs& a = d.a;
    const Pack& pk = a.pk;
    int tx, ty;
    xcd_tile(bk, tx, ty);
    const int kk = tx * FBX + threadIdx.x;
    const int i = pk.ai0 + ty * FAY + threadIdx.y;
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
    const size_t as = pk.astride;
    const int rw = d.j1 - d.j0 + 1;
#pragma unroll
    for (int c = 0; c < MC; ++c) {
        const int j = jj[c];
        if (j < 0)
            continue;
        const size_t q = static_cast<size_t>(i) * ld + j;
        const C9 co = mom_coefs(a.cv, i, j, fue[c], fuw[c], fvn[c], fvs[c]);
        a.APn[q] = co.v[8];
        if (i < d.i0 || i > d.i1 || j < d.j0 || j > d.j1)
            continue;
        const double nb =
            (co.v[0] * bc_val<IS_U>(a, i, j + 1) + co.v[1] * bc_val<IS_U>(a, i, j - 1) +
             co.v[2] * bc_val<IS_U>(a, i - 1, j) + co.v[3] * bc_val<IS_U>(a, i + 1, j)) +
            (co.v[4] * bc_val<IS_U>(a, i, j + 2) + co.v[5] * bc_val<IS_U>(a, i, j - 2) +
             co.v[6] * bc_val<IS_U>(a, i - 2, j) + co.v[7] * bc_val<IS_U>(a, i + 2, j));
        const double r0 = nb + fpt[c] - co.v[8] * bc_val<IS_U>(a, i, j);
        d.cell[static_cast<size_t>(i - d.i0) * rw + (j - d.j0)] = fabs(r0);
        float* b = d.dc + pidx(pk, c, i, kk);
#pragma unroll
        for (int k = 0; k < DC_W; ++k)
            b[k * as] = static_cast<float>(co.v[k]);
        b[DC_W * as] = static_cast<float>(d.alpha / co.v[8]);
        b[DC_R * as] = static_cast<float>(r0);
    }
    if (IS_U && a.ubc && i - pk.ai0 < 2)
        ghost_rows_u(a, i - pk.ai0, jj);
}

// One multicolour sweep of the correction d on a TI x TJ tile in LDS (FP32), the colour passes as
// in k_mom_sweep (pass k over the tile grown by SweepReach's region, so the result equals the
// colour-by-colour sweep): d = (a / AP) (r0 + nb(d)) + (1 - a) d. FIRST: d starts at 0 (nothing
// loaded). LAST: the tile's cells of the sweep range get x = x0 + d (FP64) instead of d.
struct Dc32Args {
    const float* dc; // the FP32 pack
    Pack g;          // its geometry (base unused)
    const float* src; // d before the sweep (unused when FIRST)
    float* dst;      // d after the sweep (unless LAST)
    double* x;       // LAST: the field, x0 read and x0 + d written
    int ld, nr, nc;  // d arrays: nr x ld floats, indexed like the fields
    int i0, i1, j0, j1; // sweep range
    float oma;       // 1 - a
};

template <int NCOL, int TI, int TJ, bool FIRST, bool LAST, int NT>
__device__ __forceinline__ void k_dc_sweep(const Dc32Args& sa, const Blk& bk) {
    constexpr SweepReach<NCOL> RC{};
    constexpr Ext L = RC.load;
    constexpr int LW = TJ + L.w + L.e, LH = TI + L.n + L.s;
    __shared__ float su[LH * LW];
    int tx, ty;
    xcd_tile(bk, tx, ty);
    const int ti0 = ty * TI, tj0 = tx * TJ;
    const int gi0 = ti0 - L.n, gj0 = tj0 - L.w;
    const long ld = sa.ld;
    constexpr int NL = (LH * LW + NT - 1) / NT;
    if (FIRST) {
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int e = k * NT + threadIdx.x;
            if (e < LH * LW)
                su[e] = 0.0f;
        }
    } else {
        float v[NL];
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int e = k * NT + threadIdx.x, r = e / LW, c = e - r * LW;
            const int i = gi0 + r, j = gj0 + c;
            v[k] = (e < LH * LW && i >= 0 && i < sa.nr && j >= 0 && j < sa.nc)
                       ? sa.src[static_cast<size_t>(i) * ld + j]
                       : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int e = k * NT + threadIdx.x;
            if (e < LH * LW)
                su[e] = v[k];
        }
    }
    __syncthreads();
#define SD(i, j) su[((i) - gi0) * LW + ((j) - gj0)]
    const Pack& g = sa.g;
    const size_t as = g.astride;
    constexpr int MU = RC.template per_thread<TI, TJ, NT>();
    float cf[2][MU][DC_NARR];
    int fi[2][MU], fj[2][MU];
    // The pack values of this thread's cells of pass k (loaded one pass ahead: they do not depend
    // on d). Cell u of pass k is the pass's (u * NT + tid)-th.
    auto fetch = [&](int k, float (&f)[MU][DC_NARR], int (&ci)[MU], int (&cj)[MU]) {
        const Ext P = RC.pass[k];
        const int ia = max(sa.i0, ti0 - P.n), ib = min(sa.i1, ti0 + TI - 1 + P.s);
        const int ja = max(sa.j0, tj0 - P.w), jb = min(sa.j1, tj0 + TJ - 1 + P.e);
        const int cpr = (jb - ja + NCOL) / NCOL;
        const int n = ia <= ib && ja <= jb ? (ib - ia + 1) * cpr : 0;
        const float rcp = 1.0f / static_cast<float>(max(cpr, 1)); // exact split, as in k_mom_sweep
#pragma unroll
        for (int u = 0; u < MU; ++u) {
            const int e = u * NT + threadIdx.x;
            cj[u] = -1;
            if (e >= n)
                continue;
            const int r = static_cast<int>((static_cast<float>(e) + 0.5f) * rcp);
            const int t = e - r * cpr, i = ia + r;
            const int c0 = colour_of(NCOL, i, ja);
            const int j = ja + (NCOL == 2 ? ((k - c0) & 1) : (2 * (k - c0) + 10) % 5) + NCOL * t;
            if (j > jb)
                continue;
            ci[u] = i;
            cj[u] = j;
            const float* b = sa.dc + (k * g.rows + (i - g.ai0)) * g.kp + (j - g.aj0) / NCOL;
#pragma unroll
            for (int a = 0; a < DC_NARR; ++a)
                if (NCOL != 2 || a < 4 || a >= DC_W) // two colours: no far neighbours
                    f[u][a] = b[a * as];
        }
    };
    fetch(0, cf[0], fi[0], fj[0]);
#pragma unroll
    for (int k = 0; k < NCOL; ++k) {
        if (k + 1 < NCOL)
            fetch(k + 1, cf[(k + 1) & 1], fi[(k + 1) & 1], fj[(k + 1) & 1]);
#pragma unroll
        for (int u = 0; u < MU; ++u) {
            const int j = fj[k & 1][u];
            if (j < 0)
                continue;
            const int i = fi[k & 1][u];
            const float* f = cf[k & 1][u];
            float nb = f[0] * SD(i, j + 1) + f[1] * SD(i, j - 1) + f[2] * SD(i - 1, j) +
                       f[3] * SD(i + 1, j);
            if (NCOL != 2)
                nb += f[4] * SD(i, j + 2) + f[5] * SD(i, j - 2) + f[6] * SD(i - 2, j) +
                      f[7] * SD(i + 2, j);
            SD(i, j) = f[DC_W] * (f[DC_R] + nb) + sa.oma * SD(i, j);
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < (TI * TJ + NT - 1) / NT; ++k) {
        const int e = k * NT + threadIdx.x, r = e / TJ, c = e - r * TJ;
        const int i = ti0 + r, j = tj0 + c;
        if (e >= TI * TJ || i < sa.i0 || i > sa.i1 || j < sa.j0 || j > sa.j1)
            continue;
        const size_t q = static_cast<size_t>(i) * ld + j;
        if (LAST)
            sa.x[q] = sa.x[q] + static_cast<double>(SD(i, j));
        else
            sa.dst[q] = SD(i, j);
    }
#undef SD
}

// Face form -> the coefficient pack (and AP row-major) over the assembly range.
struct ExpArgs {
    Pack fk, pk;
    MomConv cv;
    double* APn;
    int ld;
};
template <int NCOL> __global__ void __launch_bounds__(BX* BY) k_expand(ExpArgs e) {
    const Pack& pk = e.pk;
    const int j = pk.aj0 + blockIdx.x * BX + threadIdx.x;
    const int i = pk.ai0 + blockIdx.y * BY + threadIdx.y;
    if (i > pk.ai1 || j > pk.aj1)
        return;
    const int c = colour_of(NCOL, i, j);
    double pt;
    const C9 co = face_coefs<NCOL>(e.fk, e.cv, c, i, j, pt);
    double* b = pk.base + pidx(pk, c, i, (j - jfirst(pk, c, i)) / NCOL);
#pragma unroll
    for (int k = 0; k < 9; ++k)
        b[k * pk.astride] = co.v[k];
    b[9 * pk.astride] = pt;
    e.APn[static_cast<size_t>(i) * e.ld + j] = co.v[8];
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
    Pack pk; // the coefficient pack
    Pack fk; // the face pack
    MomConv cv;
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
    // FIRST only: v's ghost columns (src/momentum_v.cpp:24-34). Cells (i, bjw) and (i, bje) of
    // rows bi0..bi1 are loaded as minus their inner neighbour; none when bi1 < bi0.
    int bi0, bi1, bjw, bje;
    // FIRST only: row-major AP of the core cells, from the coefficients the residual uses (the
    // face assembly left it out: AsmArgs::ap_*), or null.
    double* apn;
};

// FIRST: a second LDS copy keeps the loaded (pre-sweep) values of the tile and the ring of R
// cells its stencils reach, and each core cell's residual |R| = |(nb + pterm) - AP x|
// (src/momentum_u.cpp:48-50) is evaluated from that copy during the cell's own colour pass, with
// the coefficients the update loads anyway. (Only that ring, not the whole halo: the two copies
// then fit two workgroups per CU.)
// FACE: the coefficients are recomputed from the face pack (face form) instead of read from the
// coefficient pack.
template <int NCOL, int TI, int TJ, bool FIRST, int NT, bool FACE>
__device__ __forceinline__ void k_mom_sweep(const SweepArgs& sa, const Blk& bk) {
    const MomArgs& m = sa.m;
    const double* src = sa.src;
    double* dst = sa.dst;
    double* cell = sa.cell;
    const int nr = sa.nr, nc = sa.nc;
    constexpr SweepReach<NCOL> RC{};
    constexpr int R = RC.R;     // stencil reach
    constexpr Ext L = RC.load;  // halo loaded around the tile
    constexpr int LW = TJ + L.w + L.e, LH = TI + L.n + L.s;
    constexpr int W0 = TJ + 2 * R, H0 = TI + 2 * R; // pre-sweep copy: tile and stencil ring
    __shared__ double su[LH * LW];
    __shared__ double su0[FIRST ? H0 * W0 : 1];
    __shared__ double res[FIRST ? TI * TJ : 1]; // |R| of the core, written out row by row
    int tx, ty;
    xcd_tile(bk, tx, ty);
    const int ti0 = ty * TI, tj0 = tx * TJ;
    const int gi0 = ti0 - L.n, gj0 = tj0 - L.w; // array cell of su[0]
    const long ld = m.ld;
    // Batched loads: every load of a chunk is in flight before the first LDS store.
    constexpr int NL = (LH * LW + NT - 1) / NT;
    {
        double v[NL];
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int e = k * NT + threadIdx.x, a = e / LW, b = e - a * LW, i = gi0 + a, j = gj0 + b;
            int jl = j;
            bool neg = false;
            if (FIRST && i >= sa.bi0 && i <= sa.bi1 && (j == sa.bjw || j == sa.bje)) {
                neg = true;
                jl = j == sa.bjw ? j + 1 : j - 1;
            }
            const double x = (e < LH * LW && i >= 0 && i < nr && j >= 0 && j < nc)
                                 ? src[static_cast<size_t>(i) * ld + jl]
                                 : 0.0;
            v[k] = neg ? -x : x;
        }
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int e = k * NT + threadIdx.x;
            if (e < LH * LW) {
                su[e] = v[k];
                if (FIRST) {
                    const int a = e / LW - (L.n - R), b = e % LW - (L.w - R);
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
    // One cell's update (and, in the core of the first sweep, its residual) with its coefficients.
    auto update = [&](int i, int j, const C9& co, double ct) {
        const double ce = co.v[0], cw = co.v[1], cn = co.v[2], cs = co.v[3];
        const double cee = co.v[4], cww = co.v[5], cnn = co.v[6], css = co.v[7], cp = co.v[8];
        if (FIRST && i >= ti0 && i < ti0 + TI && j >= tj0 && j < tj0 + TJ) {
            const double nb0 = (ce * S0(i, j + 1) + cw * S0(i, j - 1) + cn * S0(i - 1, j) +
                                cs * S0(i + 1, j)) +
                               (cee * S0(i, j + 2) + cww * S0(i, j - 2) + cnn * S0(i - 2, j) +
                                css * S0(i + 2, j));
            res[(i - ti0) * TJ + (j - tj0)] = fabs(nb0 + ct - cp * S0(i, j));
            if (sa.apn)
                sa.apn[static_cast<size_t>(i) * ld + j] = cp;
        }
        double nb = ce * SU(i, j + 1) + cw * SU(i, j - 1) + cn * SU(i - 1, j) + cs * SU(i + 1, j);
        if (NCOL != 2)
            nb = nb + (cee * SU(i, j + 2) + cww * SU(i, j - 2) + cnn * SU(i - 2, j) +
                       css * SU(i + 2, j));
        SU(i, j) = m.alpha * (nb + ct) / cp + m.one_minus_alpha * SU(i, j);
    };
    if constexpr (FACE) {
        // The face values do not depend on the field, so each thread loads those of its cells of
        // pass k + 1 before it updates its cells of pass k: their latency hides behind pass k's
        // arithmetic and barrier. Cell u of pass k is the thread's (u * NT + tid)-th of the pass.
        constexpr int MU = RC.template per_thread<TI, TJ, NT>();
        double fv[2][MU][5];
        int fi[2][MU], fj[2][MU];
        auto fetch = [&](int k, double (&f)[MU][5], int (&ci)[MU], int (&cj)[MU]) {
            const Ext P = RC.pass[k];
            const int ia = max(m.i0, ti0 - P.n), ib = min(m.i1, ti0 + TI - 1 + P.s);
            const int ja = max(m.j0, tj0 - P.w), jb = min(m.j1, tj0 + TJ - 1 + P.e);
            const int cpr = (jb - ja + NCOL) / NCOL;
            const int n = ia <= ib && ja <= jb ? (ib - ia + 1) * cpr : 0;
            // e / cpr in FP32: e < 2^16 and cpr < 2^12, so (e + 0.5) / cpr is at least 1 / (2 cpr)
            // away from an integer, far more than the rounding error.
            const float rcp = 1.0f / static_cast<float>(max(cpr, 1));
#pragma unroll
            for (int u = 0; u < MU; ++u) {
                const int e = u * NT + threadIdx.x;
                cj[u] = -1;
                if (e >= n)
                    continue;
                const int r = static_cast<int>((static_cast<float>(e) + 0.5f) * rcp);
                const int t = e - r * cpr, i = ia + r;
                const int c0 = colour_of(NCOL, i, ja);
                const int j =
                    ja + (NCOL == 2 ? ((k - c0) & 1) : (2 * (k - c0) + 10) % 5) + NCOL * t;
                if (j > jb)
                    continue;
                ci[u] = i;
                cj[u] = j;
                face_load<NCOL>(m.fk, k, i, j, f[u]);
            }
        };
        fetch(0, fv[0], fi[0], fj[0]);
#pragma unroll
        for (int k = 0; k < NCOL; ++k) {
            if (k + 1 < NCOL)
                fetch(k + 1, fv[(k + 1) & 1], fi[(k + 1) & 1], fj[(k + 1) & 1]);
#pragma unroll
            for (int u = 0; u < MU; ++u) {
                const int j = fj[k & 1][u];
                if (j < 0)
                    continue;
                const int i = fi[k & 1][u];
                const double* f = fv[k & 1][u];
                update(i, j, mom_coefs_w(m.cv, i, j, f[0], f[1], f[2], f[3]), f[4]);
            }
            __syncthreads();
        }
    }
    for (int k = 0; k < (FACE ? 0 : NCOL); ++k) {
        const Ext P = RC.pass[k];
        const int ia = max(m.i0, ti0 - P.n), ib = min(m.i1, ti0 + TI - 1 + P.s);
        const int ja = max(m.j0, tj0 - P.w), jb = min(m.j1, tj0 + TJ - 1 + P.e);
        if (ia <= ib && ja <= jb) {
            const int cpr = (jb - ja + NCOL) / NCOL;
            const int n = (ib - ia + 1) * cpr;
            for (int e = threadIdx.x; e < n; e += NT) {
                const int r = e / cpr, t = e - r * cpr, i = ia + r;
                const int c0 = colour_of(NCOL, i, ja);
                const int j = ja + (NCOL == 2 ? ((k - c0) & 1) : (2 * (k - c0) + 10) % 5) + NCOL * t;
                if (j > jb)
                    continue;
                const bool core = FIRST && i >= ti0 && i < ti0 + TI && j >= tj0 && j < tj0 + TJ;
                const double* co = pk.base + pidx(pk, k, i, (j - jfirst(pk, k, i)) / NCOL);
                C9 c;
                c.v[0] = co[0], c.v[1] = co[as], c.v[2] = co[2 * as], c.v[3] = co[3 * as];
                c.v[4] = c.v[5] = c.v[6] = c.v[7] = 0.0;
                c.v[8] = co[8 * as];
                if (NCOL != 2 || core) {
                    c.v[4] = co[4 * as];
                    c.v[5] = co[5 * as];
                    c.v[6] = co[6 * as];
                    c.v[7] = co[7 * as];
                }
                update(i, j, c, co[9 * as]);
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

// Pressure term of every cell of the assembly range [ai0, ai1] x [aj0, aj1] from the current p,
// into array `a` of pack pk (the coefficient pack's 9, or the face pack's F_PT).
__global__ void __launch_bounds__(BX* BY) k_pterm(Pack pk, int a, int ai1, int aj0,
                                                  const double* pf, int ld, long poff, double h) {
    const int j = aj0 + blockIdx.x * BX + threadIdx.x;
    const int i = pk.ai0 + blockIdx.y * BY + threadIdx.y;
    if (i > ai1 || j > pk.aj1)
        return;
    const int c = colour_of(pk.ncol, i, j);
    const size_t q = static_cast<size_t>(i) * ld + j;
    pk.base[a * pk.astride + pidx(pk, c, i, (j - jfirst(pk, c, i)) / pk.ncol)] =
        (pf[q] - pf[q + poff]) * h;
}

MomConv mom_conv(const Solver& s) {
    const sg_params& p = s.p;
    MomConv c{};
    c.H = p.H, c.NX = p.NX, c.NY = p.NY;
    c.uds = p.convection_scheme == SG_UDS;
    c.order = p.diffusion_order;
    c.dx = p.dx, c.dy = p.dy;
    c.d2 = s.diff2;
    c.d4 = s.diff4;
    return c;
}

MomArgs mom_args(Solver& s, bool is_u) {
    const sg_params& p = s.p;
    MomArgs m{};
    m.x = s.F(is_u ? SG_F_U : SG_F_V);
    m.pk = is_u ? s.pku : s.pkv;
    m.fk = is_u ? s.fku : s.fkv;
    m.cv = mom_conv(s);
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

// Face-pack geometry: the assembly range grown by a column to the west and a row to the south.
Pack make_face_pack(const sg_params& p, bool is_u) {
    Pack k = make_pack(p, is_u);
    k.aj0 -= 1;
    k.ai1 += 1;
    k.rows = k.ai1 - k.ai0 + 1;
    const int w = k.aj1 - k.aj0 + 1;
    k.kp = ((w + k.ncol - 1) / k.ncol + 15) / 16 * 16;
    k.astride = static_cast<size_t>(k.ncol) * k.rows * k.kp;
    return k;
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

void launch_assemble(Solver& s, bool is_u, const int* stop, bool with_bc, bool face,
                     bool ap_in_sweep) {
    const sg_params& p = s.p;
    AsmArgs a{};
    a.u = s.F(SG_F_U);
    a.v = s.F(SG_F_V);
    a.pf = s.F(SG_F_P);
    a.APn = s.F(is_u ? SG_F_CU_AP : SG_F_CV_AP);
    a.pk = face ? (is_u ? s.fku : s.fkv) : (is_u ? s.pku : s.pkv);
    a.ld = s.ld;
    a.h = is_u ? p.dy : p.dx;
    a.poff = is_u ? 1 : -static_cast<long>(s.ld);
    a.ubc = is_u && with_bc ? s.F(SG_F_U) : nullptr;
    a.cv = mom_conv(s);
    a.ap_i0 = 0, a.ap_i1 = -1, a.ap_j0 = 0, a.ap_j1 = -1;
    const bool skip = face && ap_in_sweep;
    if (skip) { // the sweep range (mom_args), whose first sweep writes AP
        const MomArgs m = mom_args(s, is_u);
        a.ap_i0 = m.i0, a.ap_i1 = m.i1, a.ap_j0 = m.j0, a.ap_j1 = m.j1;
    }
    s.ap_in_sweep[is_u ? 0 : 1] = skip;
    const dim3 grid(cdiv(a.pk.kp, BX), cdiv(a.pk.rows, AY));
    if (face) {
        const dim3 fg(cdiv(a.pk.aj1 - a.pk.aj0 + 1, FRX), cdiv(a.pk.rows, FRY)), fb(FRX, FRY);
        if (a.pk.ncol == 2) {
            if (is_u)
                emit<AsmArgs, k_assemble_face_rm<true, 2>, FRX * FRY>(s, fg, fb, 0, a, stop);
            else
                emit<AsmArgs, k_assemble_face_rm<false, 2>, FRX * FRY>(s, fg, fb, 0, a, stop);
        } else if (is_u) {
            emit<AsmArgs, k_assemble_face_rm<true, 5>, FRX * FRY>(s, fg, fb, 0, a, stop);
        } else {
            emit<AsmArgs, k_assemble_face_rm<false, 5>, FRX * FRY>(s, fg, fb, 0, a, stop);
        }
    } else if (is_u) {
        emit<AsmArgs, k_assemble<true>, BX * AY>(s, grid, dim3(BX, AY), 0, a, stop);
    } else {
        emit<AsmArgs, k_assemble<false>, BX * AY>(s, grid, dim3(BX, AY), 0, a, stop);
    }
    s.face[is_u ? 0 : 1] = face;
    s.pt_stale[is_u ? 0 : 1] = false;
    s.dc_stale[is_u ? 0 : 1] = false;
}

bool face_ok(const Solver& s, bool is_u) {
    const sg_params& p = s.p;
    return p.momentum_solver == SG_MOM_GS_MULTICOLOUR &&
           (is_u ? p.maxit_inner_u : p.maxit_inner_v) > 0;
}

void ensure_packed(Solver& s, bool is_u) {
    if (!s.face[is_u ? 0 : 1])
        return;
    s.face[is_u ? 0 : 1] = false;
    const Pack& pk = is_u ? s.pku : s.pkv;
    const ExpArgs e{is_u ? s.fku : s.fkv, pk, mom_conv(s),
                    s.F(is_u ? SG_F_CU_AP : SG_F_CV_AP), s.ld};
    const dim3 grid(cdiv(pk.aj1 - pk.aj0 + 1, BX), cdiv(pk.rows, BY)), block(BX, BY);
    if (pk.ncol == 2)
        hipLaunchKernelGGL(k_expand<2>, grid, block, 0, s.stream, e);
    else
        hipLaunchKernelGGL(k_expand<5>, grid, block, 0, s.stream, e);
    SG_HIP(hipGetLastError());
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
    need_coefs(s, is_u);
    refresh_pterm(s, is_u);
    ensure_packed(s, is_u);
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
// Threads of a sweep's workgroup: 512 (two workgroups per CU, the VGPR limit), except a first
// sweep whose two LDS copies and residual buffer would not let two workgroups share a CU: that
// one takes ST threads, the CU's share in one workgroup.
constexpr int SW_NT = 512;
template <int NCOL, int TI, int TJ> constexpr int first_sweep_threads() {
    constexpr SweepReach<NCOL> RC{};
    constexpr int R = RC.R;
    constexpr size_t lds = sizeof(double) * ((TI + RC.load.n + RC.load.s) * (TJ + RC.load.w + RC.load.e) +
                                             (TI + 2 * R) * (TJ + 2 * R) + TI * TJ);
    return 2 * lds <= 160 * 1024 ? SW_NT : ST;
}

template <int NCOL, bool FIRST, bool FACE, int K>
void sweep_shape(Solver& s, const SweepArgs& a, const int* stop) {
    constexpr int TI = MOM_TILES[K].ti, TJ = MOM_TILES[K].tj;
    constexpr int NT = FIRST ? first_sweep_threads<NCOL, TI, TJ>() : SW_NT;
    emit<SweepArgs, k_mom_sweep<NCOL, TI, TJ, FIRST, NT, FACE>, NT>(
        s, dim3(cdiv(s.nc, TJ), cdiv(s.nr, TI)), dim3(NT), 0, a, stop);
}

// Face form: every tile shape (tunable). Coefficient pack (the per-phase entry points after an
// explicit assembly or upload): one shape.
constexpr int PACK_TILE = 4;
template <int NCOL, bool FIRST, int... K>
void sweep_tile_f(Solver& s, const SweepArgs& a, bool face, const int* stop,
                  std::integer_sequence<int, K...>) {
    static_assert(sizeof...(K) == N_MOM_TILES);
    if (!face) {
        sweep_shape<NCOL, FIRST, false, PACK_TILE>(s, a, stop);
        return;
    }
    const int t = FIRST && s.mom_tile_first >= 0 ? s.mom_tile_first : s.mom_tile;
    ((t == K ? sweep_shape<NCOL, FIRST, true, K>(s, a, stop) : void()), ...);
}

template <int NCOL>
void sweep_tile(Solver& s, const MomArgs& m, const double* src, double* dst, double* cell,
                int first, bool face, const int* stop, bool bcv, double* apn) {
    const sg_params& p = s.p;
    SweepArgs a{m, src, dst, cell, s.nr, s.nc, first, 0, -1, p.H, p.NY + p.H + 1, apn};
    if (bcv) // the rows of launch_bc's v ghost columns
        a.bi0 = p.H, a.bi1 = p.NX + p.H - 1;
    constexpr auto shapes = std::make_integer_sequence<int, N_MOM_TILES>{};
    if (first)
        sweep_tile_f<NCOL, true>(s, a, face, stop, shapes);
    else
        sweep_tile_f<NCOL, false>(s, a, face, stop, shapes);
}
} // namespace

// One multicolour sweep of the field into dst (tile shape s.mom_tile); first: also the residual
// cells.
void launch_mom_sweep_once(Solver& s, bool is_u, const double* src, double* dst, bool first,
                           const int* stop, bool bcv) {
    const MomArgs m = mom_args(s, is_u);
    double* cell = s.rcell[is_u ? 0 : 1];
    const bool face = s.face[is_u ? 0 : 1];
    if (bcv && (is_u || !first))
        throw HipError("v's boundary condition rides only on v's first sweep");
    // The AP the face assembly left out, written by this sweep's residual pass.
    bool& pend = s.ap_in_sweep[is_u ? 0 : 1];
    double* apn = first && face && pend ? s.F(is_u ? SG_F_CU_AP : SG_F_CV_AP) : nullptr;
    if (apn)
        pend = false;
    if (m.pk.ncol == 2)
        sweep_tile<2>(s, m, src, dst, cell, first ? 1 : 0, face, stop, bcv, apn);
    else
        sweep_tile<5>(s, m, src, dst, cell, first ? 1 : 0, face, stop, bcv, apn);
}

bool bcv_in_sweep(const Solver& s) {
    return s.p.momentum_solver == SG_MOM_GS_MULTICOLOUR && s.p.maxit_inner_v > 0;
}

// sweep_u_batch / sweep_v_batch (src/momentum_u.cpp:164-171) for the two parallel solvers. With
// res_slot (multicolour only), the residual is evaluated inside the first sweep and summed (on
// flush_sums if defer) into *res_slot.
void launch_mom_sweeps(Solver& s, bool is_u, int nsweeps, const int* stop, double* res_slot,
                       bool defer, bool bcv) {
    const sg_params& p = s.p;
    need_coefs(s, is_u);
    refresh_pterm(s, is_u);
    const MomArgs m = mom_args(s, is_u);
    const bool fuse = res_slot && p.momentum_solver == SG_MOM_GS_MULTICOLOUR && nsweeps > 0;
    if (bcv && !fuse)
        throw HipError("v's boundary condition needs a first multicolour sweep");
    if (res_slot && !fuse)
        launch_mom_residual(s, is_u, res_slot, stop, defer);
    if (nsweeps <= 0)
        return;
    if (p.momentum_solver == SG_MOM_GS_MULTICOLOUR) {
        // Every sweep writes its whole destination array. An even count alternates field ->
        // scratch -> field ...; an odd count above one goes field -> scratch -> scratch2 ->
        // scratch ... -> scratch2 -> field, so that the last sweep writes the field; a single
        // sweep ends with a copy back.
        const bool three = nsweeps % 2 == 1 && nsweeps > 1;
        auto out = [&](int k) -> double* {
            if (k == nsweeps - 1 && three)
                return m.x;
            return k % 2 == 0 ? s.scratch : (three ? s.scratch2 : m.x);
        };
        for (int k = 0; k < nsweeps; ++k)
            launch_mom_sweep_once(s, is_u, k == 0 ? m.x : out(k - 1), out(k), fuse && k == 0,
                                  stop, bcv && k == 0);
        if (nsweeps == 1)
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
    ensure_packed(s, is_u);
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
        need_coefs(s, cu);
        ensure_packed(s, cu);
    }
    if (cu || cv) {
        const Pack& pk = cu ? s.pku : s.pkv;
        const int a = id - (cu ? SG_F_CU_AE : SG_F_CV_AE);
        hipLaunchKernelGGL(k_pack, dim3(cdiv(pk.aj1 - pk.aj0 + 1, BX), cdiv(pk.rows, BY)),
                           dim3(BX, BY), 0, s.stream, pk, s.F(id), s.ld, a, to_packed ? 1 : 0);
        SG_HIP(hipGetLastError());
    } else if (id == SG_F_P && to_packed) {
        s.pt_stale[0] = s.pt_stale[1] = true;
    }
}

bool mom_dc_ok(const sg_params& p) {
    return p.momentum_precision == SG_MOMP_MIXED && p.momentum_solver == SG_MOM_GS_MULTICOLOUR &&
           p.maxit_inner_u > 0 && p.maxit_inner_v > 0;
}

void need_coefs(const Solver& s, bool is_u) {
    if (s.dc_stale[is_u ? 0 : 1])
        throw std::runtime_error(
            std::string("the ") + (is_u ? "u" : "v") +
            "-momentum coefficients of the last outer iteration are not kept with "
            "momentum_precision = mixed (its sweeps use an FP32 copy); call sg_assemble_" +
            (is_u ? "u" : "v") + " first");
}

void launch_asm_dc(Solver& s, bool is_u, const int* stop, bool with_bc) {
    const sg_params& p = s.p;
    AsmDcArgs d{};
    AsmArgs& a = d.a;
    a.u = s.F(SG_F_U);
    a.v = s.F(SG_F_V);
    a.pf = s.F(SG_F_P);
    a.APn = s.F(is_u ? SG_F_CU_AP : SG_F_CV_AP);
    a.pk = is_u ? s.pku : s.pkv; // geometry only: the coefficient pack itself is not written
    a.ld = s.ld;
    a.h = is_u ? p.dy : p.dx;
    a.poff = is_u ? 1 : -static_cast<long>(s.ld);
    a.ubc = is_u && with_bc ? s.F(SG_F_U) : nullptr;
    a.cv = mom_conv(s);
    const MomArgs m = mom_args(s, is_u);
    d.dc = s.dcf[is_u ? 0 : 1];
    d.alpha = m.alpha;
    d.cell = s.rcell[is_u ? 0 : 1];
    d.i0 = m.i0, d.i1 = m.i1, d.j0 = m.j0, d.j1 = m.j1;
    const dim3 grid(cdiv(a.pk.kp, FBX), cdiv(a.pk.rows, FAY));
    if (is_u)
        emit<AsmDcArgs, k_asm_dc<true>, FBX * FAY>(s, grid, dim3(FBX, FAY), 0, d, stop);
    else
        emit<AsmDcArgs, k_asm_dc<false>, FBX * FAY>(s, grid, dim3(FBX, FAY), 0, d, stop);
    const int k = is_u ? 0 : 1;
    s.dc_stale[k] = true;
    s.face[k] = false;
    s.pt_stale[k] = false;
}

namespace {
template <int NCOL, bool FIRST, bool LAST, int K>
void dc_sweep_k(Solver& s, const Dc32Args& a, const int* stop) {
    constexpr int TI = DC_TILES[K].ti, TJ = DC_TILES[K].tj, NT = DC_TILES[K].nt;
    emit<Dc32Args, k_dc_sweep<NCOL, TI, TJ, FIRST, LAST, NT>, NT>(
        s, dim3(cdiv(s.nc, TJ), cdiv(s.nr, TI)), dim3(NT), 0, a, stop);
}

template <int NCOL, bool FIRST, bool LAST, int... K>
void dc_sweep(Solver& s, const Dc32Args& a, const int* stop, std::integer_sequence<int, K...>) {
    static_assert(sizeof...(K) == N_DC_TILES);
    ((s.dc_tile == K ? dc_sweep_k<NCOL, FIRST, LAST, K>(s, a, stop) : void()), ...);
}

template <int NCOL> void dc_sweep_any(Solver& s, const Dc32Args& a, bool first, bool last,
                                      const int* stop) {
    constexpr auto shapes = std::make_integer_sequence<int, N_DC_TILES>{};
    if (first && last)
        dc_sweep<NCOL, true, true>(s, a, stop, shapes);
    else if (first)
        dc_sweep<NCOL, true, false>(s, a, stop, shapes);
    else if (last)
        dc_sweep<NCOL, false, true>(s, a, stop, shapes);
    else
        dc_sweep<NCOL, false, false>(s, a, stop, shapes);
}
} // namespace

namespace {
Dc32Args dc_args(Solver& s, bool is_u) {
    const MomArgs m = mom_args(s, is_u);
    Dc32Args a{};
    a.dc = s.dcf[is_u ? 0 : 1];
    a.g = is_u ? s.pku : s.pkv;
    a.x = m.x;
    a.ld = s.ld, a.nr = s.nr, a.nc = s.nc;
    a.i0 = m.i0, a.i1 = m.i1, a.j0 = m.j0, a.j1 = m.j1;
    a.oma = static_cast<float>(m.one_minus_alpha);
    return a;
}
} // namespace

void launch_dc_sweep_probe(Solver& s, bool is_u, const int* stop) {
    Dc32Args a = dc_args(s, is_u);
    float* const* dd = s.dd + (is_u ? 0 : 2);
    a.src = dd[0];
    a.dst = dd[1];
    if (a.g.ncol == 2)
        dc_sweep_any<2>(s, a, false, false, stop);
    else
        dc_sweep_any<5>(s, a, false, false, stop);
}

void launch_dc_sweeps(Solver& s, bool is_u, int nsweeps, const int* stop) {
    const MomArgs m = mom_args(s, is_u);
    Dc32Args a{};
    a.dc = s.dcf[is_u ? 0 : 1];
    a.g = is_u ? s.pku : s.pkv;
    a.x = m.x;
    a.ld = s.ld, a.nr = s.nr, a.nc = s.nc;
    a.i0 = m.i0, a.i1 = m.i1, a.j0 = m.j0, a.j1 = m.j1;
    a.oma = static_cast<float>(m.one_minus_alpha);
    float* const* dd = s.dd + (is_u ? 0 : 2); // this component's pair
    for (int k = 0; k < nsweeps; ++k) {
        a.src = k > 0 ? dd[(k - 1) & 1] : nullptr;
        a.dst = dd[k & 1];
        if (m.pk.ncol == 2)
            dc_sweep_any<2>(s, a, k == 0, k == nsweeps - 1, stop);
        else
            dc_sweep_any<5>(s, a, k == 0, k == nsweeps - 1, stop);
    }
}

void refresh_pterm(Solver& s, bool is_u) {
    const int k = is_u ? 0 : 1;
    if (!s.pt_stale[k])
        return;
    s.pt_stale[k] = false;
    const Pack& pk = is_u ? s.pku : s.pkv;
    // The face pack covers the assembly range plus a column west and a row south.
    const Pack& tgt = s.face[k] ? (is_u ? s.fku : s.fkv) : pk;
    hipLaunchKernelGGL(k_pterm, dim3(cdiv(pk.aj1 - pk.aj0 + 1, BX), cdiv(pk.rows, BY)),
                       dim3(BX, BY), 0, s.stream, tgt, s.face[k] ? int(F_PT) : 9, pk.ai1, pk.aj0,
                       s.F(SG_F_P), s.ld, is_u ? 1L : -static_cast<long>(s.ld),
                       is_u ? s.p.dy : s.p.dx);
    SG_HIP(hipGetLastError());
}

} // namespace sg
