// sg_mg.hip — geometric multigrid for p' on gfx950 (src/multigrid.cpp): Galerkin 2x2
// aggregation, red-black smoother, residual, restriction, prolongation with over-correction and
// the coarsest-level solve.
//
// Two forms, one set of kernels templated on the element type:
//  - fp64 (the reference's method): every expression in the reference's order and every sum in
//    its fixed order, so p' is bit-identical to the CPU;
//  - mixed (mg_precision = mixed): the V-cycle (smoothing, residual, restriction, prolongation,
//    coarse solve) runs in FP32 on FP32 copies of the operators, inside an FP64 defect
//    correction: FP64 stays in the mean removal of b, the fine-grid defect b - A p', the update
//    p' += e and the final mean removal.
//
// Launch structure of one V-cycle:
//  - levels above the tail (more than tail_cells cells): one launch per smoother colour pass;
//    the residual is fused into the restriction (each coarse cell evaluates its 4 fine
//    residuals and sums them in the reference's order), and the prolongation is fused into the
//    first post-smoothing pass (a colour-0 update reads its colour-1 neighbours as x + omega*xc,
//    the value the separate prolongation would have stored; colour-1 cells are overwritten by the
//    next pass before anything reads them);
//  - the tail (every level with at most tail_cells cells, down to the coarsest, and its coarse
//    solve) runs in one workgroup, one launch per V-cycle, synchronising with barriers.
#include "sg_internal.hpp"
#include "sg_warm.hpp"

#include <algorithm>

namespace sg {

namespace {

// Elementwise level kernels (the mixed form's FP64 <-> FP32 conversions, the unstaged V-cycle):
// 256 x 1 workgroups (3 us per mixed outer iteration faster at 1024^2 than 64 x 4).
constexpr int BX = 256, BY = 1;
constexpr int MAXLEV = 24;

template <class T> struct Lv {
    using value_type = T;
    int nx, ny, stride;
    T *AE, *AW, *AN, *AS, *AP, *b, *x, *r;
    T* xt; // pre-smoothed iterate of the tiled V-cycle
};

// A level held in LDS, addressed through LDS pointers, so its accesses compile to LDS
// instructions (through a generic pointer they would be flat accesses).
template <class T> using Lp = __attribute__((address_space(3))) T*;
template <class T> struct LLv {
    using value_type = T;
    int nx, ny, stride;
    Lp<T> AE, AW, AN, AS, AP, b, x, r;
};

Lv<double> lv64(const MgLevel& L) {
    return {L.nx, L.ny, L.stride, L.AE, L.AW, L.AN, L.AS, L.AP, L.b, L.x, L.r, L.xt};
}
Lv<float> lv32(const MgLevel& L) {
    return {L.nx, L.ny, L.ny + 2, L.fAE, L.fAW, L.fAN, L.fAS, L.fAP, L.fb, L.fx, L.fr, L.fxt};
}

// c ? a : b by bit masking. The empty asm hides the mask's origin, so the compiler cannot fold
// the blend back into a select and then into a branch around a's computation.
__device__ __forceinline__ double blend(bool c, double a, double b) {
    long long m = -static_cast<long long>(c);
    asm volatile("" : "+v"(m));
    return __longlong_as_double((__double_as_longlong(a) & m) | (__double_as_longlong(b) & ~m));
}
__device__ __forceinline__ float blend(bool c, float a, float b) {
    int m = -static_cast<int>(c);
    asm volatile("" : "+v"(m));
    return __int_as_float((__float_as_int(a) & m) | (__float_as_int(b) & ~m));
}

// Relative tolerance of the coarsest-level solve: the reference's 1e-10 in FP64; in FP32 a
// tolerance the arithmetic can reach.
template <class T> __device__ constexpr T coarse_tol();
template <> __device__ constexpr double coarse_tol<double>() { return 1e-10; }
template <> __device__ constexpr float coarse_tol<float>() { return 1e-6f; }

// sqrt(rr) < c for a fixed c > 0, decided as the correctly rounded root decides it, mostly
// without taking the root. Outside [t (1 - g), t (1 + g)], t = c * c rounded, the root lies at
// least about g / 2 relatively below or above c, far more than the few roundings in t, in the
// bounds and in the root (g = 1e-12 in FP64, 1e-5 in FP32), so the comparison of rr decides;
// inside, or when t is outside the normal range, the root is taken. NaN reaches the root.
template <class T> struct RootBelow {
    T c, lo, hi;
    bool direct;
    __device__ explicit RootBelow(T c_) : c(c_) {
        constexpr bool d = sizeof(T) == 8;
        const T t = c * c, g = d ? T(1e-12) : T(1e-5);
        direct = !(t > (d ? T(1e-280) : T(1e-30)) && t < (d ? T(1e280) : T(1e30)));
        lo = t * (T(1) - g);
        hi = t * (T(1) + g);
    }
    __device__ bool operator()(T rr) const {
        if (!direct) {
            if (rr < lo)
                return true;
            if (rr > hi)
                return false;
        }
        return sqrt(rr) < c;
    }
};

// ---- per-cell pieces (src/multigrid.cpp) ----

// Coarse links: sums of the two fine links crossing each coarse face; AP = ae + aw + an + as
// (mg_build, :232-261).
template <class LF, class LC>
__device__ __forceinline__ void build_cell(const LF& F, const LC& C, int I, int J) {
    using T = typename LF::value_type;
    const long fs = F.stride;
    const size_t f = 2 * I * fs + 2 * J, c = static_cast<size_t>(I) * C.stride + J;
    const T ae = F.AE[f + 1] + F.AE[f + fs + 1];
    const T aw = F.AW[f] + F.AW[f + fs];
    const T an = F.AN[f] + F.AN[f + 1];
    const T as = F.AS[f + fs] + F.AS[f + fs + 1];
    C.AE[c] = ae;
    C.AW[c] = aw;
    C.AN[c] = an;
    C.AS[c] = as;
    C.AP[c] = ae + aw + an + as;
}

// Two-statement smoother update (smooth, :62-64).
template <class LV> __device__ __forceinline__ void smooth_cell(const LV& L, size_t q) {
    using T = typename LV::value_type;
    const long s = L.stride;
    const auto x = L.x;
    T t = L.b[q] / L.AP[q];
    t += (L.AE[q] * x[q + 1] + L.AW[q] * x[q - 1] + L.AS[q] * x[q + s] + L.AN[q] * x[q - s]) /
         L.AP[q];
    L.x[q] = t;
}

// b - (AP x - (E + W + S + N)) (residual, :83-84).
template <class LV>
__device__ __forceinline__ typename LV::value_type resid_cell(const LV& L, size_t q) {
    const long s = L.stride;
    const auto x = L.x;
    return L.b[q] - (L.AP[q] * x[q] - (L.AE[q] * x[q + 1] + L.AW[q] * x[q - 1] +
                                       L.AS[q] * x[q + s] + L.AN[q] * x[q - s]));
}

// ---- multi-workgroup kernels for the large levels ----

// Two levels (fine F, coarse C) and a colour or over-correction weight.
template <class T> struct FC {
    Lv<T> F, C;
    T omega;
    int col;
};

template <class T> __device__ __forceinline__ void k_mg_build(const FC<T>& a, const Blk& bk) {
    const int J = bk.x * BX + threadIdx.x, I = bk.y * BY + threadIdx.y;
    if (I < a.C.nx && J < a.C.ny)
        build_cell(a.F, a.C, I, J);
}

// One colour of the red-black smoother, colour (I + J) & 1 (smooth, :46-70).
template <class T> __device__ __forceinline__ void k_mg_smooth(const FC<T>& a, const Blk& bk) {
    const Lv<T>& L = a.F;
    const int col = a.col;
    const int I = bk.y * BY + threadIdx.y;
    const int J = ((I + col) & 1) + 2 * (bk.x * BX + threadIdx.x);
    if (I < L.nx && J < L.ny)
        smooth_cell(L, static_cast<size_t>(I) * L.stride + J);
}

// Residual and restriction fused: coarse b = r(NW) + r(NE) + r(SW) + r(SE), coarse x = 0
// (residual :73-87 then restrict_residual :90-103).
template <class T> __device__ __forceinline__ void k_mg_resid_restrict(const FC<T>& a, const Blk& bk) {
    const Lv<T>& F = a.F;
    const Lv<T>& C = a.C;
    const int J = bk.x * BX + threadIdx.x, I = bk.y * BY + threadIdx.y;
    if (I >= C.nx || J >= C.ny)
        return;
    const long fs = F.stride;
    const size_t f = 2 * I * fs + 2 * J, c = static_cast<size_t>(I) * C.stride + J;
    const T r0 = resid_cell(F, f), r1 = resid_cell(F, f + 1);
    const T r2 = resid_cell(F, f + fs), r3 = resid_cell(F, f + fs + 1);
    C.b[c] = r0 + r1 + r2 + r3;
    C.x[c] = T(0);
}

// Prolongation fused into the first post-smoothing pass (colour 0): neighbours inside the level
// are read as x + omega * xc (prolong, :106-116).
template <class T> __device__ __forceinline__ void k_mg_prolong_smooth0(const FC<T>& a, const Blk& bk) {
    const Lv<T>& F = a.F;
    const Lv<T>& C = a.C;
    const T omega = a.omega;
    const int I = bk.y * BY + threadIdx.y;
    const int J = (I & 1) + 2 * (bk.x * BX + threadIdx.x);
    if (I >= F.nx || J >= F.ny)
        return;
    const long s = F.stride;
    const size_t q = static_cast<size_t>(I) * s + J;
    const T* x = F.x;
    const T* xc = C.x;
    const long cs = C.stride;
    auto px = [&](int i, int j, size_t qq) {
        return (i >= 0 && i < F.nx && j >= 0 && j < F.ny)
                   ? x[qq] + omega * xc[static_cast<size_t>(i / 2) * cs + j / 2]
                   : x[qq];
    };
    const T xe = px(I, J + 1, q + 1), xw = px(I, J - 1, q - 1);
    const T xs = px(I + 1, J, q + s), xn = px(I - 1, J, q - s);
    T t = F.b[q] / F.AP[q];
    t += (F.AE[q] * xe + F.AW[q] * xw + F.AS[q] * xs + F.AN[q] * xn) / F.AP[q];
    F.x[q] = t;
}

// Plain prolongation (only when mg_post = 0).
template <class T> __device__ __forceinline__ void k_mg_prolong(const FC<T>& a, const Blk& bk) {
    const Lv<T>& F = a.F;
    const Lv<T>& C = a.C;
    const int j = bk.x * BX + threadIdx.x, i = bk.y * BY + threadIdx.y;
    if (i < F.nx && j < F.ny)
        F.x[static_cast<size_t>(i) * F.stride + j] +=
            a.omega * C.x[static_cast<size_t>(i / 2) * C.stride + j / 2];
}

// ---- tiled V-cycle kernels for the levels above the tail ----
//
// down: pre-smoothing of a TI x TJ tile in LDS (every colour pass of every pre-sweep; pass p runs
// over the tile grown by 2*pre - p cells, so the tile and its one-cell ring end exactly as after
// the global passes), then the residual and restriction of the tile's coarse cells (coarse x = 0)
// and, on the first V-cycle, their Galerkin operator. The iterate is read from x and written to
// xt, so no tile reads cells another tile has already updated.
// up: the tile of xt prolonged (x + omega * xc, interior cells only) into LDS, every post-smoothing
// pass (pass p over the tile grown by 2*post - 1 - p cells), written to x.

constexpr int MT_I = 32, MT_J = 64, MT_N = 256; // tile rows, columns, threads
constexpr int MT_MAXH = 8;                     // largest halo handled (2*pre + 1, 2*post)

template <class T>
__device__ __forceinline__ void tile_load(T* sx, const Lv<T>& L, const T* x, int gi0, int gj0,
                                          int LH, int LW, const Lv<T>* C, T omega) {
    constexpr int B = 8; // loads in flight per thread
    for (int base = 0; base < LH * LW; base += B * MT_N) {
        T v[B];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int e = base + k * MT_N + threadIdx.x, a = e / LW, b = e - a * LW;
            const int I = gi0 + a, J = gj0 + b;
            T val = T(0);
            if (e < LH * LW && I >= -1 && I <= L.nx && J >= -1 && J <= L.ny) {
                val = x[static_cast<long>(I) * L.stride + J];
                if (C && I >= 0 && I < L.nx && J >= 0 && J < L.ny)
                    val = val + omega * C->x[static_cast<size_t>(I / 2) * C->stride + J / 2];
            }
            v[k] = val;
        }
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int e = base + k * MT_N + threadIdx.x;
            if (e < LH * LW)
                sx[e] = v[k];
        }
    }
    __syncthreads();
}

// Colour pass `col` over rows [ia, ib], columns [ja, jb] of the level, x in LDS. Each thread
// gathers the operator of up to B cells (all loads in flight) before updating them.
template <class T>
__device__ __forceinline__ void tile_pass(T* sx, const Lv<T>& L, int gi0, int gj0, int LW, int col,
                                          int ia, int ib, int ja, int jb) {
    constexpr int B = 6;
    if (ia <= ib && ja <= jb) {
        const int cpr = (jb - ja + 2) / 2, n = (ib - ia + 1) * cpr;
        for (int base = 0; base < n; base += B * MT_N) {
            T cb[B], cp[B], ce[B], cw[B], cs[B], cn[B];
            int off[B];
#pragma unroll
            for (int k = 0; k < B; ++k) {
                const int e = base + k * MT_N + threadIdx.x;
                const int r = e / cpr, I = ia + r;
                const int J = ja + ((I + ja + col) & 1) + 2 * (e - r * cpr);
                off[k] = -1;
                if (e < n && J <= jb) {
                    const size_t q = static_cast<size_t>(I) * L.stride + J;
                    cb[k] = L.b[q];
                    cp[k] = L.AP[q];
                    ce[k] = L.AE[q];
                    cw[k] = L.AW[q];
                    cs[k] = L.AS[q];
                    cn[k] = L.AN[q];
                    off[k] = (I - gi0) * LW + (J - gj0);
                }
            }
#pragma unroll
            for (int k = 0; k < B; ++k)
                if (off[k] >= 0) {
                    T* c = sx + off[k];
                    T t = cb[k] / cp[k];
                    t += (ce[k] * c[1] + cw[k] * c[-1] + cs[k] * c[LW] + cn[k] * c[-LW]) / cp[k];
                    *c = t;
                }
        }
    }
    __syncthreads();
}

// Tiled down / up step of one level: F, C, the iterate read (xin) and written (xout), the sweep
// count (pre or post), whether to build the coarse operator, the over-correction weight.
template <class T> struct DU {
    Lv<T> F, C;
    const T* xin;
    T* xout;
    int sweeps, build;
    T omega;
    const double* bmean = nullptr; // level 0, b's mean not removed: b - *bmean / bn is read
    double bn = 0.0;
    int xzero = 0;                 // the iterate is zero (first V-cycle from p' = 0): not read
    // Mixed form, level 0 of the last V-cycle: the result is not stored in FP32 but added to the
    // FP64 p' (k_mx_update's update: x64 = (first64 ? 0 : x64) + e), stride ld64.
    double* x64 = nullptr;
    int ld64 = 0, first64 = 0;
};

template <class T> __device__ __forceinline__ void k_mg_down(const DU<T>& a, const Blk& bk) {
    const Lv<T>& F = a.F;
    const Lv<T>& C = a.C;
    const T* xin = a.xin;
    T* xout = a.xout;
    const int pre = a.sweeps, build = a.build;
    extern __shared__ __align__(16) unsigned char smem[];
    T* sx = reinterpret_cast<T*>(smem);
    const int H = 2 * pre + 1, LW = MT_J + 2 * H, LH = MT_I + 2 * H;
    int tx, ty;
    xcd_tile(bk, tx, ty);
    const int ti0 = ty * MT_I, tj0 = tx * MT_J;
    const int gi0 = ti0 - H, gj0 = tj0 - H;
    tile_load<T>(sx, F, xin, gi0, gj0, LH, LW, nullptr, T(0));
    for (int p = 0; p < 2 * pre; ++p) {
        const int d = 2 * pre - p;
        tile_pass(sx, F, gi0, gj0, LW, p & 1, max(0, ti0 - d), min(F.nx - 1, ti0 + MT_I - 1 + d),
                  max(0, tj0 - d), min(F.ny - 1, tj0 + MT_J - 1 + d));
    }
    // Residual and restriction of the tile's coarse cells; their operator on the first cycle.
    const int ci0 = ti0 / 2, cj0 = tj0 / 2;
    const int cn = min(MT_I / 2, C.nx - ci0), cm = min(MT_J / 2, C.ny - cj0);
    const long fs = F.stride;
    for (int e = threadIdx.x; e < cn * cm; e += MT_N) {
        const int I = ci0 + e / cm, J = cj0 + e % cm;
        const size_t f = 2 * I * fs + 2 * J, c = static_cast<size_t>(I) * C.stride + J;
        T r4[4];
        const size_t fq[4] = {f, f + 1, f + fs, f + fs + 1};
        const int fi[4] = {2 * I, 2 * I, 2 * I + 1, 2 * I + 1}, fj[4] = {2 * J, 2 * J + 1, 2 * J, 2 * J + 1};
        for (int k = 0; k < 4; ++k) {
            const size_t q = fq[k];
            const T* x = sx + (fi[k] - gi0) * LW + (fj[k] - gj0);
            r4[k] = F.b[q] - (F.AP[q] * x[0] - (F.AE[q] * x[1] + F.AW[q] * x[-1] +
                                                 F.AS[q] * x[LW] + F.AN[q] * x[-LW]));
        }
        C.b[c] = r4[0] + r4[1] + r4[2] + r4[3];
        C.x[c] = T(0);
        if (build)
            build_cell(F, C, I, J);
    }
#pragma unroll
    for (int k = 0; k < MT_I * MT_J / MT_N; ++k) {
        const int e = k * MT_N + threadIdx.x, I = ti0 + e / MT_J, J = tj0 + e % MT_J;
        if (I < F.nx && J < F.ny)
            xout[static_cast<size_t>(I) * F.stride + J] = sx[(I - gi0) * LW + (J - gj0)];
    }
}

template <class T> __device__ __forceinline__ void k_mg_up(const DU<T>& a, const Blk& bk) {
    const Lv<T>& F = a.F;
    const Lv<T>& C = a.C;
    const T* xpre = a.xin;
    T* xfin = a.xout;
    const T omega = a.omega;
    const int post = a.sweeps;
    extern __shared__ __align__(16) unsigned char smem[];
    T* sx = reinterpret_cast<T*>(smem);
    const int H = 2 * post, LW = MT_J + 2 * H, LH = MT_I + 2 * H;
    int tx, ty;
    xcd_tile(bk, tx, ty);
    const int ti0 = ty * MT_I, tj0 = tx * MT_J;
    const int gi0 = ti0 - H, gj0 = tj0 - H;
    tile_load<T>(sx, F, xpre, gi0, gj0, LH, LW, &C, omega);
    for (int p = 0; p < 2 * post; ++p) {
        const int d = 2 * post - 1 - p;
        tile_pass(sx, F, gi0, gj0, LW, p & 1, max(0, ti0 - d), min(F.nx - 1, ti0 + MT_I - 1 + d),
                  max(0, tj0 - d), min(F.ny - 1, tj0 + MT_J - 1 + d));
    }
#pragma unroll
    for (int k = 0; k < MT_I * MT_J / MT_N; ++k) {
        const int e = k * MT_N + threadIdx.x, I = ti0 + e / MT_J, J = tj0 + e % MT_J;
        if (I < F.nx && J < F.ny)
            xfin[static_cast<size_t>(I) * F.stride + J] = sx[(I - gi0) * LW + (J - gj0)];
    }
}

// ---- staged tiles for the levels between level 0 and the tail ----
//
// The same steps as k_mg_down / k_mg_up, but x, b and the operator of the tile and its halo are
// loaded into LDS at once (AE and AN, plus AP in FP32: the operator is symmetric, below), so the
// smoothing passes, the residual and the Galerkin build read LDS only: on these small levels a
// pass is bounded by load latency, not bandwidth. Tiles of TI rows x TS columns: on large levels
// 24 x 32 in FP64 (the four arrays of the tile and its halo fill 46 KB, so several workgroups of
// 512 threads share a CU, some loading while others smooth; 16 x 32 and 32 x 32 measure the same)
// and 32 x 32 in FP32 (512 threads); 16 x 16 on small levels (256 threads), which then still
// spread over many CUs. Each choice was the fastest measured.
template <class T, int TS, int TI = TS> constexpr int sg_n() {
    return TI != TS ? 512 : TS == 32 ? (sizeof(T) == 8 ? 1024 : 512) : 256;
}
template <class T> constexpr int BR = sizeof(T) == 8 ? 24 : 32; // rows of a large-level tile
// Levels with at least this many cells take 32 x 32 tiles.
constexpr long SG_BIG = 512L * 512;

// The p' operator is symmetric bit for bit on every level: AW(I, J) = AE(I, J - 1) and
// AS(I, J) = AN(I + 1, J). Level 0 divides the same APU / APV entry for both links of a face
// (src/pressure.cpp:21-24) and zeroes the wall links, as the zero halo is; a coarse level sums the
// same fine links for both sides of a coarse face (mg_build, src/multigrid.cpp:246-253), and levels
// exist only while both sizes are even (mg_level_dims, :143-151), so every coarse face has fine
// faces on both sides. A tile therefore stages x, b, AE and AN and reads AW and AS at the
// neighbouring cell. In FP64 it also recomputes AP = AE + AW + AN + AS, in the order of
// src/pressure.cpp:45 and src/multigrid.cpp:254 (so the same bits); the FP32 level 0 holds the
// rounded FP64 AP, so FP32 tiles stage AP as well.
template <class T> constexpr bool SG_DERIVE_AP = sizeof(T) == 8;
template <class T> constexpr int SG_NSTAGE = SG_DERIVE_AP<T> ? 4 : 5; // arrays staged per cell

template <class T> struct Stage {
    Lp<T> x, ae, an, b, ap; // ap: FP32 only
    int gi0, gj0, LW;
    __device__ __forceinline__ int at(int I, int J) const { return (I - gi0) * LW + (J - gj0); }
    __device__ __forceinline__ T aw(int q) const { return ae[q - 1]; }
    __device__ __forceinline__ T as(int q) const { return an[q + LW]; }
    __device__ __forceinline__ T app(int q) const {
        if constexpr (SG_DERIVE_AP<T>)
            return ae[q] + ae[q - 1] + an[q] + an[q + LW];
        else
            return ap[q];
    }
};

template <class T> __device__ __forceinline__ Stage<T> stage_view(int gi0, int gj0, int LH, int LW) {
    extern __shared__ __align__(16) unsigned char smem[];
    const Lp<T> base = (Lp<T>)reinterpret_cast<T*>(smem);
    const int n = LH * LW;
    return {base, base + n, base + 2 * n, base + 3 * n, SG_DERIVE_AP<T> ? nullptr : base + 4 * n,
            gi0, gj0, LW};
}

// x of the tile and its halo (zero outside the level's padded range; prolonged if C), and the
// staged operator inside the level (zero elsewhere: the walls' AW and AS read there).
// bmean: b - *bmean / bn is staged, the value b's mean removal (k_total_sub) would store, by the
// same subtraction; xzero: the iterate is zero everywhere, so it is not read; xring = 0: the ring
// around the level is zero too but not read (an iterate array whose ring holds other data).
template <class T, int NTH>
__device__ __forceinline__ void stage_load(const Stage<T>& st, const Lv<T>& L, const T* x, int LH,
                                           const Lv<T>* C, T omega, const double* bmean = nullptr,
                                           double bn = 0.0, int xzero = 0, int xring = 1) {
    constexpr int B = 4;
    const int n = LH * st.LW;
    const T bm = bmean ? static_cast<T>(*bmean / bn) : T(0);
    for (int base = 0; base < n; base += B * NTH) {
        T vx[B], ve[B], vn[B], vb[B], vp[B];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int e = base + k * NTH + threadIdx.x, a = e / st.LW, bb = e - a * st.LW;
            const int I = st.gi0 + a, J = st.gj0 + bb;
            vx[k] = ve[k] = vn[k] = vb[k] = vp[k] = T(0);
            if (e < n && I >= -1 && I <= L.nx && J >= -1 && J <= L.ny) {
                const long q = static_cast<long>(I) * L.stride + J;
                const bool inside = I >= 0 && I < L.nx && J >= 0 && J < L.ny;
                vx[k] = xzero || (!xring && !inside) ? T(0) : x[q];
                if (inside) {
                    if (C)
                        vx[k] = vx[k] + omega * C->x[static_cast<size_t>(I / 2) * C->stride + J / 2];
                    ve[k] = L.AE[q];
                    vn[k] = L.AN[q];
                    vb[k] = bmean ? L.b[q] - bm : L.b[q];
                    if constexpr (!SG_DERIVE_AP<T>)
                        vp[k] = L.AP[q];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int e = base + k * NTH + threadIdx.x;
            if (e < n) {
                st.x[e] = vx[k];
                st.ae[e] = ve[k];
                st.an[e] = vn[k];
                st.b[e] = vb[k];
                if constexpr (!SG_DERIVE_AP<T>)
                    st.ap[e] = vp[k];
            }
        }
    }
    __syncthreads();
}

// Colour pass `col` over rows [ia, ib], columns [ja, jb], everything in LDS (smooth, :62-64).
template <class T, int NTH>
__device__ __forceinline__ void stage_pass(const Stage<T>& st, int col, int ia, int ib, int ja,
                                           int jb) {
    if (ia <= ib && ja <= jb) {
        const int cpr = (jb - ja + 2) / 2, n = (ib - ia + 1) * cpr;
        const int LW = st.LW;
        for (int e = threadIdx.x; e < n; e += NTH) {
            const int r = e / cpr, I = ia + r;
            const int J = ja + ((I + ja + col) & 1) + 2 * (e - r * cpr);
            if (J > jb)
                continue;
            const int q = st.at(I, J);
            const auto x = st.x;
            const T ap = st.app(q);
            T t = st.b[q] / ap;
            t += (st.ae[q] * x[q + 1] + st.aw(q) * x[q - 1] + st.as(q) * x[q + LW] +
                  st.an[q] * x[q - LW]) /
                 ap;
            st.x[q] = t;
        }
    }
    __syncthreads();
}

// TS columns, TI rows.
template <class T, int TS, int TI = TS> __device__ __forceinline__ void k_mg_down_s(const DU<T>& a, const Blk& bk) {
    constexpr int NTH = sg_n<T, TS, TI>();
    const Lv<T>& F = a.F;
    const Lv<T>& C = a.C;
    const int pre = a.sweeps, build = a.build;
    const int H = 2 * pre + 1, LW = TS + 2 * H, LH = TI + 2 * H;
    int tx, ty;
    xcd_tile(bk, tx, ty);
    const int ti0 = ty * TI, tj0 = tx * TS;
    const Stage<T> st = stage_view<T>(ti0 - H, tj0 - H, LH, LW);
    stage_load<T, NTH>(st, F, a.xin, LH, nullptr, T(0), a.bmean, a.bn, a.xzero);
    for (int p = 0; p < 2 * pre; ++p) {
        const int d = 2 * pre - p;
        stage_pass<T, NTH>(st, p & 1, max(0, ti0 - d), min(F.nx - 1, ti0 + TI - 1 + d), max(0, tj0 - d),
                   min(F.ny - 1, tj0 + TS - 1 + d));
    }
    // Residual and restriction of the tile's coarse cells (residual :83-84, restrict_residual
    // :90-103); their operator on the first cycle (mg_build, :232-261).
    const int ci0 = ti0 / 2, cj0 = tj0 / 2;
    const int cn = min(TI / 2, C.nx - ci0), cm = min(TS / 2, C.ny - cj0);
    for (int e = threadIdx.x; e < cn * cm; e += NTH) {
        const int I = ci0 + e / cm, J = cj0 + e % cm;
        const size_t c = static_cast<size_t>(I) * C.stride + J;
        const int fq[4] = {st.at(2 * I, 2 * J), st.at(2 * I, 2 * J + 1), st.at(2 * I + 1, 2 * J),
                           st.at(2 * I + 1, 2 * J + 1)};
        T r4[4];
        for (int k = 0; k < 4; ++k) {
            const int q = fq[k];
            const auto x = st.x;
            r4[k] = st.b[q] - (st.app(q) * x[q] - (st.ae[q] * x[q + 1] + st.aw(q) * x[q - 1] +
                                                   st.as(q) * x[q + LW] + st.an[q] * x[q - LW]));
        }
        C.b[c] = r4[0] + r4[1] + r4[2] + r4[3];
        C.x[c] = T(0);
        if (build) {
            const T ae = st.ae[fq[1]] + st.ae[fq[3]];
            const T aw = st.aw(fq[0]) + st.aw(fq[2]);
            const T an = st.an[fq[0]] + st.an[fq[1]];
            const T as = st.as(fq[2]) + st.as(fq[3]);
            C.AE[c] = ae;
            C.AW[c] = aw;
            C.AN[c] = an;
            C.AS[c] = as;
            C.AP[c] = ae + aw + an + as;
        }
    }
    for (int e = threadIdx.x; e < TI * TS; e += NTH) {
        const int I = ti0 + e / TS, J = tj0 + e % TS;
        if (I < F.nx && J < F.ny)
            a.xout[static_cast<size_t>(I) * F.stride + J] = st.x[st.at(I, J)];
    }
}

template <class T, int TS, int TI = TS> __device__ __forceinline__ void k_mg_up_s(const DU<T>& a, const Blk& bk) {
    constexpr int NTH = sg_n<T, TS, TI>();
    const Lv<T>& F = a.F;
    const int post = a.sweeps;
    const int H = 2 * post, LW = TS + 2 * H, LH = TI + 2 * H;
    int tx, ty;
    xcd_tile(bk, tx, ty);
    const int ti0 = ty * TI, tj0 = tx * TS;
    const Stage<T> st = stage_view<T>(ti0 - H, tj0 - H, LH, LW);
    stage_load<T, NTH>(st, F, a.xin, LH, &a.C, a.omega, a.bmean, a.bn);
    for (int p = 0; p < 2 * post; ++p) {
        const int d = 2 * post - 1 - p;
        stage_pass<T, NTH>(st, p & 1, max(0, ti0 - d), min(F.nx - 1, ti0 + TI - 1 + d), max(0, tj0 - d),
                   min(F.ny - 1, tj0 + TS - 1 + d));
    }
    for (int e = threadIdx.x; e < TI * TS; e += NTH) {
        const int I = ti0 + e / TS, J = tj0 + e % TS;
        if (I < F.nx && J < F.ny) {
            const T v = st.x[st.at(I, J)];
            if (a.x64) {
                double& x = a.x64[static_cast<size_t>(I) * a.ld64 + J];
                x = (a.first64 ? 0.0 : x) + static_cast<double>(v);
            } else {
                a.xout[static_cast<size_t>(I) * F.stride + J] = v;
            }
        }
    }
}

// ---- red-black p' in staged tiles ----
//
// solve_pc_redblack (src/pressure.cpp:114-134): p' = 0, then maxit_inner_p sweeps of colour 0
// then colour 1 of pc_update (:70-75), which is the multigrid smoother's two-statement update
// (smooth, :62-64) on the level-0 operator. So the staged tiles above run it too: a launch loads
// p' (or nothing, the first one: p' = 0), AE, AN and b of a tile and a halo of 2 * sweeps cells
// into LDS and runs 2 * sweeps colour passes there, each over the tile grown by the passes still
// to come, instead of one launch per colour pass that streams every array for half its cells.
// The red-black solve's colour c is the cells with I + J + 1 + c even in level-0 indices
// (I = i - H, J = j - H - 1), the smoother's colour (c + 1) mod 2. AW, AS and AP are derived
// from AE and AN bit for bit (the p' operator's symmetry, as above). Launches alternate between
// p' and a second array, so every launch reads the previous launch's whole result and the last
// one writes p'; the ring around the interior is never read (it is zero by definition).
struct RbS {
    Lv<double> F;      // level 0: operator, b
    const double* xin; // p' before this launch's sweeps (unread when xzero)
    double* xout;
    int sweeps, xzero;
};
template <int TS, int TI> __device__ __forceinline__ void k_pc_rb_s(const RbS& a, const Blk& bk) {
    constexpr int NTH = sg_n<double, TS, TI>();
    const Lv<double>& F = a.F;
    const int H = 2 * a.sweeps, LW = TS + 2 * H, LH = TI + 2 * H;
    int tx, ty;
    xcd_tile(bk, tx, ty);
    const int ti0 = ty * TI, tj0 = tx * TS;
    const Stage<double> st = stage_view<double>(ti0 - H, tj0 - H, LH, LW);
    stage_load<double, NTH>(st, F, a.xin, LH, nullptr, 0.0, nullptr, 0.0, a.xzero, 0);
    for (int p = 0; p < 2 * a.sweeps; ++p) {
        const int d = 2 * a.sweeps - 1 - p;
        stage_pass<double, NTH>(st, (p + 1) & 1, max(0, ti0 - d), min(F.nx - 1, ti0 + TI - 1 + d),
                                max(0, tj0 - d), min(F.ny - 1, tj0 + TS - 1 + d));
    }
    for (int e = threadIdx.x; e < TI * TS; e += NTH) {
        const int I = ti0 + e / TS, J = tj0 + e % TS;
        if (I < F.nx && J < F.ny)
            a.xout[static_cast<size_t>(I) * F.stride + J] = st.x[st.at(I, J)];
    }
}

// ---- the tail: one workgroup for the small levels ----
//
// Levels tm..nl-1 run in one workgroup, one launch per V-cycle, with barriers between steps:
// levels tm..t0-1 in global memory (they stay in L2), levels t0..nl-1 in LDS for the whole
// launch. A level in LDS has 8 arrays (AE, AW, AN, AS, AP, b, x, r) of (nx+2) x (ny+2) elements
// with a zero halo; its operator is built in LDS from the level above (or loaded, for level 0),
// and when the workgroup starts at an LDS level (tm = t0) that level's b and x come from global
// memory and x goes back at the end. t0 = nl: no level in LDS.

template <class T> struct Tail {
    Lv<T> G[MAXLEV]; // global views of every level
    int off[MAXLEV]; // LDS element offset of each LDS level's block
    int tm, t0, nl;  // first workgroup level, first LDS level, number of levels
    int pre, post;
    T omega;
    int lds_elems;   // elements of all LDS level blocks; the row-sum scratch follows
    T* grows;        // row-sum scratch in global memory (t0 = nl)
    Ctl* ctl;        // receives the coarse-solve sweep count
};

template <class T> __device__ __forceinline__ LLv<T> lds_view(const Tail<T>& t, Lp<T> sm, int l) {
    const int nx = t.G[l].nx, ny = t.G[l].ny, st = ny + 2;
    const size_t n = static_cast<size_t>(nx + 2) * st;
    Lp<T> b = sm + t.off[l] + st + 1;
    return {nx, ny, st, b, b + n, b + 2 * n, b + 3 * n, b + 4 * n, b + 5 * n, b + 6 * n, b + 7 * n};
}

template <class LV> __device__ void wg_smooth(const LV& L, int sweeps) {
    const int half = (L.ny + 1) / 2, n = L.nx * half;
    for (int k = 0; k < sweeps; ++k)
        for (int c = 0; c < 2; ++c) {
            for (int e = threadIdx.x; e < n; e += blockDim.x) {
                const int I = e / half;
                const int J = ((I + c) & 1) + 2 * (e - I * half);
                if (J < L.ny)
                    smooth_cell(L, static_cast<size_t>(I) * L.stride + J);
            }
            __syncthreads();
        }
}

template <class LV> __device__ void wg_resid(const LV& L) {
    for (int e = threadIdx.x; e < L.nx * L.ny; e += blockDim.x) {
        const int I = e / L.ny, J = e - I * L.ny;
        const size_t q = static_cast<size_t>(I) * L.stride + J;
        L.r[q] = resid_cell(L, q);
    }
    __syncthreads();
}

// det_sum (:18-34): each row by one thread (J ascending from 0.0), then the rows by thread 0
// (I ascending from 0.0). which: 0 b, 1 b*b, 2 r*r, 3 x.
template <class LV, class T = typename LV::value_type>
__device__ T wg_det_sum(const LV& L, T* rows, int which) {
    __shared__ T total;
    for (int I = threadIdx.x; I < L.nx; I += blockDim.x) {
        T s = T(0);
        const size_t r0 = static_cast<size_t>(I) * L.stride;
        for (int J = 0; J < L.ny; ++J) {
            const size_t q = r0 + J;
            const T v = which == 0 ? L.b[q] : which == 1 ? L.b[q] * L.b[q]
                                    : which == 2 ? L.r[q] * L.r[q] : L.x[q];
            s += v;
        }
        rows[I] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        T s = T(0);
        for (int I = 0; I < L.nx; ++I)
            s += rows[I];
        total = s;
    }
    __syncthreads();
    const T v = total;
    __syncthreads();
    return v;
}

template <class LV, class P, class T = typename LV::value_type>
__device__ void wg_sub(const LV& L, P x, T m) {
    for (int e = threadIdx.x; e < L.nx * L.ny; e += blockDim.x) {
        const int I = e / L.ny, J = e - I * L.ny;
        x[static_cast<size_t>(I) * L.stride + J] -= m;
    }
    __syncthreads();
}

// coarse_solve (:120-140): mean-free b, red-black sweeps until ||r|| < tol ||b|| or 200 sweeps,
// mean-free x.
template <class LV, class T = typename LV::value_type>
__device__ int wg_coarse_solve(const LV& L, T* rows) {
    const T ncell = static_cast<T>(static_cast<double>(L.nx) * L.ny);
    wg_sub(L, L.b, wg_det_sum(L, rows, 0) / ncell);
    const T bb = wg_det_sum(L, rows, 1);
    int sweeps = 0;
    if (bb > T(0)) {
        const RootBelow<T> converged(coarse_tol<T>() * sqrt(bb)); // ||r|| < tol ||b||
        while (sweeps < 200) {
            wg_smooth(L, 1);
            ++sweeps;
            wg_resid(L);
            if (converged(wg_det_sum(L, rows, 2)))
                break;
        }
    }
    wg_sub(L, L.x, wg_det_sum(L, rows, 3) / ncell);
    return sweeps;
}

// The same coarse solve by one wave for a level of at most 64 cells (lane = cell I * ny + J):
// operator, b and x in registers, neighbours by lane shuffles, every sum by a wave-uniform chain in
// det_sum's order (each row J ascending from 0.0, then the rows ascending from 0.0). Same
// expressions in the same order as smooth_cell / resid_cell / wg_det_sum, so the same bits; no
// workgroup barrier inside the sweep loop. Ends with a barrier (other waves then see x and b).
template <class T> __device__ __forceinline__ T lane_val(T v, int l);
template <> __device__ __forceinline__ double lane_val<double>(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
template <> __device__ __forceinline__ float lane_val<float>(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

template <class LV> __device__ int wave_coarse_solve(const LV& L) {
    using T = typename LV::value_type;
    __shared__ int nsweeps;
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x, nx = L.nx, ny = L.ny;
        const bool act = lane < nx * ny;
        const int I = act ? lane / ny : 0, J = act ? lane - (lane / ny) * ny : 0;
        const size_t q = static_cast<size_t>(I) * L.stride + J;
        T ae = T(0), aw = T(0), an = T(0), as = T(0), ap = T(1), b = T(0), x = T(0);
        if (act) {
            ae = L.AE[q];
            aw = L.AW[q];
            an = L.AN[q];
            as = L.AS[q];
            ap = L.AP[q];
            b = L.b[q];
            x = L.x[q];
        }
        const bool hasE = act && J + 1 < ny, hasW = act && J > 0;
        const bool hasS = act && I + 1 < nx, hasN = act && I > 0;
        const int col = (I + J) & 1;
        auto dsum = [&](T v) {
            T tot = T(0);
            for (int i = 0; i < nx; ++i) {
                T r = T(0);
                for (int j = 0; j < ny; ++j)
                    r += lane_val<T>(v, i * ny + j);
                tot += r;
            }
            return tot;
        };
        // neighbour values (zero halo)
        auto nbs = [&](T v, T& xe, T& xw, T& xs, T& xn) {
            const T e = __shfl(v, lane + 1, 64), w = __shfl(v, lane - 1, 64);
            const T sv = __shfl(v, lane + ny, 64), nv = __shfl(v, lane - ny, 64);
            xe = hasE ? e : T(0);
            xw = hasW ? w : T(0);
            xs = hasS ? sv : T(0);
            xn = hasN ? nv : T(0);
        };
        const T ncell = static_cast<T>(static_cast<double>(nx) * ny);
        const T mb = dsum(b) / ncell;
        if (act)
            b -= mb;
        const T bb = dsum(b * b);
        int sweeps = 0;
        if (bb > T(0)) {
            const RootBelow<T> converged(coarse_tol<T>() * sqrt(bb)); // ||r|| < tol ||b||
            const T bap = b / ap; // the update's first term does not change between sweeps
            auto sweep = [&](T v) {
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    T xe, xw, xs, xn;
                    nbs(v, xe, xw, xs, xn);
                    T t = bap;
                    t += (ae * xe + aw * xw + as * xs + an * xn) / ap;
                    // A bit blend, not a select: the compiler would turn a select of a division
                    // back into a branch, and the next sweep could not interleave with the test.
                    v = blend(act && col == c, t, v);
                }
                return v;
            };
            // Sweep k + 1 runs alongside sweep k's residual test (they are independent) and is
            // dropped when the test ends the loop: the same sweeps and iterate as testing first.
            T xc = sweep(x);
            sweeps = 1;
            while (true) {
                const T xnext = sweep(xc);
                T xe, xw, xs, xn;
                nbs(xc, xe, xw, xs, xn);
                const T r = blend(act, b - (ap * xc - (ae * xe + aw * xw + as * xs + an * xn)), T(0));
                const T rr = dsum(r * r);
                if (converged(rr) || sweeps >= 200)
                    break;
                xc = xnext;
                ++sweeps;
            }
            x = xc;
        }
        const T mx = dsum(x) / ncell;
        if (act) {
            x -= mx;
            L.x[q] = x;
            L.b[q] = b;
        }
        if (lane == 0)
            nsweeps = sweeps;
    }
    __syncthreads();
    return nsweeps;
}

// The coarse solve for a level of at most 16 cells with NY columns (one 16-lane DPP row): the
// neighbours by DPP row shifts (row_shl / row_shr by 1 and by NY; lanes whose source lies outside
// the row read 0, and wall neighbours are zero anyway), and every ordered sum from the values read
// out of their lanes first, then added in det_sum's order. Same expressions in the same order as
// wave_coarse_solve.
template <int CTRL> __device__ __forceinline__ double dpp_val(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
template <int CTRL> __device__ __forceinline__ float dpp_val(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

// NX: the row count when known at compile time (the sums then have no branches), else 0.
template <int NY, int NX, class LV> __device__ int wave_coarse_solve16(const LV& L) {
    using T = typename LV::value_type;
    __shared__ int nsweeps;
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x, nx = NX > 0 ? NX : L.nx;
        const int nc = nx * NY;
        const bool act = lane < nc;
        const int I = act ? lane / NY : 0, J = act ? lane % NY : 0;
        const size_t q = static_cast<size_t>(I) * L.stride + J;
        T ae = T(0), aw = T(0), an = T(0), as = T(0), ap = T(1), b = T(0), x = T(0);
        if (act) {
            ae = L.AE[q];
            aw = L.AW[q];
            an = L.AN[q];
            as = L.AS[q];
            ap = L.AP[q];
            b = L.b[q];
            x = L.x[q];
        }
        const bool hasE = act && J + 1 < NY, hasW = act && J > 0;
        const bool hasS = act && I + 1 < nx, hasN = act && I > 0;
        const int col = (I + J) & 1;
        // Ordered sum (rows in order, each from 0.0) by DPP row shifts: lane 4i adds its row's
        // values in column order, then lane 0 adds the row sums in row order; the same additions
        // in the same order as reading every lane out first, with a shorter dependency chain.
        static_assert(NY == 4, "the DPP sums assume rows of 4 lanes");
        auto dsum = [&](T v) {
            T r = T(0) + v;
            r = r + dpp_val<0x101>(v); // row_shl:1
            r = r + dpp_val<0x102>(v);
            r = r + dpp_val<0x103>(v);
            T tot = T(0) + r;
            if (nx > 1)
                tot = tot + dpp_val<0x104>(r); // row 1's sum, from lane 4
            if (nx > 2)
                tot = tot + dpp_val<0x108>(r);
            if (nx > 3)
                tot = tot + dpp_val<0x10C>(r);
            return lane_val<T>(tot, 0);
        };
        auto nbs = [&](T v, T& xe, T& xw, T& xs, T& xn) {
            const T e = dpp_val<0x101>(v), w = dpp_val<0x111>(v);   // row_shl:1, row_shr:1
            const T sv = dpp_val<0x100 + NY>(v), nv = dpp_val<0x110 + NY>(v); // by NY
            xe = hasE ? e : T(0);
            xw = hasW ? w : T(0);
            xs = hasS ? sv : T(0);
            xn = hasN ? nv : T(0);
        };
        const T ncell = static_cast<T>(static_cast<double>(nx) * NY);
        const T mb = dsum(b) / ncell;
        if (act)
            b -= mb;
        const T bb = dsum(b * b);
        int sweeps = 0;
        if (bb > T(0)) {
            const RootBelow<T> converged(coarse_tol<T>() * sqrt(bb)); // ||r|| < tol ||b||
            const T bap = b / ap; // the update's first term does not change between sweeps
            auto sweep = [&](T v) {
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    T xe, xw, xs, xn;
                    nbs(v, xe, xw, xs, xn);
                    T t = bap;
                    t += (ae * xe + aw * xw + as * xs + an * xn) / ap;
                    // A bit blend, not a select: the compiler would turn a select of a division
                    // back into a branch, and the next sweep could not interleave with the test.
                    v = blend(act && col == c, t, v);
                }
                return v;
            };
            // Sweep k + 1 runs alongside sweep k's residual test (they are independent) and is
            // dropped when the test ends the loop: the same sweeps and iterate as testing first.
            T xc = sweep(x);
            sweeps = 1;
            while (true) {
                const T xnext = sweep(xc);
                T xe, xw, xs, xn;
                nbs(xc, xe, xw, xs, xn);
                const T r = blend(act, b - (ap * xc - (ae * xe + aw * xw + as * xs + an * xn)), T(0));
                const T rr = dsum(r * r);
                if (converged(rr) || sweeps >= 200)
                    break;
                xc = xnext;
                ++sweeps;
            }
            x = xc;
        }
        const T mx = dsum(x) / ncell;
        if (act) {
            x -= mx;
            L.x[q] = x;
            L.b[q] = b;
        }
        if (lane == 0)
            nsweeps = sweeps;
    }
    __syncthreads();
    return nsweeps;
}

template <class LF, class LC> __device__ void wg_build(const LF& F, const LC& C) {
    for (int e = threadIdx.x; e < C.nx * C.ny; e += blockDim.x)
        build_cell(F, C, e / C.ny, e % C.ny);
    __syncthreads();
}

template <class LD, class LS> __device__ void wg_copy_level(const LD& D, const LS& S, bool coefs) {
    for (int e = threadIdx.x; e < D.nx * D.ny; e += blockDim.x) {
        const int I = e / D.ny, J = e - I * D.ny;
        const size_t d = static_cast<size_t>(I) * D.stride + J, g = static_cast<size_t>(I) * S.stride + J;
        if (coefs) {
            D.AE[d] = S.AE[g];
            D.AW[d] = S.AW[g];
            D.AN[d] = S.AN[g];
            D.AS[d] = S.AS[g];
            D.AP[d] = S.AP[g];
        }
        D.b[d] = S.b[g];
        D.x[d] = S.x[g];
    }
    __syncthreads();
}

template <class T> __device__ __forceinline__ void k_mg_tail(const Tail<T>& t, const Blk&) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* sm = reinterpret_cast<T*>(smem);
    for (int e = threadIdx.x; e < t.lds_elems; e += blockDim.x)
        sm[e] = T(0);
    __syncthreads();
    T* rows = t.t0 < t.nl ? sm + t.lds_elems : t.grows;
    const Lp<T> lsm = (Lp<T>)sm;
    // Run f on level l's view: global levels through generic pointers, LDS levels through LDS
    // pointers (the combinations are separate instantiations).
    auto with = [&](int l, auto&& f) {
        if (l < t.t0)
            f(t.G[l]);
        else
            f(lds_view(t, lsm, l));
    };
    // Operators: global levels from their (global) parents, LDS levels from theirs; level 0 in
    // LDS is loaded.
    for (int l = t.tm; l < t.nl; ++l) {
        if (l == 0) {
            if (t.t0 == 0)
                with(0, [&](const auto& D) { wg_copy_level(D, t.G[0], true); });
            continue;
        }
        with(l, [&](const auto& C) {
            if (l == t.tm)
                wg_build(t.G[l - 1], C);
            else
                with(l - 1, [&](const auto& F) { wg_build(F, C); });
        });
    }
    if (t.tm == t.t0 && t.t0 > 0) // b and x of the first level
        with(t.tm, [&](const auto& D) { wg_copy_level(D, t.G[t.tm], false); });
    for (int l = t.tm; l + 1 < t.nl; ++l) {
        with(l, [&](const auto& F) {
            with(l + 1, [&](const auto& C) {
                wg_smooth(F, t.pre);
                wg_resid(F);
                for (int e = threadIdx.x; e < C.nx * C.ny; e += blockDim.x) {
                    const int I = e / C.ny, J = e - I * C.ny;
                    const long fs = F.stride;
                    const size_t f = 2 * I * fs + 2 * J, c = static_cast<size_t>(I) * C.stride + J;
                    C.b[c] = F.r[f] + F.r[f + 1] + F.r[f + fs] + F.r[f + fs + 1];
                    C.x[c] = T(0);
                }
                __syncthreads();
            });
        });
    }
    int sweeps = 0;
    with(t.nl - 1, [&](const auto& Lc) {
        const int n = Lc.nx * Lc.ny;
        if (n == 16 && Lc.ny == 4) // the coarsest level of every power-of-two grid: 4 x 4
            sweeps = wave_coarse_solve16<4, 4>(Lc);
        else if (n <= 16 && Lc.ny == 4)
            sweeps = wave_coarse_solve16<4, 0>(Lc);
        else if (n <= 64)
            sweeps = wave_coarse_solve(Lc);
        else
            sweeps = wg_coarse_solve(Lc, rows);
    });
    for (int l = t.nl - 2; l >= t.tm; --l) {
        with(l, [&](const auto& F) {
            with(l + 1, [&](const auto& C) {
                for (int e = threadIdx.x; e < F.nx * F.ny; e += blockDim.x) {
                    const int i = e / F.ny, j = e - i * F.ny;
                    F.x[static_cast<size_t>(i) * F.stride + j] +=
                        t.omega * C.x[static_cast<size_t>(i / 2) * C.stride + j / 2];
                }
                __syncthreads();
                wg_smooth(F, t.post);
            });
        });
    }
    if (t.tm == t.t0 && t.t0 < t.nl) {
        // The first level was in LDS: its x (and, if it is the coarsest, its mean-free b) back.
        const LLv<T> D = lds_view(t, lsm, t.tm);
        const Lv<T> G0 = t.G[t.tm];
        for (int e = threadIdx.x; e < D.nx * D.ny; e += blockDim.x) {
            const int I = e / D.ny, J = e - I * D.ny;
            const size_t g = static_cast<size_t>(I) * G0.stride + J,
                         q = static_cast<size_t>(I) * D.stride + J;
            G0.x[g] = D.x[q];
            if (t.tm == t.nl - 1)
                G0.b[g] = D.b[q];
        }
    }
    if (threadIdx.x == 0)
        t.ctl->coarse_sweeps = sweeps;
}

// ---- mixed precision: FP64 <-> FP32 level-0 transfers ----

struct Mix {
    int nx, ny, ld;          // level-0 interior size, FP64 stride
    const double *AE, *AW, *AN, *AS, *AP, *b; // FP64 level 0 (at interior cell 0, 0)
    double* x;               // FP64 p' (at interior cell 0, 0)
    Lv<float> f;             // FP32 level 0
    int lean;                // AW, AS, AP not assembled: derived from AE and AN (ensure_p_full)
    const double* bmean;     // b's mean not removed: b - *bmean / bn is read (as k_total_sub)
    double bn;
    int first;               // p' not zero-filled: the first update sets it (0.0 + e)
};
__device__ __forceinline__ double mix_b(const Mix& m, size_t q) {
    return m.bmean ? m.b[q] - *m.bmean / m.bn : m.b[q];
}

// Level-0 AW, AS and AP of interior cell (I, J) at q: stored, or after a lean assemble_p the
// values ensure_p_full would derive.
__device__ __forceinline__ void mix_op(const Mix& m, int I, int J, size_t q, double& aw,
                                       double& as, double& ap) {
    if (m.lean) {
        aw = J == 0 ? 0.0 : m.AE[q - 1];
        as = I == m.nx - 1 ? 0.0 : m.AN[q + m.ld];
        ap = m.AE[q] + aw + m.AN[q] + as;
    } else {
        aw = m.AW[q];
        as = m.AS[q];
        ap = m.AP[q];
    }
}

// FP32 copies of the level-0 operator; rhs = b (first step, p' = 0), x = 0.
__device__ __forceinline__ void k_mx_init(const Mix& m, const Blk& bk) {
    const int J = bk.x * BX + threadIdx.x, I = bk.y * BY + threadIdx.y;
    if (I >= m.nx || J >= m.ny)
        return;
    const size_t q = static_cast<size_t>(I) * m.ld + J, g = static_cast<size_t>(I) * m.f.stride + J;
    double aw, as, ap;
    mix_op(m, I, J, q, aw, as, ap);
    m.f.AE[g] = static_cast<float>(m.AE[q]);
    m.f.AW[g] = static_cast<float>(aw);
    m.f.AN[g] = static_cast<float>(m.AN[q]);
    m.f.AS[g] = static_cast<float>(as);
    m.f.AP[g] = static_cast<float>(ap);
    m.f.b[g] = static_cast<float>(mix_b(m, q));
    m.f.x[g] = 0.0f;
}

// FP64 defect r = b - A p' (level-0 operator) as the FP32 right-hand side; x = 0.
__device__ __forceinline__ void k_mx_defect(const Mix& m, const Blk& bk) {
    const int J = bk.x * BX + threadIdx.x, I = bk.y * BY + threadIdx.y;
    if (I >= m.nx || J >= m.ny)
        return;
    const long s = m.ld;
    const size_t q = static_cast<size_t>(I) * s + J, g = static_cast<size_t>(I) * m.f.stride + J;
    const double* x = m.x;
    double aw, as, ap;
    mix_op(m, I, J, q, aw, as, ap);
    const double r = mix_b(m, q) - (ap * x[q] - (m.AE[q] * x[q + 1] + aw * x[q - 1] +
                                                 as * x[q + s] + m.AN[q] * x[q - s]));
    m.f.b[g] = static_cast<float>(r);
    m.f.x[g] = 0.0f;
}

// The learned guess applied, in place of k_mx_init and the first k_mx_defect of a solve that starts
// from it: p' = x = alpha g, or 0 if the guess was discarded (x at the neighbours evaluated the
// same way, so they are the values p' holds); the FP32 copies of the operator; the FP32 right-hand
// side the FP64 defect b - A x (k_mx_defect's expression); FP32 x = 0.
struct GuessMix {
    GuessArgs g;
    Mix m;
};
// LY (one row per workgroup): the prediction's rows read from an LDS copy (gy_stage).
template <bool LY> __device__ __forceinline__ void k_mx_guess(const GuessMix& a, const Blk& bk) {
    static_assert(!LY || BY == 1, "one fine row per workgroup");
    __shared__ float sy[LY ? GY_ROWS * GY_MAX : 1];
    const Mix& m = a.m;
    const GuessArgs& g = a.g;
    const int J = bk.x * BX + threadIdx.x, I = bk.y * BY + threadIdx.y;
    if (I >= m.nx)
        return;
    const float* y = g.y;
    int k0 = 0;
    if constexpr (LY) {
        k0 = gy_stage(g, I, sy);
        y = sy;
    }
    if (J >= m.ny)
        return;
    const int keep = *g.flag;
    const double alpha = g.part[2 * g.nx], scale = g.st[2];
    auto row = [&](int r) { return LY ? gy_row(g, r, k0) : grow(g, r); };
    const GAxis rm = row(I - 1), ri = row(I), rp = row(I + 1);
    const GAxis cm = gcol(g, J - 1), cj = gcol(g, J), cp = gcol(g, J + 1);
    auto xat = [&](const GAxis& r, const GAxis& c) {
        return keep ? gval(y, scale, r, c) * alpha : 0.0;
    };
    const size_t q = static_cast<size_t>(I) * m.ld + J, f = static_cast<size_t>(I) * m.f.stride + J;
    const double x = xat(ri, cj);
    m.x[q] = x;
    double aw, as, ap;
    mix_op(m, I, J, q, aw, as, ap);
    m.f.AE[f] = static_cast<float>(m.AE[q]);
    m.f.AW[f] = static_cast<float>(aw);
    m.f.AN[f] = static_cast<float>(m.AN[q]);
    m.f.AS[f] = static_cast<float>(as);
    m.f.AP[f] = static_cast<float>(ap);
    const double r = mix_b(m, q) - (ap * x - (m.AE[q] * xat(ri, cp) + aw * xat(ri, cm) +
                                              as * xat(rp, cj) + m.AN[q] * xat(rm, cj)));
    m.f.b[f] = static_cast<float>(r);
    m.f.x[f] = 0.0f;
}

// p' += e (FP64 update).
__device__ __forceinline__ void k_mx_update(const Mix& m, const Blk& bk) {
    const int J = bk.x * BX + threadIdx.x, I = bk.y * BY + threadIdx.y;
    if (I >= m.nx || J >= m.ny)
        return;
    double& x = m.x[static_cast<size_t>(I) * m.ld + J];
    // From p' = 0 without the zero fill (fused mean removal): 0.0 + e, as the fill then += gives.
    x = (m.first ? 0.0 : x) + static_cast<double>(m.f.x[static_cast<size_t>(I) * m.f.stride + J]);
}

dim3 lgrid(int nx, int ny) { return dim3(cdiv(ny, BX), cdiv(nx, BY)); }

// Launch labels of the multigrid steps, so that a batch of cavities with different numbers of
// levels groups the same step of the same V-cycle: level l of V-cycle v, going down or up.
int mg_tag(int v, int up, int l) { return 1 + (v * 2 + up) * 32 + l; }
constexpr int TAG_TAIL = 31; // level slot of the tail
constexpr int TAG_BUILD = 1 << 20;

template <class T>
void smooth_passes(Solver& s, const Lv<T>& L, int sweeps, int first_col, const int* stop) {
    for (int k = first_col; k < 2 * sweeps; ++k)
        emit<FC<T>, k_mg_smooth<T>, BX * BY>(s, lgrid(L.nx, cdiv(L.ny, 2)), dim3(BX, BY), 0,
                                            FC<T>{L, L, T(0), k & 1}, stop);
}

// LDS bytes of levels t0.. (level blocks plus the row-sum scratch).
template <class T> size_t lds_bytes(const Solver& s, int t0) {
    size_t e = 0;
    int maxnx = 0;
    for (size_t l = t0; l < s.mg.size(); ++l) {
        e += 8 * static_cast<size_t>(s.mg[l].nx + 2) * (s.mg[l].ny + 2);
        maxnx = std::max(maxnx, s.mg[l].nx);
    }
    return t0 >= static_cast<int>(s.mg.size()) ? 0 : sizeof(T) * (e + maxnx + 2);
}

static long cells(const Solver& s, int l) { return static_cast<long>(s.mg[l].nx) * s.mg[l].ny; }

// First LDS level: the longest run of coarsest levels that fit the LDS budget, each with at most
// lds_cells cells; nl if not even the coarsest fits.
template <class T> int lds_start(const Solver& s) {
    const int nl = static_cast<int>(s.mg.size());
    if (lds_bytes<T>(s, nl - 1) > s.lds_budget)
        return nl;
    int t0 = nl - 1;
    while (t0 > 0 && cells(s, t0 - 1) <= s.lds_cells && lds_bytes<T>(s, t0 - 1) <= s.lds_budget)
        --t0;
    return t0;
}

// First level of the one-workgroup tail: the levels from t0 (or the coarsest) up with at most
// mid_cells cells.
int mid_start(const Solver& s, int t0) {
    int tm = std::min(t0, static_cast<int>(s.mg.size()) - 1);
    while (tm > 0 && cells(s, tm - 1) <= s.mid_cells)
        --tm;
    return tm;
}

// The tiled down/up kernels apply when their halos are small enough.
bool tiled_ok(const Solver& s) {
    return s.mg_tiled && 2 * s.p.mg_pre + 1 <= MT_MAXH && 2 * s.p.mg_post <= MT_MAXH;
}

// Coarse operators of the levels above the tail (the tail builds its own; the tiled down kernel
// builds them during the first V-cycle).
template <class T> void launch_builds(Solver& s, const Tail<T>& t, const int* stop) {
    if (tiled_ok(s))
        return;
    for (int l = 1; l < t.tm; ++l) {
        s.tag = TAG_BUILD + l;
        emit<FC<T>, k_mg_build<T>, BX * BY>(s, lgrid(t.G[l].nx, t.G[l].ny), dim3(BX, BY), 0,
                                           FC<T>{t.G[l - 1], t.G[l], T(0), 0}, stop);
    }
    s.tag = 0;
}

// The staged tile kernels and the one-workgroup tail need more than the default 64 KB of dynamic
// LDS: allow them the device's whole per-workgroup LDS. Set once per solver at set-up, never
// while the stream is being captured.
template <class T> void dyn_lds_limits(const Solver& s) {
    const int max = static_cast<int>(s.lds_budget + 1024);
    const void* fns[] = {
        reinterpret_cast<const void*>(&k_one<DU<T>, k_mg_down_s<T, 32, BR<T>>, sg_n<T, 32, BR<T>>()>),
        reinterpret_cast<const void*>(&k_one<DU<T>, k_mg_up_s<T, 32, BR<T>>, sg_n<T, 32, BR<T>>()>),
        reinterpret_cast<const void*>(&k_one<DU<T>, k_mg_down_s<T, 16>, sg_n<T, 16>()>),
        reinterpret_cast<const void*>(&k_one<DU<T>, k_mg_up_s<T, 16>, sg_n<T, 16>()>),
        reinterpret_cast<const void*>(&k_one<Tail<T>, k_mg_tail<T>, 1024>)};
    for (const void* f : fns) {
        // The limit is on static plus dynamic LDS: the tail has some static LDS of its own.
        hipFuncAttributes fa{};
        SG_HIP(hipFuncGetAttributes(&fa, f));
        SG_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   max - static_cast<int>(fa.sharedSizeBytes)));
    }
}

// LDS bytes of a staged TS x TS tile and its halo of h cells.
template <class T> size_t staged_bytes(int ts, int h, int ti = 0) {
    return SG_NSTAGE<T> * sizeof(T) * static_cast<size_t>((ti ? ti : ts) + 2 * h) * (ts + 2 * h);
}

// mg_cycles V-cycles (mg_vcycle, :299-323) on level 0 of t.
// x64 (mixed form): the FP64 p' that level 0's last up-kernel adds its result to when it is a
// staged one (first64: from p' = 0, not read); returns whether it did, else the caller updates.
template <class T>
bool launch_vcycles(Solver& s, const Tail<T>& t, const int* stop, double* x64 = nullptr,
                    int first64 = 0) {
    const sg_params& p = s.p;
    const size_t lds = lds_bytes<T>(s, t.t0);
    const bool tiled = tiled_ok(s);
    const size_t dn = sizeof(T) * (MT_I + 2 * (2 * p.mg_pre + 1)) * (MT_J + 2 * (2 * p.mg_pre + 1));
    const size_t up = sizeof(T) * (MT_I + 4 * p.mg_post) * (MT_J + 4 * p.mg_post);
    const int hdn = 2 * p.mg_pre + 1, hup = 2 * p.mg_post;
    auto big = [](const Lv<T>& L) { return static_cast<long>(L.nx) * L.ny >= SG_BIG; };
    const size_t sdn = staged_bytes<T>(32, hdn), sup = staged_bytes<T>(32, hup);
    const bool staged = tiled && s.mg_staged < t.tm && std::max(sdn, sup) <= s.lds_budget + 1024;
    // b's mean removal carried by level 0's staged kernels (FP64 only: the mixed form's FP32 b
    // is converted with it).
    const bool bfused = sizeof(T) == 8 && s.b_mean_pending;
    const double bn = static_cast<double>(p.NX) * p.NY;
    if (bfused && !(staged && s.mg_staged == 0 && t.tm >= 1))
        throw HipError("fused mean removal without staged level-0 tiles");
    bool updated = false;
    for (int cyc = 0; cyc < p.mg_cycles; ++cyc) {
        const int v = s.vcount++;
        for (int l = 0; l < t.tm; ++l) {
            const Lv<T> F = t.G[l], C = t.G[l + 1];
            s.tag = mg_tag(v, 0, l);
            if (tiled) {
                // The first cycle builds level l+1 here unless the tail builds it (l+1 = tm).
                const int build = (cyc == 0 && l + 1 < t.tm) ? 1 : 0;
                DU<T> du{F, C, F.x, F.xt, p.mg_pre, build, T(0)};
                if (bfused && l == 0) {
                    du.bmean = s.scal + S_BSUM, du.bn = bn;
                    du.xzero = cyc == 0 ? 1 : 0;
                }
                if (staged && l >= s.mg_staged && big(F))
                    emit<DU<T>, k_mg_down_s<T, 32, BR<T>>, sg_n<T, 32, BR<T>>()>(
                        s, dim3(cdiv(F.ny, 32), cdiv(F.nx, BR<T>)), dim3(sg_n<T, 32, BR<T>>()),
                        staged_bytes<T>(32, hdn, BR<T>), du, stop);
                else if (staged && l >= s.mg_staged)
                    emit<DU<T>, k_mg_down_s<T, 16>, sg_n<T, 16>()>(
                        s, dim3(cdiv(F.ny, 16), cdiv(F.nx, 16)), dim3(sg_n<T, 16>()),
                        staged_bytes<T>(16, hdn), du, stop);
                else
                emit<DU<T>, k_mg_down<T>, MT_N>(s, dim3(cdiv(F.ny, MT_J), cdiv(F.nx, MT_I)),
                                                dim3(MT_N), dn,
                                                DU<T>{F, C, F.x, F.xt, p.mg_pre, build, T(0)},
                                                stop);
            } else {
                smooth_passes(s, F, p.mg_pre, 0, stop);
                emit<FC<T>, k_mg_resid_restrict<T>, BX * BY>(s, lgrid(C.nx, C.ny), dim3(BX, BY),
                                                             0, FC<T>{F, C, T(0), 0}, stop);
            }
        }
        s.tag = mg_tag(v, 0, TAG_TAIL);
        emit<Tail<T>, k_mg_tail<T>, 1024>(s, dim3(1), dim3(s.tail_threads), lds, t, stop);
        for (int l = t.tm - 1; l >= 0; --l) {
            const Lv<T> F = t.G[l], C = t.G[l + 1];
            s.tag = mg_tag(v, 1, l);
            DU<T> du{F, C, F.xt, F.x, p.mg_post, 0, t.omega};
            if (bfused && l == 0)
                du.bmean = s.scal + S_BSUM, du.bn = bn;
            if (x64 && l == 0 && cyc + 1 == p.mg_cycles && staged && l >= s.mg_staged) {
                du.x64 = x64, du.ld64 = static_cast<int>(s.ld), du.first64 = first64;
                updated = true;
            }
            if (staged && l >= s.mg_staged && big(F)) {
                emit<DU<T>, k_mg_up_s<T, 32, BR<T>>, sg_n<T, 32, BR<T>>()>(
                    s, dim3(cdiv(F.ny, 32), cdiv(F.nx, BR<T>)), dim3(sg_n<T, 32, BR<T>>()),
                    staged_bytes<T>(32, hup, BR<T>), du, stop);
            } else if (staged && l >= s.mg_staged) {
                emit<DU<T>, k_mg_up_s<T, 16>, sg_n<T, 16>()>(
                    s, dim3(cdiv(F.ny, 16), cdiv(F.nx, 16)), dim3(sg_n<T, 16>()),
                    staged_bytes<T>(16, hup), du, stop);
            } else if (tiled) {
                emit<DU<T>, k_mg_up<T>, MT_N>(s, dim3(cdiv(F.ny, MT_J), cdiv(F.nx, MT_I)),
                                              dim3(MT_N), up,
                                              DU<T>{F, C, F.xt, F.x, p.mg_post, 0, t.omega}, stop);
            } else if (p.mg_post > 0) {
                emit<FC<T>, k_mg_prolong_smooth0<T>, BX * BY>(s, lgrid(F.nx, cdiv(F.ny, 2)),
                                                              dim3(BX, BY), 0,
                                                              FC<T>{F, C, t.omega, 0}, stop);
                smooth_passes(s, F, p.mg_post, 1, stop);
            } else {
                emit<FC<T>, k_mg_prolong<T>, BX * BY>(s, lgrid(F.nx, F.ny), dim3(BX, BY), 0,
                                                      FC<T>{F, C, t.omega, 0}, stop);
            }
        }
    }
    s.tag = 0;
    return updated;
}

template <class T> Tail<T> make_tail(Solver& s, Lv<T> (*view)(const MgLevel&)) {
    Tail<T> t{};
    const int nl = static_cast<int>(s.mg.size());
    for (int l = 0; l < nl; ++l)
        t.G[l] = view(s.mg[l]);
    t.t0 = lds_start<T>(s);
    t.tm = mid_start(s, t.t0);
    t.nl = nl;
    t.pre = s.p.mg_pre;
    t.post = s.p.mg_post;
    t.omega = static_cast<T>(s.p.mg_omega);
    t.grows = reinterpret_cast<T*>(s.rowbuf);
    t.ctl = s.ctl;
    int e = 0;
    for (int l = t.t0; l < nl; ++l) {
        t.off[l] = e;
        e += 8 * (s.mg[l].nx + 2) * (s.mg[l].ny + 2);
    }
    t.lds_elems = e;
    return t;
}

} // namespace

// Levels (src/multigrid.cpp:143-185): halve while both sides are even and the halves stay at
// least mg_coarsest (4 in the reference). Level 0 aliases the p' arrays. The mixed form adds
// FP32 copies of every level ((nx+2) x (ny+2), zero halo).
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
    if (static_cast<int>(dims.size()) > MAXLEV)
        throw ArgError("too many multigrid levels");
    const bool mixed = p.mg_precision == SG_MG_MIXED;
    s.mg.assign(dims.size(), MgLevel{});
    SG_HIP(hipMalloc(reinterpret_cast<void**>(&s.mg_r0), sizeof(double) * s.n));
    s.owned.push_back(s.mg_r0);
    SG_HIP(hipMemset(s.mg_r0, 0, sizeof(double) * s.n));
    const size_t base0 = static_cast<size_t>(p.H) * s.ld + p.H + 1;
    for (size_t l = 0; l < dims.size(); ++l) {
        MgLevel& L = s.mg[l];
        L.nx = dims[l].first;
        L.ny = dims[l].second;
        const size_t n = static_cast<size_t>(L.nx + 2) * (L.ny + 2);
        if (mixed) {
            float* blk = nullptr;
            SG_HIP(hipMalloc(reinterpret_cast<void**>(&blk), sizeof(float) * n * 9));
            SG_HIP(hipMemset(blk, 0, sizeof(float) * n * 9));
            s.owned.push_back(blk);
            float** ptr[9] = {&L.fAE, &L.fAW, &L.fAN, &L.fAS, &L.fAP, &L.fb, &L.fx, &L.fr, &L.fxt};
            for (int k = 0; k < 9; ++k)
                *ptr[k] = blk + k * n + (L.ny + 2) + 1;
        }
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
            double* xt = nullptr;
            SG_HIP(hipMalloc(reinterpret_cast<void**>(&xt), sizeof(double) * s.n));
            SG_HIP(hipMemset(xt, 0, sizeof(double) * s.n));
            s.owned.push_back(xt);
            L.xt = xt + base0;
            continue;
        }
        L.stride = L.ny + 2;
        if (mixed)
            continue; // the FP64 form's coarse levels are not used
        double* blk = nullptr;
        SG_HIP(hipMalloc(reinterpret_cast<void**>(&blk), sizeof(double) * n * 9));
        SG_HIP(hipMemset(blk, 0, sizeof(double) * n * 9));
        s.owned.push_back(blk);
        double** ptr[9] = {&L.AE, &L.AW, &L.AN, &L.AS, &L.AP, &L.b, &L.x, &L.r, &L.xt};
        for (int k = 0; k < 9; ++k)
            *ptr[k] = blk + k * n + L.stride + 1;
    }
    if (mixed)
        dyn_lds_limits<float>(s);
    else
        dyn_lds_limits<double>(s);
}

// solve_pc_mg (src/multigrid.cpp:347-357): build, p' = 0, mean-free b, mg_cycles V-cycles,
// mean-free p'. Mixed form: the same with FP32 V-cycles inside mg_dc_maxit FP64 defect
// corrections (the first with p' = 0, so its defect is b itself; with has_guess, p' holds an
// initial guess and every step starts from its FP64 defect).
void launch_mg_solve(Solver& s, const int* stop, bool has_guess) {
    const sg_params& p = s.p;
    s.vcount = 0;
    if (p.mg_precision != SG_MG_MIXED) {
        const Tail<double> t = make_tail<double>(s, lv64);
        launch_builds(s, t, stop);
        if (s.fuse_b_mean)
            launch_b_mean_fused(s, stop); // level 0 reads b - mean, p' = 0 unread
        else
            launch_remove_mean0(s, s.F(SG_F_B), stop, s.F(SG_F_PC)); // and p' = 0
        launch_vcycles(s, t, stop);
        s.b_mean_pending = false;
        finish_pc_mean(s, stop);
        return;
    }
    const Tail<float> t = make_tail<float>(s, lv32);
    const MgLevel& L0 = s.mg[0];
    const bool bfused = s.fuse_b_mean;
    Mix m{L0.nx, L0.ny, s.ld, L0.AE, L0.AW, L0.AN, L0.AS, L0.AP, L0.b, L0.x, t.G[0],
          s.p_lean ? 1 : 0, bfused ? s.scal + S_BSUM : nullptr,
          static_cast<double>(p.NX) * p.NY, 0};
    const dim3 g = lgrid(m.nx, m.ny);
    // Learned warm start: the network's guess of p' (or zero, if the guess does not lower the
    // defect) replaces the zero start.
    const bool warm = p.pressure_warm_start == SG_WARM_LEARNED && s.warm;
    if (bfused)
        launch_b_mean_fused(s, stop); // the conversions read b - mean; p' set by the first update
    else
        launch_remove_mean0(s, s.F(SG_F_B), stop,
                            !has_guess && !warm ? s.F(SG_F_PC) : nullptr); // and p' = 0
    if (warm) {
        const GuessMix gm{launch_warm_guess(s, stop, bfused), m};
        if (p.NY / gm.g.bs <= GY_MAX)
            emit<GuessMix, k_mx_guess<true>, BX * BY>(s, g, dim3(BX, BY), 0, gm, stop);
        else
            emit<GuessMix, k_mx_guess<false>, BX * BY>(s, g, dim3(BX, BY), 0, gm, stop);
    }
    else
        emit<Mix, k_mx_init, BX * BY>(s, g, dim3(BX, BY), 0, m, stop);
    has_guess = has_guess || warm;
    launch_builds(s, t, stop);
    for (int k = 0; k < p.mg_dc_maxit; ++k) {
        if (k > 0 || (has_guess && !warm)) { // the guess's defect is formed by k_mx_guess
            emit<Mix, k_mx_defect, BX * BY>(s, g, dim3(BX, BY), 0, m, stop);
        }
        m.first = bfused && k == 0 && !has_guess ? 1 : 0;
        if (!launch_vcycles(s, t, stop, m.x, m.first))
            emit<Mix, k_mx_update, BX * BY>(s, g, dim3(BX, BY), 0, m, stop);
    }
    s.b_mean_pending = false;
    finish_pc_mean(s, stop);
}

// Red-black p' in staged tiles (k_pc_rb_s): RB_SWEEPS sweeps per launch (a halo of 8 cells: the
// four staged arrays of a 24 x 32 tile fit 61 KB of LDS, the default dynamic limit, and two
// workgroups share a CU), p' last in the F(SG_F_PC) array. p' itself need not be zero beforehand
// except on its ring, which the caller's fill leaves zero.
constexpr int RB_SWEEPS = 4;
void launch_rb_staged(Solver& s, const int* stop) {
    const sg_params& p = s.p;
    const size_t o = static_cast<size_t>(p.H) * s.ld + p.H + 1;
    const Lv<double> F{p.NX,        p.NY,        static_cast<int>(s.ld), s.F(SG_F_AEP) + o,
                       s.F(SG_F_AWP) + o, s.F(SG_F_ANP) + o, s.F(SG_F_ASP) + o, s.F(SG_F_APP) + o,
                       s.F(SG_F_B) + o,   s.F(SG_F_PC) + o,  nullptr,           nullptr};
    const int n = p.maxit_inner_p, nl = (n + RB_SWEEPS - 1) / RB_SWEEPS;
    double* buf[2] = {s.F(SG_F_PC) + o, s.scratch + o};
    const double* xin = nullptr;
    constexpr int TS = 32, TI = BR<double>;
    for (int k = 0; k < nl; ++k) {
        const int sw = std::min(RB_SWEEPS, n - k * RB_SWEEPS);
        double* xout = buf[(nl - 1 - k) % 2]; // the last launch writes p'
        const RbS a{F, xin, xout, sw, k == 0 ? 1 : 0};
        emit<RbS, k_pc_rb_s<TS, TI>, sg_n<double, TS, TI>()>(
            s, dim3(cdiv(p.NY, TS), cdiv(p.NX, TI)), dim3(sg_n<double, TS, TI>()),
            staged_bytes<double>(TS, 2 * sw, TI), a, stop);
        xin = xout;
    }
}

// The same conditions as launch_vcycles' use of staged tiles on level 0, the tail starting at
// level min_tm or below. (The learned warm start reads level 0 as the mixed conversions do: the
// lean operator's derived links and b - mean.)
static bool staged_from0(Solver& s, int min_tm) {
    const sg_params& p = s.p;
    if (p.pressure_solver != SG_P_MG || s.mg.empty() || !tiled_ok(s) || s.mg_staged != 0)
        return false;
    const bool mixed = p.mg_precision == SG_MG_MIXED;
    const int hdn = 2 * p.mg_pre + 1, hup = 2 * p.mg_post;
    const int t0 = mixed ? lds_start<float>(s) : lds_start<double>(s), tm = mid_start(s, t0);
    const size_t sdn = mixed ? staged_bytes<float>(32, hdn) : staged_bytes<double>(32, hdn);
    const size_t sup = mixed ? staged_bytes<float>(32, hup) : staged_bytes<double>(32, hup);
    return tm >= min_tm && std::max(sdn, sup) <= s.lds_budget + 1024;
}

// The tail builds its first level from the level above it in global memory, so level 0's AW, AS
// and AP go unread only if the tail starts at level 2 or below. The mixed form's FP64 conversion
// and defect kernels derive the three themselves (mix_op), and so does the red-black solve in
// staged tiles.
bool mg_lean_ok(Solver& s) {
    if (s.p.pressure_solver == SG_P_GS_REDBLACK)
        return !rb_one_fits(s.p);
    return staged_from0(s, 2);
}

// Level 0 above the tail: only level 0's staged kernels (or the mixed form's conversions) read b
// and the initial p'.
bool mg_fuse_b_ok(Solver& s) { return staged_from0(s, 1); }

} // namespace sg
