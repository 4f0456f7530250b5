// sg_warm.hpp — the learned warm start's guess as the p' solve reads it: the level-0 view of the
// p' system, the interpolated and scaled network prediction, and the arguments of the kernels
// that evaluate it (sg_unet.hip: the energy sums and the decision; sg_mg.hip: the guess applied
// with the mixed solve's conversions).
#pragma once

#include "sg_internal.hpp"

namespace sg {

// Level 0 of the p' system as the warm start reads it: AW, AS and AP derived from AE and AN when
// the assembly was lean, and b - mean when b's mean removal is pending (the values the full
// assembly and k_total_sub would have stored, by the same expressions as the mixed form's
// conversions).
struct L0View {
    const double *AE, *AW, *AN, *AS, *AP, *b; // interior cell (0, 0) at index 0, stride ld
    int ld, nx, ny;
    int lean;             // AW, AS, AP not assembled
    const double* bmean;  // b's total when its mean removal is pending, else null
    double bn;
    __device__ __forceinline__ double bv(size_t q) const { return bmean ? b[q] - *bmean / bn : b[q]; }
    __device__ __forceinline__ double aw(int j, size_t q) const {
        return lean ? (j == 0 ? 0.0 : AE[q - 1]) : AW[q];
    }
    __device__ __forceinline__ double as(int i, size_t q) const {
        return lean ? (i == nx - 1 ? 0.0 : AN[q + ld]) : AS[q];
    }
    __device__ __forceinline__ double ap(int i, int j, size_t q) const {
        return lean ? AE[q] + aw(j, q) + AN[q] + as(i, q) : AP[q];
    }
};

struct GuessArgs {
    const float* y;     // prediction on the coarse grid
    const double* st;   // statistics (st[2] = scale)
    double* pc;         // level-0 p' (interior cell (0,0) at pc[0], stride ld)
    L0View l;           // level 0 of the p' system
    int ld, nx, ny, bs;
    double* part;       // per-row partials: [nx][2] (defect^2, b^2)
    int* flag;          // 1: keep the guess
    Ctl* ctl;
};

// The guess g at interior cell (i, j): scale * the prediction interpolated bilinearly from the
// coarse cell centres (clamped at the walls) to the fine cell centre; 0 outside the interior (p''s
// halo stays zero). Evaluated where it is needed (the energy sums and the final write), so it is
// never stored and read back. The interpolation is separable: each axis gives its two coarse
// neighbours and weight (GAxis, computed once per row or column where a kernel reuses it), and
// gval combines them. The block side bs is a power of two, so (k + 1/2) / bs is computed as a
// product by 1 / bs, exactly.
struct GAxis {
    int o0, o1; // offsets of the two coarse neighbours (rows: times the coarse row length)
    float w;    // weight of the second
    bool in;    // the fine index lies in the interior
};
__device__ __forceinline__ GAxis gaxis(int k, int n, int nc, float inv, int stride) {
    GAxis r;
    r.in = k >= 0 && k < n;
    const float f = (k + 0.5f) * inv - 0.5f;
    const float c = fminf(fmaxf(f, 0.0f), static_cast<float>(nc - 1));
    const int k0 = min(static_cast<int>(c), nc - 1), k1 = min(k0 + 1, nc - 1);
    r.o0 = k0 * stride;
    r.o1 = k1 * stride;
    r.w = c - k0;
    return r;
}
// y: the prediction (a.y), or the workgroup's copy of some of its rows (gy_stage) with row
// offsets taken relative to the first of them (gy_row).
__device__ __forceinline__ double gval(const float* y, double scale, const GAxis& r,
                                       const GAxis& c) {
    if (!(r.in && c.in))
        return 0.0;
    const float wi = r.w, wj = c.w;
    const float v = (1.0f - wi) * ((1.0f - wj) * y[r.o0 + c.o0] + wj * y[r.o0 + c.o1]) +
                    wi * ((1.0f - wj) * y[r.o1 + c.o0] + wj * y[r.o1 + c.o1]);
    return static_cast<double>(v) * scale;
}
__device__ __forceinline__ double gval(const GuessArgs& a, const GAxis& r, const GAxis& c) {
    return gval(a.y, a.st[2], r, c);
}
__device__ __forceinline__ GAxis grow(const GuessArgs& a, int i) {
    return gaxis(i, a.nx, a.nx / a.bs, 1.0f / a.bs, a.ny / a.bs);
}
__device__ __forceinline__ GAxis gcol(const GuessArgs& a, int j) {
    return gaxis(j, a.ny, a.ny / a.bs, 1.0f / a.bs, 1);
}
__device__ __forceinline__ double guess_at(const GuessArgs& a, int i, int j) {
    return gval(a, grow(a, i), gcol(a, j));
}

// A workgroup whose cells lie in fine rows i - 1 .. i + 1 interpolates from at most GY_ROWS
// consecutive coarse rows (from the first row of fine row i - 1 to the second of fine row i + 1:
// the rows of gaxis do not decrease with the fine row; four at bs = 1, three above). gy_stage
// copies them to sy (GY_ROWS x nyc floats, nyc <= GY_MAX) and returns the first one's index;
// gy_row(a, i, k0) is grow with its offsets relative to that row.
constexpr int GY_ROWS = 4, GY_MAX = 1024;
__device__ __forceinline__ int gy_stage(const GuessArgs& a, int i, float* sy) {
    const int nyc = a.ny / a.bs;
    const int k0 = grow(a, max(i - 1, 0)).o0 / nyc, k1 = grow(a, min(i + 1, a.nx - 1)).o1 / nyc;
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    for (int e = t; e < (k1 - k0 + 1) * nyc; e += blockDim.x * blockDim.y)
        sy[e] = a.y[static_cast<size_t>(k0) * nyc + e];
    __syncthreads();
    return k0;
}
__device__ __forceinline__ GAxis gy_row(const GuessArgs& a, int i, int k0) {
    GAxis r = grow(a, i);
    const int off = k0 * (a.ny / a.bs);
    r.o0 -= off;
    r.o1 -= off;
    return r;
}

} // namespace sg
