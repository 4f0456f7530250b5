// sg_reduce.hip — sums in the reference's fixed order: each row summed with j ascending from
// 0.0, then the row sums added in ascending row order from 0.0 (CellSums on one rank,
// src/decomp.cpp:176-183; det_sum, src/multigrid.cpp:18-34). Floating-point addition is not
// associative, so this order is what makes every sum that feeds back into a field (mean
// removals, CG alpha/beta) and every residual bit-identical to the CPU.
//
// Each row is summed by one thread, with its columns staged through LDS in 64-wide chunks so
// that global loads stay coalesced; the row sums are then added by one thread.
#include "sg_internal.hpp"

namespace sg {

namespace {

constexpr int RT = 64; // rows per workgroup (one wave)
constexpr int CW = 64; // columns per LDS chunk

__global__ void __launch_bounds__(RT)
    k_rowsum(const double* a, long pitch, int i0, int nrow, int j0, int w, double* rows,
             const int* stop) {
    if (stopped(stop))
        return;
    __shared__ double tile[RT][CW + 1];
    const int r0 = blockIdx.x * RT;
    const int t = threadIdx.x;
    double s = 0.0;
    for (int c0 = 0; c0 < w; c0 += CW) {
        const int cw = min(CW, w - c0);
        // Coalesced load: lane t reads column c0 + t of each of the block's rows.
        for (int rr = 0; rr < RT; ++rr) {
            const int row = r0 + rr;
            if (row < nrow && t < cw)
                tile[rr][t] = a[static_cast<size_t>(i0 + row) * pitch + j0 + c0 + t];
        }
        __syncthreads();
        if (r0 + t < nrow)
            for (int c = 0; c < cw; ++c)
                s += tile[t][c];
        __syncthreads();
    }
    if (r0 + t < nrow)
        rows[r0 + t] = s;
}

__global__ void k_rowtotal(const double* rows, int nrow, double* slot, const int* stop) {
    if (stopped(stop))
        return;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int r = 0; r < nrow; ++r)
            s += rows[r];
        *slot = s;
    }
}

} // namespace

void launch_ordered_sum(Solver& s, const double* a, int pitch, int i0, int i1, int j0, int j1,
                        double* slot, const int* stop) {
    const int nrow = i1 - i0 + 1, w = j1 - j0 + 1;
    if (nrow > 0 && w > 0) {
        hipLaunchKernelGGL(k_rowsum, dim3(cdiv(nrow, RT)), dim3(RT), 0, s.stream, a,
                           static_cast<long>(pitch), i0, nrow, j0, w, s.rowbuf, stop);
        SG_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_rowtotal, dim3(1), dim3(64), 0, s.stream, s.rowbuf,
                       (nrow > 0 && w > 0) ? nrow : 0, slot, stop);
    SG_HIP(hipGetLastError());
}

} // namespace sg
