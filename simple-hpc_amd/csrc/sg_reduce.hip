// sg_reduce.hip — sums in the reference's fixed order: each row summed with j ascending from
// 0.0, then the row sums added in ascending row order from 0.0 (CellSums on one rank,
// src/decomp.cpp:176-183; det_sum, src/multigrid.cpp:18-34). Floating-point addition is not
// associative, so this order is what makes every sum that feeds back into a field (mean
// removals, CG alpha/beta) and every residual bit-identical to the CPU.
//
// Two launches, no fence inside either. Rows: a workgroup of 256 threads owns 4 rows; every
// thread loads its share of a 4 x 1024 chunk at once (all loads in flight) into LDS, then one lane
// per row walks its row in order from LDS (reads a batch ahead of the adds), so each chain waits
// on FP64 adds rather than memory. Total: one workgroup loads the row sums into LDS
// the same way and one lane adds them in row order. At 1024 x 1024 each launch takes about 14 us.
// The residual sums, which feed only the stopping rule, are deferred and done together (one launch
// for the rows of all three, one for the totals) just before it.
// (A single launch that hands the row sums to the last workgroup through a ticket needs a
// device-scope fence in every workgroup; measured at the same size it took 34-48 us.)
#include "sg_internal.hpp"

namespace sg {

namespace {

constexpr int NT = 256;  // threads per workgroup
constexpr int RW = 4;    // rows per workgroup (one lane of wave 0 each)
constexpr int CW = 1024; // columns per LDS chunk
constexpr int PAD = 2;   // row padding (doubles) against bank conflicts
constexpr int LOADS = RW * CW / NT;
constexpr int TCH = 4096; // row sums per chunk of the total

// s + v[0] + ... + v[n-1] in order, from LDS: batches of 16 in two register buffers, the next
// batch's reads in flight while the current one is added, so the chain waits on the adds.
__device__ __forceinline__ double lds_chain(double s, const double* v, int n) {
    constexpr int B = 16;
    const int nb = n / B;
    double a[B], b[B];
    if (nb > 0) {
#pragma unroll
        for (int q = 0; q < B; ++q)
            a[q] = v[q];
    }
    int k = 0;
    for (; k + 1 < nb; k += 2) {
#pragma unroll
        for (int q = 0; q < B; ++q)
            b[q] = v[(k + 1) * B + q];
#pragma unroll
        for (int q = 0; q < B; ++q)
            s += a[q];
        if (k + 2 < nb) {
#pragma unroll
            for (int q = 0; q < B; ++q)
                a[q] = v[(k + 2) * B + q];
        }
#pragma unroll
        for (int q = 0; q < B; ++q)
            s += b[q];
    }
    if (k < nb) {
#pragma unroll
        for (int q = 0; q < B; ++q)
            s += a[q];
    }
    for (int c = nb * B; c < n; ++c)
        s += v[c];
    return s;
}

struct SumArgs {
    const double* a;
    long pitch;
    int i0, nrow, j0, w;
    double* rows;
    double* slot;
};

// Row sums of rows [RW*blk, RW*blk + RW) of one sum.
__device__ __forceinline__ void row_sums_block(const SumArgs& sa, int blk) {
    __shared__ double tile[RW][CW + PAD];
    const int r0 = blk * RW, t = threadIdx.x;
    double s = 0.0;
    for (int c0 = 0; c0 < sa.w; c0 += CW) {
        const int cw = min(CW, sa.w - c0);
        double v[LOADS];
#pragma unroll
        for (int k = 0; k < LOADS; ++k) {
            const int e = k * NT + t, r = e / CW, c = e - r * CW;
            v[k] = (c < cw && r0 + r < sa.nrow)
                       ? sa.a[static_cast<size_t>(sa.i0 + r0 + r) * sa.pitch + sa.j0 + c0 + c]
                       : 0.0;
        }
#pragma unroll
        for (int k = 0; k < LOADS; ++k) {
            const int e = k * NT + t, r = e / CW, c = e - r * CW;
            tile[r][c] = v[k];
        }
        __syncthreads();
        if (t < RW && r0 + t < sa.nrow)
            s = lds_chain(s, tile[t], cw);
        __syncthreads();
    }
    if (t < RW && r0 + t < sa.nrow)
        sa.rows[r0 + t] = s;
}

// The row sums of one sum added in row order into its slot.
__device__ __forceinline__ void rows_total(const SumArgs& sa) {
    __shared__ double buf[TCH];
    const int t = threadIdx.x;
    double s = 0.0;
    for (int b0 = 0; b0 < sa.nrow; b0 += TCH) {
        const int n = min(TCH, sa.nrow - b0);
        for (int k = t; k < n; k += NT)
            buf[k] = sa.rows[b0 + k];
        __syncthreads();
        if (t == 0)
            s = lds_chain(s, buf, n);
        __syncthreads();
    }
    if (t == 0)
        *sa.slot = s;
}

__device__ __forceinline__ void k_row_sums(const SumArgs& sa, const Blk& bk) {
    row_sums_block(sa, bk.x);
}

__device__ __forceinline__ void k_rows_total(const SumArgs& sa, const Blk&) { rows_total(sa); }

// Up to 3 independent sums in one launch each for the rows and the totals.
constexpr int MAXJ = 3;
struct MultiArgs {
    SumArgs j[MAXJ];
    int first[MAXJ + 1]; // first row block of each sum
    int n;
};

__device__ __forceinline__ void k_row_sums_multi(const MultiArgs& ma, const Blk& bk) {
    int t = 0;
    while (t + 1 < ma.n && bk.x >= ma.first[t + 1])
        ++t;
    row_sums_block(ma.j[t], bk.x - ma.first[t]);
}

__device__ __forceinline__ void k_rows_total_multi(const MultiArgs& ma, const Blk& bk) {
    rows_total(ma.j[bk.x]);
}

struct SlotArgs {
    double* slot;
};

__device__ __forceinline__ void k_zero_slot(const SlotArgs& a, const Blk&) {
    if (threadIdx.x == 0)
        *a.slot = 0.0;
}

} // namespace

void launch_ordered_sum(Solver& s, const double* a, int pitch, int i0, int i1, int j0, int j1,
                        double* slot, const int* stop, int lane) {
    const int nrow = i1 - i0 + 1, w = j1 - j0 + 1;
    if (nrow <= 0 || w <= 0) {
        emit<SlotArgs, k_zero_slot, 64>(s, dim3(1), dim3(64), 0, SlotArgs{slot}, stop);
    } else {
        const SumArgs sa{a, static_cast<long>(pitch), i0, nrow, j0, w,
                         s.rowbuf + static_cast<size_t>(lane) * s.nr, slot};
        emit<SumArgs, k_row_sums, NT>(s, dim3(cdiv(nrow, RW)), dim3(NT), 0, sa, stop);
        emit<SumArgs, k_rows_total, NT>(s, dim3(1), dim3(NT), 0, sa, stop);
    }
}

void defer_sum(Solver& s, int lane, const double* cell, int nrow, int w, double* slot) {
    if (nrow <= 0 || w <= 0) {
        s.deferred.push_back(DeferredSum{nullptr, 0, 0, 0, slot});
        return;
    }
    s.deferred.push_back(DeferredSum{cell, nrow, w, lane, slot});
}

void flush_sums(Solver& s, const int* stop) {
    std::vector<DeferredSum> d;
    d.swap(s.deferred);
    MultiArgs ma{};
    int blocks = 0;
    for (const DeferredSum& x : d) {
        if (!x.cell) { // empty range: the sum is 0.0
            emit<SlotArgs, k_zero_slot, 64>(s, dim3(1), dim3(64), 0, SlotArgs{x.slot}, stop);
            continue;
        }
        if (ma.n == MAXJ)
            throw ArgError("too many deferred sums");
        ma.j[ma.n] = SumArgs{x.cell, static_cast<long>(x.w), 0, x.nrow, 0, x.w,
                             s.rowbuf + static_cast<size_t>(x.lane) * s.nr, x.slot};
        ma.first[ma.n] = blocks;
        blocks += static_cast<int>(cdiv(x.nrow, RW));
        ++ma.n;
    }
    if (ma.n == 0)
        return;
    ma.first[ma.n] = blocks;
    emit<MultiArgs, k_row_sums_multi, NT>(s, dim3(blocks), dim3(NT), 0, ma, stop);
    emit<MultiArgs, k_rows_total_multi, NT>(s, dim3(ma.n), dim3(NT), 0, ma, stop);
}

} // namespace sg
