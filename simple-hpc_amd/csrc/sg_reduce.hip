// sg_reduce.hip — sums in the reference's fixed order: each row summed with j ascending from
// 0.0, then the row sums added in ascending row order from 0.0 (CellSums on one rank,
// src/decomp.cpp:176-183; det_sum, src/multigrid.cpp:18-34). Floating-point addition is not
// associative, so this order is what makes every sum that feeds back into a field (mean
// removals, CG alpha/beta) and every residual bit-identical to the CPU.
//
// One launch. A workgroup of 256 threads owns RW rows: all threads load a RW x CW chunk with
// coalesced, independent loads into LDS, then RW threads add their rows' chunk in order. The
// additions of a row are inherently sequential (~1 FP64 add per few cycles); spreading the rows
// over many workgroups keeps every CU loading while they run. The last workgroup to finish (a
// device-scope atomic ticket) adds the row sums in order and re-arms the ticket.
#include "sg_internal.hpp"

namespace sg {

namespace {

constexpr int NT = 256;  // threads per workgroup
constexpr int RW = 4;    // rows per workgroup
constexpr int CW = 1024; // columns per LDS chunk
constexpr int PAD = 2;   // row padding (doubles) against bank conflicts
constexpr int LOADS = RW * CW / NT;

// s += v[0] + v[1] + ... in order, from LDS. The reads of the next batch of 8 are issued before
// the 8 dependent additions of the current one, so LDS latency overlaps the add chain.
__device__ __forceinline__ double add_in_order(double s, const double* v, int n) {
    const int nb = n / 8;
    if (nb > 0) {
        double a[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            a[q] = v[q];
        for (int b = 1; b < nb; ++b) {
            double c[8];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                c[q] = v[8 * b + q];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                s += a[q];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                a[q] = c[q];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
            s += a[q];
    }
    for (int c = 8 * nb; c < n; ++c)
        s += v[c];
    return s;
}

struct SumArgs {
    const double* a;
    long pitch;
    int i0, nrow, j0, w;
    double* rows;
    unsigned* ticket;
    double* slot;
};

__device__ __forceinline__ void k_ordered_sum(const SumArgs& sa, const Blk& bk) {
    const double* a = sa.a;
    const long pitch = sa.pitch;
    const int i0 = sa.i0, nrow = sa.nrow, j0 = sa.j0, w = sa.w;
    double* rows = sa.rows;
    unsigned* ticket = sa.ticket;
    double* slot = sa.slot;
    __shared__ double tile[RW][CW + PAD];
    __shared__ bool last;
    const int r0 = bk.x * RW;
    const int t = threadIdx.x;
    double s = 0.0;
    for (int c0 = 0; c0 < w; c0 += CW) {
        const int cw = min(CW, w - c0);
        double v[LOADS];
#pragma unroll
        for (int k = 0; k < LOADS; ++k) {
            const int e = k * NT + t, r = e / CW, c = e - r * CW;
            v[k] = (c < cw && r0 + r < nrow)
                       ? a[static_cast<size_t>(i0 + r0 + r) * pitch + j0 + c0 + c]
                       : 0.0;
        }
#pragma unroll
        for (int k = 0; k < LOADS; ++k) {
            const int e = k * NT + t, r = e / CW, c = e - r * CW;
            tile[r][c] = v[k];
        }
        __syncthreads();
        if (t < RW && r0 + t < nrow)
            s = add_in_order(s, tile[t], cw);
        __syncthreads();
    }
    if (t < RW && r0 + t < nrow)
        rows[r0 + t] = s;
    __threadfence();
    __syncthreads();
    if (t == 0)
        last = atomicAdd(ticket, 1u) == static_cast<unsigned>(bk.nx - 1);
    __syncthreads();
    if (!last)
        return;
    __threadfence();
    // Rows ascending from 0.0, staged through LDS, added by thread 0.
    double* buf = &tile[0][0];
    const int chunk = RW * (CW + PAD);
    double total = 0.0;
    for (int b0 = 0; b0 < nrow; b0 += chunk) {
        const int n = min(chunk, nrow - b0);
        for (int k = t; k < n; k += NT)
            buf[k] = __builtin_nontemporal_load(&rows[b0 + k]);
        __syncthreads();
        if (t == 0)
            total = add_in_order(total, buf, n);
        __syncthreads();
    }
    if (t == 0) {
        *slot = total;
        *ticket = 0u;
    }
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
                        double* slot, const int* stop, int lane, hipStream_t stream) {
    const int nrow = i1 - i0 + 1, w = j1 - j0 + 1;
    hipStream_t st = stream ? stream : s.stream;
    if (nrow <= 0 || w <= 0) {
        emit<SlotArgs, k_zero_slot, 64>(s, dim3(1), dim3(64), 0, SlotArgs{slot}, stop, st);
    } else {
        const SumArgs sa{a, static_cast<long>(pitch), i0, nrow, j0, w,
                         s.rowbuf + static_cast<size_t>(lane) * s.nr, s.ticket + lane, slot};
        emit<SumArgs, k_ordered_sum, NT>(s, dim3(cdiv(nrow, RW)), dim3(NT), 0, sa, stop, st);
    }
}

void side_sum(Solver& s, int lane, const double* cell, int nrow, int w, double* slot,
              const int* stop) {
    if (s.rec) { // recorded as a side-stream launch; the batch executor forks and joins
        launch_ordered_sum(s, cell, w > 0 ? w : 1, 0, nrow - 1, 0, w - 1, slot, stop, lane, s.side);
        return;
    }
    SG_HIP(hipEventRecord(s.fork[lane - 1], s.stream));
    SG_HIP(hipStreamWaitEvent(s.side, s.fork[lane - 1], 0));
    launch_ordered_sum(s, cell, w > 0 ? w : 1, 0, nrow - 1, 0, w - 1, slot, stop, lane, s.side);
    SG_HIP(hipEventRecord(s.join[lane - 1], s.side));
}

void join_side(Solver& s) {
    if (s.rec) {
        Rec r; // join marker
        r.tag = s.tag;
        for (const Rec& q : *s.rec)
            r.occ += (q.vt == nullptr && q.tag == r.tag) ? 1 : 0;
        s.rec->push_back(std::move(r));
        return;
    }
    for (int k = 0; k < 3; ++k)
        SG_HIP(hipStreamWaitEvent(s.stream, s.join[k], 0));
}

} // namespace sg
