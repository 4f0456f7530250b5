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
// The residual sums, which feed only the stopping rule, are deferred and done together: with the
// p' solve's first mean removal when it has one (their rows in its row launch, their totals in
// its total launch, each in its own wave), else in one launch for the rows of all three and one
// for the totals and the rule.
// (A single launch that hands the row sums to the last workgroup through a ticket needs a
// device-scope fence in every workgroup; measured at the same size it took 34-48 us.)
#include "sg_internal.hpp"

namespace sg {

namespace {

constexpr int NT = 256;  // threads per workgroup
constexpr int RW = 4;    // rows per workgroup (one lane of wave 0 each)
constexpr int CW = 256;  // columns per LDS chunk
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

// Row sums of rows [RW*blk, RW*blk + RW) of one sum, CW columns at a time through two LDS
// buffers: the next chunk's loads are in flight while the lanes of wave 0 add the current one.
__device__ __forceinline__ void row_sums_block(const SumArgs& sa, int blk) {
    __shared__ double tile[2][RW][CW + PAD];
    const int r0 = blk * RW, t = threadIdx.x;
    double v[LOADS];
    auto load = [&](int c0) {
        const int cw = min(CW, sa.w - c0);
#pragma unroll
        for (int k = 0; k < LOADS; ++k) {
            const int e = k * NT + t, r = e / CW, c = e - r * CW;
            v[k] = (c < cw && r0 + r < sa.nrow)
                       ? sa.a[static_cast<size_t>(sa.i0 + r0 + r) * sa.pitch + sa.j0 + c0 + c]
                       : 0.0;
        }
    };
    auto store = [&](int q) {
#pragma unroll
        for (int k = 0; k < LOADS; ++k) {
            const int e = k * NT + t, r = e / CW, c = e - r * CW;
            tile[q][r][c] = v[k];
        }
    };
    double s = 0.0;
    const int nch = (sa.w + CW - 1) / CW;
    if (nch > 0) {
        load(0);
        store(0);
    }
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
        const int c0 = ch * CW;
        if (ch + 1 < nch)
            load(c0 + CW);
        if (t < RW && r0 + t < sa.nrow)
            s = lds_chain(s, tile[ch & 1][t], min(CW, sa.w - c0));
        if (ch + 1 < nch)
            store((ch + 1) & 1);
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

// Up to 4 independent sums in one launch each for the rows and the totals (one wave each).
constexpr int MAXJ = 4;
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

// The totals of ma's sums (one workgroup each), and the remaining workgroups zero an array
// outside its interior rows [H, NX+H-1] x columns [H+1, NY+H] (p' before a solve from zero).
struct TotRing {
    MultiArgs ma;
    double* x;
    long ld;
    int nr, nc, H, NX, NY;
};
__device__ __forceinline__ void k_totals_ring(const TotRing& a, const Blk& bk) {
    if (bk.x < a.ma.n) {
        rows_total(a.ma.j[bk.x]);
        return;
    }
    const int H = a.H, nc = a.nc, w = nc - a.NY; // columns outside the interior per row
    const int top = H * nc, bot = (a.nr - a.NX - H) * nc, total = top + bot + a.NX * w;
    const int nw = bk.nx - a.ma.n;
    for (int k = (bk.x - a.ma.n) * NT + threadIdx.x; k < total; k += nw * NT) {
        int i, j;
        if (k < top) {
            i = k / nc, j = k % nc;
        } else if (k < top + bot) {
            const int kk = k - top;
            i = a.NX + H + kk / nc, j = kk % nc;
        } else {
            const int kk = k - top - bot, c = kk % w;
            i = H + kk / w, j = c <= H ? c : a.NY + c;
        }
        a.x[static_cast<size_t>(i) * a.ld + j] = 0.0;
    }
}

// b -= total / n over the interior (what k_total_sub stores).
struct BSubArgs {
    double* b;
    long ld;
    int H, NX, NY;
    const double* tot;
    double n;
};
__global__ void __launch_bounds__(NT) k_bsub(BSubArgs a) {
    const int j = a.H + 1 + blockIdx.x * NT + threadIdx.x, i = a.H + blockIdx.y;
    if (j > a.NY + a.H)
        return;
    const double m = *a.tot / a.n;
    double* q = a.b + static_cast<size_t>(i) * a.ld + j;
    *q = *q - m;
}

// The totals of up to MAXJ sums (wave w adds sum w's row sums in row order, in parallel), then
// the stopping rule, which reads them.
constexpr int TS_CH = 1024; // row sums per LDS chunk and sum
struct TotalsStop {
    MultiArgs ma;
    StopArgs rule;
};

__device__ __forceinline__ void k_totals_stop(const TotalsStop& a, const Blk&) {
    __shared__ double buf[MAXJ][TS_CH];
    __shared__ double tot[MAXJ];
    const MultiArgs& ma = a.ma;
    const int t = threadIdx.x, w = t / 64, lane = t % 64;
    int most = 0;
#pragma unroll
    for (int x = 0; x < MAXJ; ++x)
        if (x < ma.n)
            most = max(most, ma.j[x].nrow);
    double s = 0.0;
    for (int b0 = 0; b0 < most; b0 += TS_CH) {
#pragma unroll
        for (int x = 0; x < MAXJ; ++x) {
            if (x < ma.n) {
                const int n = min(TS_CH, ma.j[x].nrow - b0);
                for (int k = t; k < n; k += NT)
                    buf[x][k] = ma.j[x].rows[b0 + k];
            }
        }
        __syncthreads();
#pragma unroll
        for (int x = 0; x < MAXJ; ++x) {
            if (x < ma.n && w == x && lane == 0) {
                const int n = min(TS_CH, ma.j[x].nrow - b0);
                if (n > 0)
                    s = lds_chain(s, buf[x], n);
            }
        }
        __syncthreads();
    }
    if (lane == 0 && w < MAXJ)
        tot[w] = s;
    __syncthreads();
    if (t == 0) {
#pragma unroll
        for (int x = 0; x < MAXJ; ++x)
            if (x < ma.n)
                *ma.j[x].slot = tot[x];
        stop_rule(a.rule);
    }
}

struct SlotArgs {
    double* slot;
};

__device__ __forceinline__ void k_stop_only(const StopArgs& a, const Blk&) {
    if (threadIdx.x == 0)
        stop_rule(a);
}

__device__ __forceinline__ void k_zero_slot(const SlotArgs& a, const Blk&) {
    if (threadIdx.x == 0)
        *a.slot = 0.0;
}

// The total of one sum's row sums, added in row order by every workgroup (no handoff between
// workgroups), then x -= total / n over the same range (mean removal, the reference's
// x(i,j) - mean); workgroup 0 also stores the total in the slot, and adds the totals of up to
// MAXJ - 1 other sums (res: residual sums whose rows were summed in the same row launch) into
// their slots, each in its own wave at the same time as its own chain.
struct MeanArgs {
    SumArgs sa;
    double* x;
    double n;
    double* zero; // also filled with 0.0 (nzero elements), or null
    size_t nzero;
    MultiArgs res;
};
constexpr int MEAN_WG = 256; // workgroups of the fused total and subtraction
constexpr int PF = 16;       // cells per thread per batch of the subtraction
constexpr int MCH = 2048;    // row sums per LDS chunk and sum
static_assert(MAXJ * 64 <= NT, "one wave per chained sum");

__device__ __forceinline__ void k_total_sub(const MeanArgs& ma, const Blk& bk) {
    __shared__ double buf[MAXJ][MCH];
    __shared__ double tot;
    const SumArgs& sa = ma.sa;
    const int t = threadIdx.x;
    // This workgroup's cells (rows bk.x, bk.x + nx, ...; columns t, t + NT, ...) in batches of
    // PF: a batch's loads are all in flight before its stores, and the first batch's before the
    // total's chain.
    const int cpr = (sa.w + NT - 1) / NT;
    const int items = ((sa.nrow - bk.x + bk.nx - 1) / bk.nx) * cpr;
    double v[PF];
    double* ptr[PF];
    bool ok[PF];
    auto load = [&](int l0) {
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const int l = l0 + k, q = l / cpr, c = t + (l - q * cpr) * NT;
            ok[k] = l < items && c < sa.w;
            ptr[k] = ma.x + static_cast<size_t>(sa.i0 + bk.x + q * bk.nx) * sa.pitch + sa.j0 + c;
            v[k] = ok[k] ? *ptr[k] : 0.0;
        }
    };
    load(0);
    if (ma.zero)
        for (size_t k = static_cast<size_t>(bk.x) * NT + t; k < ma.nzero;
             k += static_cast<size_t>(bk.nx) * NT)
            ma.zero[k] = 0.0;
    // Wave 0 chains this sum; in workgroup 0, wave 1 + x chains residual sum x.
    const int w = t / 64, lane = t % 64;
    const int nres = bk.x == 0 ? ma.res.n : 0;
    int most = sa.nrow;
    for (int x = 0; x < nres; ++x)
        most = max(most, ma.res.j[x].nrow);
    double s = 0.0;
    for (int b0 = 0; b0 < most; b0 += MCH) {
        for (int x = 0; x <= nres; ++x) {
            const SumArgs& q = x == 0 ? sa : ma.res.j[x - 1];
            const int n = min(MCH, q.nrow - b0);
            for (int k = t; k < n; k += NT)
                buf[x][k] = q.rows[b0 + k];
        }
        __syncthreads();
        if (lane == 0 && w <= nres) {
            const SumArgs& q = w == 0 ? sa : ma.res.j[w - 1];
            const int n = min(MCH, q.nrow - b0);
            if (n > 0)
                s = lds_chain(s, buf[w], n);
        }
        __syncthreads();
    }
    if (lane == 0) {
        if (w == 0) {
            tot = s;
            if (bk.x == 0)
                *sa.slot = s;
        } else if (w <= nres) {
            *ma.res.j[w - 1].slot = s;
        }
    }
    __syncthreads();
    const double m = tot / ma.n;
    for (int l0 = 0; l0 < items; l0 += PF) {
        if (l0 > 0)
            load(l0);
#pragma unroll
        for (int k = 0; k < PF; ++k)
            if (ok[k])
                *ptr[k] = v[k] - m;
    }
}

} // namespace

void launch_ordered_mean_sub(Solver& s, double* x, int pitch, int i0, int i1, int j0, int j1,
                             double n, double* slot, const int* stop, double* zero,
                             size_t nzero) {
    const int nrow = i1 - i0 + 1, w = j1 - j0 + 1;
    if (nrow <= 0 || w <= 0) {
        emit<SlotArgs, k_zero_slot, 64>(s, dim3(1), dim3(64), 0, SlotArgs{slot}, stop);
        if (zero)
            launch_fill(s, zero, nzero, 0.0, stop);
        return;
    }
    const SumArgs sa{x, static_cast<long>(pitch), i0, nrow, j0, w, s.rowbuf, slot};
    MultiArgs res{};
    if (s.ride && !s.rec && !s.deferred.empty() && s.deferred.size() < MAXJ) {
        // The deferred residual sums ride along: their rows and this sum's in one launch.
        s.ride = false;
        s.rode = true;
        MultiArgs rows{};
        int blocks = 0;
        for (const DeferredSum& d : s.deferred) {
            if (!d.cell) { // empty range: the sum is 0.0
                emit<SlotArgs, k_zero_slot, 64>(s, dim3(1), dim3(64), 0, SlotArgs{d.slot}, stop);
                continue;
            }
            rows.j[rows.n] = SumArgs{d.cell, static_cast<long>(d.w), 0, d.nrow, 0, d.w,
                                     s.rowbuf + static_cast<size_t>(d.lane) * s.nr, d.slot};
            rows.first[rows.n] = blocks;
            blocks += static_cast<int>(cdiv(d.nrow, RW));
            ++rows.n;
        }
        s.deferred.clear();
        res = rows;
        rows.j[rows.n] = sa;
        rows.first[rows.n] = blocks;
        blocks += static_cast<int>(cdiv(nrow, RW));
        ++rows.n;
        rows.first[rows.n] = blocks;
        emit<MultiArgs, k_row_sums_multi, NT>(s, dim3(blocks), dim3(NT), 0, rows, stop);
    } else {
        emit<SumArgs, k_row_sums, NT>(s, dim3(cdiv(nrow, RW)), dim3(NT), 0, sa, stop);
    }
    emit<MeanArgs, k_total_sub, NT>(s, dim3(std::min(nrow, MEAN_WG)), dim3(NT), 0,
                                    MeanArgs{sa, x, n, zero, nzero, res}, stop);
}

void launch_b_mean_fused(Solver& s, const int* stop) {
    const sg_params& p = s.p;
    const SumArgs sb{s.F(SG_F_B), static_cast<long>(s.ld), p.H, p.NX, p.H + 1, p.NY, s.rowbuf,
                     s.scal + S_BSUM};
    // The same launch structure as launch_ordered_mean_sub (one row launch, k_row_sums_multi only
    // when residual sums ride along), so batched cavities with and without the fusion still merge.
    MultiArgs rows{}, tots{};
    int blocks = 0;
    const bool riding = s.ride && !s.rec && !s.deferred.empty() && s.deferred.size() < MAXJ;
    if (riding) {
        s.ride = false;
        s.rode = true;
        for (const DeferredSum& d : s.deferred) {
            if (!d.cell) {
                emit<SlotArgs, k_zero_slot, 64>(s, dim3(1), dim3(64), 0, SlotArgs{d.slot}, stop);
                continue;
            }
            rows.j[rows.n] = SumArgs{d.cell, static_cast<long>(d.w), 0, d.nrow, 0, d.w,
                                     s.rowbuf + static_cast<size_t>(d.lane) * s.nr, d.slot};
            rows.first[rows.n] = blocks;
            blocks += static_cast<int>(cdiv(d.nrow, RW));
            ++rows.n;
        }
        s.deferred.clear();
    }
    rows.j[rows.n] = sb;
    rows.first[rows.n] = blocks;
    blocks += static_cast<int>(cdiv(p.NX, RW));
    ++rows.n;
    rows.first[rows.n] = blocks;
    if (riding)
        emit<MultiArgs, k_row_sums_multi, NT>(s, dim3(blocks), dim3(NT), 0, rows, stop);
    else
        emit<SumArgs, k_row_sums, NT>(s, dim3(cdiv(p.NX, RW)), dim3(NT), 0, sb, stop);
    tots = rows;
    constexpr int RING_WG = 8;
    emit<TotRing, k_totals_ring, NT>(
        s, dim3(tots.n + RING_WG), dim3(NT), 0,
        TotRing{tots, s.F(SG_F_PC), static_cast<long>(s.ld), s.nr, s.nc, p.H, p.NX, p.NY}, stop);
    s.b_mean_pending = true;
    s.b_raw = true;
}

void materialize_b(Solver& s) {
    if (!s.b_raw)
        return;
    s.b_raw = false;
    const sg_params& p = s.p;
    hipLaunchKernelGGL(k_bsub, dim3(cdiv(p.NY, NT), p.NX), dim3(NT), 0, s.stream,
                       BSubArgs{s.F(SG_F_B), static_cast<long>(s.ld), p.H, p.NX, p.NY,
                                s.scal + S_BSUM, static_cast<double>(p.NX) * p.NY});
    SG_HIP(hipGetLastError());
}

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

void flush_sums(Solver& s, const int* stop, const StopArgs* rule) {
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
    if (ma.n > 0) {
        ma.first[ma.n] = blocks;
        emit<MultiArgs, k_row_sums_multi, NT>(s, dim3(blocks), dim3(NT), 0, ma, stop);
        if (rule) {
            emit<TotalsStop, k_totals_stop, NT>(s, dim3(1), dim3(NT), 0, TotalsStop{ma, *rule},
                                                stop);
            return;
        }
        emit<MultiArgs, k_rows_total_multi, NT>(s, dim3(ma.n), dim3(NT), 0, ma, stop);
    }
    if (rule)
        emit<StopArgs, k_stop_only, 64>(s, dim3(1), dim3(64), 0, *rule, stop);
}

} // namespace sg
