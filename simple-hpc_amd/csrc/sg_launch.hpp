// sg_launch.hpp — one kernel body, two launch forms.
//
// Every kernel of the outer iteration is a __device__ body `void B(const A& args, const Blk& b)`
// over a trivially copyable argument struct A, where b is the workgroup's block index within the
// cavity's own grid. It is launched either
//  - for one cavity: k_one<A, B, NT>, one workgroup per block, args by value; or
//  - for a batch of cavities in one persistent launch: k_many<A, B, NT> takes a device table of
//    per-cavity args and a work table (per cavity: first work item, grid shape, stop flag); its
//    workgroups stride over the concatenated blocks of all cavities, skipping cavities whose stop
//    flag is set.
// The body is the same code in both forms, so a batched cavity computes exactly what it computes
// alone.
//
// emit() launches k_one at once, or, when the solver is recording (Solver::rec), appends the
// launch to the record list; the batch executor (sg_batch.hip) merges the records of all cavities
// into one k_many launch per label.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <exception>
#include <string>
#include <vector>

namespace sg {

// A failed HIP call ("GPU error: <call>: <message>") and an invalid argument; the C ABI turns
// them into SG_E_HIP and SG_E_ARG.
struct HipError : std::exception {
    std::string m;
    explicit HipError(std::string s) : m(std::move(s)) {}
    const char* what() const noexcept override { return m.c_str(); }
};
struct ArgError : std::exception {
    std::string m;
    explicit ArgError(std::string s) : m(std::move(s)) {}
    const char* what() const noexcept override { return m.c_str(); }
};

#define SG_HIP(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw sg::HipError(std::string("GPU error: ") + #call + ": " +                     \
                               hipGetErrorString(e_));                                         \
    } while (0)

struct Blk {
    int x, y, z;    // block index in the cavity's grid
    int nx, ny, nz; // the cavity's grid
};

// Per-cavity entry of a batched launch.
struct Work {
    int start;      // first work item (prefix sum of the preceding cavities' blocks)
    int gx, gy, gz; // the cavity's grid
    const int* stop;
};

// XCD-aware tile order. The workgroups of a launch are dispatched round-robin over the 8 XCDs
// (linear block id mod 8), each XCD with its own L2. Renumbering the tiles so that XCD k owns the
// k-th contiguous run of tiles (row-major) keeps the halos that neighbouring tiles share inside one
// L2 instead of fetching them from HBM once per XCD. A bijection on the grid: any block order
// gives the same results.
constexpr int NXCD = 8;
__device__ __forceinline__ void xcd_tile(const Blk& b, int& tx, int& ty) {
    const int n = b.nx * b.ny, l = b.y * b.nx + b.x;
    const int q = n / NXCD, r = n % NXCD, k = l % NXCD;
    const int t = k * q + (k < r ? k : r) + l / NXCD;
    tx = t % b.nx;
    ty = t / b.nx;
}

template <class A, void (*B)(const A&, const Blk&), int NT>
__global__ void __launch_bounds__(NT) k_one(A a, const int* stop) {
    if (*stop != 0)
        return;
    const Blk b{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y),
                static_cast<int>(blockIdx.z), static_cast<int>(gridDim.x),
                static_cast<int>(gridDim.y),  static_cast<int>(gridDim.z)};
    B(a, b);
}

// The tables are read through the constant address space: every workgroup-uniform load of an
// argument becomes a scalar load, as for kernel arguments, instead of occupying vector registers.
template <class T> using Const = const __attribute__((address_space(4))) T*;

template <class A, void (*B)(const A&, const Blk&), int NT>
__global__ void __launch_bounds__(NT) k_many(const A* tab_, const Work* w_, int ncav, int total) {
    const Const<A> tab = (Const<A>)tab_;
    const Const<Work> w = (Const<Work>)w_;
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        int lo = 0, hi = ncav - 1; // last cavity whose first item is <= item
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (w[mid].start <= item)
                lo = mid;
            else
                hi = mid - 1;
        }
        // Uniform across the workgroup (one flag per cavity), so the body's barriers stay uniform.
        if (__builtin_amdgcn_readfirstlane(*w[lo].stop) == 0) {
            const int gx = w[lo].gx, gy = w[lo].gy;
            int local = item - w[lo].start;
            Blk b;
            b.nx = gx;
            b.ny = gy;
            b.nz = w[lo].gz;
            b.x = local % gx;
            local /= gx;
            b.y = local % gy;
            b.z = local / gy;
            B(*(const A*)(tab + lo), b);
        }
        __syncthreads(); // the next item's body reuses the shared memory
    }
}

struct Rec;
struct Group;

// Typed operations of one (A, B, NT) instantiation.
struct OpVT {
    const char* name;
    void (*prepare)(Group& g, std::vector<void*>& owned, int cus); // device tables, grid size
    void (*launch)(const Group& g, hipStream_t st);
};

// One recorded launch.
struct Rec {
    const OpVT* vt = nullptr;
    std::vector<unsigned char> args;
    dim3 grid, block;
    size_t lds = 0;
    const int* stop = nullptr;
    int tag = 0; // structural label set by the enqueuing code (multigrid level and cycle)
    int occ = 0; // occurrence of (vt, tag) in the cavity's record list
};

// A batched launch: the records of several cavities with one label.
struct Group {
    const OpVT* vt = nullptr;
    std::vector<const Rec*> m;
    void* tab = nullptr;
    Work* work = nullptr;
    int total = 0, grid = 0;
    dim3 block;
    size_t lds = 0;
};

template <class A, void (*B)(const A&, const Blk&), int NT> struct Op {
    static void prepare(Group& g, std::vector<void*>& owned, int cus) {
        const int n = static_cast<int>(g.m.size());
        std::vector<A> tab(n);
        std::vector<Work> work(n);
        int start = 0;
        g.lds = 0;
        g.block = g.m[0]->block;
        for (int k = 0; k < n; ++k) {
            const Rec& r = *g.m[k];
            std::memcpy(&tab[k], r.args.data(), sizeof(A));
            work[k] = Work{start, static_cast<int>(r.grid.x), static_cast<int>(r.grid.y),
                           static_cast<int>(r.grid.z), r.stop};
            start += static_cast<int>(r.grid.x * r.grid.y * r.grid.z);
            g.lds = g.lds > r.lds ? g.lds : r.lds;
        }
        g.total = start;
        void* dt = nullptr;
        void* dw = nullptr;
        SG_HIP(hipMalloc(&dt, sizeof(A) * n));
        owned.push_back(dt);
        SG_HIP(hipMalloc(&dw, sizeof(Work) * n));
        owned.push_back(dw);
        SG_HIP(hipMemcpy(dt, tab.data(), sizeof(A) * n, hipMemcpyHostToDevice));
        SG_HIP(hipMemcpy(dw, work.data(), sizeof(Work) * n, hipMemcpyHostToDevice));
        g.tab = dt;
        g.work = static_cast<Work*>(dw);
        const void* fn = reinterpret_cast<const void*>(&k_many<A, B, NT>);
        if (g.lds > 65536)
            SG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(g.lds)));
        int per_cu = 0;
        const int threads = static_cast<int>(g.block.x * g.block.y * g.block.z);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, g.lds) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        const long slots = static_cast<long>(per_cu) * cus;
        g.grid = static_cast<int>(g.total < slots ? g.total : slots);
        if (g.grid < 1)
            g.grid = 1;
    }
    static void launch(const Group& g, hipStream_t st) {
        hipLaunchKernelGGL((k_many<A, B, NT>), dim3(g.grid), g.block, g.lds, st,
                           static_cast<const A*>(g.tab), g.work,
                           static_cast<int>(g.m.size()), g.total);
    }
    static const OpVT vt;
};
template <class A, void (*B)(const A&, const Blk&), int NT>
const OpVT Op<A, B, NT>::vt{"k_many", &Op<A, B, NT>::prepare, &Op<A, B, NT>::launch};

} // namespace sg
