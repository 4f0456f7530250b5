// sg_unet.hip — the learned warm start of the p' solve: a small U-Net run and trained on gfx950
// with bfloat16 MFMA kernels (FP32 accumulation), and its use ahead of the multigrid solve.
//
// Network (no counterpart in the reference): 4 levels of 16, 32, 64, 128 channels, two 3x3
// convolutions (zero padding, bias, ReLU) per level, 2x2 average pooling down, nearest 2x
// upsampling and channel concatenation with the encoder features up, a 1x1 head to one channel.
// Input: 3 channels (mass imbalance, a_P, Re) zero-padded to 8. Activations and weights are
// bfloat16 (round to nearest even), NHWC; accumulators, biases, the head, gradients of weights,
// the loss and Adam's state are FP32.
//
// Convolution as implicit GEMM on v_mfma_f32_16x16x32_bf16: rows = output pixels, columns =
// output channels, K = 9 * C_in ordered (tap, channel), so the 8 K-values a lane feeds are 8
// consecutive channels of one input pixel (one 16-byte load). A wave computes 16 pixels x 64
// channels (4 accumulators). The input gradient is the same kernel on the output gradient with the
// flipped, transposed weights; the weight gradient is a GEMM over pixels (rows = output channels,
// columns = (tap, input channel)), split over pixel ranges whose partial sums are added in a
// fixed order, so training is deterministic.
//
// Warm start: the p' system is reduced to a coarse grid (level L: 2^L x 2^L blocks): block sums of
// b and block means of a_P, normalised by rms(block b) and mean(block a_P) (deterministic
// fixed-order sums). The network predicts the block means of the converged p' in those units;
// the guess g (the prediction times the scale, interpolated bilinearly to the fine cells) enters
// scaled by the energy-minimising factor b.g / g.A.g, which is never farther from the solution
// than zero in the A-norm (a degenerate guess is discarded: p' starts from zero, the plain solve),
// and the mixed-precision multigrid solve then starts from it. The stopping rule of the outer loop is untouched: a run ends only when the
// SIMPLE residuals meet the tolerance.
#include "sg_internal.hpp"
#include "sg_warm.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

namespace sg {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

constexpr int IN_CH = 3, IN_PAD = 8;
constexpr int LV[4] = {16, 32, 64, 128};
constexpr int NCONV = 14;

struct Spec {
    int cin, cout, ci_real;
};
// The 14 convolutions in execution order (input channels as stored).
constexpr Spec SPEC[NCONV] = {
    {IN_PAD, 16, IN_CH}, {16, 16, 16},   {16, 32, 16},   {32, 32, 32},   {32, 64, 32},
    {64, 64, 64},        {64, 128, 64},  {128, 128, 128}, {192, 64, 192}, {64, 64, 64},
    {96, 32, 96},        {32, 32, 32},   {48, 16, 48},   {16, 16, 16}};

int pad32(int k) { return (k + 31) / 32 * 32; }

__device__ __forceinline__ u16 f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<u16>(u >> 16);
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float(static_cast<unsigned>(h) << 16); }

__device__ __forceinline__ bf16x8 ld8(const u16* p) {
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}

// ---- weights ----

// FP32 master (co, ci_real, ky, kx) -> forward weights [co][kpad] (k = tap * cin + ci) and input-
// gradient weights [ci][kpadt] (k = tap' * cout + co, tap' = 8 - tap), zero elsewhere.
__global__ void k_wprep(const float* master, int cin, int cout, int ci_real, int kpad, int kpadt,
                        u16* wf, u16* wt) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int nf = cout * kpad;
    if (e < nf) {
        const int co = e / kpad, k = e - co * kpad;
        float v = 0.0f;
        if (k < 9 * cin) {
            const int tap = k / cin, ci = k - tap * cin;
            if (ci < ci_real)
                v = master[(static_cast<size_t>(co) * ci_real + ci) * 9 + tap];
        }
        wf[e] = f2bf(v);
    }
    const int nt = cin * kpadt;
    if (wt && e < nt) {
        const int ci = e / kpadt, k = e - ci * kpadt;
        float v = 0.0f;
        if (k < 9 * cout && ci < ci_real) {
            const int tp = k / cout, co = k - tp * cout;
            v = master[(static_cast<size_t>(co) * ci_real + ci) * 9 + (8 - tp)];
        }
        wt[e] = f2bf(v);
    }
}

// ---- convolution (forward and input gradient) ----

// Where a convolution's input operand comes from. X: a stored activation (x, ldx). The other
// forms compute the input of an inference pass on the fly from what the preceding layers stored,
// by the same expressions as the kernels they replace (so the same bits), saving the launch and
// the store:
//  - POOL: the 2x2 average pooling of x, an image of twice the sides (k_pool's sums and scale);
//  - CAT: the concatenation of the nearest 2x upsampling of x2 (ld2, channels [0, csplit), from an
//    image of half the sides: k_up's copy) and x (channels [csplit, cin), offset by csplit);
//  - FEAT: the network input of the warm start from the coarse features (k_feat_input_f32's
//    values, rounded to bfloat16).
enum { SRC_X = 0, SRC_POOL = 1, SRC_CAT = 2, SRC_FEAT = 3 };

struct ConvArgs {
    const u16* x;
    int ldx;         // input view: pixel stride (channels of the buffer)
    const u16* w;    // [cout][kpad]
    int kpad;
    const float* bias; // may be null
    u16* y;
    int ldy;         // output view
    int n, h, wd, cin, cout;
    int relu;
    const u16* zero; // 16 zero bytes: the operand of taps outside the image and steps past K
    const u16* x2;   // SRC_CAT: the upsampled channels' source
    int ld2, csplit;
    const float *cb, *ca; // SRC_FEAT: coarse block sums of b, block means of a_P
    const double* st;     // SRC_FEAT: their statistics
    float lre;            // SRC_FEAT: log10(Re / 1000)
    const float *hw, *hb; // head fused into the epilogue (hw null: none)
    float* yh;
};

// One lane's raw operand of one K step (8 input channels of one pixel and tap) as loaded; the
// arithmetic of the pooled and feature forms happens at use, so the loads of a group stay in
// flight while the previous group's MFMAs run.
template <int SRC> struct Raw {
    bf16x8 v;
};
template <> struct Raw<SRC_POOL> {
    uint4 v[4]; // the 2x2 source pixels: (2i, 2j), (2i, 2j+1), (2i+1, 2j), (2i+1, 2j+1)
};
template <> struct Raw<SRC_FEAT> {
    float b, a;
    int ok;
};

struct PixPos {
    int pn, pi, pj; // image, row, column of the lane's output pixel
    bool pv;        // the pixel exists
};

template <int SRC>
__device__ __forceinline__ Raw<SRC> fetch(const ConvArgs& a, const PixPos& px, int s, int oct,
                                          int K, float rc) {
    // kk / cin in FP32: kk < 2^12 and cin <= 2^8, so (kk + 0.5) / cin lies at least 1 / (2 cin)
    // from an integer, far more than the rounding error.
    const int kk = s * 32 + oct * 8;
    const int tap = static_cast<int>((static_cast<float>(kk) + 0.5f) * rc);
    const int ci = kk - tap * a.cin;
    const int dr = (tap * 11) >> 5; // tap / 3 for tap < 9
    const int ii = px.pi + dr - 1, jj = px.pj + (tap - 3 * dr) - 1;
    const bool ok = px.pv & (kk < K) & (ii >= 0) & (ii < a.h) & (jj >= 0) & (jj < a.wd);
    // Element offsets fit 32 bits (the host checks the sources' sizes against 2^31).
    Raw<SRC> r;
    if constexpr (SRC == SRC_X) {
        r.v = ld8(ok ? a.x + ((px.pn * a.h + ii) * a.wd + jj) * a.ldx + ci : a.zero);
    } else if constexpr (SRC == SRC_CAT) {
        const u16* p =
            ci < a.csplit
                ? a.x2 + ((px.pn * (a.h >> 1) + (ii >> 1)) * (a.wd >> 1) + (jj >> 1)) * a.ld2 + ci
                : a.x + ((px.pn * a.h + ii) * a.wd + jj) * a.ldx + (ci - a.csplit);
        r.v = ld8(ok ? p : a.zero);
    } else if constexpr (SRC == SRC_POOL) {
        const int w2 = 2 * a.wd;
        const u16* p = a.x + ((px.pn * 2 * a.h + 2 * ii) * w2 + 2 * jj) * a.ldx + ci;
        const int o[4] = {0, a.ldx, w2 * a.ldx, (w2 + 1) * a.ldx};
#pragma unroll
        for (int k = 0; k < 4; ++k)
            r.v[k] = *reinterpret_cast<const uint4*>(ok ? p + o[k] : a.zero);
    } else {
        const int e = (px.pn * a.h + ii) * a.wd + jj;
        const float* zf = reinterpret_cast<const float*>(a.zero);
        r.b = *(ok ? a.cb + e : zf);
        r.a = *(ok ? a.ca + e : zf);
        r.ok = ok;
    }
    return r;
}

__device__ __forceinline__ unsigned pool2(unsigned q0, unsigned q1, unsigned q2, unsigned q3) {
    auto one = [](unsigned a, unsigned b, unsigned c, unsigned d) {
        const float s = (__uint_as_float(a) + __uint_as_float(b)) +
                        (__uint_as_float(c) + __uint_as_float(d));
        return static_cast<unsigned>(f2bf(s * 0.25f));
    };
    const unsigned lo = one(q0 << 16, q1 << 16, q2 << 16, q3 << 16);
    const unsigned hi = one(q0 & 0xFFFF0000u, q1 & 0xFFFF0000u, q2 & 0xFFFF0000u, q3 & 0xFFFF0000u);
    return lo | (hi << 16);
}

template <int SRC>
__device__ __forceinline__ bf16x8 value(const Raw<SRC>& r, double st0, double st1, float lre) {
    if constexpr (SRC == SRC_X || SRC == SRC_CAT) {
        return r.v;
    } else if constexpr (SRC == SRC_POOL) {
        const uint4 o{pool2(r.v[0].x, r.v[1].x, r.v[2].x, r.v[3].x),
                      pool2(r.v[0].y, r.v[1].y, r.v[2].y, r.v[3].y),
                      pool2(r.v[0].z, r.v[1].z, r.v[2].z, r.v[3].z),
                      pool2(r.v[0].w, r.v[1].w, r.v[2].w, r.v[3].w)};
        return __builtin_bit_cast(bf16x8, o);
    } else {
        uint4 o{0u, 0u, 0u, 0u};
        if (r.ok) {
            const float v0 = static_cast<float>(static_cast<double>(r.b) / st0);
            const float v1 = static_cast<float>(static_cast<double>(r.a) / st1 - 1.0);
            o.x = static_cast<unsigned>(f2bf(v0)) | (static_cast<unsigned>(f2bf(v1)) << 16);
            o.y = static_cast<unsigned>(f2bf(lre));
        }
        return __builtin_bit_cast(bf16x8, o);
    }
}

constexpr int CV_T = 256; // 4 waves; a workgroup covers 64 output channels

// NB = column blocks (16 output channels each) per wave, PIXW = pixel blocks (16 pixels each) per
// workgroup. PIXW = 4 (large images): each wave computes 16 pixels x 16 NB channels and the
// workgroup 64 pixels (the A fragment is loaded once for NB MFMAs). PIXW = 1 (small images): the
// four waves of a workgroup take 16 pixels x 16 channels each (4x more waves, 4x shorter
// dependent MFMA chains). The host picks NB so that 16 NB divides the output channels.
// The K loop runs in groups of D steps, double-buffered in registers: the loads of group g + 1
// are issued before the MFMAs of group g, so a wave waits on one load latency per group instead
// of one per step. Every load is unconditional (an operand outside the image or past K is read
// from a zero buffer: the choice is made on the address, so nothing waits on the data before its
// MFMA). The MFMAs run in the same order on the same operands as one step at a time: the same
// accumulated bits, whatever the operand's source form (SRC).
// HEAD (NB = 1 and 16 output channels, so a wave holds every channel of its 16 pixels): the
// network's head, yh = sum over channels c in order of bf16(out_c) * hw[c], plus hb, is formed in
// the epilogue from the stored values (k_head's sum), gathered across the 16 lanes of a pixel
// row by shuffles.
// KS = 4 (PIXW = 1, inference on small images, where a wave's chain of K steps is a chain of
// memory latencies): the four waves of a workgroup share one 16 x 16 output tile and take a
// quarter of the K steps each; the quarters' sums are added in wave order through LDS. The same
// products summed in another order: not the bits of KS = 1.
template <int NB, int PIXW, int SRC, bool HEAD, int KS>
__global__ void __launch_bounds__(CV_T) k_conv(ConvArgs a) {
    static_assert((PIXW == 4 && (NB == 1 || NB == 2 || NB == 4) && KS == 1) ||
                      (PIXW == 1 && NB == 1 && (KS == 1 || KS == 4)),
                  "launch forms");
    static_assert(!HEAD || NB == 1, "the head needs all 16 channels of a pixel in one wave");
    constexpr int D = (NB == 4 || SRC == SRC_POOL) ? 2 : 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int m = lane & 15, oct = lane >> 4;
    const int npix = a.n * a.h * a.wd;
    const int pix0 = PIXW == 4 ? (blockIdx.x * 4 + wave) * 16 : blockIdx.x * 16;
    const int co0 = PIXW == 4   ? blockIdx.y * 16 * NB
                    : KS == 4 ? blockIdx.y * 16
                              : blockIdx.y * 64 + wave * 16;
    if (KS == 1 && co0 >= a.cout)
        return;
    PixPos px{0, 0, 0, pix0 + m < npix};
    if (px.pv) {
        const int p = pix0 + m;
        px.pj = p % a.wd;
        const int r = p / a.wd;
        px.pi = r % a.h;
        px.pn = r / a.h;
    }
    double st0 = 1.0, st1 = 1.0;
    if constexpr (SRC == SRC_FEAT) {
        st0 = a.st[0];
        st1 = a.st[1];
    }
    f4 acc[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
        acc[c] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    const int K = 9 * a.cin, nk = a.kpad >> 5;
    // This wave's steps [s0, s0 + len).
    const int s0 = KS == 1 ? 0 : wave * nk / KS, len = KS == 1 ? nk : (wave + 1) * nk / KS - s0;
    const u16* wrow[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
        wrow[c] = a.w + static_cast<size_t>(co0 + c * 16 + m) * a.kpad + oct * 8;
    const float rc = 1.0f / static_cast<float>(a.cin);
    Raw<SRC> A0[D], A1[D];
    bf16x8 B0[D][NB], B1[D][NB];
    auto load_group = [&](int g, Raw<SRC> (&A)[D], bf16x8 (&B)[D][NB]) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int s = s0 + min(g * D + d, len - 1);
            A[d] = fetch<SRC>(a, px, s, oct, K, rc);
#pragma unroll
            for (int c = 0; c < NB; ++c)
                B[d][c] = ld8(wrow[c] + s * 32);
        }
    };
    auto mma_group = [&](int g, const Raw<SRC> (&A)[D], const bf16x8 (&B)[D][NB]) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (g * D + d >= len)
                break;
            const bf16x8 av = value<SRC>(A[d], st0, st1, a.lre);
#pragma unroll
            for (int c = 0; c < NB; ++c)
                acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, B[d][c], acc[c], 0, 0, 0);
        }
    };
    // The next group is loaded whether or not it exists (its steps clamp to the last one), so
    // every path into an MFMA has the same loads in flight and the waits count only its own.
    if (len > 0) {
        const int ng = (len + D - 1) / D;
        load_group(0, A0, B0);
        for (int g = 0; g < ng; g += 2) {
            load_group(g + 1, A1, B1);
            mma_group(g, A0, B0);
            if (g + 1 >= ng)
                break;
            load_group(g + 2, A0, B0);
            mma_group(g + 1, A1, B1);
        }
    }
    if constexpr (KS > 1) {
        __shared__ f4 red[KS - 1][64];
        if (wave > 0)
            red[wave - 1][lane] = acc[0];
        __syncthreads();
        if (wave > 0)
            return;
#pragma unroll
        for (int w = 0; w < KS - 1; ++w)
            acc[0] += red[w][lane];
    }
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        const int co = co0 + c * 16 + m;
        const float b = a.bias ? a.bias[co] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int q = pix0 + 4 * oct + i;
            float z = acc[c][i] + b;
            if (a.relu)
                z = fmaxf(z, 0.0f);
            const u16 zb = f2bf(z);
            if (q < npix)
                a.y[static_cast<size_t>(q) * a.ldy + co] = zb;
            if constexpr (HEAD) {
                // Every lane of the 16 sharing oct forms the sum of pixel q; lane m == i stores it.
                const float zf = bf2f(zb);
                float h = 0.0f;
#pragma unroll
                for (int k = 0; k < 16; ++k)
                    h += __shfl(zf, 16 * oct + k) * a.hw[k];
                if (m == i && q < npix)
                    a.yh[q] = h + *a.hb;
            }
        }
    }
}

// ---- weight gradient ----

struct WgArgs {
    const u16* x;
    int ldx, cin;
    const u16* dz;
    int ldz, cout;
    float* part; // [splits][cout][9*cin]
    int n, h, wd, chunk;
};

// Weight gradient of one convolution: dW[co][k] = sum over pixels q of dz[q][co] * x(q moved by
// tap(k))[ci(k)], k = tap * cin + ci, a GEMM whose reduction runs over pixels. The MFMA wants 8
// consecutive pixels per lane for each operand, while the activations are stored channel-
// contiguous (NHWC), so a workgroup stages each batch of WG_PX pixels in LDS transposed:
//  - loads are 16 bytes (8 channels of one pixel; a group of 8 k lies in one tap, cin being a
//    multiple of 8), four pixels per thread, transposed in registers and stored as 8-byte runs of
//    four pixels of one channel;
//  - the pixel -> (image, row, column) split is done once per pixel and batch, not per element;
//  - each wave then runs WG_PX / 32 MFMAs (16 x 16 x 32) on its 16 k, A and B one 16-byte LDS
//    read each.
// A workgroup computes 16 output channels x 64 k over one pixel split (its partial sums in fixed
// pixel order); k_wreduce adds the splits in order, so training stays deterministic.
// grid: x = cdiv(9*cin, 64), y = cout / 16, z = pixel splits. Needs x, dz 16-byte aligned and
// ldx, ldz, cin multiples of 8 (checked on the host).
constexpr int WG_PX = 128;       // pixels per batch
constexpr int WG_LD = WG_PX + 8; // LDS row (u16): rows start 16-byte aligned

// Four pixels' 8 channels (v[t]: channels 0..7 of pixel t) to 8 rows of 4 pixels: row c of the
// LDS tile gets (v[0].c, v[1].c, v[2].c, v[3].c) at column col.
__device__ __forceinline__ void wg_store_quad(u16* tile, int row0, int col, const uint4 (&v)[4]) {
    const unsigned* w0 = reinterpret_cast<const unsigned*>(&v[0]);
    const unsigned* w1 = reinterpret_cast<const unsigned*>(&v[1]);
    const unsigned* w2 = reinterpret_cast<const unsigned*>(&v[2]);
    const unsigned* w3 = reinterpret_cast<const unsigned*>(&v[3]);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        // channel 2e: low halves; channel 2e + 1: high halves
        const uint2 lo = make_uint2((w0[e] & 0xFFFFu) | (w1[e] << 16), (w2[e] & 0xFFFFu) | (w3[e] << 16));
        const uint2 hi = make_uint2((w0[e] >> 16) | (w1[e] & 0xFFFF0000u),
                                    (w2[e] >> 16) | (w3[e] & 0xFFFF0000u));
        *reinterpret_cast<uint2*>(tile + (row0 + 2 * e) * WG_LD + col) = lo;
        *reinterpret_cast<uint2*>(tile + (row0 + 2 * e + 1) * WG_LD + col) = hi;
    }
}

__global__ void __launch_bounds__(256) k_wgrad(WgArgs a) {
    __shared__ __align__(16) u16 sdz[16 * WG_LD];
    __shared__ __align__(16) u16 sx[64 * WG_LD];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = lane & 15, oct = lane >> 4;
    const int K = 9 * a.cin;
    const int co0 = blockIdx.y * 16, kc0 = blockIdx.x * 64;
    const int npix = a.n * a.h * a.wd, hw = a.h * a.wd;
    const int q0b = blockIdx.z * a.chunk, q1 = min(npix, q0b + a.chunk);
    // Staging roles. x: thread = (pixel quad xq 0..31, k group xg 0..7); dz (wave 0 only):
    // thread = (pixel quad 0..31, channel half 0..1).
    const int xq = threadIdx.x & 31, xg = threadIdx.x >> 5;
    const int kg = kc0 + 8 * xg;
    const bool kval = kg < K;
    int ci = 0, dy = 0, dx = 0;
    if (kval) {
        const int tap = kg / a.cin;
        ci = kg - tap * a.cin;
        dy = tap / 3 - 1;
        dx = tap % 3 - 1;
    }
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int qb = q0b; qb < q1; qb += WG_PX) {
        {
            uint4 v[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int q = qb + 4 * xq + t;
                v[t] = make_uint4(0, 0, 0, 0);
                if (kval && q < q1) {
                    const int nn = q / hw, r = q - nn * hw, i = r / a.wd, j = r - i * a.wd;
                    const int ii = i + dy, jj = j + dx;
                    if (ii >= 0 && ii < a.h && jj >= 0 && jj < a.wd)
                        v[t] = *reinterpret_cast<const uint4*>(
                            a.x + (static_cast<size_t>(nn) * hw + ii * a.wd + jj) * a.ldx + ci);
                }
            }
            wg_store_quad(sx, 8 * xg, 4 * xq, v);
        }
        if (threadIdx.x < 64) {
            const int dq = threadIdx.x >> 1, half = threadIdx.x & 1;
            uint4 v[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int q = qb + 4 * dq + t;
                v[t] = q < q1 ? *reinterpret_cast<const uint4*>(a.dz + static_cast<size_t>(q) * a.ldz +
                                                               co0 + 8 * half)
                              : make_uint4(0, 0, 0, 0);
            }
            wg_store_quad(sdz, 8 * half, 4 * dq, v);
        }
        __syncthreads();
#pragma unroll
        for (int g = 0; g < WG_PX / 32; ++g) {
            const bf16x8 A = ld8(sdz + m * WG_LD + g * 32 + oct * 8);
            const bf16x8 B = ld8(sx + (wave * 16 + m) * WG_LD + g * 32 + oct * 8);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int cr = co0 + 4 * oct + i;
        const int kr = kc0 + wave * 16 + m;
        if (cr < a.cout && kr < K)
            a.part[(static_cast<size_t>(blockIdx.z) * a.cout + cr) * K + kr] = acc[i];
    }
}

// Sum of the splits in order, into the master-layout gradient (co, ci_real, tap).
__global__ void k_wreduce(const float* part, int splits, int cout, int cin, int ci_real, float* g) {
    const int K = 9 * cin;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= cout * K)
        return;
    float s = 0.0f;
    for (int z = 0; z < splits; ++z)
        s += part[static_cast<size_t>(z) * cout * K + e];
    const int co = e / K, k = e - co * K, tap = k / cin, ci = k - tap * cin;
    if (ci < ci_real)
        g[(static_cast<size_t>(co) * ci_real + ci) * 9 + tap] = s;
}

// Bias gradient: sum over pixels, one workgroup per channel, fixed-order tree.
__global__ void __launch_bounds__(256) k_bgrad(const u16* dz, int ldz, int npix, float* g) {
    __shared__ float sh[256];
    const int co = blockIdx.x;
    float s = 0.0f;
    for (int q = threadIdx.x; q < npix; q += 256)
        s += bf2f(dz[static_cast<size_t>(q) * ldz + co]);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        g[co] = sh[0];
}

// ---- elementwise pieces (views: base pointer, pixel stride) ----

// d *= (out > 0), c channels.
__global__ void k_mask(u16* d, int ldd, const u16* out, int ldo, int npix, int c) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= npix * c)
        return;
    const int q = e / c, ch = e - q * c;
    if (!(bf2f(out[static_cast<size_t>(q) * ldo + ch]) > 0.0f))
        d[static_cast<size_t>(q) * ldd + ch] = 0;
}

// 2x2 average pooling: x (n, h, w) -> y (n, h/2, w/2).
__global__ void k_pool(const u16* x, int ldx, u16* y, int ldy, int n, int h, int w, int c) {
    const int ho = h / 2, wo = w / 2;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * ho * wo * c)
        return;
    const int ch = e % c, q = e / c, j = q % wo, r = q / wo, i = r % ho, nn = r / ho;
    const size_t b = (static_cast<size_t>(nn * h + 2 * i) * w + 2 * j) * ldx + ch;
    const float s = (bf2f(x[b]) + bf2f(x[b + ldx])) + (bf2f(x[b + static_cast<size_t>(w) * ldx]) +
                                                      bf2f(x[b + static_cast<size_t>(w + 1) * ldx]));
    y[static_cast<size_t>(q) * ldy + ch] = f2bf(s * 0.25f);
}

// Nearest 2x upsampling: x (n, h, w) -> y (n, 2h, 2w).
__global__ void k_up(const u16* x, int ldx, u16* y, int ldy, int n, int h, int w, int c) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int H = 2 * h, W = 2 * w;
    if (e >= n * H * W * c)
        return;
    const int ch = e % c, q = e / c, j = q % W, r = q / W, i = r % H, nn = r / H;
    y[static_cast<size_t>(q) * ldy + ch] =
        x[(static_cast<size_t>(nn * h + i / 2) * w + j / 2) * ldx + ch];
}

// Upsampling's gradient: sum of the 4 children, (n, 2h, 2w) -> (n, h, w).
__global__ void k_up_back(const u16* d, int ldd, u16* y, int ldy, int n, int h, int w, int c) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * h * w * c)
        return;
    const int W = 2 * w;
    const int ch = e % c, q = e / c, j = q % w, r = q / w, i = r % h, nn = r / h;
    const size_t b = (static_cast<size_t>(nn * 2 * h + 2 * i) * W + 2 * j) * ldd + ch;
    const float s = (bf2f(d[b]) + bf2f(d[b + ldd])) + (bf2f(d[b + static_cast<size_t>(W) * ldd]) +
                                                      bf2f(d[b + static_cast<size_t>(W + 1) * ldd]));
    y[static_cast<size_t>(q) * ldy + ch] = f2bf(s);
}

// Pooling's gradient plus the skip gradient: y(n, 2h, 2w) = d(parent) / 4 + skip.
__global__ void k_pool_back_add(const u16* d, int ldd, const u16* skip, int lds, u16* y, int ldy,
                                int n, int h, int w, int c) {
    const int H = 2 * h, W = 2 * w;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * H * W * c)
        return;
    const int ch = e % c, q = e / c, j = q % W, r = q / W, i = r % H, nn = r / H;
    const float v = bf2f(d[(static_cast<size_t>(nn * h + i / 2) * w + j / 2) * ldd + ch]) * 0.25f +
                    bf2f(skip[static_cast<size_t>(q) * lds + ch]);
    y[static_cast<size_t>(q) * ldy + ch] = f2bf(v);
}

// Head: y = d0 . hw + hb (FP32 output).
__global__ void k_head(const u16* d0, int ld, const float* hw, const float* hb, float* y, int npix) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npix)
        return;
    float s = 0.0f;
    for (int c = 0; c < LV[0]; ++c)
        s += bf2f(d0[static_cast<size_t>(q) * ld + c]) * hw[c];
    y[q] = s + *hb;
}

// Loss pieces: r = y - t, dy = 2 r / npix (per pixel), per-workgroup partial sums of r^2.
__global__ void __launch_bounds__(256) k_loss(const float* y, const float* t, float* dy, int npix,
                                              float* part) {
    __shared__ float sh[256];
    const int q = blockIdx.x * 256 + threadIdx.x;
    float r2 = 0.0f;
    if (q < npix) {
        const float r = y[q] - t[q];
        dy[q] = 2.0f * r / static_cast<float>(npix);
        r2 = r * r;
    }
    sh[threadIdx.x] = r2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[blockIdx.x] = sh[0];
}

// Head backward: dd0 = bf16(dy * hw); per-workgroup partials of d0 * dy (16) and dy (1).
__global__ void __launch_bounds__(256) k_head_back(const u16* d0, int ld, const float* dy,
                                                   const float* hw, u16* dd0, int ldd, int npix,
                                                   float* part) {
    __shared__ float sh[256][LV[0] + 1];
    const int q = blockIdx.x * 256 + threadIdx.x;
    float g = 0.0f;
    if (q < npix)
        g = dy[q];
    for (int c = 0; c < LV[0]; ++c) {
        float v = 0.0f;
        if (q < npix) {
            dd0[static_cast<size_t>(q) * ldd + c] = f2bf(g * hw[c]);
            v = bf2f(d0[static_cast<size_t>(q) * ld + c]) * g;
        }
        sh[threadIdx.x][c] = v;
    }
    sh[threadIdx.x][LV[0]] = g;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int c = 0; c <= LV[0]; ++c)
                sh[threadIdx.x][c] += sh[threadIdx.x + w][c];
        __syncthreads();
    }
    if (threadIdx.x <= LV[0])
        part[blockIdx.x * (LV[0] + 1) + threadIdx.x] = sh[0][threadIdx.x];
}

// Sum of nb partial rows of width w in order (one thread per column).
__global__ void k_sum_rows(const float* part, int nb, int w, float* out) {
    const int c = threadIdx.x;
    if (c >= w)
        return;
    float s = 0.0f;
    for (int b = 0; b < nb; ++b)
        s += part[static_cast<size_t>(b) * w + c];
    out[c] = s;
}

// Adam (FP32 master, bias-corrected).
__global__ void k_adam(float* w, const float* g, float* m1, float* m2, int n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n)
        return;
    const float gi = g[e];
    const float a = b1 * m1[e] + (1.0f - b1) * gi;
    const float b = b2 * m2[e] + (1.0f - b2) * gi * gi;
    m1[e] = a;
    m2[e] = b;
    w[e] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// f32 NHWC (8 channels) -> bf16 input view.
__global__ void k_to_bf(const float* x, u16* y, int n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n)
        y[e] = f2bf(x[e]);
}

unsigned g1(long n, int t = 256) { return static_cast<unsigned>((n + t - 1) / t); }

} // namespace

// ---- the network object ----

struct UNet {
    int nparams = 0;
    int woff[NCONV] = {}, boff[NCONV] = {}, hwoff = 0, hboff = 0;
    int kpad[NCONV] = {}, kpadt[NCONV] = {};
    float *master = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    u16 *wf[NCONV] = {}, *wt[NCONV] = {};
    u16* zero16 = nullptr; // 16 zero bytes (alloc zero-fills): the convolution's padding operand
    int step = 0;
    int device = 0; // the device the network's arrays live on (current at creation)
    hipStream_t stream = nullptr;
    std::vector<void*> owned;
    // activation workspaces by (n, h, w)
    struct Ctx {
        int n = 0, h = 0, w = 0;
        u16* x_in = nullptr;
        u16 *t0, *cat0, *d0a, *d0, *p0, *t1, *cat1, *d1a, *d1, *p1, *t2, *cat2, *d2a, *d2, *p2,
            *t3, *bt;
        float *y = nullptr, *t = nullptr, *dy = nullptr, *xf = nullptr, *part = nullptr;
        float* lsum = nullptr; // loss: sum of squared errors
        // gradients
        u16 *g_d0, *g_d0a, *g_cat0, *g_d1, *g_d1a, *g_cat1, *g_d2, *g_d2a, *g_cat2, *g_bt, *g_t3,
            *g_p2, *g_e2, *g_t2, *g_p1, *g_e1, *g_t1, *g_p0, *g_e0, *g_t0;
        float* wpart = nullptr; // weight-gradient splits
        size_t wpart_elems = 0;
        std::vector<void*> owned;
        ~Ctx() {
            for (void* q : owned)
                (void)hipFree(q);
        }
    };
    std::map<std::tuple<int, int, int>, std::unique_ptr<Ctx>> ctx;

    template <class T> T* alloc(std::vector<void*>& list, size_t n) {
        void* q = nullptr;
        SG_HIP(hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1)));
        SG_HIP(hipMemset(q, 0, sizeof(T) * std::max<size_t>(n, 1)));
        list.push_back(q);
        return static_cast<T*>(q);
    }

    UNet() {
        int o = 0;
        for (int k = 0; k < NCONV; ++k) {
            woff[k] = o;
            o += SPEC[k].cout * SPEC[k].ci_real * 9;
            boff[k] = o;
            o += SPEC[k].cout;
            kpad[k] = pad32(9 * SPEC[k].cin);
            kpadt[k] = pad32(9 * SPEC[k].cout);
        }
        hwoff = o;
        o += LV[0];
        hboff = o;
        o += 1;
        nparams = o;
        SG_HIP(hipGetDevice(&device));
        SG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        master = alloc<float>(owned, nparams);
        grad = alloc<float>(owned, nparams);
        m1 = alloc<float>(owned, nparams);
        m2 = alloc<float>(owned, nparams);
        for (int k = 0; k < NCONV; ++k) {
            wf[k] = alloc<u16>(owned, static_cast<size_t>(SPEC[k].cout) * kpad[k]);
            wt[k] = alloc<u16>(owned, static_cast<size_t>(SPEC[k].cin) * kpadt[k]);
        }
        zero16 = alloc<u16>(owned, 8);
    }
    ~UNet() {
        if (stream)
            (void)hipStreamSynchronize(stream);
        ctx.clear();
        for (void* q : owned)
            (void)hipFree(q);
        if (stream)
            (void)hipStreamDestroy(stream);
    }

    void prep_weights(hipStream_t st) {
        for (int k = 0; k < NCONV; ++k) {
            const Spec& s = SPEC[k];
            const long n = std::max<long>(static_cast<long>(s.cout) * kpad[k],
                                          static_cast<long>(s.cin) * kpadt[k]);
            hipLaunchKernelGGL(k_wprep, dim3(g1(n)), dim3(256), 0, st, master + woff[k], s.cin,
                               s.cout, s.ci_real, kpad[k], kpadt[k], wf[k], wt[k]);
            SG_HIP(hipGetLastError());
        }
    }

    Ctx& get_ctx(int n, int h, int w) {
        auto key = std::make_tuple(n, h, w);
        auto it = ctx.find(key);
        if (it != ctx.end())
            return *it->second;
        if (h % 8 || w % 8 || h < 8 || w < 8 || n < 1)
            throw ArgError("U-Net input sides must be positive multiples of 8");
        auto c = std::make_unique<Ctx>();
        Ctx& x = *c;
        x.n = n;
        x.h = h;
        x.w = w;
        auto& L = x.owned;
        const size_t P0 = static_cast<size_t>(n) * h * w, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
        x.x_in = alloc<u16>(L, P0 * IN_PAD);
        x.xf = alloc<float>(L, P0 * IN_PAD);
        x.t0 = alloc<u16>(L, P0 * 16);
        x.cat0 = alloc<u16>(L, P0 * 48);
        x.d0a = alloc<u16>(L, P0 * 16);
        x.d0 = alloc<u16>(L, P0 * 16);
        x.p0 = alloc<u16>(L, P1 * 16);
        x.t1 = alloc<u16>(L, P1 * 32);
        x.cat1 = alloc<u16>(L, P1 * 96);
        x.d1a = alloc<u16>(L, P1 * 32);
        x.d1 = alloc<u16>(L, P1 * 32);
        x.p1 = alloc<u16>(L, P2 * 32);
        x.t2 = alloc<u16>(L, P2 * 64);
        x.cat2 = alloc<u16>(L, P2 * 192);
        x.d2a = alloc<u16>(L, P2 * 64);
        x.d2 = alloc<u16>(L, P2 * 64);
        x.p2 = alloc<u16>(L, P3 * 64);
        x.t3 = alloc<u16>(L, P3 * 128);
        x.bt = alloc<u16>(L, P3 * 128);
        x.y = alloc<float>(L, P0);
        x.t = alloc<float>(L, P0);
        x.dy = alloc<float>(L, P0);
        x.part = alloc<float>(L, (P0 / 256 + 2) * (LV[0] + 1));
        x.lsum = alloc<float>(L, 1);
        x.g_d0 = alloc<u16>(L, P0 * 16);
        x.g_d0a = alloc<u16>(L, P0 * 16);
        x.g_cat0 = alloc<u16>(L, P0 * 48);
        x.g_d1 = alloc<u16>(L, P1 * 32);
        x.g_d1a = alloc<u16>(L, P1 * 32);
        x.g_cat1 = alloc<u16>(L, P1 * 96);
        x.g_d2 = alloc<u16>(L, P2 * 64);
        x.g_d2a = alloc<u16>(L, P2 * 64);
        x.g_cat2 = alloc<u16>(L, P2 * 192);
        x.g_bt = alloc<u16>(L, P3 * 128);
        x.g_t3 = alloc<u16>(L, P3 * 128);
        x.g_p2 = alloc<u16>(L, P3 * 64);
        x.g_e2 = alloc<u16>(L, P2 * 64);
        x.g_t2 = alloc<u16>(L, P2 * 64);
        x.g_p1 = alloc<u16>(L, P2 * 32);
        x.g_e1 = alloc<u16>(L, P1 * 32);
        x.g_t1 = alloc<u16>(L, P1 * 32);
        x.g_p0 = alloc<u16>(L, P1 * 16);
        x.g_e0 = alloc<u16>(L, P0 * 16);
        x.g_t0 = alloc<u16>(L, P0 * 16);
        // weight-gradient splits: largest cout * 9 * cin over the layers, times the split count
        size_t mx = 0;
        for (int k = 0; k < NCONV; ++k)
            mx = std::max<size_t>(mx, static_cast<size_t>(SPEC[k].cout) * 9 * SPEC[k].cin);
        x.wpart_elems = mx * wsplits(P0);
        x.wpart = alloc<float>(L, x.wpart_elems);
        Ctx& ref = *c;
        ctx.emplace(key, std::move(c));
        return ref;
    }

    // The forward pass's activations of an (n, h, w) input, carved from the block at base
    // (256-byte aligned pieces, nothing allocated). Returns the bytes used; with base null only
    // counts them. The context's input x_in is set; its y is the caller's output.
    static size_t carve_forward(Ctx& x, unsigned char* base, int n, int h, int w) {
        size_t off = 0;
        auto take = [&](u16*& q, size_t elems) {
            q = base ? reinterpret_cast<u16*>(base + off) : nullptr;
            off += (elems * sizeof(u16) + 255) / 256 * 256;
        };
        const size_t P0 = static_cast<size_t>(n) * h * w, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
        x.n = n, x.h = h, x.w = w;
        take(x.x_in, P0 * IN_PAD);
        take(x.t0, P0 * 16);
        take(x.cat0, P0 * 48);
        take(x.d0a, P0 * 16);
        take(x.d0, P0 * 16);
        take(x.p0, P1 * 16);
        take(x.t1, P1 * 32);
        take(x.cat1, P1 * 96);
        take(x.d1a, P1 * 32);
        take(x.d1, P1 * 32);
        take(x.p1, P2 * 32);
        take(x.t2, P2 * 64);
        take(x.cat2, P2 * 192);
        take(x.d2a, P2 * 64);
        take(x.d2, P2 * 64);
        take(x.p2, P3 * 64);
        take(x.t3, P3 * 128);
        take(x.bt, P3 * 128);
        return off;
    }

    // Pixel splits of the weight gradient at npix pixels (each split a multiple of 32 pixels).
    static int wsplits(size_t npix) { return static_cast<int>(std::min<size_t>(64, std::max<size_t>(1, npix / 1024))); }

    // A convolution's input operand: a stored view (x, ldx) or one of the computed forms (SRC_*).
    struct In {
        const u16* x = nullptr;
        int ldx = 0;
        int src = SRC_X;
        const u16* x2 = nullptr; // SRC_CAT
        int ld2 = 0, csplit = 0;
        const float *cb = nullptr, *ca = nullptr; // SRC_FEAT
        const double* st = nullptr;
        float lre = 0.0f;
    };
    static In view(const u16* x, int ldx) { return In{x, ldx}; }
    static In pooled(const u16* x, int ldx) { return In{x, ldx, SRC_POOL}; }
    static In cat(const u16* up, int ldup, int cup, const u16* x, int ldx) {
        return In{x, ldx, SRC_CAT, up, ldup, cup};
    }

    // split: small images take the split-K form (inference only: training keeps one order), where
    // the unsplit form would have fewer workgroups than two per CU (at 64 x 64 and below; larger
    // images fill the GPU without it and measured slower with it).
    template <int SRC, bool HEAD>
    static void launch_conv(hipStream_t st, const ConvArgs& a, long npix, bool split) {
        // 64 pixels per workgroup (each wave up to 64 channels) when that fills the GPU, else 16
        // pixels x 64 channels (or, split, 16 pixels x 16 channels with K split over the waves).
        if (static_cast<long>(g1(npix, 64)) * cdiv(a.cout, 64) >= 1024) {
            if constexpr (!HEAD) {
                if (a.cout % 64 == 0) {
                    hipLaunchKernelGGL((k_conv<4, 4, SRC, false, 1>),
                                       dim3(g1(npix, 64), a.cout / 64), dim3(CV_T), 0, st, a);
                    return;
                }
                if (a.cout % 32 == 0) {
                    hipLaunchKernelGGL((k_conv<2, 4, SRC, false, 1>),
                                       dim3(g1(npix, 64), a.cout / 32), dim3(CV_T), 0, st, a);
                    return;
                }
            }
            hipLaunchKernelGGL((k_conv<1, 4, SRC, HEAD, 1>), dim3(g1(npix, 64), a.cout / 16),
                               dim3(CV_T), 0, st, a);
        } else if (split && static_cast<long>(g1(npix, 16)) * cdiv(a.cout, 64) < 512) {
            hipLaunchKernelGGL((k_conv<1, 1, SRC, HEAD, 4>), dim3(g1(npix, 16), a.cout / 16),
                               dim3(CV_T), 0, st, a);
        } else {
            hipLaunchKernelGGL((k_conv<1, 1, SRC, HEAD, 1>), dim3(g1(npix, 16), cdiv(a.cout, 64)),
                               dim3(CV_T), 0, st, a);
        }
    }

    // Convolution k (dgrad: its input gradient) on the n x h x w image of `in`, output view
    // (y, ldy); yh: the head fused into the epilogue (the last convolution of an inference pass);
    // split: the split-K form on small images (inference).
    void conv(hipStream_t st, int k, const In& in, u16* y, int ldy, int n, int h, int w, bool dgrad,
              float* yh = nullptr, bool split = false) {
        const Spec& s = SPEC[k];
        ConvArgs a{};
        a.x = in.x;
        a.ldx = in.ldx;
        a.y = y;
        a.ldy = ldy;
        a.n = n;
        a.h = h;
        a.wd = w;
        if (!dgrad) {
            a.w = wf[k];
            a.kpad = kpad[k];
            a.bias = master + boff[k];
            a.cin = s.cin;
            a.cout = s.cout;
            a.relu = 1;
        } else {
            a.w = wt[k];
            a.kpad = kpadt[k];
            a.bias = nullptr;
            a.cin = s.cout;
            a.cout = s.cin;
            a.relu = 0;
        }
        const long npix = static_cast<long>(n) * h * w;
        a.zero = zero16;
        a.x2 = in.x2;
        a.ld2 = in.ld2;
        a.csplit = in.csplit;
        a.cb = in.cb;
        a.ca = in.ca;
        a.st = in.st;
        a.lre = in.lre;
        auto al16 = [](const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; };
        const long lim = 1L << 31;
        bool ok = a.cout % 16 == 0 && a.cin % 8 == 0 && ldy % 8 == 0 && a.kpad % 32 == 0 &&
                  9 * a.cin <= a.kpad && npix * ldy < lim;
        switch (in.src) {
        case SRC_X:
            ok = ok && in.x && al16(in.x) && in.ldx % 8 == 0 && npix * in.ldx < lim;
            break;
        case SRC_POOL: // the source image has twice the sides
            ok = ok && in.x && al16(in.x) && in.ldx % 8 == 0 && 4 * npix * in.ldx < lim;
            break;
        case SRC_CAT: // channels [0, csplit) from the half-size image x2
            ok = ok && in.x && in.x2 && al16(in.x) && al16(in.x2) && in.ldx % 8 == 0 &&
                 in.ld2 % 8 == 0 && in.csplit % 8 == 0 && in.csplit > 0 && in.csplit < a.cin &&
                 h % 2 == 0 && w % 2 == 0 && npix * in.ldx < lim && npix * in.ld2 < lim;
            break;
        case SRC_FEAT:
            ok = ok && in.cb && in.ca && in.st && a.cin == IN_PAD && !dgrad;
            break;
        default:
            ok = false;
        }
        if (!ok)
            throw ArgError("convolution: 16-byte aligned views, input channels in multiples of 8, "
                           "output channels in multiples of 16");
        if (yh) {
            if (a.cout != 16 || in.src != SRC_X)
                throw ArgError("convolution: the fused head needs 16 output channels");
            a.hw = master + hwoff;
            a.hb = master + hboff;
            a.yh = yh;
            launch_conv<SRC_X, true>(st, a, npix, split);
        } else {
            switch (in.src) {
            case SRC_X:
                launch_conv<SRC_X, false>(st, a, npix, split);
                break;
            case SRC_POOL:
                launch_conv<SRC_POOL, false>(st, a, npix, split);
                break;
            case SRC_CAT:
                launch_conv<SRC_CAT, false>(st, a, npix, split);
                break;
            default:
                launch_conv<SRC_FEAT, false>(st, a, npix, split);
            }
        }
        SG_HIP(hipGetLastError());
    }
    void conv(hipStream_t st, int k, const u16* x, int ldx, u16* y, int ldy, int n, int h, int w,
              bool dgrad) {
        conv(st, k, view(x, ldx), y, ldy, n, h, w, dgrad);
    }

    // The forward pass for training: every layer's input is stored (the backward pass reads
    // them), pooling and upsampling are launches of their own.
    void forward_train(hipStream_t st, Ctx& c) {
        const int n = c.n, h = c.h, w = c.w;
        const int h1 = h / 2, w1 = w / 2, h2 = h / 4, w2 = w / 4, h3 = h / 8, w3 = w / 8;
        auto ew = [&](long cnt) { return dim3(g1(cnt)); };
        conv(st, 0, c.x_in, IN_PAD, c.t0, 16, n, h, w, false);
        conv(st, 1, c.t0, 16, c.cat0 + 32, 48, n, h, w, false); // e0
        hipLaunchKernelGGL(k_pool, ew(static_cast<long>(n) * h1 * w1 * 16), dim3(256), 0, st,
                           c.cat0 + 32, 48, c.p0, 16, n, h, w, 16);
        conv(st, 2, c.p0, 16, c.t1, 32, n, h1, w1, false);
        conv(st, 3, c.t1, 32, c.cat1 + 64, 96, n, h1, w1, false); // e1
        hipLaunchKernelGGL(k_pool, ew(static_cast<long>(n) * h2 * w2 * 32), dim3(256), 0, st,
                           c.cat1 + 64, 96, c.p1, 32, n, h1, w1, 32);
        conv(st, 4, c.p1, 32, c.t2, 64, n, h2, w2, false);
        conv(st, 5, c.t2, 64, c.cat2 + 128, 192, n, h2, w2, false); // e2
        hipLaunchKernelGGL(k_pool, ew(static_cast<long>(n) * h3 * w3 * 64), dim3(256), 0, st,
                           c.cat2 + 128, 192, c.p2, 64, n, h2, w2, 64);
        conv(st, 6, c.p2, 64, c.t3, 128, n, h3, w3, false);
        conv(st, 7, c.t3, 128, c.bt, 128, n, h3, w3, false);
        hipLaunchKernelGGL(k_up, ew(static_cast<long>(n) * h2 * w2 * 128), dim3(256), 0, st, c.bt,
                           128, c.cat2, 192, n, h3, w3, 128);
        conv(st, 8, c.cat2, 192, c.d2a, 64, n, h2, w2, false);
        conv(st, 9, c.d2a, 64, c.d2, 64, n, h2, w2, false);
        hipLaunchKernelGGL(k_up, ew(static_cast<long>(n) * h1 * w1 * 64), dim3(256), 0, st, c.d2,
                           64, c.cat1, 96, n, h2, w2, 64);
        conv(st, 10, c.cat1, 96, c.d1a, 32, n, h1, w1, false);
        conv(st, 11, c.d1a, 32, c.d1, 32, n, h1, w1, false);
        hipLaunchKernelGGL(k_up, ew(static_cast<long>(n) * h * w * 32), dim3(256), 0, st, c.d1, 32,
                           c.cat0, 48, n, h1, w1, 32);
        conv(st, 12, c.cat0, 48, c.d0a, 16, n, h, w, false);
        conv(st, 13, c.d0a, 16, c.d0, 16, n, h, w, false);
        const long P0 = static_cast<long>(n) * h * w;
        hipLaunchKernelGGL(k_head, dim3(g1(P0)), dim3(256), 0, st, c.d0, 16, master + hwoff,
                           master + hboff, c.y, static_cast<int>(P0));
        SG_HIP(hipGetLastError());
    }

    // The forward pass for inference in 14 launches. Pooling and upsampling are folded into the
    // next convolution's operand loads (the pooled and upsampled images are never stored) and the
    // head into the last convolution's epilogue, by the same expressions; small images take the
    // split-K convolutions (their sums in another order than the training pass's). With `in` of
    // form SRC_FEAT the input is computed from the coarse features instead of read from x_in.
    void forward(hipStream_t st, Ctx& c, const In* in = nullptr) {
        const int n = c.n, h = c.h, w = c.w;
        const int h1 = h / 2, w1 = w / 2, h2 = h / 4, w2 = w / 4, h3 = h / 8, w3 = w / 8;
        auto cv = [&](int k, const In& x, u16* y, int ldy, int hh, int ww, float* yh = nullptr) {
            conv(st, k, x, y, ldy, n, hh, ww, false, yh, true);
        };
        cv(0, in ? *in : view(c.x_in, IN_PAD), c.t0, 16, h, w);
        cv(1, view(c.t0, 16), c.cat0 + 32, 48, h, w); // e0
        cv(2, pooled(c.cat0 + 32, 48), c.t1, 32, h1, w1);
        cv(3, view(c.t1, 32), c.cat1 + 64, 96, h1, w1); // e1
        cv(4, pooled(c.cat1 + 64, 96), c.t2, 64, h2, w2);
        cv(5, view(c.t2, 64), c.cat2 + 128, 192, h2, w2); // e2
        cv(6, pooled(c.cat2 + 128, 192), c.t3, 128, h3, w3);
        cv(7, view(c.t3, 128), c.bt, 128, h3, w3);
        cv(8, cat(c.bt, 128, 128, c.cat2 + 128, 192), c.d2a, 64, h2, w2);
        cv(9, view(c.d2a, 64), c.d2, 64, h2, w2);
        cv(10, cat(c.d2, 64, 64, c.cat1 + 64, 96), c.d1a, 32, h1, w1);
        cv(11, view(c.d1a, 32), c.d1, 32, h1, w1);
        cv(12, cat(c.d1, 32, 32, c.cat0 + 32, 48), c.d0a, 16, h, w);
        cv(13, view(c.d0a, 16), c.d0, 16, h, w, c.y);
    }

    // Weight and bias gradients of conv k (input view x, masked output gradient dz).
    void wgrad(hipStream_t st, Ctx& c, int k, const u16* x, int ldx, const u16* dz, int ldz, int n,
               int h, int w) {
        const Spec& s = SPEC[k];
        const long npix = static_cast<long>(n) * h * w;
        const int splits = wsplits(npix);
        int chunk = static_cast<int>((npix + splits - 1) / splits);
        chunk = (chunk + 31) / 32 * 32;
        WgArgs a{x, ldx, s.cin, dz, ldz, s.cout, c.wpart, n, h, w, chunk};
        if ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(dz)) % 16 != 0 ||
            ldx % 8 != 0 || ldz % 8 != 0 || s.cin % 8 != 0 || s.cout % 16 != 0)
            throw ArgError("weight gradient: views must be 16-byte aligned with channel counts "
                           "in multiples of 8 (16 for outputs)");
        hipLaunchKernelGGL(k_wgrad, dim3(cdiv(9 * s.cin, 64), s.cout / 16, splits),
                           dim3(256), 0, st, a);
        SG_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_wreduce, dim3(g1(static_cast<long>(s.cout) * 9 * s.cin)), dim3(256),
                           0, st, c.wpart, splits, s.cout, s.cin, s.ci_real, grad + woff[k]);
        SG_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_bgrad, dim3(s.cout), dim3(256), 0, st, dz, ldz,
                           static_cast<int>(npix), grad + boff[k]);
        SG_HIP(hipGetLastError());
    }

    void mask(hipStream_t st, u16* d, int ldd, const u16* out, int ldo, long npix, int ch) {
        hipLaunchKernelGGL(k_mask, dim3(g1(npix * ch)), dim3(256), 0, st, d, ldd, out, ldo,
                           static_cast<int>(npix), ch);
        SG_HIP(hipGetLastError());
    }

    // Layer k backward: mask its output gradient in place, weight gradients, input gradient.
    void layer_back(hipStream_t st, Ctx& c, int k, u16* gout, int ldg, const u16* out, int ldo,
                    const u16* x, int ldx, u16* gin, int ldgin, int n, int h, int w) {
        const long npix = static_cast<long>(n) * h * w;
        mask(st, gout, ldg, out, ldo, npix, SPEC[k].cout);
        wgrad(st, c, k, x, ldx, gout, ldg, n, h, w);
        if (gin)
            conv(st, k, gout, ldg, gin, ldgin, n, h, w, true);
    }

    // Loss (mean squared error over all pixels) and every gradient into grad.
    void backward(hipStream_t st, Ctx& c) {
        const int n = c.n, h = c.h, w = c.w;
        const int h1 = h / 2, w1 = w / 2, h2 = h / 4, w2 = w / 4, h3 = h / 8, w3 = w / 8;
        const long P0 = static_cast<long>(n) * h * w;
        const int nb = static_cast<int>(g1(P0));
        hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, st, c.y, c.t, c.dy,
                           static_cast<int>(P0), c.part);
        SG_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(64), 0, st, c.part, nb, 1, c.lsum);
        SG_HIP(hipGetLastError());
        auto ew = [&](long cnt) { return dim3(g1(cnt)); };
        hipLaunchKernelGGL(k_head_back, dim3(nb), dim3(256), 0, st, c.d0, 16, c.dy,
                           master + hwoff, c.g_d0, 16, static_cast<int>(P0), c.part);
        SG_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(64), 0, st, c.part, nb, LV[0] + 1,
                           grad + hwoff);
        SG_HIP(hipGetLastError());
        layer_back(st, c, 13, c.g_d0, 16, c.d0, 16, c.d0a, 16, c.g_d0a, 16, n, h, w);
        layer_back(st, c, 12, c.g_d0a, 16, c.d0a, 16, c.cat0, 48, c.g_cat0, 48, n, h, w);
        hipLaunchKernelGGL(k_up_back, ew(static_cast<long>(n) * h1 * w1 * 32), dim3(256), 0, st,
                           c.g_cat0, 48, c.g_d1, 32, n, h1, w1, 32);
        layer_back(st, c, 11, c.g_d1, 32, c.d1, 32, c.d1a, 32, c.g_d1a, 32, n, h1, w1);
        layer_back(st, c, 10, c.g_d1a, 32, c.d1a, 32, c.cat1, 96, c.g_cat1, 96, n, h1, w1);
        hipLaunchKernelGGL(k_up_back, ew(static_cast<long>(n) * h2 * w2 * 64), dim3(256), 0, st,
                           c.g_cat1, 96, c.g_d2, 64, n, h2, w2, 64);
        layer_back(st, c, 9, c.g_d2, 64, c.d2, 64, c.d2a, 64, c.g_d2a, 64, n, h2, w2);
        layer_back(st, c, 8, c.g_d2a, 64, c.d2a, 64, c.cat2, 192, c.g_cat2, 192, n, h2, w2);
        hipLaunchKernelGGL(k_up_back, ew(static_cast<long>(n) * h3 * w3 * 128), dim3(256), 0, st,
                           c.g_cat2, 192, c.g_bt, 128, n, h3, w3, 128);
        layer_back(st, c, 7, c.g_bt, 128, c.bt, 128, c.t3, 128, c.g_t3, 128, n, h3, w3);
        layer_back(st, c, 6, c.g_t3, 128, c.t3, 128, c.p2, 64, c.g_p2, 64, n, h3, w3);
        hipLaunchKernelGGL(k_pool_back_add, ew(static_cast<long>(n) * h2 * w2 * 64), dim3(256), 0,
                           st, c.g_p2, 64, c.g_cat2 + 128, 192, c.g_e2, 64, n, h3, w3, 64);
        layer_back(st, c, 5, c.g_e2, 64, c.cat2 + 128, 192, c.t2, 64, c.g_t2, 64, n, h2, w2);
        layer_back(st, c, 4, c.g_t2, 64, c.t2, 64, c.p1, 32, c.g_p1, 32, n, h2, w2);
        hipLaunchKernelGGL(k_pool_back_add, ew(static_cast<long>(n) * h1 * w1 * 32), dim3(256), 0,
                           st, c.g_p1, 32, c.g_cat1 + 64, 96, c.g_e1, 32, n, h2, w2, 32);
        layer_back(st, c, 3, c.g_e1, 32, c.cat1 + 64, 96, c.t1, 32, c.g_t1, 32, n, h1, w1);
        layer_back(st, c, 2, c.g_t1, 32, c.t1, 32, c.p0, 16, c.g_p0, 16, n, h1, w1);
        hipLaunchKernelGGL(k_pool_back_add, ew(static_cast<long>(n) * h * w * 16), dim3(256), 0,
                           st, c.g_p0, 16, c.g_cat0 + 32, 48, c.g_e0, 16, n, h1, w1, 16);
        layer_back(st, c, 1, c.g_e0, 16, c.cat0 + 32, 48, c.t0, 16, c.g_t0, 16, n, h, w);
        layer_back(st, c, 0, c.g_t0, 16, c.t0, 16, c.x_in, IN_PAD, nullptr, 0, n, h, w);
    }

    float loss_of(Ctx& c) {
        const long P0 = static_cast<long>(c.n) * c.h * c.w;
        float s = 0.0f;
        SG_HIP(hipMemcpyAsync(&s, c.lsum, sizeof(float), hipMemcpyDeviceToHost, stream));
        SG_HIP(hipStreamSynchronize(stream));
        return s / static_cast<float>(P0);
    }

    void upload_x(Ctx& c, const float* x) {
        const size_t n = static_cast<size_t>(c.n) * c.h * c.w * IN_PAD;
        SG_HIP(hipMemcpyAsync(c.xf, x, sizeof(float) * n, hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_to_bf, dim3(g1(static_cast<long>(n))), dim3(256), 0, stream, c.xf,
                           c.x_in, static_cast<int>(n));
        SG_HIP(hipGetLastError());
    }
};

// ---- forward pass without allocation (sg_unet_forward_noalloc) ----

// Every intermediate activation of one forward pass of at most NOALLOC_PIXELS pixels (nimg * h *
// w), bfloat16: 374 bytes per pixel and 256-byte alignment of each of the 18 pieces, 6.2 MB at
// the maximum. One copy per GPU, held until the process ends.
constexpr long NOALLOC_PIXELS = 16384;
constexpr size_t NOALLOC_BYTES = 7u << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[MAX_DEVICES] = {}; // device address of `scratch` on each GPU

// The device address of `scratch` on the current device, looked up once per device with that
// device current (each GPU holds its own copy at its own address). The array's name is only a
// host-side stand-in: kernels get this address.
unsigned char* scratch_here() {
    int dev = 0;
    SG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES)
        throw ArgError("device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* q = nullptr;
        SG_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<unsigned char*>(q);
    }
    return g_scratch[dev];
}
} // namespace

// ---- warm start inside the p' solve ----

namespace {

struct FeatArgs {
    L0View l;           // level 0
    int bs;             // block side
    float* cb;          // coarse block sums of b [nxc*nyc]
    float* ca;          // coarse block means of a_P
    double* rows;       // block row segment sums of b and a_P, interleaved [nx*nyc][2]
};

// Coarse block sums of b and block means of a_P in two launches: the sum of every block row
// segment, then the sum of a block's row sums in row order (one thread per coarse cell). Rows: a
// workgroup takes FC consecutive columns of one fine row, every thread loads its share of them
// (consecutive threads, consecutive columns) into LDS, then one lane per segment adds its bs
// values in column order from LDS.
constexpr int FC = 1024; // columns per k_feat_rows workgroup (a multiple of every block side)
constexpr int FT = 256;  // its threads
__global__ void __launch_bounds__(FT) k_feat_rows(FeatArgs a) {
    __shared__ double sb[FC], sa[FC];
    const L0View& l = a.l;
    const int nyc = l.ny / a.bs;
    const int i = blockIdx.y, c0 = blockIdx.x * FC, w = min(FC, l.ny - c0);
    const size_t q0 = static_cast<size_t>(i) * l.ld + c0;
    double vb[FC / FT], va[FC / FT];
#pragma unroll
    for (int k = 0; k < FC / FT; ++k) {
        const int c = k * FT + threadIdx.x;
        vb[k] = c < w ? l.bv(q0 + c) : 0.0;
        va[k] = c < w ? l.ap(i, c0 + c, q0 + c) : 0.0;
    }
#pragma unroll
    for (int k = 0; k < FC / FT; ++k) {
        sb[k * FT + threadIdx.x] = vb[k];
        sa[k * FT + threadIdx.x] = va[k];
    }
    __syncthreads();
    for (int s = threadIdx.x; s < w / a.bs; s += FT) {
        double tb = 0.0, ta = 0.0;
        for (int c = s * a.bs; c < (s + 1) * a.bs; ++c) {
            tb += sb[c];
            ta += sa[c];
        }
        const size_t e = static_cast<size_t>(i) * nyc + c0 / a.bs + s;
        a.rows[2 * e] = tb;
        a.rows[2 * e + 1] = ta;
    }
}

__global__ void k_feat_blocks(FeatArgs a) {
    const int nyc = a.l.ny / a.bs, nxc = a.l.nx / a.bs;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nxc * nyc)
        return;
    const int I = e / nyc, J = e - I * nyc;
    double sb = 0.0, sa = 0.0;
    for (int i0 = 0; i0 < a.bs; i0 += 16) {
        double2 v[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const size_t r = static_cast<size_t>(I * a.bs + i0 + t) * nyc + J;
            v[t] = i0 + t < a.bs ? reinterpret_cast<const double2*>(a.rows)[r] : double2{0.0, 0.0};
        }
#pragma unroll
        for (int t = 0; t < 16; ++t)
            if (i0 + t < a.bs) {
                sb += v[t].x;
                sa += v[t].y;
            }
    }
    a.cb[e] = static_cast<float>(sb);
    a.ca[e] = static_cast<float>(sa / (a.bs * a.bs));
}

// Deterministic sums of cb^2 and ca over n cells: one workgroup, fixed-order tree.
__global__ void __launch_bounds__(1024) k_feat_stats(const float* cb, const float* ca, int n,
                                                     double* out) {
    __shared__ double s1[1024], s2[1024];
    double a = 0.0, b = 0.0;
    // Each thread adds its elements in index order; eight loads are in flight at a time.
    int e = threadIdx.x;
    for (; e + 7 * 1024 < n; e += 8 * 1024) {
        float vb[8], va[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            vb[q] = cb[e + q * 1024];
            va[q] = ca[e + q * 1024];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            a += static_cast<double>(vb[q]) * vb[q];
            b += va[q];
        }
    }
    for (; e < n; e += 1024) {
        a += static_cast<double>(cb[e]) * cb[e];
        b += ca[e];
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double sb = sqrt(s1[0] / n), ma = s2[0] / n;
        out[0] = sb > 0.0 ? sb : 1.0;
        out[1] = ma > 0.0 ? ma : 1.0;
        // Scale of the prediction: b / a_P, times the coarse cell count (the inverse operator's
        // growth with the grid), over 1024.
        out[2] = out[0] / out[1] * (static_cast<double>(n) / 1024.0);
    }
}

// The network input (8 channels) from the coarse fields and the statistics, in FP32 for the host
// (training data); an inference pass computes the same values in its first convolution's operand
// loads (SRC_FEAT) and rounds them to bfloat16 there.
__global__ void k_feat_input_f32(const float* cb, const float* ca, int n, const double* st,
                                 float lre, float* x) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n)
        return;
    const float v[IN_PAD] = {static_cast<float>(cb[e] / st[0]),
                             static_cast<float>(ca[e] / st[1] - 1.0), lre, 0, 0, 0, 0, 0};
    for (int c = 0; c < IN_PAD; ++c)
        x[static_cast<size_t>(e) * IN_PAD + c] = v[c];
}

// The guess enters scaled by alpha = b.g / g.A.g, the minimiser of the energy
// E(alpha g) = alpha^2 g.A.g / 2 - alpha b.g, whose minimum -(b.g)^2 / (2 g.A.g) is never above
// E(0) = 0: since E(x) - E(p') = |x - p'|_A^2 / 2, the scaled guess is never farther from the
// solution than zero in the A-norm. It is discarded (p' = 0) when g.A.g is not positive or alpha is
// not finite. The row partials are added by a fixed tree (each thread its consecutive share in
// order, then pairwise in LDS): deterministic, and latency-bound on log2 of the thread count
// rather than on a chain as long as the grid (the learned form is pinned by tolerance, not bits).
constexpr int GD_T = 1024;
__device__ __forceinline__ void k_guess_decide(const GuessArgs& a, const Blk&) {
    __shared__ double sg[GD_T], sb[GD_T];
    const int t = threadIdx.x, per = (a.nx + GD_T - 1) / GD_T;
    double gag = 0.0, bg = 0.0;
    for (int q0 = t * per; q0 < min(a.nx, (t + 1) * per); q0 += 8) {
        double2 v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            v[q] = q0 + q < min(a.nx, (t + 1) * per)
                       ? reinterpret_cast<const double2*>(a.part)[q0 + q]
                       : double2{0.0, 0.0};
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            gag += v[q].x;
            bg += v[q].y;
        }
    }
    sg[t] = gag;
    sb[t] = bg;
    __syncthreads();
    for (int w = GD_T / 2; w > 0; w >>= 1) {
        if (t < w) {
            sg[t] += sg[t + w];
            sb[t] += sb[t + w];
        }
        __syncthreads();
    }
    if (t != 0)
        return;
    gag = sg[0];
    bg = sb[0];
    const double alpha = gag > 0.0 ? bg / gag : 0.0;
    const int keep = (gag > 0.0 && isfinite(alpha) && alpha != 0.0) ? 1 : 0;
    *a.flag = keep;
    a.part[2 * a.nx] = keep ? alpha : 0.0;
    if (keep)
        ++a.ctl->warm_ok;
    else
        ++a.ctl->warm_rejected;
}

// Per row: sums of g (A g) and b g for the guess g (one workgroup per row, fixed-order tree). A
// thread takes the columns j = t, t + 256, ... in order, four at a time with all their operator
// loads issued before the arithmetic. LY: the prediction's rows this row interpolates from are
// read from an LDS copy (gy_stage).
template <bool LY> __device__ __forceinline__ void k_guess_rows(const GuessArgs& a, const Blk& bk) {
    __shared__ double sd[256], sb[256];
    __shared__ float sy[LY ? GY_ROWS * GY_MAX : 1];
    const int i = bk.x;
    const long s = a.ld;
    const L0View& l = a.l;
    const float* y = a.y;
    int k0 = 0;
    if constexpr (LY) {
        k0 = gy_stage(a, i, sy);
        y = sy;
    }
    const double scale = a.st[2];
    auto row = [&](int r) { return LY ? gy_row(a, r, k0) : grow(a, r); };
    const GAxis rm = row(i - 1), ri = row(i), rp = row(i + 1); // uniform
    double d2 = 0.0, b2 = 0.0;
    for (int c0 = 0; c0 < a.ny; c0 += 4 * 256) {
        double ap[4], ae[4], aw[4], as[4], an[4], bv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int j = c0 + k * 256 + threadIdx.x;
            const size_t q = static_cast<size_t>(i) * s + min(j, a.ny - 1);
            ap[k] = l.ap(i, min(j, a.ny - 1), q);
            ae[k] = l.AE[q];
            aw[k] = l.aw(min(j, a.ny - 1), q);
            as[k] = l.as(i, q);
            an[k] = l.AN[q];
            bv[k] = l.bv(q);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int j = c0 + k * 256 + threadIdx.x;
            if (j < a.ny) {
                const GAxis cm = gcol(a, j - 1), cj = gcol(a, j), cp = gcol(a, j + 1);
                const double x = gval(y, scale, ri, cj);
                const double ag = ap[k] * x - (ae[k] * gval(y, scale, ri, cp) +
                                               aw[k] * gval(y, scale, ri, cm) +
                                               as[k] * gval(y, scale, rp, cj) +
                                               an[k] * gval(y, scale, rm, cj));
                d2 += x * ag;
                b2 += bv[k] * x;
            }
        }
    }
    sd[threadIdx.x] = d2;
    sb[threadIdx.x] = b2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sd[threadIdx.x] += sd[threadIdx.x + w];
            sb[threadIdx.x] += sb[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.part[2 * i] = sd[0];
        a.part[2 * i + 1] = sb[0];
    }
}

} // namespace

// Per-solver state of the warm start.
struct Warm {
    UNet* net = nullptr;
    int level = 2, bs = 4, nxc = 0, nyc = 0;
    float *cb = nullptr, *ca = nullptr;
    double* st = nullptr;
    double* part = nullptr;
    double* rows = nullptr;
    int* flag = nullptr;
    float* xf = nullptr;
    std::vector<void*> owned;
    ~Warm() {
        for (void* q : owned)
            (void)hipFree(q);
    }
};

void warm_free(Solver& s) {
    delete s.warm;
    s.warm = nullptr;
}

// Level 0 as it stands: lean when assemble_p wrote only AE, AN and b; b - mean when bfused (b's
// mean removal carried by the multigrid solve: its total is in S_BSUM).
static L0View l0_view(Solver& s, bool bfused) {
    const MgLevel& L0 = s.mg[0];
    return {L0.AE,  L0.AW,  L0.AN, L0.AS, L0.AP,           L0.b,
            s.ld,   s.p.NX, s.p.NY, s.p_lean ? 1 : 0, bfused ? s.scal + S_BSUM : nullptr,
            static_cast<double>(s.p.NX) * s.p.NY};
}

static void feat_launch(Solver& s, hipStream_t st, bool bfused) {
    Warm& w = *s.warm;
    FeatArgs f{l0_view(s, bfused), w.bs, w.cb, w.ca, w.rows};
    const int nc = w.nxc * w.nyc;
    if (w.bs > FC || FC % w.bs != 0)
        throw ArgError("warm start: the block side must divide the feature row chunk");
    hipLaunchKernelGGL(k_feat_rows, dim3(cdiv(s.p.NY, FC), s.p.NX), dim3(FT), 0, st, f);
    SG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_feat_blocks, dim3(g1(nc)), dim3(256), 0, st, f);
    SG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_feat_stats, dim3(1), dim3(1024), 0, st, w.cb, w.ca, nc, w.st);
    SG_HIP(hipGetLastError());
}

void warm_attach(Solver& s, UNet* net, int level) {
    if (s.p.pressure_solver != SG_P_MG)
        throw ArgError("the learned warm start needs pressure_solver = mg");
    if (level < -1 || level > 8)
        throw ArgError("warm-start level must be -1 (automatic) or in 0..8");
    if (level < 0) {
        // The coarsest level whose grid is still at least 64 x 64 on both sides: the size of the
        // coarse grids the network was trained on (32^2 to 128^2).
        level = 0;
        const int nmin = std::min(s.p.NX, s.p.NY);
        while (level < 8 && (nmin >> (level + 1)) >= 64 && s.p.NX % (8 << (level + 1)) == 0 &&
               s.p.NY % (8 << (level + 1)) == 0)
            ++level;
    }
    const int bs = 1 << level;
    if (s.p.NX % (8 * bs) || s.p.NY % (8 * bs))
        throw ArgError("warm start: N_cells_x and N_cells_y must be multiples of 8 * 2^level");
    warm_free(s);
    s.warm = new Warm;
    Warm& w = *s.warm;
    w.net = net;
    w.level = level;
    w.bs = bs;
    w.nxc = s.p.NX / bs;
    w.nyc = s.p.NY / bs;
    const size_t nc = static_cast<size_t>(w.nxc) * w.nyc;
    auto al = [&](size_t bytes) {
        void* q = nullptr;
        SG_HIP(hipMalloc(&q, std::max<size_t>(bytes, 16)));
        SG_HIP(hipMemset(q, 0, std::max<size_t>(bytes, 16)));
        w.owned.push_back(q);
        return q;
    };
    w.cb = static_cast<float*>(al(nc * sizeof(float)));
    w.ca = static_cast<float*>(al(nc * sizeof(float)));
    w.st = static_cast<double*>(al(4 * sizeof(double)));
    w.part = static_cast<double*>(al((2 * static_cast<size_t>(s.p.NX) + 1) * sizeof(double)));
    w.flag = static_cast<int*>(al(sizeof(int)));
    w.rows = static_cast<double*>(al(2 * sizeof(double) * s.p.NX * static_cast<size_t>(w.nyc)));
    w.xf = static_cast<float*>(al(nc * IN_PAD * sizeof(float)));
    if (net)
        net->get_ctx(1, w.nxc, w.nyc); // workspace exists before any capture
    drop_graphs(s);
}

// The guess up to its decision: features, network, the energy sums and the acceptance test
// (enqueued on s.stream; every kernel is stream-ordered, so the sequence can be captured). The
// caller applies it (p' = alpha g or 0) with its own level-0 conversions (k_mx_guess).
GuessArgs launch_warm_guess(Solver& s, const int* stop, bool bfused) {
    Warm& w = *s.warm;
    if (!w.net)
        throw ArgError("pressure_warm_start = learned but no network is attached");
    if (s.rec)
        throw ArgError("the learned warm start does not run in a batch");
    UNet& net = *w.net;
    UNet::Ctx& c = net.get_ctx(1, w.nxc, w.nyc);
    feat_launch(s, s.stream, bfused);
    UNet::In in;
    in.src = SRC_FEAT;
    in.cb = w.cb;
    in.ca = w.ca;
    in.st = w.st;
    in.lre = static_cast<float>(std::log10(s.p.Re / 1000.0));
    net.forward(s.stream, c, &in);
    const MgLevel& L0 = s.mg[0];
    GuessArgs g{c.y,   w.st,   L0.x, l0_view(s, bfused), s.ld, s.p.NX, s.p.NY,
                w.bs,  w.part, w.flag, s.ctl};
    if (w.nyc <= GY_MAX)
        emit<GuessArgs, k_guess_rows<true>, 256>(s, dim3(s.p.NX), dim3(256), 0, g, stop);
    else
        emit<GuessArgs, k_guess_rows<false>, 256>(s, dim3(s.p.NX), dim3(256), 0, g, stop);
    emit<GuessArgs, k_guess_decide, GD_T>(s, dim3(1), dim3(GD_T), 0, g, stop);
    return g;
}

} // namespace sg

using namespace sg;

namespace {
template <class F> int uguard(F f) {
    try {
        f();
        return SG_OK;
    } catch (const HipError& e) {
        set_error(e.what());
        (void)hipGetLastError(); // reported here, not by the next call
        return SG_E_HIP;
    } catch (const ArgError& e) {
        set_error(e.what());
        return SG_E_ARG;
    } catch (const std::bad_alloc&) {
        set_error("out of host or device memory");
        return SG_E_NOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SG_E_STATE;
    }
}
UNet& N(sg_unet* h) {
    if (!h)
        throw ArgError("null network");
    return *reinterpret_cast<UNet*>(h);
}
} // namespace

extern "C" {

int sg_unet_nparams(int* n) {
    return uguard([&] {
        if (!n)
            throw ArgError("null pointer");
        UNet tmp;
        *n = tmp.nparams;
    });
}

int sg_unet_create(sg_unet** out) {
    return uguard([&] {
        if (!out)
            throw ArgError("null pointer");
        auto* net = new UNet;
        try {
            scratch_here(); // the no-allocation forward pass's scratch on this device
        } catch (...) {
            delete net;
            throw;
        }
        *out = reinterpret_cast<sg_unet*>(net);
    });
}

int sg_unet_destroy(sg_unet* h) {
    delete reinterpret_cast<UNet*>(h);
    return SG_OK;
}

int sg_unet_set_params(sg_unet* h, const float* theta, int n) {
    return uguard([&] {
        UNet& u = N(h);
        if (!theta || n != u.nparams)
            throw ArgError("parameter vector of the wrong size");
        SG_HIP(hipMemcpyAsync(u.master, theta, sizeof(float) * n, hipMemcpyHostToDevice, u.stream));
        SG_HIP(hipMemsetAsync(u.m1, 0, sizeof(float) * n, u.stream));
        SG_HIP(hipMemsetAsync(u.m2, 0, sizeof(float) * n, u.stream));
        u.step = 0;
        u.prep_weights(u.stream);
        SG_HIP(hipStreamSynchronize(u.stream));
    });
}

int sg_unet_get_params(sg_unet* h, float* theta, int n) {
    return uguard([&] {
        UNet& u = N(h);
        if (!theta || n != u.nparams)
            throw ArgError("parameter vector of the wrong size");
        SG_HIP(hipMemcpyAsync(theta, u.master, sizeof(float) * n, hipMemcpyDeviceToHost, u.stream));
        SG_HIP(hipStreamSynchronize(u.stream));
    });
}

int sg_unet_forward(sg_unet* h, int nimg, int ht, int wd, const float* x, float* y) {
    return uguard([&] {
        UNet& u = N(h);
        if (!x || !y)
            throw ArgError("null pointer");
        UNet::Ctx& c = u.get_ctx(nimg, ht, wd);
        u.upload_x(c, x);
        u.forward(u.stream, c);
        SG_HIP(hipMemcpyAsync(y, c.y, sizeof(float) * nimg * ht * wd, hipMemcpyDeviceToHost,
                              u.stream));
        SG_HIP(hipStreamSynchronize(u.stream));
    });
}

int sg_unet_forward_noalloc(sg_unet* h, int nimg, int ht, int wd, const float* x, float* y,
                            void* stream) {
    return uguard([&] {
        UNet& u = N(h);
        if (!x || !y)
            throw ArgError("null pointer");
        if (ht % 8 || wd % 8 || ht < 8 || wd < 8 || nimg < 1)
            throw ArgError("U-Net input sides must be positive multiples of 8");
        const long npix = static_cast<long>(nimg) * ht * wd;
        if (npix > NOALLOC_PIXELS)
            throw ArgError("sg_unet_forward_noalloc: nimg * h * w = " + std::to_string(npix) +
                           " pixels exceeds the scratch maximum of " +
                           std::to_string(NOALLOC_PIXELS));
        int dev = 0;
        SG_HIP(hipGetDevice(&dev));
        if (dev != u.device)
            throw ArgError("sg_unet_forward_noalloc: the current device is not the network's");
        UNet::Ctx c;
        if (UNet::carve_forward(c, nullptr, nimg, ht, wd) > NOALLOC_BYTES)
            throw ArgError("sg_unet_forward_noalloc: activations exceed the scratch area");
        UNet::carve_forward(c, scratch_here(), nimg, ht, wd);
        c.y = y;
        const hipStream_t st = static_cast<hipStream_t>(stream);
        const long nin = npix * IN_PAD;
        hipLaunchKernelGGL(k_to_bf, dim3(g1(nin)), dim3(256), 0, st, x, c.x_in,
                           static_cast<int>(nin));
        SG_HIP(hipGetLastError());
        u.forward(st, c);
    });
}

int sg_unet_loss_grad(sg_unet* h, int nimg, int ht, int wd, const float* x, const float* t,
                      float* loss, float* grad) {
    return uguard([&] {
        UNet& u = N(h);
        if (!x || !t)
            throw ArgError("null pointer");
        UNet::Ctx& c = u.get_ctx(nimg, ht, wd);
        u.upload_x(c, x);
        SG_HIP(hipMemcpyAsync(c.t, t, sizeof(float) * nimg * ht * wd, hipMemcpyHostToDevice,
                              u.stream));
        u.forward_train(u.stream, c);
        u.backward(u.stream, c);
        if (loss)
            *loss = u.loss_of(c);
        if (grad) {
            SG_HIP(hipMemcpyAsync(grad, u.grad, sizeof(float) * u.nparams, hipMemcpyDeviceToHost,
                                  u.stream));
            SG_HIP(hipStreamSynchronize(u.stream));
        }
    });
}

int sg_unet_train_step(sg_unet* h, int nimg, int ht, int wd, const float* x, const float* t,
                       float lr, float* loss) {
    return uguard([&] {
        UNet& u = N(h);
        if (!x || !t || !(lr > 0.0f))
            throw ArgError("null pointer or non-positive learning rate");
        UNet::Ctx& c = u.get_ctx(nimg, ht, wd);
        u.upload_x(c, x);
        SG_HIP(hipMemcpyAsync(c.t, t, sizeof(float) * nimg * ht * wd, hipMemcpyHostToDevice,
                              u.stream));
        u.forward_train(u.stream, c);
        u.backward(u.stream, c);
        ++u.step;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, static_cast<float>(u.step));
        const float c2 = 1.0f - std::pow(b2, static_cast<float>(u.step));
        hipLaunchKernelGGL(k_adam, dim3(g1(u.nparams)), dim3(256), 0, u.stream, u.master, u.grad,
                           u.m1, u.m2, u.nparams, lr, b1, b2, 1e-8f, c1, c2);
        SG_HIP(hipGetLastError());
        u.prep_weights(u.stream);
        if (loss)
            *loss = u.loss_of(c);
        else
            SG_HIP(hipStreamSynchronize(u.stream));
    });
}

int sg_unet_mfma_probe(sg_unet* h, int nimg, int ht, int wd, int reps, double* seconds,
                       double* flops) {
    return uguard([&] {
        UNet& u = N(h);
        UNet::Ctx& c = u.get_ctx(nimg, ht, wd);
        hipEvent_t e0, e1;
        SG_HIP(hipEventCreate(&e0));
        SG_HIP(hipEventCreate(&e1));
        u.forward(u.stream, c);
        SG_HIP(hipEventRecord(e0, u.stream));
        for (int r = 0; r < std::max(reps, 1); ++r)
            u.forward(u.stream, c);
        SG_HIP(hipEventRecord(e1, u.stream));
        SG_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        SG_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (seconds)
            *seconds = 1e-3 * ms / std::max(reps, 1);
        if (flops) {
            double f = 0.0;
            const double P0 = static_cast<double>(nimg) * ht * wd;
            const double lvl[NCONV] = {1, 1, 4, 4, 16, 16, 64, 64, 16, 16, 4, 4, 1, 1};
            for (int k = 0; k < NCONV; ++k)
                f += 2.0 * 9 * SPEC[k].ci_real * SPEC[k].cout * P0 / lvl[k];
            *flops = f;
        }
    });
}

int sg_attach_warm_start(sg_solver* sh, sg_unet* net, int level) {
    return uguard([&] {
        if (!sh)
            throw ArgError("null solver");
        Solver& s = *reinterpret_cast<Solver*>(sh);
        materialize_pc(s);
        materialize_b(s);
        warm_attach(s, reinterpret_cast<UNet*>(net), level);
    });
}

int sg_warm_features(sg_solver* sh, int level, float* x, double* scale) {
    return uguard([&] {
        if (!sh || !x)
            throw ArgError("null pointer");
        Solver& s = *reinterpret_cast<Solver*>(sh);
        materialize_pc(s);
        materialize_b(s);
        if (!s.warm || s.warm->level != level) {
            UNet* keep = s.warm ? s.warm->net : nullptr;
            warm_attach(s, keep, level);
        }
        Warm& w = *s.warm;
        const int nc = w.nxc * w.nyc;
        feat_launch(s, s.stream, false); // b materialized above
        const float lre = static_cast<float>(std::log10(s.p.Re / 1000.0));
        hipLaunchKernelGGL(k_feat_input_f32, dim3(g1(nc)), dim3(256), 0, s.stream, w.cb, w.ca, nc,
                           w.st, lre, w.xf);
        SG_HIP(hipGetLastError());
        double st[3];
        SG_HIP(hipMemcpyAsync(x, w.xf, sizeof(float) * nc * IN_PAD, hipMemcpyDeviceToHost,
                              s.stream));
        SG_HIP(hipMemcpyAsync(st, w.st, sizeof st, hipMemcpyDeviceToHost, s.stream));
        SG_HIP(hipStreamSynchronize(s.stream));
        if (scale)
            *scale = st[2];
    });
}

int sg_warm_level(sg_solver* sh, int* level) {
    return uguard([&] {
        if (!sh || !level)
            throw ArgError("null pointer");
        const Solver& s = *reinterpret_cast<Solver*>(sh);
        *level = s.warm ? s.warm->level : -1;
    });
}

int sg_warm_stats(sg_solver* sh, int* accepted, int* rejected) {
    return uguard([&] {
        if (!sh)
            throw ArgError("null solver");
        Solver& s = *reinterpret_cast<Solver*>(sh);
        materialize_pc(s);
        materialize_b(s);
        Ctl c;
        SG_HIP(hipMemcpyAsync(&c, s.ctl, sizeof c, hipMemcpyDeviceToHost, s.stream));
        SG_HIP(hipStreamSynchronize(s.stream));
        if (accepted)
            *accepted = c.warm_ok;
        if (rejected)
            *rejected = c.warm_rejected;
    });
}

} // extern "C"
