// SVF surrogate network on the device: BF16 convolutions on MFMA with FP32
// accumulation, forward and backward, plus the elementwise layers, the loss
// and the optimiser. Activations are NHWC bf16 with channel counts that are
// multiples of 16 (16 channels = one K chunk of a 32x32x16 MFMA).
//
//   k_conv_t        implicit-GEMM 3x3 / 1x1 convolution, M = pixels, N = output
//                   channels, K = taps x input channels. Inputs: an optional
//                   half-resolution tensor read through nearest upsampling and
//                   an optional full-resolution skip tensor (channel concat);
//                   epilogue: bias, ReLU or sigmoid, outputs split at a channel
//                   (the data gradient of a concat layer lands in two tensors).
//                   Also runs the data gradient with flipped weights.
//   k_conv_wgrad    weight gradient, M = output channels, N = input channels,
//                   K = pixels split into chunks; partial sums per chunk
//   k_reduce_chunks deterministic sum of the chunk partials (weights and biases)
//   small kernels   2x2 average pool (fwd/bwd), upsampling bwd, ReLU bwd,
//                   sigmoid-MSE head loss, Adam, weight flip, input packing
//
// MFMA operand layout (v_mfma_f32_32x32x16_bf16, wave64): lane l holds
// A[row l % 32][k 8 * (l / 32) + j] and B[k 8 * (l / 32) + j][col l % 32],
// j = 0..7; accumulator i of lane l is C[8 * (i / 4) + 4 * (l / 32) + i % 4][l % 32].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sw_internal.h"

#define set_err swi_set_err
#define hip_check swi_hip_check
#define S swi_stream

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {

__device__ __forceinline__ float bf2f(__bf16 v) { return (float)v; }
__device__ __forceinline__ __bf16 f2bf(float v) { return (__bf16)v; }   // round to nearest even

struct ConvArgs {
    const __bf16 *in1;     // c1 channels; half resolution when up1 (read through nearest upsampling)
    const __bf16 *in2;     // c2 channels at full resolution (skip), may be null when c2 == 0
    const __bf16 *wt;      // [cout][ks * ks][c1 + c2]
    const float *bias;     // [cout] or null
    __bf16 *out1;          // output channels [0, o1) (NHWC with o1 channels)
    __bf16 *out2;          // output channels [o1, cout) (NHWC with cout - o1 channels), may be null
    float *outf;           // f32 output [pixel][cout] (head), may be null
    int c1, c2, up1, o1, cout, n, h, w, act;   // act: 0 none, 1 ReLU, 2 sigmoid
};

// Implicit-GEMM convolution, LDS tiled: a block computes BM = 32 * WAVES_M * WM pixels x
// BN = 32 * WAVES_N * WN output channels. K runs over (tap, 16-channel chunk) pairs; a stage of KST
// chunks is gathered from global memory into registers while the previous stage's MFMAs run, then
// written to the other half of a double-buffered LDS tile (A: [chunk][pixel][k half], B:
// [chunk][co][k half], 16 bytes per entry). Each wave issues KST x WM x WN MFMAs per barrier.
template <int KS, int WAVES_M, int WM, int WN, int KST>
__global__ void __launch_bounds__(256) k_conv_t(ConvArgs a) {
    constexpr int WAVES_N = 4 / WAVES_M;
    constexpr int BM = 32 * WAVES_M * WM, BN = 32 * WAVES_N * WN;
    constexpr int AV = (2 * BM + 255) / 256;     // A entries per thread per chunk
    static_assert(2 * BM % 256 == 0 && 2 * BN <= 256, "tile shape");
    __shared__ bf16x8 sA[2][KST][2 * BM];
    __shared__ bf16x8 sB[2][KST][2 * BN];
    const int tid = threadIdx.x, l = tid & 63, wave = tid >> 6;
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;
    const long M = (long)a.n * a.h * a.w;
    const long m0 = (long)blockIdx.x * BM;
    const int n0 = blockIdx.y * BN;
    const int cin = a.c1 + a.c2;
    const int cpt = cin / 16;                    // chunks per tap
    const int nchunk = KS * KS * cpt;
    const int hh = a.h >> 1, wh = a.w >> 1;
    // A loader: entry e = tid + 256 v -> pixel e / 2, k half e % 2
    int ax[AV], ay[AV], aimg[AV];
    bool amv[AV];
#pragma unroll
    for (int v = 0; v < AV; ++v) {
        const int e = tid + 256 * v;
        const long am = m0 + (e >> 1);
        amv[v] = am < M;
        const long amm = amv[v] ? am : 0;
        ax[v] = (int)(amm % a.w);
        ay[v] = (int)((amm / a.w) % a.h);
        aimg[v] = (int)(amm / ((long)a.w * a.h));
    }
    const int ah = tid & 1;
    const bool ldB = tid < 2 * BN;
    const int bco = n0 + (tid >> 1);
    const bool bcv = ldB && bco < a.cout;
    bf16x8 ga[KST][AV], gb[KST];
    auto gload = [&](int st) {
#pragma unroll
        for (int c = 0; c < KST; ++c) {
            const int kc = st * KST + c;
            const bool kv = kc < nchunk;
            const int tap = kv ? kc / cpt : 0;
            const int cb = kv ? (kc % cpt) * 16 : 0;
            const int ch = cb + 8 * ah;
#pragma unroll
            for (int v = 0; v < AV; ++v) {
                bf16x8 av = (bf16x8){};
                const int yy = ay[v] + tap / KS - KS / 2, xx = ax[v] + tap % KS - KS / 2;
                if (kv && amv[v] && (unsigned)yy < (unsigned)a.h && (unsigned)xx < (unsigned)a.w) {
                    if (ch < a.c1) {
                        const long p = a.up1 ? (((long)aimg[v] * hh + (yy >> 1)) * wh + (xx >> 1))
                                             : (((long)aimg[v] * a.h + yy) * a.w + xx);
                        av = *(const bf16x8 *)(a.in1 + p * a.c1 + ch);
                    } else {
                        const long p = ((long)aimg[v] * a.h + yy) * a.w + xx;
                        av = *(const bf16x8 *)(a.in2 + p * a.c2 + (ch - a.c1));
                    }
                }
                ga[c][v] = av;
            }
            bf16x8 bv = (bf16x8){};
            if (kv && bcv) bv = *(const bf16x8 *)(a.wt + ((long)bco * KS * KS + tap) * cin + cb + 8 * (tid & 1));
            gb[c] = bv;
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int c = 0; c < KST; ++c) {
#pragma unroll
            for (int v = 0; v < AV; ++v) sA[buf][c][tid + 256 * v] = ga[c][v];
            if (ldB) sB[buf][c][tid] = gb[c];
        }
    };
    f32x16 acc[WM][WN];
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) acc[i][j] = (f32x16){};
    const int nst = (nchunk + KST - 1) / KST;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1;
        if (st + 1 < nst) gload(st + 1);             // in flight during this stage's MFMAs
#pragma unroll
        for (int c = 0; c < KST; ++c) {
            if (st * KST + c >= nchunk) break;        // uniform
            bf16x8 av[WM], bv[WN];
#pragma unroll
            for (int i = 0; i < WM; ++i) av[i] = sA[buf][c][((wm * WM + i) * 32 + (l & 31)) * 2 + (l >> 5)];
#pragma unroll
            for (int j = 0; j < WN; ++j) bv[j] = sB[buf][c][((wn * WN + j) * 32 + (l & 31)) * 2 + (l >> 5)];
#pragma unroll
            for (int i = 0; i < WM; ++i)
#pragma unroll
                for (int j = 0; j < WN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        if (st + 1 < nst) lstore(buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < WN; ++j) {
        const int co = n0 + (wn * WN + j) * 32 + (l & 31);
        if (co >= a.cout) continue;
        const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < WM; ++i) {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const long p = m0 + (wm * WM + i) * 32 + 8 * (q >> 2) + 4 * (l >> 5) + (q & 3);
                if (p >= M) continue;
                float v = acc[i][j][q] + b;
                if (a.act == 1) v = v > 0.f ? v : 0.f;
                else if (a.act == 2) v = 1.f / (1.f + __expf(-v));
                if (a.outf) a.outf[p * a.cout + co] = v;
                if (co < a.o1) {
                    if (a.out1) a.out1[p * a.o1 + co] = f2bf(v);
                } else if (a.out2) {
                    a.out2[p * (a.cout - a.o1) + (co - a.o1)] = f2bf(v);
                }
            }
        }
    }
}

// Direct 3x3 convolution, input stationary: a block owns an 8 x 32 pixel tile of one image and
// BN = 32 * WN output channels. Per 16-channel input chunk it stages the 10 x 34 input halo tile
// and the chunk's weights of all 9 taps in LDS (each input cell is fetched once instead of once
// per tap), then every wave (two image rows = 64 pixels) issues 9 x 2 x WN MFMAs from LDS.
template <int WN>
__global__ void __launch_bounds__(256) k_conv_d(ConvArgs a) {
    constexpr int TH = 8, TW = 32, HH = TH + 2, HW2 = TW + 2, BN = 32 * WN;
    __shared__ bf16x8 sX[HH * HW2 * 2];         // [halo cell][k half]
    __shared__ bf16x8 sW[9 * BN * 2];           // [tap][co][k half]
    const int tid = threadIdx.x, l = tid & 63, wave = tid >> 6;
    const int tiles_x = (a.w + TW - 1) / TW, tiles_y = (a.h + TH - 1) / TH;
    const int t = blockIdx.x;
    const int img = t / (tiles_x * tiles_y);
    const int y0 = ((t / tiles_x) % tiles_y) * TH, x0 = (t % tiles_x) * TW;
    const int n0 = blockIdx.y * BN;
    const int cin = a.c1 + a.c2;
    const int hh = a.h >> 1, wh = a.w >> 1;
    f32x16 acc[2][WN];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) acc[i][j] = (f32x16){};
    for (int cb = 0; cb < cin; cb += 16) {
        // halo tile: HH * HW2 cells x 2 halves = 680 16-byte entries
        for (int e = tid; e < HH * HW2 * 2; e += 256) {
            const int cell = e >> 1, half = e & 1;
            const int yy = y0 - 1 + cell / HW2, xx = x0 - 1 + cell % HW2;
            bf16x8 v = (bf16x8){};
            if ((unsigned)yy < (unsigned)a.h && (unsigned)xx < (unsigned)a.w) {
                const int ch = cb + 8 * half;
                if (ch < a.c1) {
                    const long p = a.up1 ? (((long)img * hh + (yy >> 1)) * wh + (xx >> 1))
                                         : (((long)img * a.h + yy) * a.w + xx);
                    v = *(const bf16x8 *)(a.in1 + p * a.c1 + ch);
                } else {
                    const long p = ((long)img * a.h + yy) * a.w + xx;
                    v = *(const bf16x8 *)(a.in2 + p * a.c2 + (ch - a.c1));
                }
            }
            sX[e] = v;
        }
        // weights: 9 taps x BN co x 2 halves
        for (int e = tid; e < 9 * BN * 2; e += 256) {
            const int half = e & 1, co_l = (e >> 1) % BN, tap = (e >> 1) / BN;
            const int co = n0 + co_l;
            bf16x8 v = (bf16x8){};
            if (co < a.cout) v = *(const bf16x8 *)(a.wt + ((long)co * 9 + tap) * cin + cb + 8 * half);
            sW[e] = v;
        }
        __syncthreads();
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;
            bf16x8 av[2], bv[WN];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int ry = wave * 2 + i;          // tile row of MFMA row tile i
                av[i] = sX[((ry + dy) * HW2 + (l & 31) + dx) * 2 + (l >> 5)];
            }
#pragma unroll
            for (int j = 0; j < WN; ++j) bv[j] = sW[(tap * BN + j * 32 + (l & 31)) * 2 + (l >> 5)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < WN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < WN; ++j) {
        const int co = n0 + j * 32 + (l & 31);
        if (co >= a.cout) continue;
        const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int yy = y0 + wave * 2 + i;
            if (yy >= a.h) continue;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int xx = x0 + 8 * (q >> 2) + 4 * (l >> 5) + (q & 3);
                if (xx >= a.w) continue;
                const long p = ((long)img * a.h + yy) * a.w + xx;
                float v = acc[i][j][q] + b;
                if (a.act == 1) v = v > 0.f ? v : 0.f;
                else if (a.act == 2) v = 1.f / (1.f + __expf(-v));
                if (a.outf) a.outf[p * a.cout + co] = v;
                if (co < a.o1) {
                    if (a.out1) a.out1[p * a.o1 + co] = f2bf(v);
                } else if (a.out2) {
                    a.out2[p * (a.cout - a.o1) + (co - a.o1)] = f2bf(v);
                }
            }
        }
    }
}

// weight gradient partials: ws[chunk][co][tap][ci] = sum over the chunk's pixels of
// dz[p][co] * x[p + tap offset][ci]; one wave per (co tile, ci tile, tap)
struct WgradArgs {
    const __bf16 *dz;      // [pixel][cout]
    const __bf16 *in1;     // as in the forward pass
    const __bf16 *in2;
    float *ws;             // [nchunk][cout][ks * ks][cin]
    int c1, c2, up1, cout, n, h, w, ks, chunk, nchunk;
};

template <int KS>
__global__ void __launch_bounds__(256) k_conv_wgrad(WgradArgs a) {
    const int l = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cin = a.c1 + a.c2;
    const int co_t = (a.cout + 31) / 32, ci_t = (cin + 31) / 32;
    const int item = blockIdx.y * 4 + wave;
    if (item >= co_t * ci_t * KS * KS) return;
    const int tap = item % (KS * KS);
    const int cit = (item / (KS * KS)) % ci_t;
    const int cot = item / (KS * KS * ci_t);
    const int chunk = blockIdx.x;
    const long M = (long)a.n * a.h * a.w;
    const long p0 = (long)chunk * a.chunk;
    const long p1 = p0 + a.chunk < M ? p0 + a.chunk : M;
    const int co = cot * 32 + (l & 31);
    const int ci = cit * 32 + (l & 31);
    const int half = l >> 5;
    const int dy = tap / KS - KS / 2, dx = tap % KS - KS / 2;
    const int hh = a.h >> 1, wh = a.w >> 1;
    f32x16 acc = (f32x16){};
    // pixel coordinates of this lane's first pixel, advanced by 16 per step without divisions
    long pq = p0 + 8 * half;
    int px = (int)(pq % a.w), py = (int)((pq / a.w) % a.h), pimg = (int)(pq / ((long)a.w * a.h));
    for (long pb = p0; pb < p1; pb += 16) {
        bf16x8 av, bv;
        int x = px, y = py, img = pimg;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const long p = pb + 8 * half + j;
            float dzv = 0.f, xv = 0.f;
            if (p < p1) {
                if (co < a.cout) dzv = bf2f(a.dz[p * a.cout + co]);
                const int yy = y + dy, xx = x + dx;
                if (ci < cin && (unsigned)yy < (unsigned)a.h && (unsigned)xx < (unsigned)a.w) {
                    if (ci < a.c1) {
                        const long q = a.up1 ? (((long)img * hh + (yy >> 1)) * wh + (xx >> 1))
                                             : (((long)img * a.h + yy) * a.w + xx);
                        xv = bf2f(a.in1[q * a.c1 + ci]);
                    } else {
                        const long q = ((long)img * a.h + yy) * a.w + xx;
                        xv = bf2f(a.in2[q * a.c2 + (ci - a.c1)]);
                    }
                }
            }
            av[j] = f2bf(dzv);
            bv[j] = f2bf(xv);
            if (++x == a.w) { x = 0; if (++y == a.h) { y = 0; ++img; } }
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
        // advance the lane's first pixel by 16
        for (int k = 0; k < 16; ++k)
            if (++px == a.w) { px = 0; if (++py == a.h) { py = 0; ++pimg; } }
    }
    if (ci >= cin) return;
    float *o = a.ws + (long)chunk * a.cout * KS * KS * cin;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int r = cot * 32 + 8 * (i >> 2) + 4 * half + (i & 3);
        if (r < a.cout) o[((long)r * KS * KS + tap) * cin + ci] = acc[i];
    }
}

// bias gradient partials: ws[chunk][co] = sum over the chunk's pixels of dz[p][co]; the block's
// threads split the pixels (lane group g sums pixels p0 + g, p0 + g + groups, ...) and combine in LDS
__global__ void __launch_bounds__(256) k_bias_partial(const __bf16 *__restrict__ dz, float *__restrict__ ws, long M,
                                                       int cout, int chunk) {
    __shared__ float red[256];
    const int co = blockIdx.y * 32 + (threadIdx.x & 31);
    const int g = threadIdx.x >> 5;              // 8 pixel groups
    const long p0 = (long)blockIdx.x * chunk;
    const long p1 = p0 + chunk < M ? p0 + chunk : M;
    float s = 0.f;
    if (co < cout)
        for (long p = p0 + g; p < p1; p += 8) s += bf2f(dz[p * cout + co]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (g == 0 && co < cout) {
        float t = 0.f;
        for (int k = 0; k < 8; ++k) t += red[k * 32 + (threadIdx.x & 31)];
        ws[(long)blockIdx.x * cout + co] = t;
    }
}

// out[i] = sum over chunks of ws[chunk][i], in chunk order
__global__ void __launch_bounds__(256) k_reduce_chunks(const float *__restrict__ ws, float *__restrict__ out, long n,
                                                        int nchunk) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += ws[(long)c * n + i];
    out[i] = s;
}

// w_t[ci][kk - 1 - tap][co] = w[co][tap][ci]: the data gradient of a convolution is a
// convolution of dz with the spatially flipped, transposed kernel
__global__ void __launch_bounds__(256) k_flip(const __bf16 *__restrict__ w, __bf16 *__restrict__ wt, int cout, int cin,
                                               int kk) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)cout * kk * cin;
    if (i >= n) return;
    const int ci = (int)(i % cin);
    const int tap = (int)((i / cin) % kk);
    const int co = (int)(i / ((long)cin * kk));
    wt[((long)ci * kk + (kk - 1 - tap)) * cout + co] = w[i];
}

__global__ void __launch_bounds__(256) k_pool2(const __bf16 *__restrict__ in, __bf16 *__restrict__ out, int n, int h,
                                                int w, int c) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ho = h >> 1, wo = w >> 1;
    const long total = (long)n * ho * wo * c;
    if (i >= total) return;
    const int ch = (int)(i % c);
    const long p = i / c;
    const int xo = (int)(p % wo), yo = (int)((p / wo) % ho), img = (int)(p / ((long)wo * ho));
    const long b = (((long)img * h + 2 * yo) * w + 2 * xo) * c + ch;
    const float s = (bf2f(in[b]) + bf2f(in[b + c])) + (bf2f(in[b + (long)w * c]) + bf2f(in[b + (long)w * c + c]));
    out[i] = f2bf(0.25f * s);
}

// din = add + 0.25 * dout broadcast over the 2x2 window (add may be null)
__global__ void __launch_bounds__(256) k_pool2_bwd(const __bf16 *__restrict__ dout, const __bf16 *__restrict__ add,
                                                    __bf16 *__restrict__ din, int n, int h, int w, int c) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)n * h * w * c;
    if (i >= total) return;
    const int ch = (int)(i % c);
    const long p = i / c;
    const int x = (int)(p % w), y = (int)((p / w) % h), img = (int)(p / ((long)w * h));
    const long q = (((long)img * (h >> 1) + (y >> 1)) * (w >> 1) + (x >> 1)) * c + ch;
    float v = 0.25f * bf2f(dout[q]);
    if (add) v += bf2f(add[i]);
    din[i] = f2bf(v);
}

// dhalf[y][x] = sum of dfull over the 2x2 block it was copied to
__global__ void __launch_bounds__(256) k_up2_bwd(const __bf16 *__restrict__ dfull, __bf16 *__restrict__ dhalf, int n,
                                                  int h, int w, int c) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ho = h >> 1, wo = w >> 1;
    const long total = (long)n * ho * wo * c;
    if (i >= total) return;
    const int ch = (int)(i % c);
    const long p = i / c;
    const int xo = (int)(p % wo), yo = (int)((p / wo) % ho), img = (int)(p / ((long)wo * ho));
    const long b = (((long)img * h + 2 * yo) * w + 2 * xo) * c + ch;
    const float s = (bf2f(dfull[b]) + bf2f(dfull[b + c])) + (bf2f(dfull[b + (long)w * c]) +
                                                             bf2f(dfull[b + (long)w * c + c]));
    dhalf[i] = f2bf(s);
}

// dz = dy where y > 0 (y is the ReLU output)
__global__ void __launch_bounds__(256) k_relu_bwd(const __bf16 *__restrict__ dy, const __bf16 *__restrict__ y,
                                                   __bf16 *__restrict__ dz, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dz[i] = bf2f(y[i]) > 0.f ? dy[i] : f2bf(0.f);
}

// sigmoid-MSE head: out holds sigmoid(z) [pixel][cs] (f32); labels [img][core][core][nl] f32 for
// the core window [c0, c0 + core) of every image. dz = d(mean squared error)/dz as bf16 [pixel][cs]
// (zero outside the core and for channels >= nl); per-block partial losses in part[block].
__global__ void __launch_bounds__(256) k_head_loss(const float *__restrict__ out, const float *__restrict__ lab,
                                                    __bf16 *__restrict__ dz, float *__restrict__ part, int n, int h,
                                                    int w, int cs, int nl, int c0, int core, float inv_count) {
    __shared__ float red[256];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)n * h * w * cs;
    float e2 = 0.f;
    if (i < total) {
        const int ch = (int)(i % cs);
        const long p = i / cs;
        const int x = (int)(p % w), y = (int)((p / w) % h), img = (int)(p / ((long)w * h));
        const int cy = y - c0, cx = x - c0;
        float g = 0.f;
        if (ch < nl && (unsigned)cy < (unsigned)core && (unsigned)cx < (unsigned)core) {
            const float s = out[i];
            const float d = s - lab[(((long)img * core + cy) * core + cx) * nl + ch];
            e2 = d * d;
            g = 2.f * d * inv_count * s * (1.f - s);
        }
        dz[i] = f2bf(g);
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] * inv_count;
}

// Adam on f32 master weights; writes the bf16 copy the convolutions read
__global__ void __launch_bounds__(256) k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                                               const float *__restrict__ g, __bf16 *__restrict__ pb, long n, float lr,
                                               float b1, float b2, float eps, float c1, float c2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (pb) pb[i] = f2bf(pi);
}

// network input: the composed march heights {a, vegdsm, vegdsm2} of a site (float4 per cell) at
// window origins org[img] = (row, col) -> bf16 [img][hw][hw][16], channels 0-2 = height / 50 m
// (0 outside the site)
__global__ void __launch_bounds__(256) k_pack_input(const float4 *__restrict__ hp, const int2 *__restrict__ org,
                                                     __bf16 *__restrict__ x, int n, int hw, int rows, int cols) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)n * hw * hw;
    if (p >= total) return;
    const int xx = (int)(p % hw), yy = (int)((p / hw) % hw), img = (int)(p / ((long)hw * hw));
    const int r = org[img].x + yy, c = org[img].y + xx;
    float f0 = 0.f, f1 = 0.f, f2 = 0.f;
    if ((unsigned)r < (unsigned)rows && (unsigned)c < (unsigned)cols) {
        const float4 h = hp[(long)r * cols + c];
        f0 = h.x * 0.02f;
        f1 = h.y * 0.02f;
        f2 = h.z * 0.02f;
    }
    bf16x8 lo = (bf16x8){}, hi = (bf16x8){};
    lo[0] = f2bf(f0);
    lo[1] = f2bf(f1);
    lo[2] = f2bf(f2);
    *(bf16x8 *)(x + p * 16) = lo;
    *(bf16x8 *)(x + p * 16 + 8) = hi;
}

// training labels: fields [16][rows][cols] (SVF field planes) at the cores org + c0 -> [img][core][core][nl]
// with field index map fm[0..nl)
__global__ void __launch_bounds__(256) k_gather_labels(const float *__restrict__ fields, const int2 *__restrict__ org,
                                                        const int32_t *__restrict__ fm, float *__restrict__ lab, int n,
                                                        int core, int c0, int nl, int rows, int cols) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)n * core * core * nl;
    if (i >= total) return;
    const int ch = (int)(i % nl);
    const long p = i / nl;
    const int xx = (int)(p % core), yy = (int)((p / core) % core), img = (int)(p / ((long)core * core));
    const int r = org[img].x + c0 + yy, c = org[img].y + c0 + xx;
    float v = 0.f;
    if ((unsigned)r < (unsigned)rows && (unsigned)c < (unsigned)cols)
        v = fields[(long)fm[ch] * rows * cols + (long)r * cols + c];
    lab[i] = v;
}

// surrogate output cores back into field planes: out [img][hw][hw][cs] (f32) cores
// [c0, c0 + core) -> fields[fm[ch]] at org + c0 (inside the site only)
__global__ void __launch_bounds__(256) k_scatter_cores(const float *__restrict__ out, const int2 *__restrict__ org,
                                                        const int32_t *__restrict__ fm, float *__restrict__ fields,
                                                        int n, int hw, int cs, int core, int c0, int nl, int rows,
                                                        int cols) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)n * core * core * nl;
    if (i >= total) return;
    const int ch = (int)(i % nl);
    const long p = i / nl;
    const int xx = (int)(p % core), yy = (int)((p / core) % core), img = (int)(p / ((long)core * core));
    const int r = org[img].x + c0 + yy, c = org[img].y + c0 + xx;
    if ((unsigned)r >= (unsigned)rows || (unsigned)c >= (unsigned)cols) return;
    fields[(long)fm[ch] * rows * cols + (long)r * cols + c] =
        out[(((long)img * hw + c0 + yy) * hw + c0 + xx) * cs + ch];
}

inline int blocks(long n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------------------
// op-list executor (sw_cnn_fwd / sw_cnn_bwd)
// ---------------------------------------------------------------------------
static int launch_conv(const sw_cnn_op_t &o, const __bf16 *w, void *stream) {
    ConvArgs a;
    a.in1 = (const __bf16 *)o.in0;
    a.in2 = (const __bf16 *)o.in1;
    a.wt = w;
    a.bias = (const float *)o.bias;
    a.out1 = (__bf16 *)o.out0;
    a.out2 = (__bf16 *)o.out1;
    a.outf = (float *)o.outf;
    a.c1 = o.c1; a.c2 = o.c2; a.up1 = o.up; a.o1 = o.o1; a.cout = o.cout;
    a.n = o.n; a.h = o.h; a.w = o.w; a.act = o.act;
    const long M = (long)o.n * o.h * o.w;
    if (o.ks == 3) {
        // direct convolution: 8 x 32 pixel tiles x 32 * WN channels
        const int wn = o.cout > 32 ? 2 : 1;
        const long tiles = (long)o.n * ((o.h + 7) / 8) * ((o.w + 31) / 32);
        const dim3 g((unsigned)tiles, (unsigned)blocks(o.cout, 32 * wn));
        if (wn == 2) hipLaunchKernelGGL((k_conv_d<2>), g, dim3(256), 0, S(stream), a);
        else hipLaunchKernelGGL((k_conv_d<1>), g, dim3(256), 0, S(stream), a);
        return hip_check(hipGetLastError(), "k_conv_d");
    }
    // block tiles: 128 pixels x 128 channels, 256 x 64 or 256 x 32 (narrow layers), 4 K chunks per stage
    const int bn = o.cout > 64 ? 128 : (o.cout > 32 ? 64 : 32);
    const int bm = bn == 128 ? 128 : 256;
    const dim3 grid((unsigned)blocks(M, bm), (unsigned)blocks(o.cout, bn));
    if (o.ks == 3) {
        if (bn == 128) hipLaunchKernelGGL((k_conv_t<3, 2, 2, 2, 4>), grid, dim3(256), 0, S(stream), a);
        else if (bn == 64) hipLaunchKernelGGL((k_conv_t<3, 4, 2, 2, 4>), grid, dim3(256), 0, S(stream), a);
        else hipLaunchKernelGGL((k_conv_t<3, 4, 2, 1, 4>), grid, dim3(256), 0, S(stream), a);
    } else {
        if (bn == 128) hipLaunchKernelGGL((k_conv_t<1, 2, 2, 2, 4>), grid, dim3(256), 0, S(stream), a);
        else if (bn == 64) hipLaunchKernelGGL((k_conv_t<1, 4, 2, 2, 4>), grid, dim3(256), 0, S(stream), a);
        else hipLaunchKernelGGL((k_conv_t<1, 4, 2, 1, 4>), grid, dim3(256), 0, S(stream), a);
    }
    return hip_check(hipGetLastError(), "k_conv_t");
}

static int check_conv(const sw_cnn_op_t &o, const char *who) {
    if (o.ks != 1 && o.ks != 3) return set_err(SW_E_ARG, who);
    if (o.c1 <= 0 || o.c1 % 16 || o.c2 < 0 || o.c2 % 16 || o.cout <= 0 || o.n <= 0 || o.h <= 0 || o.w <= 0)
        return set_err(SW_E_ARG, who);
    if (o.up && (o.h % 2 || o.w % 2)) return set_err(SW_E_ARG, who);
    if (!o.in0 || (o.c2 && !o.in1) || !o.weight) return set_err(SW_E_ARG, who);
    return SW_OK;
}

static int cnn_run(const sw_cnn_op_t *ops, int nops, void *stream) {
    if (!ops || nops < 0) return set_err(SW_E_ARG, "sw_cnn: NULL op list");
    for (int k = 0; k < nops; ++k) {
        const sw_cnn_op_t &o = ops[k];
        int rc = SW_OK;
        const long pix = (long)o.n * o.h * o.w;
        switch (o.kind) {
        case SW_CNN_CONV:
            if ((rc = check_conv(o, "sw_cnn: conv: bad shape (channels multiple of 16, ks 1 or 3)"))) return rc;
            if (!o.out0 && !o.outf) return set_err(SW_E_ARG, "sw_cnn: conv without output");
            rc = launch_conv(o, (const __bf16 *)o.weight, stream);
            break;
        case SW_CNN_CONV_BWD_DATA: {
            // dz (c1 = cout of the forward layer) convolved with the flipped kernel into dx (cout = its cin)
            if ((rc = check_conv(o, "sw_cnn: conv data gradient: bad shape"))) return rc;
            if (!o.ws) return set_err(SW_E_ARG, "sw_cnn: conv data gradient needs a workspace for the flipped kernel");
            const long nw = (long)o.cout * o.ks * o.ks * o.c1;
            hipLaunchKernelGGL(k_flip, dim3(blocks(nw, 256)), dim3(256), 0, S(stream), (const __bf16 *)o.weight,
                               (__bf16 *)o.ws, o.c1, o.cout, o.ks * o.ks);
            if ((rc = hip_check(hipGetLastError(), "k_flip"))) return rc;
            rc = launch_conv(o, (const __bf16 *)o.ws, stream);
            break;
        }
        case SW_CNN_CONV_BWD_WEIGHT: {
            // in0/in1/c1/c2/up: the forward inputs; dz: the output gradient; gw/gb: gradients (overwritten)
            if ((rc = check_conv(o, "sw_cnn: conv weight gradient: bad shape"))) return rc;
            if (!o.dz || !o.gw || !o.ws || o.nchunk <= 0) return set_err(SW_E_ARG, "sw_cnn: weight gradient arguments");
            const int cin = o.c1 + o.c2;
            const long per = (long)o.cout * o.ks * o.ks * cin;
            const int chunk = (int)(((pix + o.nchunk - 1) / o.nchunk + 15) / 16 * 16);
            const int nchunk = (int)((pix + chunk - 1) / chunk);
            WgradArgs a;
            a.dz = (const __bf16 *)o.dz; a.in1 = (const __bf16 *)o.in0; a.in2 = (const __bf16 *)o.in1;
            a.ws = (float *)o.ws; a.c1 = o.c1; a.c2 = o.c2; a.up1 = o.up; a.cout = o.cout;
            a.n = o.n; a.h = o.h; a.w = o.w; a.ks = o.ks; a.chunk = chunk; a.nchunk = nchunk;
            const int items = ((o.cout + 31) / 32) * ((cin + 31) / 32) * o.ks * o.ks;
            const dim3 grid((unsigned)nchunk, (unsigned)blocks(items, 4));
            if (o.ks == 3) hipLaunchKernelGGL((k_conv_wgrad<3>), grid, dim3(256), 0, S(stream), a);
            else hipLaunchKernelGGL((k_conv_wgrad<1>), grid, dim3(256), 0, S(stream), a);
            if ((rc = hip_check(hipGetLastError(), "k_conv_wgrad"))) return rc;
            hipLaunchKernelGGL(k_reduce_chunks, dim3(blocks(per, 256)), dim3(256), 0, S(stream), (const float *)o.ws,
                               (float *)o.gw, per, nchunk);
            if ((rc = hip_check(hipGetLastError(), "k_reduce_chunks"))) return rc;
            if (o.gb) {
                float *wsb = (float *)o.ws + (long)nchunk * per;
                hipLaunchKernelGGL(k_bias_partial, dim3((unsigned)nchunk, (unsigned)blocks(o.cout, 32)), dim3(256),
                                   0, S(stream), (const __bf16 *)o.dz, wsb, pix, o.cout, chunk);
                if ((rc = hip_check(hipGetLastError(), "k_bias_partial"))) return rc;
                hipLaunchKernelGGL(k_reduce_chunks, dim3(blocks(o.cout, 256)), dim3(256), 0, S(stream), wsb,
                                   (float *)o.gb, (long)o.cout, nchunk);
                rc = hip_check(hipGetLastError(), "k_reduce_chunks");
            }
            break;
        }
        case SW_CNN_POOL2:
            if (o.h % 2 || o.w % 2 || !o.in0 || !o.out0) return set_err(SW_E_ARG, "sw_cnn: pool2 arguments");
            hipLaunchKernelGGL(k_pool2, dim3(blocks(pix / 4 * o.c1, 256)), dim3(256), 0, S(stream),
                               (const __bf16 *)o.in0, (__bf16 *)o.out0, o.n, o.h, o.w, o.c1);
            rc = hip_check(hipGetLastError(), "k_pool2");
            break;
        case SW_CNN_POOL2_BWD:
            if (o.h % 2 || o.w % 2 || !o.in0 || !o.out0) return set_err(SW_E_ARG, "sw_cnn: pool2 bwd arguments");
            hipLaunchKernelGGL(k_pool2_bwd, dim3(blocks(pix * o.c1, 256)), dim3(256), 0, S(stream),
                               (const __bf16 *)o.in0, (const __bf16 *)o.in1, (__bf16 *)o.out0, o.n, o.h, o.w, o.c1);
            rc = hip_check(hipGetLastError(), "k_pool2_bwd");
            break;
        case SW_CNN_UP2_BWD:
            if (o.h % 2 || o.w % 2 || !o.in0 || !o.out0) return set_err(SW_E_ARG, "sw_cnn: up2 bwd arguments");
            hipLaunchKernelGGL(k_up2_bwd, dim3(blocks(pix / 4 * o.c1, 256)), dim3(256), 0, S(stream),
                               (const __bf16 *)o.in0, (__bf16 *)o.out0, o.n, o.h, o.w, o.c1);
            rc = hip_check(hipGetLastError(), "k_up2_bwd");
            break;
        case SW_CNN_RELU_BWD:
            if (!o.in0 || !o.in1 || !o.out0) return set_err(SW_E_ARG, "sw_cnn: relu bwd arguments");
            hipLaunchKernelGGL(k_relu_bwd, dim3(blocks(pix * o.c1, 256)), dim3(256), 0, S(stream),
                               (const __bf16 *)o.in0, (const __bf16 *)o.in1, (__bf16 *)o.out0, pix * o.c1);
            rc = hip_check(hipGetLastError(), "k_relu_bwd");
            break;
        case SW_CNN_HEAD_LOSS: {
            // outf: sigmoid outputs [pix][c1]; in0: labels (f32); out0: dz (bf16); ws: per-block losses
            if (!o.outf || !o.in0 || !o.out0 || !o.ws || o.core <= 0 || o.nl <= 0 || o.nl > o.c1)
                return set_err(SW_E_ARG, "sw_cnn: head loss arguments");
            const long total = pix * o.c1;
            const float inv = 1.f / ((float)o.n * o.core * o.core * o.nl);
            hipLaunchKernelGGL(k_head_loss, dim3(blocks(total, 256)), dim3(256), 0, S(stream), (const float *)o.outf,
                               (const float *)o.in0, (__bf16 *)o.out0, (float *)o.ws, o.n, o.h, o.w, o.c1, o.nl,
                               o.c0, o.core, inv);
            rc = hip_check(hipGetLastError(), "k_head_loss");
            break;
        }
        default:
            return set_err(SW_E_ARG, "sw_cnn: unknown op kind");
        }
        if (rc) return rc;
    }
    return SW_OK;
}

extern "C" int sw_cnn_fwd(const sw_cnn_op_t *ops, int nops, void *stream) { return cnn_run(ops, nops, stream); }
extern "C" int sw_cnn_bwd(const sw_cnn_op_t *ops, int nops, void *stream) { return cnn_run(ops, nops, stream); }

extern "C" int sw_cnn_bwd_scratch_bytes(const sw_cnn_op_t *op, int64_t *bytes) {
    if (!op || !bytes) return set_err(SW_E_ARG, "sw_cnn_bwd_scratch_bytes: NULL pointer");
    const long pix = (long)op->n * op->h * op->w;
    const int cin = op->c1 + op->c2;
    if (op->kind == SW_CNN_CONV_BWD_WEIGHT) {
        const long per = (long)op->cout * op->ks * op->ks * cin;
        const int chunk = (int)(((pix + op->nchunk - 1) / op->nchunk + 15) / 16 * 16);
        const long nchunk = (pix + chunk - 1) / chunk;
        *bytes = (int64_t)(nchunk * (per + op->cout) * 4);
    } else if (op->kind == SW_CNN_CONV_BWD_DATA) {
        *bytes = (int64_t)op->cout * op->ks * op->ks * op->c1 * 2;
    } else {
        *bytes = 0;
    }
    return SW_OK;
}

extern "C" int sw_adam(float *p, float *m, float *v, const float *g, uint16_t *p_bf16, int64_t n, float lr, float b1,
                       float b2, float eps, int step, void *stream) {
    if (!p || !m || !v || !g || n < 0 || step < 1) return set_err(SW_E_ARG, "sw_adam: arguments");
    if (n == 0) return SW_OK;
    float c1 = 1.f, c2 = 1.f;
    for (int i = 0; i < step; ++i) { c1 *= b1; c2 *= b2; }
    hipLaunchKernelGGL(k_adam, dim3(blocks(n, 256)), dim3(256), 0, S(stream), p, m, v, g, (__bf16 *)p_bf16, (long)n, lr,
                       b1, b2, eps, 1.f - c1, 1.f - c2);
    return hip_check(hipGetLastError(), "k_adam");
}

extern "C" int sw_cnn_pack_input(const void *hpack, const int32_t *origins, uint16_t *x, int n, int hw, int rows,
                                 int cols, void *stream) {
    if (!hpack || !origins || !x || n <= 0 || hw <= 0) return set_err(SW_E_ARG, "sw_cnn_pack_input: arguments");
    const long total = (long)n * hw * hw;
    hipLaunchKernelGGL(k_pack_input, dim3(blocks(total, 256)), dim3(256), 0, S(stream), (const float4 *)hpack,
                       (const int2 *)origins, (__bf16 *)x, n, hw, rows, cols);
    return hip_check(hipGetLastError(), "k_pack_input");
}

extern "C" int sw_cnn_gather_labels(const float *fields, const int32_t *origins, const int32_t *field_map,
                                    float *labels, int n, int core, int c0, int nl, int rows, int cols, void *stream) {
    if (!fields || !origins || !field_map || !labels || n <= 0 || core <= 0 || nl <= 0)
        return set_err(SW_E_ARG, "sw_cnn_gather_labels: arguments");
    const long total = (long)n * core * core * nl;
    hipLaunchKernelGGL(k_gather_labels, dim3(blocks(total, 256)), dim3(256), 0, S(stream), fields,
                       (const int2 *)origins, field_map, labels, n, core, c0, nl, rows, cols);
    return hip_check(hipGetLastError(), "k_gather_labels");
}

extern "C" int sw_cnn_scatter_cores(const float *out, const int32_t *origins, const int32_t *field_map, float *fields,
                                    int n, int hw, int cs, int core, int c0, int nl, int rows, int cols,
                                    void *stream) {
    if (!out || !origins || !field_map || !fields || n <= 0 || core <= 0 || nl <= 0 || nl > cs)
        return set_err(SW_E_ARG, "sw_cnn_scatter_cores: arguments");
    const long total = (long)n * core * core * nl;
    hipLaunchKernelGGL(k_scatter_cores, dim3(blocks(total, 256)), dim3(256), 0, S(stream), out,
                       (const int2 *)origins, field_map, fields, n, hw, cs, core, c0, nl, rows, cols);
    return hip_check(hipGetLastError(), "k_scatter_cores");
}
