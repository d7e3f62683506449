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
    int r0, rn;            // output region [r0, r0 + rn)^2 of every image (k_conv_n16 only); rn 0: all
    const float4 *site;    // k_conv_n16: input = window org[img] of the site's composed heights (c1 == 16)
    const int2 *org;
    int srows, scols;
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

// Direct 3x3 convolution with 16 output channels (the full-resolution layers and the data
// gradients into 16-channel tensors), output stationary on v_mfma_f32_16x16x32_bf16 with the
// weights as the row operand: C[co][pixel], so a lane ends with 4 consecutive channels of one
// pixel (8-byte stores, 512 contiguous bytes per wave store). K pairs two taps of one 16-channel
// chunk (lane groups 0-1: tap t, groups 2-3: tap t + 1; the ninth tap pairs with zeros). A block
// owns an 8 x 32 pixel tile; each wave two rows as four 16-pixel column tiles.
//   pool (POOL): also writes the 2x2 average of the stored bf16 outputs to pool_out, in the
//   summation order of k_pool2 (bit-identical to a separate pool op); needs even h, w.
//   head (HEAD): runs the 1x1 sigmoid head (16 -> 16 channels) on the bf16-rounded outputs with
//   v_mfma_f32_16x16x16bf16_1k, whose B operand layout (k = 4 * (lane / 16) + j) is this
//   kernel's C layout, and writes the f32 sigmoid to outf; out1 may then be null.
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct Conv16Args {
    ConvArgs c;
    __bf16 *pool_out;        // (n, h/2, w/2, 16) or null
    const __bf16 *head_w;    // [16][16] (co2, ci) or null
    const float *head_b;     // [16] or null
};

template <int NT, bool POOL, bool HEAD>
__global__ void __launch_bounds__(256) k_conv_n16(Conv16Args ca) {
    constexpr int TH = 8, TW = 32, HH = TH + 2, HW2 = TW + 2;
    static_assert(!HEAD || NT == 1, "the head follows a 16-channel layer");
    const ConvArgs &a = ca.c;
    __shared__ bf16x8 sX[HH * HW2 * 2];         // [halo cell][k half]
    const int tid = threadIdx.x, l = tid & 63, wave = tid >> 6, g = l >> 4, r16 = l & 15;
    // output rows / columns [ro, ro + rh) x [ro, ro + rw): the image, or the op's square region
    const int ro = a.rn ? a.r0 : 0, rh = a.rn ? a.rn : a.h, rw = a.rn ? a.rn : a.w;
    const int tiles_x = (rw + TW - 1) / TW, tiles_y = (rh + TH - 1) / TH;
    const int t = blockIdx.x;
    const int img = t / (tiles_x * tiles_y);
    const int y0 = ro + ((t / tiles_x) % tiles_y) * TH, x0 = ro + (t % tiles_x) * TW;
    const int ylim = ro + rh, xlim = ro + rw;     // within the image (checked by the launcher)
    const int cin = a.c1 + a.c2;
    const int hh = a.h >> 1, wh = a.w >> 1;
    const int co0 = blockIdx.y * 16 * NT;         // first output channel of the block
    f32x4 acc[4][NT];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[i][j] = (f32x4){};
    for (int cb = 0; cb < cin; cb += 16) {
        for (int e = tid; e < HH * HW2 * 2; e += 256) {
            const int cell = e >> 1, half = e & 1;
            const int yy = y0 - 1 + cell / HW2, xx = x0 - 1 + cell % HW2;
            bf16x8 v = (bf16x8){};
            if ((unsigned)yy < (unsigned)a.h && (unsigned)xx < (unsigned)a.w) {
                const int ch = cb + 8 * half;
                if (a.site) {
                    // the window's composed heights / 50 m in channels 0-2 (k_pack_input's values)
                    const int r = a.org[img].x + yy, c = a.org[img].y + xx;
                    if (half == 0 && (unsigned)r < (unsigned)a.srows && (unsigned)c < (unsigned)a.scols) {
                        const float4 hq = a.site[(long)r * a.scols + c];
                        v[0] = f2bf(hq.x * 0.02f);
                        v[1] = f2bf(hq.y * 0.02f);
                        v[2] = f2bf(hq.z * 0.02f);
                    }
                } else if (ch < a.c1) {
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
        // weight fragments of the 5 tap pairs: lane (co = co0 + 16 j + r16, group g) holds taps 2p + (g >> 1)
        bf16x8 wf[5][NT];
#pragma unroll
        for (int pr = 0; pr < 5; ++pr) {
            const int tap = 2 * pr + (g >> 1);
#pragma unroll
            for (int j = 0; j < NT; ++j)
                wf[pr][j] = (tap < 9)
                    ? *(const bf16x8 *)(a.wt + ((long)(co0 + 16 * j + r16) * 9 + tap) * cin + cb + 8 * (g & 1))
                    : (bf16x8){};
        }
        __syncthreads();
#pragma unroll
        for (int pr = 0; pr < 5; ++pr) {
            const int tap = 2 * pr + (g >> 1);
            const int dy = tap < 9 ? tap / 3 : 0, dx = tap < 9 ? tap % 3 : 0;
#pragma unroll
            for (int pt = 0; pt < 4; ++pt) {
                const int ry = wave * 2 + (pt >> 1), cx = 16 * (pt & 1) + r16;
                const bf16x8 xv = sX[((ry + dy) * HW2 + cx + dx) * 2 + (g & 1)];
#pragma unroll
                for (int j = 0; j < NT; ++j)
                    acc[pt][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[pr][j], xv, acc[pt][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // epilogue: lane holds channels co0 + 16 j + 4 g + (0..3) of pixel (y0 + ry, x0 + cx)
    float bias[NT][4];
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) bias[j][i] = a.bias ? a.bias[co0 + 16 * j + 4 * g + i] : 0.f;
    float hb[4];
    s16x4 hw[1];
    if (HEAD) {
#pragma unroll
        for (int i = 0; i < 4; ++i) hb[i] = ca.head_b ? ca.head_b[4 * g + i] : 0.f;
        // head A operand: row co2 = r16, k = ci = 4 g + j
        hw[0] = *(const s16x4 *)(ca.head_w + r16 * 16 + 4 * g);
    }
    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
    float st[4][NT][4];     // stored (bf16-rounded) values, for the pool
#pragma unroll
    for (int pt = 0; pt < 4; ++pt) {
        const int ry = wave * 2 + (pt >> 1), cx = 16 * (pt & 1) + r16;
        const int yy = y0 + ry, xx = x0 + cx;
        const bool ok = yy < ylim && xx < xlim;
        const long p = ((long)img * a.h + yy) * a.w + xx;
        __bf16 ob[NT][4];
#pragma unroll
        for (int j = 0; j < NT; ++j) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = acc[pt][j][i] + bias[j][i];
                if (a.act == 1) v = v > 0.f ? v : 0.f;
                else if (a.act == 2) v = 1.f / (1.f + __expf(-v));
                ob[j][i] = f2bf(v);
                st[pt][j][i] = bf2f(ob[j][i]);
            }
            if (ok && a.out1)
                *(bf16x4 *)(a.out1 + p * a.cout + co0 + 16 * j + 4 * g) =
                    (bf16x4){ob[j][0], ob[j][1], ob[j][2], ob[j][3]};
        }
        if (HEAD) {
            s16x4 xb;
#pragma unroll
            for (int i = 0; i < 4; ++i) xb[i] = (short)__builtin_bit_cast(unsigned short, ob[0][i]);
            f32x4 h = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(hw[0], xb, (f32x4){}, 0, 0, 0);
            float4 o;
            o.x = 1.f / (1.f + __expf(-(h[0] + hb[0])));
            o.y = 1.f / (1.f + __expf(-(h[1] + hb[1])));
            o.z = 1.f / (1.f + __expf(-(h[2] + hb[2])));
            o.w = 1.f / (1.f + __expf(-(h[3] + hb[3])));
            if (ok) *(float4 *)(a.outf + p * 16 + 4 * g) = o;
        }
    }
    if (POOL) {
        // 2x2 windows: rows ry (pt 0/1) and ry + 1 (pt 2/3) of this lane, columns x and x + 1 (lane ^ 1)
#pragma unroll
        for (int pc = 0; pc < 2; ++pc) {
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                float nb_t[4], nb_b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    nb_t[i] = __shfl_xor(st[pc][j][i], 1);
                    nb_b[i] = __shfl_xor(st[pc + 2][j][i], 1);
                }
                const int cx = 16 * pc + r16;
                const int yo = (y0 + wave * 2) >> 1, xo = (x0 + cx) >> 1;
                if ((r16 & 1) == 0 && yo < (a.h >> 1) && xo < (a.w >> 1)) {
                    __bf16 pb[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        pb[i] = f2bf(0.25f * ((st[pc][j][i] + nb_t[i]) + (st[pc + 2][j][i] + nb_b[i])));
                    const long q = ((long)img * (a.h >> 1) + yo) * (a.w >> 1) + xo;
                    *(bf16x4 *)(ca.pool_out + q * a.cout + co0 + 16 * j + 4 * g) = (bf16x4){pb[0], pb[1], pb[2], pb[3]};
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

// gate: mean absolute difference between the network output and the exact fields over the sampled
// pixels of each window (pixels pix[seg[w] .. seg[w + 1]) of window w, per pixel the mean over the nl
// fields), one block per window with a fixed-order reduction (deterministic)
// per window over its sampled pixels: MAX = false, the mean over pixels of the mean absolute
// difference over the nl fields; MAX = true, the largest absolute difference of any field at any
// sampled pixel (fixed reduction order either way)
template <bool MAX>
__global__ void __launch_bounds__(256) k_gate_error(const float *__restrict__ out, const float *__restrict__ fields,
                                                     const int2 *__restrict__ org, const int32_t *__restrict__ pix,
                                                     const int32_t *__restrict__ seg, const int32_t *__restrict__ fm,
                                                     float *__restrict__ werr, int hw, int cs, int nl, int rows,
                                                     int cols) {
    __shared__ float red[256];
    const int w = blockIdx.x;
    const int s0 = seg[w], s1 = seg[w + 1];
    const long n = (long)rows * cols;
    float acc = 0.f;
    for (int i = s0 + (int)threadIdx.x; i < s1; i += 256) {
        const int p = pix[i];
        const int r = p / cols, c = p - r * cols;
        const long o = (((long)w * hw + (r - org[w].x)) * hw + (c - org[w].y)) * cs;
        float e = 0.f;
        for (int k = 0; k < nl; ++k) {
            const float d = fabsf(out[o + k] - fields[(long)fm[k] * n + p]);
            e = MAX ? fmaxf(e, d) : e + d;
        }
        acc = MAX ? fmaxf(acc, e) : acc + e / (float)nl;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            red[threadIdx.x] = MAX ? fmaxf(red[threadIdx.x], red[threadIdx.x + st]) : red[threadIdx.x] + red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) werr[w] = s1 > s0 ? (MAX ? red[0] : red[0] / (float)(s1 - s0)) : 0.f;
}

// gate: network output cores into the field planes except at sampled pixels (which keep the exact
// values), plus SVFtotal = svf - (1 - svfveg) * 0.97 from the written fields 0 and 5
__global__ void __launch_bounds__(256) k_gate_write(const float *__restrict__ out, const int2 *__restrict__ org,
                                                     const int32_t *__restrict__ fm, const uint8_t *__restrict__ sampled,
                                                     float *__restrict__ fields, int nwin, int hw, int cs, int core,
                                                     int c0, int nl, int rows, int cols) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)nwin * core * core;
    if (i >= total) return;
    const int xx = (int)(i % core), yy = (int)((i / core) % core), img = (int)(i / ((long)core * core));
    const int r = org[img].x + c0 + yy, c = org[img].y + c0 + xx;
    if ((unsigned)r >= (unsigned)rows || (unsigned)c >= (unsigned)cols) return;
    const long n = (long)rows * cols, p = (long)r * cols + c;
    if (sampled[p]) return;
    const float *o = out + (((long)img * hw + c0 + yy) * hw + c0 + xx) * cs;
    for (int k = 0; k < nl; ++k) fields[(long)fm[k] * n + p] = o[k];
    fields[15 * n + p] = o[0] - (1.f - o[5]) * 0.97f;
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
    a.r0 = o.kind == SW_CNN_CONV ? o.c0 : 0;
    a.rn = o.kind == SW_CNN_CONV ? o.core : 0;
    a.site = o.kind == SW_CNN_CONV ? (const float4 *)o.site : nullptr;
    a.org = (const int2 *)o.site_origins;
    a.srows = o.site_rows;
    a.scols = o.site_cols;
    const long M = (long)o.n * o.h * o.w;
    const bool fused = o.pool_out || o.head_w;
    const int th = a.rn ? a.rn : o.h, tw = a.rn ? a.rn : o.w;
    const long tiles8x32 = (long)o.n * ((th + 7) / 8) * ((tw + 31) / 32);
    // the channel-row kernel stages its input once per 32 output channels: it wins for 16 output
    // channels, for shallow inputs and for grids the 64-channel tiles of k_conv_d leave underfilled
    // (under 512 blocks); deep inputs with many output channels keep k_conv_d (measured per layer)
    const bool rows16 = o.cout == 16 || fused || a.site || o.c1 + o.c2 <= 32 ||
                        tiles8x32 * ((o.cout + 63) / 64) < 512;
    if (a.site && (o.c1 != 16 || o.c2 || o.up || !a.org || o.site_rows <= 0 || o.site_cols <= 0 || o.ks != 3 ||
                   o.cout % 16 || o.o1 < o.cout || o.outf))
        return set_err(SW_E_ARG, "sw_cnn: a site-input conv needs 16 input channels, origins, the site size and a 3x3 kernel");
    if (a.rn && (a.r0 < 0 || a.r0 + a.rn > o.h || a.r0 + a.rn > o.w || o.pool_out || o.ks != 3 ||
                 o.cout % 16 || o.o1 < o.cout || (o.outf && !o.head_w)))
        return set_err(SW_E_ARG, "sw_cnn: an output region needs a 3x3 conv with a whole output, no pool, inside the image");
    if (o.ks == 3 && (rows16 || a.rn) && o.cout % 16 == 0 && o.o1 >= o.cout && !o.outf == !o.head_w &&
        (!o.head_w || o.cout == 16)) {
        // channels as MFMA rows: 8 x 32 pixel tiles x 16 (cout 16) or 32 output channels per block on
        // 16x16x32 MFMAs, optional pool / head epilogues
        Conv16Args ca;
        ca.c = a;
        ca.pool_out = (__bf16 *)o.pool_out;
        ca.head_w = (const __bf16 *)o.head_w;
        ca.head_b = o.head_b;
        const int nt = o.cout % 32 == 0 ? 2 : 1;
        const dim3 g((unsigned)tiles8x32, (unsigned)(o.cout / (16 * nt)));
        if (nt == 2) {
            if (o.pool_out) hipLaunchKernelGGL((k_conv_n16<2, true, false>), g, dim3(256), 0, S(stream), ca);
            else hipLaunchKernelGGL((k_conv_n16<2, false, false>), g, dim3(256), 0, S(stream), ca);
        } else if (o.pool_out && o.head_w) {
            hipLaunchKernelGGL((k_conv_n16<1, true, true>), g, dim3(256), 0, S(stream), ca);
        } else if (o.pool_out) {
            hipLaunchKernelGGL((k_conv_n16<1, true, false>), g, dim3(256), 0, S(stream), ca);
        } else if (o.head_w) {
            hipLaunchKernelGGL((k_conv_n16<1, false, true>), g, dim3(256), 0, S(stream), ca);
        } else {
            hipLaunchKernelGGL((k_conv_n16<1, false, false>), g, dim3(256), 0, S(stream), ca);
        }
        return hip_check(hipGetLastError(), "k_conv_n16");
    }
    if (fused) return set_err(SW_E_ARG, "sw_cnn: pool / head epilogues need a 3x3 conv without a split output");
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
    if ((!o.in0 && !(o.kind == SW_CNN_CONV && o.site)) || (o.c2 && !o.in1) || !o.weight) return set_err(SW_E_ARG, who);
    if (o.pool_out && (o.h % 2 || o.w % 2)) return set_err(SW_E_ARG, who);
    if (o.head_w && !o.outf) return set_err(SW_E_ARG, who);
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

extern "C" int sw_gate_error(const float *out, const float *fields, const int32_t *org, const int32_t *pix, const int32_t *seg,
                  const int32_t *fm, float *werr, int nwin, int hw, int cs, int nl, int rows, int cols, void *stream) {
    if (!out || !fields || !org || !pix || !seg || !fm || !werr) return set_err(SW_E_ARG, "sw_gate_error: NULL pointer");
    if (nwin < 0 || hw <= 0 || cs <= 0 || nl <= 0 || nl > cs || rows <= 0 || cols <= 0)
        return set_err(SW_E_ARG, "sw_gate_error: bad sizes");
    if (nwin == 0) return SW_OK;
    hipLaunchKernelGGL(k_gate_error<false>, dim3(nwin), dim3(256), 0, S(stream), out, fields, (const int2 *)org, pix,
                       seg, fm, werr, hw, cs, nl, rows, cols);
    return hip_check(hipGetLastError(), "k_gate_error");
}

extern "C" int sw_gate_max_error(const float *out, const float *fields, const int32_t *org, const int32_t *pix,
                                 const int32_t *seg, const int32_t *fm, float *werr, int nwin, int hw, int cs, int nl,
                                 int rows, int cols, void *stream) {
    if (!out || !fields || !org || !pix || !seg || !fm || !werr)
        return set_err(SW_E_ARG, "sw_gate_max_error: NULL pointer");
    if (nwin < 0 || hw <= 0 || cs <= 0 || nl <= 0 || nl > cs || rows <= 0 || cols <= 0)
        return set_err(SW_E_ARG, "sw_gate_max_error: bad sizes");
    if (nwin == 0) return SW_OK;
    hipLaunchKernelGGL(k_gate_error<true>, dim3(nwin), dim3(256), 0, S(stream), out, fields, (const int2 *)org, pix,
                       seg, fm, werr, hw, cs, nl, rows, cols);
    return hip_check(hipGetLastError(), "k_gate_error<max>");
}

extern "C" int sw_gate_write(const float *out, const int32_t *org, const int32_t *fm, const uint8_t *sampled, float *fields,
                  int nwin, int hw, int cs, int core, int c0, int nl, int rows, int cols, void *stream) {
    if (!out || !org || !fm || !sampled || !fields) return set_err(SW_E_ARG, "sw_gate_write: NULL pointer");
    if (nwin < 0 || core <= 0 || c0 < 0 || c0 + core > hw || nl < 6 || nl > cs || rows <= 0 || cols <= 0)
        return set_err(SW_E_ARG, "sw_gate_write: bad sizes");
    if (nwin == 0) return SW_OK;
    const long total = (long)nwin * core * core;
    hipLaunchKernelGGL(k_gate_write, dim3(blocks(total, 256)), dim3(256), 0, S(stream), out, (const int2 *)org, fm,
                       sampled, fields, nwin, hw, cs, core, c0, nl, rows, cols);
    return hip_check(hipGetLastError(), "k_gate_write");
}
