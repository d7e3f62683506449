// libsolweig_hip -- SOLWEIG full-solve kernels for MI355X (gfx950, wave64).
//
// Kernels (one thread per output pixel, waves laid along image rows so each
// march step of a wave reads one contiguous, coalesced row segment):
//   k_svf_march   153-patch sky march -> packed mask cube      (shadow.py:166-322)
//   k_svf_fold    exact annulus fold -> 15 SVF fields + total   (shadow.py:545-613)
//   k_sun_step    sun march + wall tail -> GVF walk sources     (solweig.py:1079-1224, 121-134)
//   k_flux_step   GVF walk + Kdown/Kup/Kside + Ldown/Lside -> Tmrt -> UTCI
//                 (solweig.py:75-402, 451-832, 1439-1648, 2101-2299;
//                  calculate_utci.py:259-330)
//
// Float discipline: built with -ffp-contract=off and without fast-math, so
// each arithmetic operation rounds once, as the reference's elementwise
// torch ops do. Transcendentals are evaluated in double and rounded to
// float. Heights must be finite and non-negative (checked by the host),
// which makes max/min NaN semantics moot and the early exits exact.
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "sw_internal.h"

#define SW_BUILD_FLAGS "hipcc -O3 --offload-arch=gfx950 -ffp-contract=off -fno-fast-math -packed-fp32-ops(off)"

static thread_local char g_err[512];

int swi_set_err(int code, const char *msg) {
    snprintf(g_err, sizeof(g_err), "%s", msg);
    return code;
}

int swi_hip_check(hipError_t e, const char *where) {
    if (e != hipSuccess) {
        snprintf(g_err, sizeof(g_err), "%s: %s", where, hipGetErrorString(e));
        return SW_E_HIP;
    }
    return SW_OK;
}

#define set_err swi_set_err
#define hip_check swi_hip_check
#define S swi_stream

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float d_sin(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float d_cos(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float d_tan(float x) { return (float)tan((double)x); }
__device__ __forceinline__ float d_atan(float x) { return (float)atan((double)x); }
__device__ __forceinline__ float d_acos(float x) { return (float)acos((double)x); }
__device__ __forceinline__ float d_asin(float x) { return (float)asin((double)x); }
__device__ __forceinline__ float d_exp(float x) { return (float)exp((double)x); }
__device__ __forceinline__ float d_log(float x) { return (float)log((double)x); }
__device__ __forceinline__ float d_pow(float x, double e) { return (float)pow((double)x, e); }
// x^4, x^5, x^6 of a float, rounded once to float from a double product
// chain: x^2 is exact in double, each further product rounds at 2^-53,
// far below the float rounding step.
struct Pow456 {
    float p4, p5, p6;
};
__device__ __forceinline__ Pow456 pow456(float x) {
    const double d = (double)x;
    const double d2 = d * d;
    const double d4 = d2 * d2;
    Pow456 r;
    r.p4 = (float)d4;
    r.p5 = (float)(d4 * d);
    r.p6 = (float)(d4 * d2);
    return r;
}

// Map a launch-order block index to a tile so that consecutive tiles land on
// the same XCD (blocks are dealt round-robin over the 8 XCDs, each with its
// own L2): tile = (b % 8) * (nb / 8) + b / 8. The grid is launched as a
// multiple of 8 blocks (grid_xcd), which makes the map a bijection on
// [0, nb); tiles past the image are idle.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int per = nb >> 3;
    return (b & 7) * per + (b >> 3);
}

__device__ __forceinline__ bool wave_any(bool v) {
    return __ballot(v) != 0ull;
}

// ---------------------------------------------------------------------------
// heights
// ---------------------------------------------------------------------------
__global__ void k_pack_heights(const float *__restrict__ a, const float *__restrict__ v1,
                               const float *__restrict__ v2, float4 *__restrict__ hp, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) hp[i] = make_float4(a[i], v1[i], v2[i], 0.f);
}

// ---------------------------------------------------------------------------
// scene compose on the device (utci_process.py:667-679, 738-744) and the
// tile statistics: order-preserving float keys reduced with atomicMax.
// stats[0] = key(amax), stats[1] = key(hmax), stats[2] = flags.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float funkey(uint32_t k) {
    return (k & 0x80000000u) ? __uint_as_float(k & 0x7FFFFFFFu) : __uint_as_float(~k);
}

__global__ void k_stats_init(uint32_t *stats) {
    if (threadIdx.x < 4) stats[threadIdx.x] = 0u;
}

__global__ void __launch_bounds__(256)
k_compose(const float *__restrict__ dsm, const float *__restrict__ dem, const float *__restrict__ cdsm,
          const float *__restrict__ aspect_deg, const float *__restrict__ lc, float4 *__restrict__ hp,
          float *__restrict__ vegdsm2_out, float *__restrict__ bu, uint8_t *__restrict__ cls,
          float *__restrict__ aspect_rad, uint32_t *__restrict__ stats, long n) {
    __shared__ uint32_t s_amax, s_hmax, s_flags;
    if (threadIdx.x == 0) { s_amax = 0u; s_hmax = 0u; s_flags = 0u; }
    __syncthreads();
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float a = dsm[i], d = dem[i];
        float t1 = cdsm[i];
        if (t1 < 0.f) t1 = 0.f;
        const float vegdem = t1 + d;
        const float vegdem2 = t1 * 0.25f + d;
        const float bush = (vegdem2 * vegdem == 0.f) ? vegdem : 0.f;
        float v1 = t1 + a;
        if (v1 == a) v1 = 0.f;
        float v2 = t1 * 0.25f + a;
        if (v2 == a) v2 = 0.f;
        hp[i] = make_float4(a, v1, v2, 0.f);
        vegdsm2_out[i] = v2;
        bu[i] = (a - d < 2.f) ? 1.f : 0.f;
        aspect_rad[i] = (aspect_deg[i] * 3.14159265358979323846f) / 180.0f;
        uint32_t flags = 0u;
        uint8_t k = 0;
        if (lc) {
            float g = lc[i];
            if (g < 1.f || g > 7.f) g = 6.f;
            if (g == 3.f || g == 4.f) g = 5.f;
            if (g != (float)(int)g) flags |= 8u;  // non-integer land-cover code
            k = (uint8_t)(int)g;
        }
        cls[i] = k;
        if (!(isfinite(a) && isfinite(d) && isfinite(t1))) flags |= 1u;
        if (bush > 0.f) flags |= 4u;
        atomicMax(&s_amax, fkey(fmaxf(a, vegdem)));
        atomicMax(&s_hmax, fkey(fmaxf(a, fmaxf(v1, v2))));
        if (flags) atomicOr(&s_flags, flags);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicMax(&stats[0], s_amax);
        atomicMax(&stats[1], s_hmax);
        if (s_flags) atomicOr(&stats[2], s_flags);
    }
}

// per-patch executed step count: the march runs while the previous step had
// amax >= dz, so the first step whose dz exceeds amax is the last one run;
// one wave per patch
__global__ void __launch_bounds__(64)
k_svf_counts(const int4 *__restrict__ tab, const int32_t *__restrict__ offs, uint32_t *__restrict__ stats,
             int32_t *__restrict__ counts) {
    const int p = blockIdx.x;
    const float amax = funkey(stats[0]);
    const int o = offs[p], len = offs[p + 1] - o;
    int c = len;
    for (int base = 0; base < len; base += 64) {
        const int s = base + (int)threadIdx.x;
        const bool stop = (s < len) && !(amax >= __int_as_float(tab[o + s].z));
        const unsigned long long m = __ballot(stop);
        if (m) {
            c = base + __ffsll((long long)m);   // 1-based index of the first stop step = count
            break;
        }
    }
    if (threadIdx.x == 0) {
        counts[p] = c;
        if (p == 0) stats[3] = 0u;   // negative-height flag, set by k_pad_heights
    }
}

// Heights padded by SW_SVF_PAD cells on every side for the march, 12 bytes
// per cell {dsm, vegdsm, vegdsm2}; pad cells hold {0, 0, -0}. For a
// non-negative target and dz >= 0 a zero-height source has the same effect
// as the reference's zero-filled out-of-grid temporary, and the sign bit of
// vegdsm2 marks "outside" for the exact out-of-grid exit (grid cells carry
// +0 there: -0 and +0 act alike in every use). Every table step moves at
// most one cell and a lane stops after its first outside step at s >= 1, so
// its reads reach at most 1 + SVF_CHUNK cells past the grid, inside the
// SW_SVF_PAD border. stats[3] gets bit 1 when any
// height is negative (the march then checks bounds).
__global__ void __launch_bounds__(256)
k_pad_heights(const float4 *__restrict__ hp, float *__restrict__ hpad, uint32_t *__restrict__ stats, int rows,
              int cols) {
    __shared__ uint32_t s_neg;
    if (threadIdx.x == 0) s_neg = 0u;
    __syncthreads();
    const int W = cols + 2 * SW_SVF_PAD, H = rows + 2 * SW_SVF_PAD;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (long)W * H) {
        const int pr = (int)(i / W), pc = (int)(i - (long)pr * W);
        const int r = pr - SW_SVF_PAD, c = pc - SW_SVF_PAD;
        float x = 0.f, y = 0.f, z = -0.f;
        if ((unsigned)r < (unsigned)rows && (unsigned)c < (unsigned)cols) {
            const float4 v = hp[(long)r * cols + c];
            x = v.x;
            y = v.y;
            z = v.z + 0.f;   // -0 -> +0, every other value unchanged
            if (x < 0.f || y < 0.f || z < 0.f) s_neg = 1u;
        }
        hpad[3 * i] = x;
        hpad[3 * i + 1] = y;
        hpad[3 * i + 2] = z;
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_neg) atomicOr(&stats[3], 1u);
}

// Maxima of max(building, canopy) over SW_SVF_COARSE x SW_SVF_COARSE cells of the padded grid,
// stored behind the padded heights (the march's chunk-skip test)
__global__ void __launch_bounds__(256)
k_pad_coarse(const float *__restrict__ hpad, float *__restrict__ coarse, int H, int W, int Hc, int Wc) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)Hc * Wc) return;
    const int R = (int)(i / Wc), C = (int)(i - (long)R * Wc);
    float m = 0.f;
    for (int dr = 0; dr < SW_SVF_COARSE; ++dr) {
        const int r = R * SW_SVF_COARSE + dr;
        if (r >= H) break;
        for (int dc = 0; dc < SW_SVF_COARSE; ++dc) {
            const int c = C * SW_SVF_COARSE + dc;
            if (c >= W) break;
            const long q = 3 * ((long)r * W + c);
            m = fmaxf(m, fmaxf(hpad[q], hpad[q + 1]));
        }
    }
    coarse[i] = m;
}

struct H3 {
    float x, y, z;
};

typedef unsigned int sw_u3 __attribute__((ext_vector_type(3)));

// one padded cell through a buffer resource: range-checked by the hardware
// (an offset past the buffer reads zeros instead of faulting)
__device__ __forceinline__ H3 ld3(__amdgpu_buffer_rsrc_t rs, uint32_t byte_off) {
    const sw_u3 v = __builtin_amdgcn_raw_buffer_load_b96(rs, byte_off, 0, 0);
    H3 h;
    h.x = __uint_as_float(v.x);
    h.y = __uint_as_float(v.y);
    h.z = __uint_as_float(v.z);
    return h;
}

__device__ __forceinline__ bool outside(const H3 &h) { return (__float_as_uint(h.z) >> 31) != 0u; }

// one coarse maximum through a buffer resource (an offset past the array reads 0)
__device__ __forceinline__ float ldc(__amdgpu_buffer_rsrc_t rs, uint32_t byte_off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, byte_off, 0, 0));
}

// ---------------------------------------------------------------------------
// SVF patch march. Semantics of shadow(): out-of-bounds sources read an
// exact 0, the first step carries the vegetation condition and the
// target-local trunk gate and clears the accumulator, and a step runs only
// while the previous step had amax >= dz (per-patch counts computed on the
// device by k_svf_counts). Tables: int4 {dx, dy, dz bits, byte offset
// 12 * (dx * W + dy)} with W = cols + 2 * SW_SVF_PAD, the padded row pitch.
//
// State per lane: f (running max of a - dz), sh = f > a, vs in {0, 1}
// (vegsh of the reference, never negative), vbb = "vbshvegsh sum > 0".
// Exact early exits (every later step leaves the outputs unchanged):
// sh && !vs; for a >= 0 after the first step, a source outside the grid
// (all later ones are outside too) or hmax - dz <= a (no later source rises
// above the target).
// ---------------------------------------------------------------------------
struct SvfLane {
    bool sh, vs, vbb;
};

// f = max(a, tmp_1 .. tmp_s) exceeds a exactly when some tmp_i does, so the
// running maximum reduces to an OR of comparisons
// (bitwise & | on the booleans: no short-circuit branches in the hot loop)
__device__ __forceinline__ void svf_first(SvfLane &st, float at, float veg2t, float ha, float hv, float hv2,
                                          float dz) {
    const float tmp = ha - dz, tv = hv - dz, tv2 = hv2 - dz;
    st.sh = tmp > at;
    st.vs = (tv > at) & !(tv2 > at) & !st.sh;
    float fv = tv - tmp;
    if (fv <= 0.f) fv = 1000.f;
    st.vs = st.vs | (fv < dz);
    st.vs = st.vs & (veg2t > at);
    st.vbb = false;
}

__device__ __forceinline__ void svf_next(SvfLane &st, float at, float ha, float hv, float hv2, float dz) {
    const float tmp = ha - dz, tv = hv - dz, tv2 = hv2 - dz;
    st.sh = st.sh | (tmp > at);
    st.vs = (st.vs | ((tv > at) & !(tv2 > at))) & !st.sh;
    st.vbb = st.vbb | st.vs;
}

__device__ __forceinline__ uint32_t svf_out(const SvfLane &st) {
    const uint32_t out_sh = st.sh ? 0u : 1u;
    const uint32_t out_vs = st.vs ? 0u : 1u;
    const uint32_t out_vb = 1u - (st.vbb ? 1u : 0u) + (st.vs ? 1u : 0u);
    return out_sh | (out_vs << 1) | (out_vb << 2);
}

// fast path: every height >= 0 and dz >= 0, sources read from the padded
// buffer; chunks of SVF_CHUNK steps keep that many gathers in flight per
// lane. The exit tests run on the chunk's last step: each exit condition,
// once true, holds for every later step (the fixed point sh && !vegsh, a
// ray that left the grid stays out, dz grows along the patch), so testing
// the last step of a chunk finds the exit the per-step test would have
// found in that chunk, and the steps run past it change no output.
#define SVF_CHUNK 8                 // default steps per gather chunk (autotuned: 4, 8 or 12)
//
// Chunk skip: when every cell a chunk's steps read has max(building, canopy) - dz <= a (tested on the
// coarse maxima of the cells covering the chunk's bounding box against the chunk's smallest dz; float
// subtraction is monotone, so each step's h - dz <= that bound), no step of the chunk sets sh or a
// vegetation condition, and the chunk's whole effect is vs = vs && !sh, vbb = vbb || vs, which is
// applied without reading the cells. The exit test then uses the last step's cell coordinates
// (within the pad, an out-of-grid cell is exactly one whose z reads -0).
template <int CHUNK>
__device__ __forceinline__ uint32_t march_svf_fast(__amdgpu_buffer_rsrc_t rs, uint32_t tpb, float at,
                                                   float veg2t, bool valid, const int4 *__restrict__ T,
                                                   int count, float hmax, __amdgpu_buffer_rsrc_t crs,
                                                   int pr, int pc, int rows, int cols, int Hc, int Wc) {
    static_assert(CHUNK + 1 <= SW_SVF_PAD, "march chunk reads must stay inside the pad");
    SvfLane st;
    st.sh = false; st.vs = false; st.vbb = false;
    bool active = valid && count > 0;
    if (active) {
        const int4 e = T[0];
        const H3 h = ld3(rs, tpb + (uint32_t)e.w);
        svf_first(st, at, veg2t, h.x, h.y, h.z, __int_as_float(e.z));
        active = !(st.sh & !st.vs);
    }
    int s = 1;
    for (; s + CHUNK <= count; s += CHUNK) {
        if (!wave_any(active)) break;
        int4 e[CHUNK];
#pragma unroll
        for (int k = 0; k < CHUNK; ++k) e[k] = T[s + k];
        // the chunk's offset bounding box and smallest dz, the same for every lane (the tables move
        // monotonically in x and y with non-decreasing dz, so the first and last steps bound them)
        const int xl = min(e[0].x, e[CHUNK - 1].x), xh = max(e[0].x, e[CHUNK - 1].x);
        const int yl = min(e[0].y, e[CHUNK - 1].y), yh = max(e[0].y, e[CHUNK - 1].y);
        const float dzl = __int_as_float(e[0].z), dl = __int_as_float(e[CHUNK - 1].z);
        bool skip = false;
        if (active) {
            // coarse cells past the grid read 0 (range-checked loads), which the steps they cover
            // would read as well
            const int R0 = (pr + xl) >> 3, R1 = (pr + xh) >> 3, C0 = (pc + yl) >> 3, C1 = (pc + yh) >> 3;
            if (R1 - R0 <= 1 && C1 - C0 <= 1) {
                const uint32_t o00 = (uint32_t)(R0 * Wc + C0) * 4u, o01 = (uint32_t)(R0 * Wc + C1) * 4u;
                const uint32_t o10 = (uint32_t)(R1 * Wc + C0) * 4u, o11 = (uint32_t)(R1 * Wc + C1) * 4u;
                const bool inb = R0 >= 0 && C0 >= 0 && R1 < Hc && C1 < Wc;
                const float m = fmaxf(fmaxf(ldc(crs, o00), ldc(crs, o01)), fmaxf(ldc(crs, o10), ldc(crs, o11)));
                skip = inb && (m - dzl <= at);
            }
        }
        // uniform branches: a wave whose active lanes all skip issues no gather and no step update
        const bool need = active && !skip;
        if (wave_any(need)) {
            if (need) {
                H3 h[CHUNK];
#pragma unroll
                for (int k = 0; k < CHUNK; ++k) h[k] = ld3(rs, tpb + (uint32_t)e[k].w);
#pragma unroll
                for (int k = 0; k < CHUNK; ++k) svf_next(st, at, h[k].x, h[k].y, h[k].z, __int_as_float(e[k].z));
                active = !((st.sh & !st.vs) | outside(h[CHUNK - 1]) | (hmax - dl <= at));
            }
        }
        if (wave_any(active && skip)) {
            if (active && skip) {
                st.vs = st.vs & !st.sh;
                st.vbb = st.vbb | st.vs;
                const int lr = pr + e[CHUNK - 1].x - SW_SVF_PAD, lc = pc + e[CHUNK - 1].y - SW_SVF_PAD;
                const bool out = ((unsigned)lr >= (unsigned)rows) | ((unsigned)lc >= (unsigned)cols);
                active = !((st.sh & !st.vs) | out | (hmax - dl <= at));
            }
        }
    }
    for (; s < count; ++s) {
        if (!wave_any(active)) break;
        const int4 e = T[s];
        if (active) {
            const H3 h = ld3(rs, tpb + (uint32_t)e.w);
            const float dz = __int_as_float(e.z);
            svf_next(st, at, h.x, h.y, h.z, dz);
            active = !((st.sh & !st.vs) | outside(h) | (hmax - dz <= at));
        }
    }
    return svf_out(st);
}

// checked path (some height is negative): out-of-grid sources read exact
// zeros as in the reference; the out-of-grid exit applies to a >= 0 only
__device__ __forceinline__ uint32_t march_svf_checked(__amdgpu_buffer_rsrc_t rs, uint32_t tpb, int r, int c,
                                                      int rows, int cols, float at, float veg2t, bool valid,
                                                      const int4 *__restrict__ T, int count, float hmax) {
    SvfLane st;
    st.sh = false; st.vs = false; st.vbb = false;
    bool active = valid;
    const bool exit_ok = at >= 0.f;
    for (int s = 0; s < count; ++s) {
        if (!wave_any(active)) break;
        const int4 e = T[s];
        if (active) {
            const bool inb = ((unsigned)(r + e.x) < (unsigned)rows) && ((unsigned)(c + e.y) < (unsigned)cols);
            const float dz = __int_as_float(e.z);
            float tmp = 0.f, tv = 0.f, tv2 = 0.f;
            if (inb) {
                const H3 h = ld3(rs, tpb + (uint32_t)e.w);
                tmp = h.x - dz;
                tv = h.y - dz;
                tv2 = h.z - dz;
            }
            // the update on (tmp, tv, tv2) directly: a source value of 0 is not h - dz
            if (s == 0) {
                st.sh = tmp > at;
                st.vs = (tv > at) && !(tv2 > at) && !st.sh;
                float fv = tv - tmp;
                if (fv <= 0.f) fv = 1000.f;
                if (fv < dz) st.vs = true;
                st.vs = st.vs && (veg2t > at);
                st.vbb = false;
                if (st.sh && !st.vs) active = false;
            } else {
                st.sh = st.sh || (tmp > at);
                st.vs = (st.vs || ((tv > at) && !(tv2 > at))) && !st.sh;
                st.vbb = st.vbb || st.vs;
                if ((st.sh && !st.vs) || (exit_ok && (!inb || (hmax - dz) <= at))) active = false;
            }
        }
    }
    return svf_out(st);
}

// grid: x = 16 x 16 pixel tiles (XCD-remapped; each wave marches an 8 x 8
// square, which keeps more lanes busy than a row segment because exit
// steps are spatially correlated), y = 32-patch word group.
#define SVF_TILE 16
// SHAPE: the pixels of a wave within the 16 x 16 tile: 0 an 8 x 8 square, 1 a 4 x 16 row strip, 2 a
// 16 x 4 column strip (exits are spatially correlated, so the shape sets how long a wave keeps lanes
// marching). CHUNK: steps per gather chunk of the fast path. Neither changes any output.
template <int CHUNK, int SHAPE>
__global__ void __launch_bounds__(256)
k_svf_march(const float *__restrict__ hpad, const int32_t *__restrict__ offs, const int4 *__restrict__ tab,
            const int32_t *__restrict__ counts, uint32_t *__restrict__ masks,
            const uint32_t *__restrict__ stats, int rows, int cols, const int32_t *__restrict__ tiles,
            int ntiles, int split) {
    const long n = (long)rows * cols;
    const float hmax = funkey(stats[1]);
    const bool checked = stats[3] != 0u;
    const int tiles_x = (cols + SVF_TILE - 1) / SVF_TILE;
    int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tiles) {                               // listed tiles only (uniform per block)
        if (tile >= ntiles) return;
        tile = tiles[tile];
        if (tile < 0 || tile >= tiles_x * ((rows + SVF_TILE - 1) / SVF_TILE)) return;
    }
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    const int ry = SHAPE == 0 ? (wv >> 1) * 8 + (ln >> 3) : (SHAPE == 1 ? wv * 4 + (ln >> 4) : (ln >> 2));
    const int cx = SHAPE == 0 ? (wv & 1) * 8 + (ln & 7) : (SHAPE == 1 ? (ln & 15) : wv * 4 + (ln & 3));
    const int r = (tile / tiles_x) * SVF_TILE + ry;
    const int c = (tile % tiles_x) * SVF_TILE + cx;
    const bool valid = (r < rows) && (c < cols);
    const int W = cols + 2 * SW_SVF_PAD;
    const uint32_t nbytes = (uint32_t)((long)(rows + 2 * SW_SVF_PAD) * W * 12);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void *)hpad, (short)0, (int)nbytes,
                                                                          0x00020000);
    const uint32_t tpb = valid ? (uint32_t)(((long)(r + SW_SVF_PAD) * W + (c + SW_SVF_PAD)) * 12) : 0u;
    const int Hc = (rows + 2 * SW_SVF_PAD + SW_SVF_COARSE - 1) / SW_SVF_COARSE;
    const int Wc = (W + SW_SVF_COARSE - 1) / SW_SVF_COARSE;
    const __amdgpu_buffer_rsrc_t crs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(hpad + 3 * (long)(rows + 2 * SW_SVF_PAD) * W), (short)0, (int)((long)Hc * Wc * 4), 0x00020000);
    const long t = valid ? (long)r * cols + c : 0;
    float at = 0.f, veg2t = 0.f;
    if (valid) {
        const H3 h = ld3(rs, tpb);
        at = h.x;
        veg2t = h.z;
    }
    // split 1: one block per mask word (32 patches); split 4: one block per byte of a word (8 patches,
    // short tile lists: four times the blocks, a quarter of the latency)
    const int w = blockIdx.y / split, sub = blockIdx.y - w * split;
    const int p0 = w * 32;
    const int pb = p0 + sub * (32 / split);
    const int p1 = min(pb + 32 / split, SW_NPATCH);
    uint32_t bsh = 0, bvs = 0, bv1 = 0, bv2 = 0;
    for (int p = pb; p < p1; ++p) {
        // the zero-height pad stands for the reference's zero source only when
        // dz >= 0; float32 tan(90 deg) makes the zenith patch's dz negative
        const bool chk = checked || (counts[p] > 0 && __int_as_float(tab[offs[p]].z) < 0.f);
        const uint32_t m = chk
            ? march_svf_checked(rs, tpb, r, c, rows, cols, at, veg2t, valid, tab + offs[p], counts[p], hmax)
            : march_svf_fast<CHUNK>(rs, tpb, at, veg2t, valid, tab + offs[p], counts[p], hmax, crs,
                                    r + SW_SVF_PAD, c + SW_SVF_PAD, rows, cols, Hc, Wc);
        const uint32_t bit = 1u << (p - p0);
        if (m & 1u) bsh |= bit;
        if (m & 2u) bvs |= bit;
        const uint32_t vbv = m >> 2;
        if (vbv == 1u) bv1 |= bit;
        if (vbv == 2u) bv2 |= bit;
    }
    if (valid && split == 1) {
        masks[(long)(0 * SW_MASK_WORDS + w) * n + t] = bsh;
        masks[(long)(1 * SW_MASK_WORDS + w) * n + t] = bvs;
        masks[(long)(2 * SW_MASK_WORDS + w) * n + t] = bv1;
        masks[(long)(3 * SW_MASK_WORDS + w) * n + t] = bv2;
    } else if (valid && split == 2) {
        // this block's half of each word (little-endian: bits 16 sub .. 16 sub + 15)
        uint16_t *mh = (uint16_t *)masks;
        mh[((long)(0 * SW_MASK_WORDS + w) * n + t) * 2 + sub] = (uint16_t)(bsh >> (16 * sub));
        mh[((long)(1 * SW_MASK_WORDS + w) * n + t) * 2 + sub] = (uint16_t)(bvs >> (16 * sub));
        mh[((long)(2 * SW_MASK_WORDS + w) * n + t) * 2 + sub] = (uint16_t)(bv1 >> (16 * sub));
        mh[((long)(3 * SW_MASK_WORDS + w) * n + t) * 2 + sub] = (uint16_t)(bv2 >> (16 * sub));
    } else if (valid) {
        // this block's byte of each word (little-endian: bits 8 sub .. 8 sub + 7)
        uint8_t *mb = (uint8_t *)masks;
        mb[((long)(0 * SW_MASK_WORDS + w) * n + t) * 4 + sub] = (uint8_t)(bsh >> (8 * sub));
        mb[((long)(1 * SW_MASK_WORDS + w) * n + t) * 4 + sub] = (uint8_t)(bvs >> (8 * sub));
        mb[((long)(2 * SW_MASK_WORDS + w) * n + t) * 4 + sub] = (uint8_t)(bv1 >> (8 * sub));
        mb[((long)(3 * SW_MASK_WORDS + w) * n + t) * 4 + sub] = (uint8_t)(bv2 >> (8 * sub));
    }
}

// launch of the march with a configuration (chunk 4 / 8 / 12, shape 0 / 1 / 2, split 1 / 2 / 4)
static void launch_march(dim3 grid, hipStream_t st, int chunk, int shape, const float *hpad, const int32_t *offs,
                         const int4 *tab, const int32_t *counts, uint32_t *masks, const uint32_t *stats, int rows,
                         int cols, const int32_t *tiles, int ntiles, int split) {
#define SW_MARCH_CASE(C, SH)                                                                                  \
    if (chunk == C && shape == SH) {                                                                         \
        hipLaunchKernelGGL((k_svf_march<C, SH>), grid, dim3(256), 0, st, hpad, offs, tab, counts, masks, stats, \
                           rows, cols, tiles, ntiles, split);                                                \
        return;                                                                                              \
    }
    SW_MARCH_CASE(8, 0) SW_MARCH_CASE(4, 0) SW_MARCH_CASE(12, 0)
    SW_MARCH_CASE(8, 1) SW_MARCH_CASE(4, 1) SW_MARCH_CASE(12, 1)
    SW_MARCH_CASE(8, 2) SW_MARCH_CASE(4, 2) SW_MARCH_CASE(12, 2)
#undef SW_MARCH_CASE
}

static int check_march_config(const sw_march_config_t *cfg, const char *who) {
    if (!cfg) return SW_OK;
    const bool ok = (cfg->chunk == 4 || cfg->chunk == 8 || cfg->chunk == 12) && cfg->shape >= 0 && cfg->shape <= 2 &&
                    (cfg->split == 1 || cfg->split == 2 || cfg->split == 4);
    return ok ? SW_OK : set_err(SW_E_ARG, who);
}

__device__ __forceinline__ void load_masks(const uint32_t *__restrict__ masks, long n, long t,
                                           uint32_t *msh, uint32_t *mvs, uint32_t *mv1, uint32_t *mv2) {
#pragma unroll
    for (int w = 0; w < SW_MASK_WORDS; ++w) {
        msh[w] = masks[(long)(0 * SW_MASK_WORDS + w) * n + t];
        mvs[w] = masks[(long)(1 * SW_MASK_WORDS + w) * n + t];
        mv1[w] = masks[(long)(2 * SW_MASK_WORDS + w) * n + t];
        mv2[w] = masks[(long)(3 * SW_MASK_WORDS + w) * n + t];
    }
}

// one patch of the fold with NA annuli (fully unrolled, weights in SGPRs)
template <int NA>
__device__ __forceinline__ void fold_patch(float sh, float vs, float vb, const float *__restrict__ wi,
                                           const float *__restrict__ wa, int dm, float &s0, float &v0, float &b0,
                                           float &s1, float &v1, float &b1, float &s2, float &v2, float &b2,
                                           float &s3, float &v3, float &b3, float &s4, float &v4, float &b4) {
    float W[NA], A[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) { W[k] = wi[k]; A[k] = wa[k]; }
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        s0 = __builtin_fmaf(sh, W[k], s0); v0 = __builtin_fmaf(vs, W[k], v0); b0 = __builtin_fmaf(vb, W[k], b0);
    }
#define SW_FOLD_DIR(BIT, S, V, B)                                                                   \
    if (dm & BIT) {                                                                              \
        _Pragma("unroll") for (int k = 0; k < NA; ++k) {                                          \
            S = __builtin_fmaf(sh, A[k], S); V = __builtin_fmaf(vs, A[k], V); B = __builtin_fmaf(vb, A[k], B); \
        }                                                                                        \
    }
    SW_FOLD_DIR(1, s1, v1, b1)
    SW_FOLD_DIR(2, s2, v2, b2)
    SW_FOLD_DIR(4, s3, v3, b3)
    SW_FOLD_DIR(8, s4, v4, b4)
#undef SW_FOLD_DIR
}

// ---------------------------------------------------------------------------
// SVF fold, literal per-annulus order (bit-exact with the reference's sums)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void fold_pixel(const uint32_t *__restrict__ masks, const float *__restrict__ veg2,
                                           const float *__restrict__ w_iso, const float *__restrict__ w_aniso,
                                           const int32_t *__restrict__ ring, const int32_t *__restrict__ nann,
                                           const int32_t *__restrict__ dirm, float *__restrict__ out, long n,
                                           long t) {
    uint32_t msh[SW_MASK_WORDS], mvs[SW_MASK_WORDS], mv1[SW_MASK_WORDS], mv2[SW_MASK_WORDS];
    load_masks(masks, n, t, msh, mvs, mv1, mv2);
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f;
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f, v4 = 0.f;
    float b0 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f, b4 = 0.f;
    // sh, vegsh are 0/1 and vbsh is 0/1/2, so every weight x mask product is
    // exact and fma(mask, weight, sum) rounds once, like "sum + weight * mask".
    // Each sum keeps the reference's (patch, annulus) order; the direction
    // windows are wave-uniform per patch, so each runs as its own loop.
    for (int p = 0; p < SW_NPATCH; ++p) {
        const int w = p >> 5;
        const uint32_t bit = 1u << (p & 31);
        const float sh = (msh[w] & bit) ? 1.f : 0.f;
        const float vs = (mvs[w] & bit) ? 1.f : 0.f;
        const float vb = (mv1[w] & bit) ? 1.f : ((mv2[w] & bit) ? 2.f : 0.f);
        const int i = ring[p], dm = dirm[p];
        const float *wi = w_iso + i * 12, *wa = w_aniso + i * 12;
        if (nann[i] == 12) {
            fold_patch<12>(sh, vs, vb, wi, wa, dm, s0, v0, b0, s1, v1, b1, s2, v2, b2, s3, v3, b3, s4, v4, b4);
        } else {
            for (int k = 0; k < nann[i]; ++k) {
                s0 = __builtin_fmaf(sh, wi[k], s0); v0 = __builtin_fmaf(vs, wi[k], v0); b0 = __builtin_fmaf(vb, wi[k], b0);
            }
            if (dm & 1) for (int k = 0; k < nann[i]; ++k) { s1 = __builtin_fmaf(sh, wa[k], s1); v1 = __builtin_fmaf(vs, wa[k], v1); b1 = __builtin_fmaf(vb, wa[k], b1); }
            if (dm & 2) for (int k = 0; k < nann[i]; ++k) { s2 = __builtin_fmaf(sh, wa[k], s2); v2 = __builtin_fmaf(vs, wa[k], v2); b2 = __builtin_fmaf(vb, wa[k], b2); }
            if (dm & 4) for (int k = 0; k < nann[i]; ++k) { s3 = __builtin_fmaf(sh, wa[k], s3); v3 = __builtin_fmaf(vs, wa[k], v3); b3 = __builtin_fmaf(vb, wa[k], b3); }
            if (dm & 8) for (int k = 0; k < nann[i]; ++k) { s4 = __builtin_fmaf(sh, wa[k], s4); v4 = __builtin_fmaf(vs, wa[k], v4); b4 = __builtin_fmaf(vb, wa[k], b4); }
        }
    }
    const float cst = 3.0459e-004f;
    s2 = s2 + cst;
    s3 = s3 + cst;
    s0 = s0 > 1.f ? 1.f : s0; s1 = s1 > 1.f ? 1.f : s1; s2 = s2 > 1.f ? 1.f : s2;
    s3 = s3 > 1.f ? 1.f : s3; s4 = s4 > 1.f ? 1.f : s4;
    const float last = (veg2[t] == 0.f) ? cst : 0.f;
    v2 = v2 + last; v3 = v3 + last; b2 = b2 + last; b3 = b3 + last;
    v0 = v0 > 1.f ? 1.f : v0; v1 = v1 > 1.f ? 1.f : v1; v2 = v2 > 1.f ? 1.f : v2;
    v3 = v3 > 1.f ? 1.f : v3; v4 = v4 > 1.f ? 1.f : v4;
    b0 = b0 > 1.f ? 1.f : b0; b1 = b1 > 1.f ? 1.f : b1; b2 = b2 > 1.f ? 1.f : b2;
    b3 = b3 > 1.f ? 1.f : b3; b4 = b4 > 1.f ? 1.f : b4;
    out[0 * n + t] = s0; out[1 * n + t] = s1; out[2 * n + t] = s2; out[3 * n + t] = s3; out[4 * n + t] = s4;
    out[5 * n + t] = v0; out[6 * n + t] = v1; out[7 * n + t] = v2; out[8 * n + t] = v3; out[9 * n + t] = v4;
    out[10 * n + t] = b0; out[11 * n + t] = b1; out[12 * n + t] = b2; out[13 * n + t] = b3; out[14 * n + t] = b4;
    out[15 * n + t] = s0 - (1.f - v0) * (1.f - 0.03f);
}

__global__ void __launch_bounds__(256)
k_svf_fold(const uint32_t *__restrict__ masks, const float *__restrict__ veg2,
           const float *__restrict__ w_iso, const float *__restrict__ w_aniso,
           const int32_t *__restrict__ ring, const int32_t *__restrict__ nann,
           const int32_t *__restrict__ dirm, float *__restrict__ out, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    fold_pixel(masks, veg2, w_iso, w_aniso, ring, nann, dirm, out, n, t);
}

// the fold over listed 16 x 16 tiles
__global__ void __launch_bounds__(256)
k_svf_fold_tiles(const uint32_t *__restrict__ masks, const float *__restrict__ veg2,
                 const float *__restrict__ w_iso, const float *__restrict__ w_aniso,
                 const int32_t *__restrict__ ring, const int32_t *__restrict__ nann,
                 const int32_t *__restrict__ dirm, float *__restrict__ out, int rows, int cols,
                 const int32_t *__restrict__ tiles, int ntiles) {
    const int g = xcd_tile(blockIdx.x, gridDim.x);
    if (g >= ntiles) return;
    const int tiles_x = (cols + SVF_TILE - 1) / SVF_TILE;
    const int tile = tiles[g];
    if (tile < 0 || tile >= tiles_x * ((rows + SVF_TILE - 1) / SVF_TILE)) return;
    const int r = (tile / tiles_x) * SVF_TILE + (threadIdx.x >> 4);
    const int c = (tile % tiles_x) * SVF_TILE + (threadIdx.x & 15);
    if (r >= rows || c >= cols) return;
    fold_pixel(masks, veg2, w_iso, w_aniso, ring, nann, dirm, out, (long)rows * cols, (long)r * cols + c);
}

__global__ void k_unpack(const uint32_t *__restrict__ masks, float *__restrict__ shm,
                         float *__restrict__ vsm, float *__restrict__ vbm, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * SW_NPATCH) return;
    const long t = i / SW_NPATCH;
    const int p = (int)(i - t * SW_NPATCH);
    const int w = p >> 5;
    const uint32_t bit = 1u << (p & 31);
    if (shm) shm[i] = (masks[(long)(0 * SW_MASK_WORDS + w) * n + t] & bit) ? 1.f : 0.f;
    if (vsm) vsm[i] = (masks[(long)(1 * SW_MASK_WORDS + w) * n + t] & bit) ? 1.f : 0.f;
    if (vbm) {
        const bool one = masks[(long)(2 * SW_MASK_WORDS + w) * n + t] & bit;
        const bool two = masks[(long)(3 * SW_MASK_WORDS + w) * n + t] & bit;
        vbm[i] = one ? 1.f : (two ? 2.f : 0.f);
    }
}

// single-patch march with float outputs (reference native ABI form)
// Pack the reference's float cubes (rows, cols, 153) into the mask planes;
// flags bit 1: a value outside {0, 1} (sh, vegsh) or {0, 1, 2} (vbsh).
__global__ void k_pack_masks(const float *__restrict__ shm, const float *__restrict__ vsm,
                             const float *__restrict__ vbm, uint32_t *__restrict__ masks,
                             uint32_t *__restrict__ flags, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float *a = shm + t * SW_NPATCH, *b = vsm + t * SW_NPATCH, *c = vbm + t * SW_NPATCH;
    bool bad = false;
    for (int w = 0; w < SW_MASK_WORDS; ++w) {
        uint32_t m0 = 0u, m1 = 0u, m2 = 0u, m3 = 0u;
        const int pend = min(w * 32 + 32, SW_NPATCH);
        for (int p = w * 32; p < pend; ++p) {
            const uint32_t bit = 1u << (p & 31);
            const float x = a[p], y = b[p], z = c[p];
            if (x == 1.f) m0 |= bit; else bad |= (x != 0.f);
            if (y == 1.f) m1 |= bit; else bad |= (y != 0.f);
            if (z == 1.f) m2 |= bit;
            else if (z == 2.f) m3 |= bit;
            else bad |= (z != 0.f);
        }
        masks[(long)(0 * SW_MASK_WORDS + w) * n + t] = m0;
        masks[(long)(1 * SW_MASK_WORDS + w) * n + t] = m1;
        masks[(long)(2 * SW_MASK_WORDS + w) * n + t] = m2;
        masks[(long)(3 * SW_MASK_WORDS + w) * n + t] = m3;
    }
    if (bad) atomicOr(flags, 1u);
}

// Tile statistics of packed heights with a caller-given amax:
// stats = {key(amax), key(max of the three planes), flags}; flags 1 marks a
// non-finite height. stats must be initialised by k_stats_init.
__global__ void __launch_bounds__(256)
k_height_stats(const float4 *__restrict__ hp, uint32_t *__restrict__ stats, long n) {
    __shared__ uint32_t s_hmax, s_flags;
    if (threadIdx.x == 0) { s_hmax = 0u; s_flags = 0u; }
    __syncthreads();
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float4 h = hp[i];
        if (!(isfinite(h.x) && isfinite(h.y) && isfinite(h.z))) atomicOr(&s_flags, 1u);
        else atomicMax(&s_hmax, fkey(fmaxf(h.x, fmaxf(h.y, h.z))));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicMax(&stats[1], s_hmax);
        if (s_flags) atomicOr(&stats[2], s_flags);
    }
}

__global__ void k_set_key(uint32_t *stats, float v) { stats[0] = fkey(v); }

__global__ void k_march_svf_single(const float *__restrict__ a, const float *__restrict__ v1,
                                   const float *__restrict__ v2, const int32_t *__restrict__ dx,
                                   const int32_t *__restrict__ dy, const float *__restrict__ dz,
                                   float *__restrict__ sh, float *__restrict__ vegsh,
                                   float *__restrict__ vbsh, int rows, int cols, int count) {
    const long n = (long)rows * cols;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < n;
    const int r = valid ? (int)(t / cols) : 0;
    const int c = valid ? (int)(t - (long)r * cols) : 0;
    const float at = valid ? a[t] : 0.f;
    const float veg2t = valid ? v2[t] : 0.f;
    float f = at;
    int s_sh = 0, vs = 0, vb = 0;
    for (int s = 0; s < count && valid; ++s) {
        const int sr = r + dx[s], sc = c + dy[s];
        const float dzs = dz[s];
        float tmp = 0.f, tv = 0.f, tv2 = 0.f;
        if (((unsigned)sr < (unsigned)rows) && ((unsigned)sc < (unsigned)cols)) {
            const long q = (long)sr * cols + sc;
            tv = v1[q] - dzs;
            tv2 = v2[q] - dzs;
            tmp = a[q] - dzs;
        }
        f = fmaxf(f, tmp);
        s_sh = (f > at) ? 1 : 0;
        const int g2 = (tv > at ? 1 : 0) - (tv2 > at ? 1 : 0);
        vs = vs > g2 ? vs : g2;
        if (vs * s_sh > 0) vs = 0;
        vb = vs + vb;
        if (s == 0) {
            float fv = tv - tmp;
            if (fv <= 0.f) fv = 1000.f;
            if (fv < dzs) vs = 1;
            vs = vs * ((veg2t > at) ? 1 : 0);
            vb = 0;
        }
    }
    if (!valid) return;
    int b = vb > 0 ? 1 : vb;
    b = b - vs;
    const int vsf = vs > 0 ? 1 : vs;
    sh[t] = (float)(1 - s_sh);
    vegsh[t] = (float)(1 - vsf);
    vbsh[t] = (float)(1 - b);
}

// sun march core shared by the single-step ABI and the day step kernel.
struct SunOut {
    int sh, vs, vb;
    float shvo, shvoveg;  // f - a and the vegetation shadow volume (before * vegsh)
    int stop;             // TRACK: the step after which the lane left the march (count if it never left)
};

template <bool EXITS, bool TRACK = false>
__device__ __forceinline__ SunOut march_sun_pixel(const float4 *__restrict__ hp, int rows, int cols, int r,
                                                  int c, float at, float vegt, bool valid, bool wall,
                                                  const int4 *__restrict__ tab, int count, float amax,
                                                  float hmax) {
    float f = at, svv = vegt;
    int sh = 0, vs = 0, vb = 0;
    int stop = count;
    bool active = valid;
    const bool exit_ok = EXITS && (at >= 0.f) && (vegt >= 0.f);
    for (int s = 0; s < count; ++s) {
        if (EXITS) {
            if (!wave_any(active)) break;
        }
        if (active) {
            const int4 e = tab[s];
            const int sr = r + e.x, sc = c + e.y;
            const float dzs = __int_as_float(e.z), dzp = __int_as_float(e.w);
            const bool inb = ((unsigned)sr < (unsigned)rows) && ((unsigned)sc < (unsigned)cols);
            float tmp = 0.f, tv = 0.f, tv2 = 0.f, lf = 0.f, lg = 0.f;
            if (inb) {
                const float4 h = hp[(long)sr * cols + sc];
                tv = h.y - dzs;
                tv2 = h.z - dzs;
                tmp = h.x - dzs;
                lf = h.y - dzp;
                lg = h.z - dzp;
            }
            f = fmaxf(f, tmp);
            svv = fmaxf(svv, tv);
            sh = (f > at) ? 1 : 0;
            int v2 = (tv > at ? 1 : 0) + (tv2 > at ? 1 : 0) + (lf > at ? 1 : 0) + (lg > at ? 1 : 0);
            if (v2 == 4) v2 = 0;
            if (v2 > 0) v2 = 1;
            vs = vs > v2 ? vs : v2;
            if (vs * sh > 0) vs = 0;
            vb = vb + vs;
            if (EXITS) {
                if (exit_ok && !inb) active = false;
                else if (!wall && sh == 1 && vs == 0) active = false;
                else if (!wall && exit_ok && (hmax - dzp) <= at) active = false;
                if (TRACK && !active) stop = s;
            }
        }
        if (!(amax >= __int_as_float(tab[s].z))) break;
    }
    SunOut o;
    o.sh = sh; o.vs = vs; o.vb = vb; o.shvo = f - at; o.shvoveg = svv; o.stop = stop;
    return o;
}

// The vegetation shadow volume over the steps a lane skipped after its exit (steps stop + 1 ..
// under the same amax rule). The exits are exact for every plane the day loop reads; the
// reference's wallshve plane on a pixel without a wall is (shvoveg - a) * 0, a zero whose sign
// still records whether the complete running max reached the pixel's height.
__device__ __forceinline__ float sun_veg_volume_rest(const float4 *__restrict__ hp, int rows, int cols, int r,
                                                     int c, float svv, const int4 *__restrict__ tab, int count,
                                                     int stop, float amax) {
    for (int s = stop + 1; s < count; ++s) {
        if (!(amax >= __int_as_float(tab[s - 1].z))) break;
        const int4 e = tab[s];
        const int sr = r + e.x, sc = c + e.y;
        float tv = 0.f;
        if (((unsigned)sr < (unsigned)rows) && ((unsigned)sc < (unsigned)cols))
            tv = hp[(long)sr * cols + sc].y - __int_as_float(e.z);
        svv = fmaxf(svv, tv);
    }
    return svv;
}

// ---------------------------------------------------------------------------
// static per-pixel fields of the day loop
// ---------------------------------------------------------------------------
__global__ void k_tile_static(const float *__restrict__ fields, float *__restrict__ svfbuveg,
                              float *__restrict__ tan_asvf, float *__restrict__ tan_sa, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float svf = fields[t], svfveg = fields[5 * n + t];
    svfbuveg[t] = svf - (1.0f - svfveg) * 0.97f;
    tan_asvf[t] = d_tan(d_acos(sqrtf(svf)));
    float tmp = svf + svfveg - 1.0f;
    if (tmp < 0.f) tmp = 0.f;
    tan_sa[t] = d_tan(d_asin(d_exp(d_log(1.0f - tmp) / 2.0f)));
}

// ---------------------------------------------------------------------------
// day step: sun march + wall tail -> walk sources {Lup bits, flags}
// flags: bit0 sh, bit1 vegsh (march outputs), bit2 sunwall, bit3 ground,
//        bits 4..6 land-cover class.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float pow4(float x) {
    const double d = (double)x;
    const double d2 = d * d;
    return (float)(d2 * d2);
}

// Walk-source tile in LDS: a 16 x 16 pixel tile plus the walk halo. Once every wave of the block has
// walked, the step's diffuse-sky patch terms (4 float2 per patch, one per (sh, vegsh) state) take
// its place.
#define FLUX_TX 16
#define FLUX_TY 16

static size_t flux_lds_bytes(int walk_halo) {
    if (walk_halo < 0) return 0;
    const size_t tile = (size_t)(FLUX_TX + 2 * walk_halo) * (FLUX_TY + 2 * walk_halo);   // one flag byte per cell
    const size_t table = (size_t)SW_NPATCH * 4 * sizeof(float2);   // overlays the tile after the walk
    return tile > table ? tile : table;
}

// a / c from the correctly rounded reciprocal y = 1 / c with one Markstein correction: the
// correctly rounded quotient for every finite a with 2^-96 <= |a| < 2^120 (checked for every such
// float32 a and c in {0.9, 2, 3, 9, 11, 12, 18, 22, 23, 44, 45}); a = +-0 gives +0. Three
// instructions instead of the ten of an IEEE division. The operands here are sums of per-cell
// float32 radiation and albedo terms: zero, or far inside that range.
__device__ __forceinline__ float div_rc(float a, float c, float y) {
    const float q = a * y;
    return __builtin_fmaf(__builtin_fmaf(-q, c, a), y, q);
}

// one 256-pixel block of the sun step; `tile` is the block's index within its tile.
// PLANES also writes the reference's eight output planes of the march and wall tail
// (vegsh, sh, vbshvegsh, wallsh, wallsun, wallshve, facesh, facesun) to planes[q * n + t].
template <bool TILED, bool PLANES = false>
__device__ __forceinline__ void sun_step_body(const float4 *__restrict__ hp, const float *__restrict__ walls,
                                              const float *__restrict__ aspect, const uint8_t *__restrict__ cls,
                                              const float *__restrict__ bu, const sw_step_t *__restrict__ steps,
                                              const sw_model_t *__restrict__ model, const int4 *__restrict__ suntab,
                                              const uint32_t *__restrict__ stats, uint2 *__restrict__ src, int step,
                                              int rows, int cols, int tile, float *__restrict__ planes = nullptr) {
    const long n = (long)rows * cols;
    const float amax = funkey(stats[0]);
    const float hmax = funkey(stats[1]);
    long t;
    bool valid;
    int r, c;
    if (TILED) {
        // a 16 x 16 tile of the flux grid, 8 x 8 pixels per wave
        const int tiles_x = (cols + FLUX_TX - 1) / FLUX_TX;
        const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
        const int rr = (tile / tiles_x) * FLUX_TY + (wv >> 1) * 8 + (ln >> 3);
        const int cc = (tile % tiles_x) * FLUX_TX + (wv & 1) * 8 + (ln & 7);
        valid = rr < rows && cc < cols;
        r = valid ? rr : 0;
        c = valid ? cc : 0;
        t = (long)r * cols + c;
    } else {
        t = (long)tile * 256 + threadIdx.x;
        valid = (t < n);
        r = valid ? (int)(t / cols) : 0;
        c = valid ? (int)(t - (long)r * cols) : 0;
    }
    const sw_step_t *sp = steps + step;
    float at = 0.f, vegt = 0.f, w = 0.f;
    if (valid) {
        const float4 h = hp[t];
        at = h.x;
        vegt = h.y;
        w = walls[t];
    }
    const SunOut o = march_sun_pixel<true, PLANES>(hp, rows, cols, r, c, at, vegt, valid, w > 0.f,
                                                   suntab + sp->sun_off, sp->sun_count, amax, hmax);
    if (!valid) return;
    // finalisation (solweig.py:1198-1219)
    const int shv = 1 - o.sh;
    const int vsf = o.vs > 0 ? 1 : o.vs;
    float shvoveg = o.shvoveg;
    if (PLANES && o.stop < sp->sun_count)
        shvoveg = sun_veg_volume_rest(hp, rows, cols, r, c, shvoveg, suntab + sp->sun_off, sp->sun_count, o.stop,
                                      amax);
    const float svv = (shvoveg - at) * (float)vsf;
    const int vso = 1 - vsf;
    const float wallbol = (w > 0.f) ? 1.f : 0.f;
    const float asp = aspect[t];
    float facesh;
    if (sp->facesh_branch == 0) facesh = (float)((asp < sp->azilow) || (asp >= sp->azihigh)) - wallbol + 1.f;
    else facesh = (float)((asp > sp->azilow) || (asp <= sp->azihigh)) * -1.f + 1.f;
    float wallsun = w - o.shvo;
    if (wallsun < 0.f) wallsun = 0.f;
    if (facesh == 1.f) wallsun = 0.f;
    const float wallsh = w - wallsun;
    float wallshve = svv * wallbol;
    wallshve = wallshve - wallsh;
    if (wallshve < 0.f) wallshve = 0.f;
    wallsun = wallsun - wallshve;
    if (wallsun < 0.f) wallsun = 0.f;
    if (PLANES) {
        // the tail's remaining outputs (solweig.py:1207, 1221); the walk reads only wallsun
        const float facesun = ((facesh + wallbol == 1.f) && (w > 0.f)) ? 1.f : 0.f;
        const float wallshve_out = (wallshve > w) ? w : wallshve;
        int vb = o.vb > 0 ? 1 : o.vb;
        vb = vb - o.vs;
        planes[t] = (float)vso;
        planes[n + t] = (float)shv;
        planes[2 * n + t] = (float)(1 - vb);
        planes[3 * n + t] = wallsh;
        planes[4 * n + t] = wallsun;
        planes[5 * n + t] = wallshve_out;
        planes[6 * n + t] = facesh;
        planes[7 * n + t] = facesun;
    }
    // walk sources (sunonsurface_2018a prologue)
    const float b = bu[t];
    const int k = cls[t];
    const float shadow = (float)shv - (1.f - (float)vso) * sp->one_m_psi;
    const float se = model->sbc * model->emis_class[k];
    const float x = sp->tg_class[k] * shadow + sp->ta + 273.15f;
    const float lup = se * pow4(x) - se * sp->tak4;
    const float q = wallsun / w * b;
    const uint32_t sunwall = (q == 1.f) ? 1u : 0u;
    // flag bits 4-10 {sh, vegsh, sunwall, ground, class x3}: the walk reads
    // (flags & 0x7F0) as the byte offset of its 16-byte decode entry
    const uint32_t flags = ((uint32_t)shv | ((uint32_t)vso << 1) | (sunwall << 2) |
                            ((b == 1.f ? 1u : 0u) << 3) | ((uint32_t)k << 4)) << 4;
    src[t] = make_uint2(__float_as_uint(lup), flags);
}

template <bool PLANES>
__global__ void __launch_bounds__(256)
k_sun_step(const float4 *__restrict__ hp, const float *__restrict__ walls,
           const float *__restrict__ aspect, const uint8_t *__restrict__ cls,
           const float *__restrict__ bu, const sw_step_t *__restrict__ steps,
           const sw_model_t *__restrict__ model, const int4 *__restrict__ suntab,
           const uint32_t *__restrict__ stats, uint2 *__restrict__ src, int step, int rows, int cols,
           float *__restrict__ planes) {
    // 16 x 16 pixel tiles, 8 x 8 pixels per wave (exits are spatially correlated)
    const int g = xcd_tile(blockIdx.x, gridDim.x);
    if (g >= ((cols + FLUX_TX - 1) / FLUX_TX) * ((rows + FLUX_TY - 1) / FLUX_TY)) return;
    sun_step_body<true, PLANES>(hp, walls, aspect, cls, bu, steps, model, suntab, stats, src, step, rows, cols, g,
                                planes);
}

// the sun step over listed 16 x 16 tiles
__global__ void __launch_bounds__(256)
k_sun_tiles(const float4 *__restrict__ hp, const float *__restrict__ walls,
            const float *__restrict__ aspect, const uint8_t *__restrict__ cls,
            const float *__restrict__ bu, const sw_step_t *__restrict__ steps,
            const sw_model_t *__restrict__ model, const int4 *__restrict__ suntab,
            const uint32_t *__restrict__ stats, uint2 *__restrict__ src, int step, int rows, int cols,
            const int32_t *__restrict__ tiles, int ntiles) {
    const int g = xcd_tile(blockIdx.x, gridDim.x);
    if (g >= ntiles) return;
    const int tile = tiles[g];
    if (tile < 0 || tile >= ((cols + FLUX_TX - 1) / FLUX_TX) * ((rows + FLUX_TY - 1) / FLUX_TY)) return;
    sun_step_body<true>(hp, walls, aspect, cls, bu, steps, model, suntab, stats, src, step, rows, cols, tile);
}

// The batch kernels read the item table through this mirror of
// sw_batch_item_t whose pointers are declared in the global address space:
// a pointer loaded as a plain (flat) pointer would make every access through
// it a flat load, and the step records would lose their scalar loads.
#define SW_GLOBAL __attribute__((address_space(1)))
struct BatchTileDev {
    int32_t rows, cols;
    const SW_GLOBAL float4 *hpack;
    const SW_GLOBAL float *walls, *aspect_rad, *buildings;
    const SW_GLOBAL uint8_t *cls;
    const SW_GLOBAL uint32_t *stats, *masks;
    const SW_GLOBAL float *svfbuveg, *tan_asvf, *tan_svfalfa;
    const SW_GLOBAL uint32_t *walkface;
    SW_GLOBAL uint2 *src;
    SW_GLOBAL float *state;
};
struct BatchItemDev {
    BatchTileDev tile;
    const SW_GLOBAL sw_step_t *steps;
    const SW_GLOBAL int4 *sun_tab;
    SW_GLOBAL float *out;
    int32_t sun_block0, flux_block0;
};
static_assert(sizeof(BatchItemDev) == sizeof(sw_batch_item_t), "batch item mirror");
static_assert(offsetof(BatchItemDev, steps) == offsetof(sw_batch_item_t, steps), "batch item mirror");
static_assert(offsetof(BatchItemDev, out) == offsetof(sw_batch_item_t, out), "batch item mirror");
static_assert(offsetof(BatchItemDev, flux_block0) == offsetof(sw_batch_item_t, flux_block0), "batch item mirror");
static_assert(offsetof(BatchTileDev, state) == offsetof(sw_tile_t, state), "batch tile mirror");

// Batched launches: one grid covers the blocks of every item, item k owning
// launch-order tiles [block0_k, block0_k + blocks_k) after the XCD remap.
// The owner is found by binary search over the items' (ascending) block0.
template <int FIELD>
__device__ __forceinline__ int batch_owner(const BatchItemDev *__restrict__ items, int nitems, int g) {
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int b0 = FIELD == 0 ? items[mid].sun_block0 : items[mid].flux_block0;
        if (b0 <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256)
k_sun_batch(const BatchItemDev *__restrict__ items, int nitems, int total, const sw_model_t *__restrict__ model,
            int step) {
    const int g = xcd_tile(blockIdx.x, gridDim.x);
    if (g >= total) return;
    const BatchItemDev *it = items + batch_owner<0>(items, nitems, g);
    const BatchTileDev &T = it->tile;
    sun_step_body<true>((const float4 *)T.hpack, (const float *)T.walls, (const float *)T.aspect_rad,
                        (const uint8_t *)T.cls, (const float *)T.buildings, (const sw_step_t *)it->steps, model,
                        (const int4 *)it->sun_tab, (const uint32_t *)T.stats, (uint2 *)T.src, step, T.rows, T.cols,
                        g - it->sun_block0);
}

// ---------------------------------------------------------------------------
// UTCI
// ---------------------------------------------------------------------------
__device__ __forceinline__ float utci_point(float ta, float rh, float tmrt, float va) {
    if (ta <= -999.f || rh <= -999.f || va <= -999.f || tmrt <= -999.f) return -999.f;
    const float tk = ta + 273.15f;
    float es = 2.7150305f * d_log(tk);
    es = es + -2.8365744E3f * (1.f / (tk * tk));
    es = es + -6.028076559E3f * (1.f / tk);
    es = es + 1.954263612E1f * 1.f;
    es = es + -2.737830188E-2f * tk;
    es = es + 1.6261698E-5f * (tk * tk);
    es = es + 7.0229056E-10f * (tk * tk * tk);
    es = es + -1.8680009E-13f * pow456(tk).p4;
    es = d_exp(es) * 0.01f;
    const float ehPa = es * rh / 100.f;
    const float dtm = tmrt - ta;
    const float pa = ehPa / 10.0f;
    const Pow456 qt = pow456(ta), qv = pow456(va), qd = pow456(dtm), qp = pow456(pa);
    const float ta1 = ta, ta2 = ta * ta, ta3 = ta * ta * ta, ta4 = qt.p4, ta5 = qt.p5, ta6 = qt.p6;
    const float va1 = va, va2 = va * va, va3 = va * va * va, va4 = qv.p4, va5 = qv.p5, va6 = qv.p6;
    const float dt1 = dtm, dt2 = dtm * dtm, dt3 = dtm * dtm * dtm, dt4 = qd.p4, dt5 = qd.p5, dt6 = qd.p6;
    const float pa1 = pa, pa2 = pa * pa, pa3 = pa * pa * pa, pa4 = qp.p4, pa5 = qp.p5, pa6 = qp.p6;
    float u = ta;
#include "utci_poly.inc"
    (void)ta6; (void)va6; (void)dt6; (void)pa6;
    return u;
}

// ---------------------------------------------------------------------------
// WBGT (utci_process.py:1021-1033, black globe calculate_wbgt.py:717-761)
// ---------------------------------------------------------------------------
// Black-globe temperature from the closed-form root of the globe heat balance
// Tg^4 + a Tg = (Tmrt + 273.15)^4 + a (Ta + 273.15), a = h_cg / (eps sigma),
// in float32 with one rounding per operation; powers are correctly rounded
// (through double). Large a overflows n to inf and Tg to -inf, as float32
// arithmetic does. The clamps let NaN through. WBGT blends the step's
// f32(0.7 Tw) with the globe (and air) temperature by the sun/shade test
// shadow < 0.1.
__device__ __forceinline__ float d_pow(float x, float y) { return (float)pow((double)x, (double)y); }

__device__ __forceinline__ float wbgt_point(float tmrt, float shadow, float ta, float w07, float a) {
    const float b = d_pow(tmrt + 273.15f, 4.f) + a * (ta + 273.15f);
    const float m = 9.f * (a * a);
    const float n = 27.f * d_pow(a, 4.f);
    const float p = 256.f * ((b * b) * b);
    const float E = d_pow(m + 1.73205f * sqrtf(n + p), 0.333333343f);
    const float Q = 3.4943f * b;
    const float k = 0.381571f * E - Q / E;
    const float ks = (k < 1.17549435e-38f) ? 1.17549435e-38f : k;
    float inner = 2.f * a / sqrtf(ks) - k;
    inner = (inner < 0.f) ? 0.f : inner;
    const float kz = (k < 0.f) ? 0.f : k;
    const float tg = 0.5f * sqrtf(inner) - 0.5f * sqrtf(kz) - 273.15f;
    return (shadow < 0.1f) ? w07 + 0.3f * tg : (w07 + 0.2f * tg) + 0.1f * ta;
}

__global__ void k_utci(const float *__restrict__ ta, const float *__restrict__ rh,
                       const float *__restrict__ tmrt, const float *__restrict__ va,
                       float *__restrict__ out, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = utci_point(ta[i], rh[i], tmrt[i], va[i]);
}

// ---------------------------------------------------------------------------
// flux step
// ---------------------------------------------------------------------------
// Pass A over the 153 patches (Kside_veg_v2022a patch sums, solweig.py:564-832,
// and the first define_patch_characteristics pass, solweig.py:1439-1648).
struct PassA {
    float aniLum, KsideD, Kref_sun, Kref_sh, Kref_veg;
    float Ldown_sky, Ldown_veg, Ldown_sun, Ldown_sh;
    float Lside_sky, Lside_veg, Lside_sun, Lside_sh;
};

// The diffuse-sky terms of a day step per patch and (sh, vegsh) state st = sh | vegsh << 1:
// {diffsh * lv, diffsh * lumchi * angl * ster}, diffsh = sh - (1 - vegsh) * 0.97 (solweig.py:2069-2073,
// Kside_veg_v2022a), each product rounded in the reference's order. One LDS read per patch
// replaces the state select and four products.
__device__ __forceinline__ void diffuse_table(const sw_step_t *__restrict__ sp, float2 *__restrict__ tab) {
    const float d00 = 0.f - (1.f - 0.f) * 0.97f, d10 = 1.f - (1.f - 0.f) * 0.97f;  // diffsh for vegsh == 0
    for (int i = threadIdx.x; i < SW_NPATCH * 4; i += blockDim.x) {
        const int p = i >> 2, st = i & 3;
        const float diffsh = (st & 2) ? ((st & 1) ? 1.f : 0.f) : ((st & 1) ? d10 : d00);
        const sw_patch_k_t &q = sp->pk[p];
        tab[i] = make_float2(diffsh * q.lv, diffsh * q.lumchi * q.angl * q.ster);
    }
}

template <bool DAY, bool ROLLED>
__device__ __forceinline__ void patch_pass_a(const sw_step_t *__restrict__ sp, const uint32_t *msh,
                                             const uint32_t *mvs, const uint32_t *mv1, const uint32_t *mv2,
                                             float hsvf, const float2 *__restrict__ s_diff, PassA &o) {
    float aniLum = 0.f, KsideD = 0.f, Kref_sun = 0.f, Kref_sh = 0.f, Kref_veg = 0.f;
    float Ldown_sky = 0.f, Ldown_veg = 0.f, Ldown_sun = 0.f, Ldown_sh = 0.f;
    float Lside_sky = 0.f, Lside_veg = 0.f, Lside_sun = 0.f, Lside_sh = 0.f;
#pragma unroll
    for (int w = 0; w < SW_MASK_WORDS; ++w) {
        const uint32_t ash = w == 0 ? msh[0] : w == 1 ? msh[1] : w == 2 ? msh[2] : w == 3 ? msh[3] : msh[4];
        const uint32_t avs = w == 0 ? mvs[0] : w == 1 ? mvs[1] : w == 2 ? mvs[2] : w == 3 ? mvs[3] : mvs[4];
        const uint32_t av1 = w == 0 ? mv1[0] : w == 1 ? mv1[1] : w == 2 ? mv1[2] : w == 3 ? mv1[3] : mv1[4];
        const uint32_t av2 = w == 0 ? mv2[0] : w == 1 ? mv2[1] : w == 2 ? mv2[2] : w == 3 ? mv2[3] : mv2[4];
        // per-word patch booleans, bit k = patch w * 32 + k: tvs = vegsh == 0 or vbsh == 0;
        // tsh = (1 - sh) * vbsh == 1; tsky = sh and vegsh
        const uint32_t btvs = ~avs | ~(av1 | av2), btsh = ~ash & av1, btsky = ash & avs;
        const int pend = min(w * 32 + 32, SW_NPATCH);
        auto patch = [&](int p) {
            const uint32_t k = (uint32_t)(p - w * 32);
            const float tvs = (float)__builtin_amdgcn_ubfe(btvs, k, 1u);
            const float tsky = (float)__builtin_amdgcn_ubfe(btsky, k, 1u);
            const bool tsh = __builtin_amdgcn_ubfe(btsh, k, 1u) != 0u;
            const sw_patch_k_t q = sp->pk[p];
            const sw_patch_l_t u = sp->pl[p];
            // sum + c * b with b in {0, 1}: fma(b, c, sum) rounds once, like the reference
            if (DAY) {
                const uint32_t st = __builtin_amdgcn_ubfe(ash, k, 1u) | (__builtin_amdgcn_ubfe(avs, k, 1u) << 1);
                const float2 dv = s_diff[4 * p + (int)st];
                aniLum = aniLum + dv.x;
                KsideD = KsideD + dv.y;
                Kref_veg = __builtin_fmaf(tvs, q.k_sh, Kref_veg);
            }
            Ldown_sky = __builtin_fmaf(tsky, q.lsky_down, Ldown_sky);
            Lside_sky = __builtin_fmaf(tsky, q.lsky_side, Lside_sky);
            Lside_veg = __builtin_fmaf(tvs, q.lv_side, Lside_veg);
            Ldown_veg = __builtin_fmaf(tvs, q.lv_down, Ldown_veg);
            // wall-patch terms need tsh; a wave whose lanes all have tsh == 0 adds
            // only zeros there and skips them
            if (wave_any(tsh)) {
                if (DAY) {
                    const float x = hsvf + q.yi_k;
                    Kref_sun = __builtin_fmaf((tsh && x < q.sun_lt) ? 1.f : 0.f, q.k_sun, Kref_sun);
                    Kref_sh = __builtin_fmaf((tsh && x > q.sun_gt) ? 1.f : 0.f, q.k_sh, Kref_sh);
                }
                // sun-facing window: sunlit / shaded by the threshold test; elsewhere every
                // shaded wall patch counts as shaded and none as sunlit
                const float xl = hsvf + u.yi_l;
                const bool ls = q.l_sun != 0;
                const float sunlit = (tsh && ls && (xl < q.sun_lt)) ? 1.f : 0.f;
                const float shaded = (tsh && (!ls || (xl > q.sun_gt))) ? 1.f : 0.f;
                Lside_sun = __builtin_fmaf(sunlit, u.ls_side, Lside_sun);
                Lside_sh = __builtin_fmaf(shaded, u.lh_side, Lside_sh);
                Ldown_sun = __builtin_fmaf(sunlit, u.ls_down, Ldown_sun);
                Ldown_sh = __builtin_fmaf(shaded, u.lh_down, Ldown_sh);
            }
        };
        // ROLLED (the walking kernel, which holds both passes): no unrolling. Unrolled, its
        // iterations keep several patches' scalar constants live at once; the register
        // allocator then spills SGPRs to VGPR lanes and holds 114 VGPRs (4 waves per SIMD).
        // Rolled, it needs 68 VGPRs (7 waves per SIMD), spills nothing and runs 7 % faster.
        // The night-only kernel keeps 4 (its rolled form is slower).
        if (ROLLED) {
#pragma unroll 1
            for (int p = w * 32; p < pend; ++p) patch(p);
        } else {
#pragma unroll 4
            for (int p = w * 32; p < pend; ++p) patch(p);
        }
    }
    o.aniLum = aniLum; o.KsideD = KsideD; o.Kref_sun = Kref_sun; o.Kref_sh = Kref_sh; o.Kref_veg = Kref_veg;
    o.Ldown_sky = Ldown_sky; o.Ldown_veg = Ldown_veg; o.Ldown_sun = Ldown_sun; o.Ldown_sh = Ldown_sh;
    o.Lside_sky = Lside_sky; o.Lside_veg = Lside_veg; o.Lside_sun = Lside_sun; o.Lside_sh = Lside_sh;
}

// Ground-view walk of one azimuth (sunonsurface_2018a, solweig.py:136-288).
// Every boolean of the walk is a float 0 or 1, as in the reference: the
// running minimum of building flags is a product f, the "wall seen" flag is
// a maximum, and "wall or building seen" (bubw) equals 1 - f because f never
// rises again, so its count is the step count minus the sum of f. Sums that add a constant (walls seen, their longwave and albedo
// terms) are kept as counts k and read from the sequential-sum tables
// s_tlw[k] / s_tal[k] afterwards, which is the same float32 value the
// step-by-step additions produce. The other sums use fma(flag, value, sum)
// with flag exactly 0 or 1: the product is exact, so the single rounding
// equals the reference's "sum + value * flag". A source that left the grid
// keeps the last in-grid temporaries (stale border); INTERIOR tiles have
// every source in the grid. Once every lane of the wave has met a building
// (flag product 0) the remaining steps only add to the counts, in closed
// form.
struct WalkAcc {
    float kb;                   // steps with a wall seen (bub): an exact small integer
    float fs;                   // sum of the flag product f; steps with bubw = steps - fs
    float wlupsh, walbsh, walbnosh;
};

template <bool INTERIOR>
__device__ __forceinline__ void walk_azimuth(const uint8_t *__restrict__ s_src, const float4 *__restrict__ s_dec,
                                             const sw_model_t *__restrict__ md, int j, int nsec, int nfirst,
                                             int lbase, int lw, int r, int c, int rows, int cols, float fb,
                                             WalkAcc &acc, WalkAcc &snap) {
    float f = fb, bub = 0.f;
    float tl = 0.f, ta = 0.f, tn = 0.f, tb = 0.f, tw = 0.f;
    bool live = true;
    acc.kb = 0.f; acc.fs = 0.f; acc.wlupsh = 0.f; acc.walbsh = 0.f; acc.walbnosh = 0.f;
    const int32_t *wdx = md->walk_dx + j * SW_WALK_MAX, *wdy = md->walk_dy + j * SW_WALK_MAX;
#define SW_WALK_STEP(S)                                                                      \
    {                                                                                        \
        const int ddx = wdx[S], ddy = wdy[S];                                                \
        if (!INTERIOR)                                                                       \
            live = live && ((unsigned)(r + ddx) < (unsigned)rows) && ((unsigned)(c + ddy) < (unsigned)cols); \
        if (INTERIOR || live) {                                                              \
            const float4 d = s_dec[s_src[lbase + ddx * lw + ddy]];                          \
            tl = d.x; ta = d.y; tn = fabsf(d.z); tw = d.w;                                   \
            tb = (__float_as_int(d.z) >= 0) ? 1.f : 0.f;                                     \
        }                                                                                    \
        f = f * tb;                                                                          \
        acc.wlupsh = __builtin_fmaf(f, tl, acc.wlupsh);                                      \
        acc.walbsh = __builtin_fmaf(f, ta, acc.walbsh);                                      \
        acc.walbnosh = __builtin_fmaf(f, tn, acc.walbnosh);                                  \
        bub = fmaxf(bub, tw * f);                                                            \
        acc.kb += bub;                                                                       \
        acc.fs += f;                                                                         \
    }
    int s = 0;
    if (INTERIOR) {
        // Every source is in the tile, so the reads of a step depend on nothing computed in the
        // walk: four steps' reads are issued together (two dependent LDS reads per step, whose
        // latency otherwise stalls the wave), then the four steps are accumulated in order. After
        // the wave's flag products are all 0, a further step adds exactly bub to kb and +0 to every
        // other sum (0 * finite source), which is what the closed-form finish below adds for it,
        // so testing for the exit once per group changes no result.
#define SW_WALK_ACC(E, D)                                                                    \
    {                                                                                        \
        f = (__float_as_int(D.z) >= 0) ? f : 0.f;                                            \
        acc.wlupsh = __builtin_fmaf(f, D.x, acc.wlupsh);                                     \
        acc.walbsh = __builtin_fmaf(f, D.y, acc.walbsh);                                     \
        acc.walbnosh = __builtin_fmaf(f, fabsf(D.z), acc.walbnosh);                          \
        bub = fmaxf(bub, D.w * f);                                                           \
        acc.kb += bub;                                                                       \
        acc.fs += f;                                                                         \
    }
#define SW_WALK_GROUP(S)                                                                     \
    {                                                                                        \
        const uint32_t e0 = s_src[lbase + wdx[S] * lw + wdy[S]];                             \
        const uint32_t e1 = s_src[lbase + wdx[S + 1] * lw + wdy[S + 1]];                     \
        const uint32_t e2 = s_src[lbase + wdx[S + 2] * lw + wdy[S + 2]];                     \
        const uint32_t e3 = s_src[lbase + wdx[S + 3] * lw + wdy[S + 3]];                     \
        const float4 d0 = s_dec[e0], d1 = s_dec[e1], d2 = s_dec[e2], d3 = s_dec[e3];          \
        SW_WALK_ACC(e0, d0) SW_WALK_ACC(e1, d1) SW_WALK_ACC(e2, d2) SW_WALK_ACC(e3, d3)      \
    }
        for (; s + 4 <= nfirst; s += 4) SW_WALK_GROUP(s)
        for (; s < nfirst; ++s) SW_WALK_STEP(s)
        snap = acc;
        for (; s + 4 <= nsec; s += 4) {
            if (!wave_any(f != 0.f)) break;
            SW_WALK_GROUP(s)
        }
        for (; s < nsec; ++s) {
            if (!wave_any(f != 0.f)) break;
            SW_WALK_STEP(s)
        }
#undef SW_WALK_GROUP
#undef SW_WALK_ACC
    } else {
        for (; s < nfirst; ++s) SW_WALK_STEP(s)
        snap = acc;
#pragma unroll 2
        for (; s < nsec; ++s) {
            if (!wave_any(f != 0.f)) break;
            SW_WALK_STEP(s)
        }
    }
#undef SW_WALK_STEP
    // every lane has f == 0 here (or the walk is complete): bub stays and f
    // adds nothing for the nsec - s steps left
    acc.kb += bub * (float)(nsec - s);
}

// WALK = false instantiates a night-only kernel (no walk tile, no shortwave
// code, fewer registers); it refuses a day step with NaN outputs.
// one 16 x 16 tile of the flux step; `tile` is the block's tile index within its tile grid
template <bool WALK>
__device__ __forceinline__ void flux_body(const uint32_t *__restrict__ masks, const uint2 *__restrict__ src,
                                          const float *__restrict__ svfbuveg, const float *__restrict__ tan_asvf,
                                          const float *__restrict__ tan_sa, const uint32_t *__restrict__ walkface,
                                          const float *__restrict__ bu, const uint8_t *__restrict__ cls,
                                          const sw_step_t *__restrict__ steps, const sw_model_t *__restrict__ model,
                                          float *__restrict__ state, float *__restrict__ out, int out_mask, int step,
                                          int halo, int rows, int cols, int tile) {
    extern __shared__ uint2 s_src[];
    __shared__ float s_alb[SW_NCLASS];
    __shared__ float4 s_dec[128];                 // walk source decode by flag bits 0-6: {alb * shadow, alb, ground, sunwall}
    __shared__ float s_tlw[SW_WALK_MAX + 1];      // k-fold sequential sums of lwall
    __shared__ float s_tal[SW_WALK_MAX + 1];      // k-fold sequential sums of albedo_b
    const sw_step_t *sp = steps + step;
    const sw_model_t *md = model;
    const int tiles_x = (cols + FLUX_TX - 1) / FLUX_TX;
    const int tiles_y = (rows + FLUX_TY - 1) / FLUX_TY;
    if (tile >= tiles_x * tiles_y) return;  // whole block idle (uniform)
    const int r0 = (tile / tiles_x) * FLUX_TY, c0 = (tile % tiles_x) * FLUX_TX;
    // each wave covers an 8 x 8 square of the 16 x 16 tile
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    const int ty = (wv >> 1) * 8 + (ln >> 3), tx = (wv & 1) * 8 + (ln & 7);
    const int r = r0 + ty, c = c0 + tx;
    const bool valid = (r < rows) && (c < cols);
    const long n = (long)rows * cols;
    const long t = valid ? (long)r * cols + c : 0;
    const bool day = WALK && sp->is_day != 0;
    const int lw = FLUX_TX + 2 * halo, lh = FLUX_TY + 2 * halo;
    const float omp = sp->one_m_psi;
    if (threadIdx.x < SW_NCLASS) s_alb[threadIdx.x] = md->alb_class[threadIdx.x];
    // halo < walk_halo (incl. the tile-less launch, halo = -1) refuses a day step below
    const bool staged = day && halo >= md->walk_halo;
    if (staged) {
        // stage the walk sources of the tile + halo (out-of-grid cells are
        // never read: the walk tracks liveness on global coordinates)
        for (int i = threadIdx.x; i < lw * lh; i += blockDim.x) {
            const int ly = i / lw, lx = i - ly * lw;
            const int gr = r0 - halo + ly, gc = c0 - halo + lx;
            uint32_t v = 0u;
            if ((unsigned)gr < (unsigned)rows && (unsigned)gc < (unsigned)cols)
                v = ((const uint32_t *)src)[2 * ((long)gr * cols + gc) + 1];
            ((uint8_t *)s_src)[i] = (uint8_t)(v >> 4);
        }
        if (threadIdx.x < 128) {
            // decode of a walk source's flag byte {sh, vegsh, sunwall, ground, class}: its Lup source
            // (the sun step's expression, so the same bits), albedo x shadow, albedo signed by the
            // ground flag (the sign bit carries it, also for albedo 0), sunwall
            const int shb = threadIdx.x & 1, vsb = (threadIdx.x >> 1) & 1, k = (threadIdx.x >> 4) & 7;
            const float al = md->alb_class[k];
            const float shd = (float)shb - (1.f - (float)vsb) * omp;
            const float se = md->sbc * md->emis_class[k];
            const float xx = sp->tg_class[k] * shd + sp->ta + 273.15f;
            const float tl = se * pow4(xx) - se * sp->tak4;
            s_dec[threadIdx.x] = make_float4(tl, al * shd, ((threadIdx.x >> 3) & 1) ? al : -al,
                                             (float)((threadIdx.x >> 2) & 1));
        }
        if (threadIdx.x == 128) {
            float a1 = 0.f, a2 = 0.f;
            const float lwall = sp->lwall, albb = md->albedo_b;
            s_tlw[0] = 0.f; s_tal[0] = 0.f;
            for (int k = 1; k <= SW_WALK_MAX; ++k) {
                a1 = a1 + lwall;
                a2 = a2 + albb;
                s_tlw[k] = a1;
                s_tal[k] = a2;
            }
        }
    }
    __syncthreads();
    // lanes off the grid stay to the end (the block meets a barrier after the walk); they read
    // pixel 0 and store nothing
    if ((!WALK && sp->is_day) || (day && !staged)) {
        // the shared tile is smaller than the walk reach: refuse loudly (uniform over the block)
        int k = 0;
        if (valid)
            for (int b = 1; b <= 128; b <<= 1)
                if (out_mask & b) out[(long)(k++) * n + t] = __builtin_nanf("");
        return;
    }
    const float b_t = bu[t];
    const int k_t = cls[t];
    const float se_t = md->sbc * md->emis_class[k_t];
    float Kdown = 0.f, Kup = 0.f, KupE = 0.f, KupS = 0.f, KupW = 0.f, KupN = 0.f, Kside = 0.f;
    float Lup, LupE, LupS, LupW, LupN, shadow = 0.f;
    float F_sh = 0.f, galb = 0.f, gnosh = 0.f, galbE = 0.f, gnoshE = 0.f, galbS = 0.f, gnoshS = 0.f,
          galbW = 0.f, gnoshW = 0.f, galbN = 0.f, gnoshN = 0.f;
    if (day) {
        const uint2 me = src[t];
        const uint32_t fl_t = me.y;
        const float lup_t = __uint_as_float(me.x);
        shadow = (float)((fl_t >> 4) & 1u) - (1.f - (float)((fl_t >> 5) & 1u)) * omp;
        const float alb_t = s_alb[k_t];
        // ---- ground view factors (gvf_2018a): 18 azimuths x nsec steps ----
        float gL = 0.f, gA = 0.f, gN = 0.f, gLE = 0.f, gAE = 0.f, gNE = 0.f, gLS = 0.f, gAS = 0.f, gNS = 0.f;
        float gLW = 0.f, gAW = 0.f, gNW = 0.f, gLN = 0.f, gAN = 0.f, gNN = 0.f;
        const int nsec = md->nsec;
        const float first = md->first, second = md->second;
        int nfirst = 0;
        while (nfirst < nsec && (float)(nfirst + 1) <= first) ++nfirst;
        const float fp1 = first + 1.f, sp1 = second + 1.f;
        const float r_fp1 = 1.f / fp1, r_sp1 = 1.f / sp1, r_first = 1.f / first, r_second = 1.f / second;
        const float r_09 = 1.f / 0.9f, r_18 = 1.f / 18.f, r_9 = 1.f / 9.f;
        const uint32_t wf = walkface[t];
        const float nb = b_t * -1.f + 1.f;
        const float fb = b_t;                    // buildings is 0/1
        const int lbase = (ty + halo) * lw + (tx + halo);
        const int wh = md->walk_halo;
        const bool interior = (r0 >= wh) && (r0 + FLUX_TY + wh <= rows) && (c0 >= wh) && (c0 + FLUX_TX + wh <= cols);
        for (int j = 0; j < SW_WALK_AZ; ++j) {
            WalkAcc A, F;
            if (interior)
                walk_azimuth<true>((const uint8_t *)s_src, s_dec, md, j, nsec, nfirst, lbase, lw, r, c, rows, cols,
                                   fb, A, F);
            else
                walk_azimuth<false>((const uint8_t *)s_src, s_dec, md, j, nsec, nfirst, lbase, lw, r, c, rows,
                                    cols, fb, A, F);
            const int kb = (int)A.kb, kw = nsec - (int)A.fs, fkb = (int)F.kb, fkw = nfirst - (int)F.fs;
            const float wwall = A.kb, wlwall = s_tlw[kb], walbwall = s_tal[kb];
            const float walbwallnosh = s_tal[kw];
            const float f_wwall = F.kb, f_wlwall = s_tlw[fkb], f_walbwall = s_tal[fkb];
            const float f_walbwallnosh = s_tal[fkw];
            // keep == 1  <=>  wwall == second and facesh == 0 (bit j of walkface)
            const bool keep1 = (wwall == second) && ((wf >> j) & 1u);
            const float wsi1 = (f_wwall > 0.f) ? 1.f : 0.f;
            const float wi1 = (f_walbwallnosh > 0.f) ? 1.f : 0.f;
            const float wsi2 = (wwall > 0.f) ? 1.f : 0.f;
            const float wi2 = (walbwallnosh > 0.f) ? 1.f : 0.f;
            const float lw2 = keep1 ? 0.f : wlwall;
            const float aw2 = keep1 ? 0.f : walbwall;
            // the reference's (x / c1) * w + (y / c2) * (1 - w) with w in {0, 1}: one of the two
            // products is a zero added to the other, so only the selected quotient is formed
            const float gl1 = wsi1 != 0.f ? div_rc(f_wlwall + F.wlupsh, fp1, r_fp1) : div_rc(F.wlupsh, first, r_first);
            const float gl2 = wsi2 != 0.f ? div_rc(lw2 + A.wlupsh, sp1, r_sp1) : div_rc(A.wlupsh, second, r_second);
            const float ga1 = wsi1 != 0.f ? div_rc(f_walbwall + F.walbsh, fp1, r_fp1) : div_rc(F.walbsh, first, r_first);
            const float ga2 = wsi2 != 0.f ? div_rc(aw2 + A.walbsh, sp1, r_sp1) : div_rc(A.walbsh, second, r_second);
            const float gn1 = wi1 != 0.f ? div_rc(f_walbwallnosh + F.walbnosh, fp1, r_fp1)
                                         : div_rc(F.walbnosh, first, r_first);
            const float gn2 = div_rc(wi2 != 0.f ? walbwallnosh + A.walbnosh : A.walbnosh, second, r_second);
            float gLj = div_rc(gl1 * 0.5f + gl2 * 0.4f, 0.9f, r_09);
            gLj = gLj + lup_t * nb;
            float gAj = div_rc(ga1 * 0.5f + ga2 * 0.4f, 0.9f, r_09);
            gAj = gAj + alb_t * nb * shadow;
            float gNj = div_rc(gn1 * 0.5f + gn2 * 0.4f, 0.9f, r_09);
            gNj = gNj * b_t + alb_t * nb;
            gL += gLj; gA += gAj; gN += gNj;
            const int dm = md->walk_dir[j];
            if (dm & 1) { gLE += gLj; gAE += gAj; gNE += gNj; }
            if (dm & 2) { gLS += gLj; gAS += gAj; gNS += gNj; }
            if (dm & 4) { gLW += gLj; gAW += gAj; gNW += gNj; }
            if (dm & 8) { gLN += gLj; gAN += gAj; gNN += gNj; }
        }
        const float lup_air = se_t * sp->tak4;
        const float gLup = div_rc(gL, 18.f, r_18) + lup_air;
        galb = div_rc(gA, 18.f, r_18); gnosh = div_rc(gN, 18.f, r_18);
        const float gLupE = div_rc(gLE, 9.f, r_9) + lup_air;
        galbE = div_rc(gAE, 9.f, r_9); gnoshE = div_rc(gNE, 9.f, r_9);
        const float gLupS = div_rc(gLS, 9.f, r_9) + lup_air;
        galbS = div_rc(gAS, 9.f, r_9); gnoshS = div_rc(gNS, 9.f, r_9);
        const float gLupW = div_rc(gLW, 9.f, r_9) + lup_air;
        galbW = div_rc(gAW, 9.f, r_9); gnoshW = div_rc(gNW, 9.f, r_9);
        const float gLupN = div_rc(gLN, 9.f, r_9) + lup_air;
        galbN = div_rc(gAN, 9.f, r_9); gnoshN = div_rc(gNN, 9.f, r_9);
        // ---- thermal delay (TsWaveDelay_2015a x 6, shared weights) ----
        const float tg0[6] = {gLup, gLupE, gLupS, gLupW, gLupN, sp->tg_class[k_t] * shadow + sp->ta};
        float L[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            float T1 = sp->tw_first ? tg0[q] : state[(long)q * n + t];
            if (sp->tw_update) {
                T1 = tg0[q] * sp->tw_omw + T1 * sp->tw_w;
                L[q] = T1;
            } else {
                L[q] = tg0[q] * sp->tw_omw + T1 * sp->tw_w;
            }
            if (valid) state[(long)q * n + t] = T1;
        }
        Lup = L[0]; LupE = L[1]; LupS = L[2]; LupW = L[3]; LupN = L[4];
        // ---- cylinder wedge F_sh (cylindric_wedge) ----
        const float tan_a = tan_sa[t], tb = sp->tan_zen;
        const float xa = 1.f - 2.f / (tan_a * tb);
        const float ha = 2.f / (tan_a * tb);
        const float ba = 1.f / tan_a;
        const float hkil = 2.f * ba * ha;
        float ukil = 0.f;
        if (xa < 0.f) {
            const float qa = tb / 2.f;
            const float Za = sqrtf(ba * ba - (qa * qa) / 4.f);
            const float phi = d_atan(Za / qa);
            const float cphi = d_cos(phi);
            const float A = (d_sin(phi) - phi * cphi) / (1.f - cphi);
            ukil = 2.f * ba * xa * A;
        }
        const float Ssurf = hkil + ukil;
        const float two_pi = 6.28318530717958647692f;
        F_sh = (two_pi * ba - Ssurf) / (two_pi * ba);
        if (F_sh != F_sh) F_sh = 0.5f;
    } else {
        const float x = sp->ta + 273.15f;
        Lup = se_t * pow4(x);
        LupE = LupS = LupW = LupN = Lup;
    }
    // ---- pass A over the 153 patches: diffuse sky + Kside (day) and the
    // first define_patch_characteristics pass (every step). Products of
    // step and patch constants come precomputed (k_sh, k_sun, lv_*, ls_*,
    // lh_*) and are added where the patch's boolean factors are 1, which is
    // the value the full product takes (adding +0 elsewhere changes no sum
    // that starts at +0 and only grows); each accumulator keeps the
    // reference's patch order ----
    if (staged) {
        // every wave of the block is done with the walk tile: the step's diffuse-sky table replaces it
        __syncthreads();
        diffuse_table(sp, (float2 *)s_src);
        __syncthreads();
    }
    uint32_t msh[SW_MASK_WORDS], mvs[SW_MASK_WORDS], mv1[SW_MASK_WORDS], mv2[SW_MASK_WORDS];
    load_masks(masks, n, t, msh, mvs, mv1, mv2);
    const float hsvf = tan_asvf[t];
    PassA pa;
    if (day) patch_pass_a<true, WALK>(sp, msh, mvs, mv1, mv2, hsvf, (const float2 *)s_src, pa);
    else patch_pass_a<false, WALK>(sp, msh, mvs, mv1, mv2, hsvf, nullptr, pa);
    const float aniLum = pa.aniLum, KsideD = pa.KsideD, Kref_sun = pa.Kref_sun, Kref_sh = pa.Kref_sh,
                Kref_veg = pa.Kref_veg;
    const float Ldown_sky = pa.Ldown_sky, Ldown_veg = pa.Ldown_veg, Ldown_sun = pa.Ldown_sun, Ldown_sh = pa.Ldown_sh;
    const float Lside_sky = pa.Lside_sky, Lside_veg = pa.Lside_veg, Lside_sun = pa.Lside_sun, Lside_sh = pa.Lside_sh;
    float Ldown_ref = 0.f, Lside_ref = 0.f;
    if (day) {
        const float dRad = aniLum * sp->radD;
        const float sb = 1.f - svfbuveg[t];
        const float mix = sp->radG * (1.f - F_sh) + sp->radD * F_sh;
        Kdown = sp->radI * shadow * sp->sinalt + dRad + md->albedo_b * sb * mix;
        const float kdiff = sp->radD * svfbuveg[t] + md->albedo_b * sb * mix;
        Kup = galb * sp->radI * sp->sinalt + kdiff * gnosh;
        KupE = galbE * sp->radI * sp->sinalt + kdiff * gnoshE;
        KupS = galbS * sp->radI * sp->sinalt + kdiff * gnoshS;
        KupW = galbW * sp->radI * sp->sinalt + kdiff * gnoshW;
        KupN = galbN * sp->radI * sp->sinalt + kdiff * gnoshN;
        const float KsideI = shadow * sp->radI * sp->cosalt_k;
        Kside = KsideI + KsideD + Kref_sun + Kref_sh + Kref_veg;
    }
    // ---- pass B: reflected longwave ----
    const float pi_f = 3.14159265358979323846f;
    const float refl = ((Ldown_sky + Lup) * (1.f - md->ewall) * 0.5f) / pi_f;
#pragma unroll
    for (int w = 0; w < SW_MASK_WORDS; ++w) {
        // reflected term uses patches where any of sh, vegsh, vbsh is 0
        const uint32_t tr = ~((w == 0 ? msh[0] : w == 1 ? msh[1] : w == 2 ? msh[2] : w == 3 ? msh[3] : msh[4]) &
                        (w == 0 ? mvs[0] : w == 1 ? mvs[1] : w == 2 ? mvs[2] : w == 3 ? mvs[3] : mvs[4]) &
                        ((w == 0 ? mv1[0] : w == 1 ? mv1[1] : w == 2 ? mv1[2] : w == 3 ? mv1[3] : mv1[4]) |
                         (w == 0 ? mv2[0] : w == 1 ? mv2[1] : w == 2 ? mv2[2] : w == 3 ? mv2[3] : mv2[4])));
        const int pend = min(w * 32 + 32, SW_NPATCH);
#pragma unroll 4
        for (int p = w * 32; p < pend; ++p) {
            const float trf = (float)__builtin_amdgcn_ubfe(tr, (uint32_t)(p - w * 32), 1u);
            const sw_patch_k_t q = sp->pk[p];
            // sum + ((refl * ster) * c) * trf with trf in {0, 1}: the product rounded in the
            // reference's order, then added by an fma whose flag product is exact
            const float rs = refl * q.ster;
            Lside_ref = __builtin_fmaf(trf, rs * q.cos_alt, Lside_ref);
            Ldown_ref = __builtin_fmaf(trf, rs * q.sin_alt, Ldown_ref);
        }
    }
    const float Lside = Lside_sky + Lside_veg + Lside_sh + Lside_sun + Lside_ref;
    const float Ldown = Ldown_sky + Ldown_veg + Ldown_sh + Ldown_sun + Ldown_ref;
    // ---- Tmrt (standing cylinder, anisotropic sky) ----
    const float Keast = KupE * 0.5f, Kwest = KupW * 0.5f, Knorth = KupN * 0.5f, Ksouth = KupS * 0.5f;
    const float Least = LupE * 0.5f, Lwest = LupW * 0.5f, Lnorth = LupN * 0.5f, Lsouth = LupS * 0.5f;
    const float Sstr = md->absK * (Kside * md->Fcyl + (Kdown + Kup) * md->Fup + (Knorth + Keast + Ksouth + Kwest) * md->Fside)
                     + md->absL * ((Ldown + Lup) * md->Fup + Lside * md->Fcyl + (Lnorth + Least + Lsouth + Lwest) * md->Fside);
    const float tmrt = sqrtf(sqrtf(Sstr / (md->absL * md->sbc))) - 273.2f;
    if (!valid) return;
    int k = 0;
    if (out_mask & 1) out[(long)(k++) * n + t] = tmrt;
    if (out_mask & 2) {
        const float u = (b_t == 1.f) ? utci_point(sp->ta_mat, sp->rh, tmrt, sp->va) : __builtin_nanf("");
        out[(long)(k++) * n + t] = u;
    }
    if (out_mask & 4) out[(long)(k++) * n + t] = shadow;
    if (out_mask & 8) out[(long)(k++) * n + t] = Kdown;
    if (out_mask & 16) out[(long)(k++) * n + t] = Kup;
    if (out_mask & 32) out[(long)(k++) * n + t] = Ldown;
    if (out_mask & 64) out[(long)(k++) * n + t] = Lup;
    if (out_mask & 128) out[(long)(k++) * n + t] = wbgt_point(tmrt, shadow, sp->ta_mat, sp->wbgt_w, sp->wbgt_a);
}

// 8 waves per SIMD (63 VGPRs, a few spilled; the walk waits on LDS and gains from the extra wave)
template <bool WALK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
k_flux_step(const uint32_t *__restrict__ masks, const uint2 *__restrict__ src,
            const float *__restrict__ svfbuveg, const float *__restrict__ tan_asvf,
            const float *__restrict__ tan_sa, const uint32_t *__restrict__ walkface,
            const float *__restrict__ bu, const uint8_t *__restrict__ cls,
            const sw_step_t *__restrict__ steps, const sw_model_t *__restrict__ model,
            float *__restrict__ state, float *__restrict__ out, int out_mask, int step, int halo,
            int rows, int cols) {
    flux_body<WALK>(masks, src, svfbuveg, tan_asvf, tan_sa, walkface, bu, cls, steps, model, state, out, out_mask,
                    step, halo, rows, cols, xcd_tile(blockIdx.x, gridDim.x));
}

template <bool WALK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
k_flux_tiles(const uint32_t *__restrict__ masks, const uint2 *__restrict__ src,
             const float *__restrict__ svfbuveg, const float *__restrict__ tan_asvf,
             const float *__restrict__ tan_sa, const uint32_t *__restrict__ walkface,
             const float *__restrict__ bu, const uint8_t *__restrict__ cls,
             const sw_step_t *__restrict__ steps, const sw_model_t *__restrict__ model,
             float *__restrict__ state, float *__restrict__ out, int out_mask, int step, int halo,
             int rows, int cols, const int32_t *__restrict__ tiles, int ntiles) {
    const int g = xcd_tile(blockIdx.x, gridDim.x);
    if (g >= ntiles) return;
    const int tile = tiles[g];
    if (tile < 0) return;                      // flux_body skips tiles past the grid
    flux_body<WALK>(masks, src, svfbuveg, tan_asvf, tan_sa, walkface, bu, cls, steps, model, state, out, out_mask,
                    step, halo, rows, cols, tile);
}

template <bool WALK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
k_flux_batch(const BatchItemDev *__restrict__ items, int nitems, int total, const sw_model_t *__restrict__ model,
             int out_mask, int out_step, int step, int halo) {
    const int g = xcd_tile(blockIdx.x, gridDim.x);
    if (g >= total) return;
    const BatchItemDev *it = items + batch_owner<1>(items, nitems, g);
    const BatchTileDev &T = it->tile;
    const long n = (long)T.rows * T.cols;
    SW_GLOBAL float *out = it->out + (long)out_step * __builtin_popcount((unsigned)out_mask) * n;
    flux_body<WALK>((const uint32_t *)T.masks, (const uint2 *)T.src, (const float *)T.svfbuveg,
                    (const float *)T.tan_asvf, (const float *)T.tan_svfalfa, (const uint32_t *)T.walkface,
                    (const float *)T.buildings, (const uint8_t *)T.cls, (const sw_step_t *)it->steps, model,
                    (float *)T.state, (float *)out, out_mask, step, halo, T.rows, T.cols, g - it->flux_block0);
}

// walk face-shading bits: bit j set where facesh of walk azimuth j is 0
__global__ void k_walk_face(const float *__restrict__ walls, const float *__restrict__ aspect,
                            const sw_model_t *__restrict__ md, uint32_t *__restrict__ out, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float wallbol = (walls[t] > 0.f) ? 1.f : 0.f;
    const float asp = aspect[t];
    uint32_t bits = 0;
    for (int j = 0; j < SW_WALK_AZ; ++j) {
        const float lo = md->walk_lo[j], hi = md->walk_hi[j];
        float facesh;
        if (md->walk_branch[j] == 0) facesh = (float)((asp < lo) || (asp >= hi)) - wallbol + 1.f;
        else facesh = (float)((asp > lo) || (asp <= hi)) * -1.f + 1.f;
        if (facesh == 0.f) bits |= 1u << j;
    }
    out[t] = bits;
}

// single sun march with float outputs (reference native ABI form)
__global__ void k_march_wh23_single(const float *__restrict__ a, const float *__restrict__ v1,
                                    const float *__restrict__ v2, const int32_t *__restrict__ dx,
                                    const int32_t *__restrict__ dy, const float *__restrict__ dz,
                                    const float *__restrict__ dzp, float *__restrict__ sh,
                                    float *__restrict__ vegsh, float *__restrict__ vbsh, int rows, int cols,
                                    int count) {
    const long n = (long)rows * cols;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int r = (int)(t / cols), c = (int)(t - (long)r * cols);
    const float at = a[t];
    float f = at;
    int s_sh = 0, vs = 0, vb = 0;
    for (int s = 0; s < count; ++s) {
        const int sr = r + dx[s], sc = c + dy[s];
        const float dzs = dz[s], dzq = dzp[s];
        float tmp = 0.f, tv = 0.f, tv2 = 0.f, lf = 0.f, lg = 0.f;
        if (((unsigned)sr < (unsigned)rows) && ((unsigned)sc < (unsigned)cols)) {
            const long q = (long)sr * cols + sc;
            tv = v1[q] - dzs;
            tv2 = v2[q] - dzs;
            tmp = a[q] - dzs;
            lf = v1[q] - dzq;
            lg = v2[q] - dzq;
        }
        f = fmaxf(f, tmp);
        s_sh = (f > at) ? 1 : 0;
        int g2 = (tv > at ? 1 : 0) + (tv2 > at ? 1 : 0) + (lf > at ? 1 : 0) + (lg > at ? 1 : 0);
        if (g2 == 4) g2 = 0;
        if (g2 > 0) g2 = 1;
        vs = vs > g2 ? vs : g2;
        if (vs * s_sh > 0) vs = 0;
        vb = vb + vs;
    }
    int b = vb > 0 ? 1 : vb;
    b = b - vs;
    const int vsf = vs > 0 ? 1 : vs;
    sh[t] = (float)(1 - s_sh);
    vegsh[t] = (float)(1 - vsf);
    vbsh[t] = (float)(1 - b);
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static inline int grid_for(long n, int block) { return (int)((n + block - 1) / block); }

extern "C" {

int sw_abi_version(void) { return SW_ABI_VERSION; }
const char *sw_last_error(void) { return g_err; }
const char *sw_build_flags(void) { return SW_BUILD_FLAGS; }

int sw_device_count(int *count) {
    if (!count) return set_err(SW_E_ARG, "sw_device_count: count is NULL");
    return hip_check(hipGetDeviceCount(count), "hipGetDeviceCount");
}

int sw_sync(void *stream) {
    if (stream) return hip_check(hipStreamSynchronize(S(stream)), "hipStreamSynchronize");
    return hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

int swi_check_shape(int rows, int cols, const char *who) {
    if (rows <= 0 || cols <= 0 || (long)rows * cols > (1L << 31) - 1) {
        char buf[160];
        snprintf(buf, sizeof(buf), "%s: invalid grid %d x %d", who, rows, cols);
        return set_err(SW_E_ARG, buf);
    }
    return SW_OK;
}

#define check_shape swi_check_shape

int sw_pack_heights(const float *a, const float *v1, const float *v2, void *hpack, int rows, int cols,
                    void *stream) {
    int rc = check_shape(rows, cols, "sw_pack_heights");
    if (rc) return rc;
    if (!a || !v1 || !v2 || !hpack) return set_err(SW_E_ARG, "sw_pack_heights: NULL pointer");
    const long n = (long)rows * cols;
    hipLaunchKernelGGL(k_pack_heights, dim3(grid_for(n, 256)), dim3(256), 0, S(stream), a, v1, v2,
                       (float4 *)hpack, n);
    return hip_check(hipGetLastError(), "k_pack_heights");
}

int sw_compose(const float *dsm, const float *dem, const float *cdsm, const float *aspect_deg,
               const float *landcover, void *hpack, float *vegdsm2, float *buildings, uint8_t *cls,
               float *aspect_rad, uint32_t *stats, int rows, int cols, void *stream) {
    int rc = check_shape(rows, cols, "sw_compose");
    if (rc) return rc;
    if (!dsm || !dem || !cdsm || !aspect_deg || !hpack || !vegdsm2 || !buildings || !cls || !aspect_rad || !stats)
        return set_err(SW_E_ARG, "sw_compose: NULL pointer");
    const long n = (long)rows * cols;
    hipLaunchKernelGGL(k_stats_init, dim3(1), dim3(64), 0, S(stream), stats);
    rc = hip_check(hipGetLastError(), "k_stats_init");
    if (rc) return rc;
    hipLaunchKernelGGL(k_compose, dim3(grid_for(n, 256)), dim3(256), 0, S(stream), dsm, dem, cdsm, aspect_deg,
                       landcover, (float4 *)hpack, vegdsm2, buildings, cls, aspect_rad, stats, n);
    return hip_check(hipGetLastError(), "k_compose");
}

int64_t sw_svf_workspace_floats(int rows, int cols) {
    if (rows <= 0 || cols <= 0) return 0;
    const int64_t H = rows + 2 * SW_SVF_PAD, W = cols + 2 * SW_SVF_PAD;
    const int64_t Hc = (H + SW_SVF_COARSE - 1) / SW_SVF_COARSE, Wc = (W + SW_SVF_COARSE - 1) / SW_SVF_COARSE;
    return 3 * H * W + Hc * Wc;
}

// the coarse maxima behind the padded heights
static int pad_coarse(void *hpad, int rows, int cols, void *stream) {
    const int H = rows + 2 * SW_SVF_PAD, W = cols + 2 * SW_SVF_PAD;
    const int Hc = (H + SW_SVF_COARSE - 1) / SW_SVF_COARSE, Wc = (W + SW_SVF_COARSE - 1) / SW_SVF_COARSE;
    float *coarse = (float *)hpad + 3 * (long)H * W;
    hipLaunchKernelGGL(k_pad_coarse, dim3(grid_for((long)Hc * Wc, 256)), dim3(256), 0, S(stream),
                       (const float *)hpad, coarse, H, W, Hc, Wc);
    return hip_check(hipGetLastError(), "k_pad_coarse");
}

int sw_svf_march(const void *hpack, const int32_t *offs, const int32_t *tab, int32_t *counts, void *hpad,
                 uint32_t *masks, uint32_t *stats, int rows, int cols, void *stream) {
    return sw_svf_march_cfg(hpack, offs, tab, counts, hpad, masks, stats, rows, cols, nullptr, stream);
}

int sw_svf_march_cfg(const void *hpack, const int32_t *offs, const int32_t *tab, int32_t *counts, void *hpad,
                     uint32_t *masks, uint32_t *stats, int rows, int cols, const sw_march_config_t *cfg,
                     void *stream) {
    int rc = check_shape(rows, cols, "sw_svf_march");
    if (rc) return rc;
    if ((rc = check_march_config(cfg, "sw_svf_march: march configuration (chunk 4/8/12, shape 0-2, split 1/2/4)")))
        return rc;
    if (rc) return rc;
    if (!hpack || !offs || !tab || !counts || !hpad || !masks || !stats)
        return set_err(SW_E_ARG, "sw_svf_march: NULL pointer");
    const long padded = (long)(rows + 2 * SW_SVF_PAD) * (cols + 2 * SW_SVF_PAD);
    if (padded * 12 >= (1L << 31)) return set_err(SW_E_ARG, "sw_svf_march: tile too large for the padded buffer");
    hipLaunchKernelGGL(k_svf_counts, dim3(SW_NPATCH), dim3(64), 0, S(stream), (const int4 *)tab, offs, stats,
                       counts);
    rc = hip_check(hipGetLastError(), "k_svf_counts");
    if (rc) return rc;
    hipLaunchKernelGGL(k_pad_heights, dim3(grid_for(padded, 256)), dim3(256), 0, S(stream), (const float4 *)hpack,
                       (float *)hpad, stats, rows, cols);
    rc = hip_check(hipGetLastError(), "k_pad_heights");
    if (rc) return rc;
    rc = pad_coarse(hpad, rows, cols, stream);
    if (rc) return rc;
    const long tiles = (long)((rows + SVF_TILE - 1) / SVF_TILE) * ((cols + SVF_TILE - 1) / SVF_TILE);
    const int chunk = cfg ? cfg->chunk : SVF_CHUNK, shape = cfg ? cfg->shape : 0, split = cfg ? cfg->split : 1;
    dim3 grid((unsigned)(((tiles + 7) / 8) * 8), SW_MASK_WORDS * split);
    launch_march(grid, S(stream), chunk, shape, (const float *)hpad, offs, (const int4 *)tab, counts, masks, stats,
                 rows, cols, (const int32_t *)nullptr, 0, split);
    return hip_check(hipGetLastError(), "k_svf_march");
}

// ---------------------------------------------------------------------------
// scratch-free SVF sweep: the march workspace (step counts, padded heights) in
// a static device array instead of caller buffers. Each GPU holds its own copy
// of the array; host code addresses it only through the pointer that
// hipGetSymbolAddress returns with that GPU current, kept per device.
// ---------------------------------------------------------------------------
#define SW_SCRATCH_COUNTS_BYTES 1024   // counts[153] at the start, padded heights after
#define SW_SCRATCH_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[SW_SVF_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[SW_SCRATCH_MAX_DEVICES];

// the current device's address of `scratch`, looked up once per device
static int scratch_ptr(void **out) {
    int dev = -1;
    int rc = hip_check(hipGetDevice(&dev), "sw_svf_sweep_noalloc: hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= SW_SCRATCH_MAX_DEVICES)
        return set_err(SW_E_UNSUPPORTED, "sw_svf_sweep_noalloc: device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "sw_svf_sweep_noalloc: hipGetSymbolAddress");
        if (rc) return rc;
        if (!p) return set_err(SW_E_HIP, "sw_svf_sweep_noalloc: hipGetSymbolAddress gave a NULL address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SW_OK;
}

int sw_svf_sweep_noalloc(const void *hpack, uint32_t *stats, const int32_t *offs, const int32_t *tab,
                         const float *vegdsm2, const float *w_iso, const float *w_aniso, const int32_t *ring,
                         const int32_t *nann, const int32_t *dirm, uint32_t *masks, float *fields, int rows,
                         int cols, const sw_march_config_t *cfg, void *stream) {
    int rc = check_shape(rows, cols, "sw_svf_sweep_noalloc");
    if (rc) return rc;
    if (!hpack || !stats || !offs || !tab || !vegdsm2 || !w_iso || !w_aniso || !ring || !nann || !dirm || !masks ||
        !fields)
        return set_err(SW_E_ARG, "sw_svf_sweep_noalloc: NULL pointer");
    const long long need = (long long)sw_svf_workspace_floats(rows, cols) * 4 + SW_SCRATCH_COUNTS_BYTES;
    if (need > (long long)SW_SVF_SCRATCH_BYTES) {
        char buf[200];
        snprintf(buf, sizeof(buf),
                 "sw_svf_sweep_noalloc: a %d x %d tile needs %lld bytes of march workspace, above the %lld-byte "
                 "static scratch (use sw_svf_sweep with caller buffers)",
                 rows, cols, need, (long long)SW_SVF_SCRATCH_BYTES);
        return set_err(SW_E_ARG, buf);
    }
    void *base = nullptr;
    if ((rc = scratch_ptr(&base))) return rc;
    // both parts are rewritten in full before they are read: k_svf_counts writes all 153 counts,
    // k_pad_heights every padded cell and k_pad_coarse every coarse maximum, so the previous call's
    // contents never leak into this one
    int32_t *counts = (int32_t *)base;
    void *hpad = (unsigned char *)base + SW_SCRATCH_COUNTS_BYTES;
    return sw_svf_sweep_cfg(hpack, stats, offs, tab, vegdsm2, w_iso, w_aniso, ring, nann, dirm, counts, hpad, masks,
                            fields, rows, cols, cfg, stream);
}

int sw_svf_sweep_tiles(const void *hpack, uint32_t *stats, const int32_t *offs, const int32_t *tab,
                       const float *vegdsm2, const float *w_iso, const float *w_aniso, const int32_t *ring,
                       const int32_t *nann, const int32_t *dirm, int32_t *counts, void *hpad, uint32_t *masks,
                       float *fields, const int32_t *tiles, int ntiles, int rows, int cols, void *stream) {
    int rc = check_shape(rows, cols, "sw_svf_sweep_tiles");
    if (rc) return rc;
    if (!hpack || !stats || !offs || !tab || !vegdsm2 || !w_iso || !w_aniso || !ring || !nann || !dirm || !counts ||
        !hpad || !masks || !fields || (ntiles > 0 && !tiles))
        return set_err(SW_E_ARG, "sw_svf_sweep_tiles: NULL pointer");
    if (ntiles < 0) return set_err(SW_E_ARG, "sw_svf_sweep_tiles: negative tile count");
    if (ntiles == 0) return SW_OK;
    const long padded = (long)(rows + 2 * SW_SVF_PAD) * (cols + 2 * SW_SVF_PAD);
    if (padded * 12 >= (1L << 31)) return set_err(SW_E_ARG, "sw_svf_sweep_tiles: tile too large");
    hipLaunchKernelGGL(k_svf_counts, dim3(SW_NPATCH), dim3(64), 0, S(stream), (const int4 *)tab, offs, stats,
                       counts);
    rc = hip_check(hipGetLastError(), "k_svf_counts");
    if (rc) return rc;
    hipLaunchKernelGGL(k_pad_heights, dim3(grid_for(padded, 256)), dim3(256), 0, S(stream), (const float4 *)hpack,
                       (float *)hpad, stats, rows, cols);
    rc = hip_check(hipGetLastError(), "k_pad_heights");
    if (rc) return rc;
    rc = pad_coarse(hpad, rows, cols, stream);
    if (rc) return rc;
    const unsigned g = (unsigned)(((ntiles + 7) / 8) * 8);
    // short lists (under 8 blocks per CU at one word per block) split each word over four blocks
    const int split = (long)g * SW_MASK_WORDS < 2048 ? 4 : 1;
    launch_march(dim3(g, SW_MASK_WORDS * split), S(stream), SVF_CHUNK, 0, (const float *)hpad, offs,
                 (const int4 *)tab, counts, masks, stats, rows, cols, tiles, ntiles, split);
    rc = hip_check(hipGetLastError(), "k_svf_march");
    if (rc) return rc;
    hipLaunchKernelGGL(k_svf_fold_tiles, dim3(g), dim3(256), 0, S(stream), masks, vegdsm2, w_iso, w_aniso, ring,
                       nann, dirm, fields, rows, cols, tiles, ntiles);
    return hip_check(hipGetLastError(), "k_svf_fold_tiles");
}

int sw_solve_day_tiles(const sw_tile_t *tile, const sw_step_t *steps, const sw_model_t *model,
                       const int32_t *sun_tab, const int32_t *is_day, int t0, int t1, float *out, int out_mask,
                       int walk_halo, int reset_state, const int32_t *sun_tiles, int nsun,
                       const int32_t *flux_tiles, int nflux, void *stream) {
    if (!tile || !steps || !model || !sun_tab || !is_day || !out || (nsun > 0 && !sun_tiles) ||
        (nflux > 0 && !flux_tiles))
        return set_err(SW_E_ARG, "sw_solve_day_tiles: NULL pointer");
    int rc = check_shape(tile->rows, tile->cols, "sw_solve_day_tiles");
    if (rc) return rc;
    if (t0 < 0 || t1 < t0) return set_err(SW_E_ARG, "sw_solve_day_tiles: invalid step range");
    if (out_mask <= 0 || out_mask > 255) return set_err(SW_E_ARG, "sw_solve_day_tiles: out_mask selects nothing");
    if (walk_halo < 0 || walk_halo > 48) return set_err(SW_E_ARG, "sw_solve_day_tiles: walk halo out of range");
    if (nsun < 0 || nflux < 0) return set_err(SW_E_ARG, "sw_solve_day_tiles: negative tile count");
    const sw_tile_t &T = *tile;
    const int rows = T.rows, cols = T.cols;
    const long n = (long)rows * cols;
    const long stride = (long)__builtin_popcount((unsigned)out_mask) * n;
    if (reset_state) {
        rc = sw_fill_f32(T.state, 0.f, 6 * (int64_t)n, stream);
        if (rc) return rc;
    }
    if (nflux == 0) return SW_OK;
    const size_t lds = flux_lds_bytes(walk_halo);
    if (lds > 65536) {
        rc = hip_check(hipFuncSetAttribute((const void *)k_flux_tiles<true>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                       "hipFuncSetAttribute");
        if (rc) return rc;
    }
    const unsigned gs = (unsigned)(((nsun + 7) / 8) * 8), gf = (unsigned)(((nflux + 7) / 8) * 8);
    for (int i = t0; i < t1; ++i) {
        float *o = out + (long)(i - t0) * stride;
        if (is_day[i - t0]) {
            if (nsun > 0) {
                hipLaunchKernelGGL(k_sun_tiles, dim3(gs), dim3(256), 0, S(stream), (const float4 *)T.hpack, T.walls,
                                   T.aspect_rad, T.cls, T.buildings, steps, model, (const int4 *)sun_tab, T.stats,
                                   (uint2 *)T.src, i, rows, cols, sun_tiles, nsun);
                rc = hip_check(hipGetLastError(), "k_sun_tiles");
                if (rc) return rc;
            }
            hipLaunchKernelGGL(k_flux_tiles<true>, dim3(gf), dim3(256), lds, S(stream), T.masks,
                               (const uint2 *)T.src, T.svfbuveg, T.tan_asvf, T.tan_svfalfa, T.walkface, T.buildings,
                               T.cls, steps, model, T.state, o, out_mask, i, walk_halo, rows, cols, flux_tiles, nflux);
            rc = hip_check(hipGetLastError(), "k_flux_tiles");
        } else {
            hipLaunchKernelGGL(k_flux_tiles<false>, dim3(gf), dim3(256), 0, S(stream), T.masks, (const uint2 *)T.src,
                               T.svfbuveg, T.tan_asvf, T.tan_svfalfa, T.walkface, T.buildings, T.cls, steps, model,
                               T.state, o, out_mask, i, -1, rows, cols, flux_tiles, nflux);
            rc = hip_check(hipGetLastError(), "k_flux_tiles<night>");
        }
        if (rc) return rc;
    }
    return SW_OK;
}

int sw_svf_fold(const uint32_t *masks, const float *vegdsm2, const float *w_iso, const float *w_aniso,
                const int32_t *ring, const int32_t *nann, const int32_t *dirm, float *fields, int rows,
                int cols, void *stream) {
    int rc = check_shape(rows, cols, "sw_svf_fold");
    if (rc) return rc;
    if (!masks || !vegdsm2 || !w_iso || !w_aniso || !ring || !nann || !dirm || !fields)
        return set_err(SW_E_ARG, "sw_svf_fold: NULL pointer");
    const long n = (long)rows * cols;
    hipLaunchKernelGGL(k_svf_fold, dim3(grid_for(n, 256)), dim3(256), 0, S(stream), masks, vegdsm2, w_iso,
                       w_aniso, ring, nann, dirm, fields, n);
    return hip_check(hipGetLastError(), "k_svf_fold");
}

int sw_unpack_masks(const uint32_t *masks, float *shmat, float *vegshmat, float *vbshmat, int rows, int cols,
                    void *stream) {
    int rc = check_shape(rows, cols, "sw_unpack_masks");
    if (rc) return rc;
    if (!masks) return set_err(SW_E_ARG, "sw_unpack_masks: NULL masks");
    const long n = (long)rows * cols;
    hipLaunchKernelGGL(k_unpack, dim3(grid_for(n * SW_NPATCH, 256)), dim3(256), 0, S(stream), masks, shmat,
                       vegshmat, vbshmat, n);
    return hip_check(hipGetLastError(), "k_unpack");
}

int sw_pack_masks(const float *shmat, const float *vegshmat, const float *vbshmat, uint32_t *masks,
                  uint32_t *flags, int rows, int cols, void *stream) {
    int rc = check_shape(rows, cols, "sw_pack_masks");
    if (rc) return rc;
    if (!shmat || !vegshmat || !vbshmat || !masks || !flags) return set_err(SW_E_ARG, "sw_pack_masks: NULL pointer");
    const long n = (long)rows * cols;
    hipLaunchKernelGGL(k_pack_masks, dim3(grid_for(n, 256)), dim3(256), 0, S(stream), shmat, vegshmat, vbshmat,
                       masks, flags, n);
    return hip_check(hipGetLastError(), "k_pack_masks");
}

int sw_height_stats(const void *hpack, float amax, uint32_t *stats, int rows, int cols, void *stream) {
    int rc = check_shape(rows, cols, "sw_height_stats");
    if (rc) return rc;
    if (!hpack || !stats) return set_err(SW_E_ARG, "sw_height_stats: NULL pointer");
    const long n = (long)rows * cols;
    hipLaunchKernelGGL(k_stats_init, dim3(1), dim3(64), 0, S(stream), stats);
    rc = hip_check(hipGetLastError(), "k_stats_init");
    if (rc) return rc;
    hipLaunchKernelGGL(k_set_key, dim3(1), dim3(1), 0, S(stream), stats, amax);
    rc = hip_check(hipGetLastError(), "k_set_key");
    if (rc) return rc;
    hipLaunchKernelGGL(k_height_stats, dim3(grid_for(n, 256)), dim3(256), 0, S(stream), (const float4 *)hpack,
                       stats, n);
    return hip_check(hipGetLastError(), "k_height_stats");
}

int sw_set_amax(uint32_t *stats, float amax, void *stream) {
    if (!stats) return set_err(SW_E_ARG, "sw_set_amax: NULL pointer");
    hipLaunchKernelGGL(k_set_key, dim3(1), dim3(1), 0, S(stream), stats, amax);
    return hip_check(hipGetLastError(), "k_set_key");
}

int sw_run_march_svf_shadow(const float *a, const float *vegdem, const float *vegdem2, const int32_t *dxs,
                            const int32_t *dys, const float *dzs, float *sh, float *vegsh, float *vbsh,
                            int rows, int cols, int count, void *stream) {
    int rc = check_shape(rows, cols, "sw_run_march_svf_shadow");
    if (rc) return rc;
    if (count < 0) return set_err(SW_E_ARG, "sw_run_march_svf_shadow: negative count");
    const long n = (long)rows * cols;
    hipLaunchKernelGGL(k_march_svf_single, dim3(grid_for(n, 256)), dim3(256), 0, S(stream), a, vegdem,
                       vegdem2, dxs, dys, dzs, sh, vegsh, vbsh, rows, cols, count);
    return hip_check(hipGetLastError(), "k_march_svf_single");
}

int sw_run_march_wallheight23(const float *a, const float *vegdem, const float *vegdem2, const int32_t *dxs,
                              const int32_t *dys, const float *dzs, const float *dzprevs, float *sh,
                              float *vegsh, float *vbsh, int rows, int cols, int count, void *stream) {
    int rc = check_shape(rows, cols, "sw_run_march_wallheight23");
    if (rc) return rc;
    if (count < 0) return set_err(SW_E_ARG, "sw_run_march_wallheight23: negative count");
    const long n = (long)rows * cols;
    hipLaunchKernelGGL(k_march_wh23_single, dim3(grid_for(n, 256)), dim3(256), 0, S(stream), a, vegdem,
                       vegdem2, dxs, dys, dzs, dzprevs, sh, vegsh, vbsh, rows, cols, count);
    return hip_check(hipGetLastError(), "k_march_wh23_single");
}

int sw_tile_static(const float *fields, const float *walls, const float *aspect_rad, const sw_model_t *model,
                   float *svfbuveg, float *tan_asvf, float *tan_svfalfa, uint32_t *walkface, int rows,
                   int cols, void *stream) {
    int rc = check_shape(rows, cols, "sw_tile_static");
    if (rc) return rc;
    if (!fields || !walls || !aspect_rad || !model || !svfbuveg || !tan_asvf || !tan_svfalfa || !walkface)
        return set_err(SW_E_ARG, "sw_tile_static: NULL pointer");
    const long n = (long)rows * cols;
    hipLaunchKernelGGL(k_tile_static, dim3(grid_for(n, 256)), dim3(256), 0, S(stream), fields, svfbuveg,
                       tan_asvf, tan_svfalfa, n);
    rc = hip_check(hipGetLastError(), "k_tile_static");
    if (rc) return rc;
    hipLaunchKernelGGL(k_walk_face, dim3(grid_for(n, 256)), dim3(256), 0, S(stream), walls, aspect_rad, model,
                       walkface, n);
    return hip_check(hipGetLastError(), "k_walk_face");
}

static int sun_step(const void *hpack, const float *walls, const float *aspect_rad, const uint8_t *cls,
                    const float *buildings, const sw_step_t *steps, const sw_model_t *model, const int32_t *sun_tab,
                    const uint32_t *stats, uint32_t *src, float *planes, int step, int rows, int cols, void *stream,
                    const char *who) {
    int rc = check_shape(rows, cols, who);
    if (rc) return rc;
    char buf[96];
    if (!hpack || !walls || !aspect_rad || !cls || !buildings || !steps || !model || !sun_tab || !stats || !src) {
        snprintf(buf, sizeof(buf), "%s: NULL pointer", who);
        return set_err(SW_E_ARG, buf);
    }
    if (step < 0) {
        snprintf(buf, sizeof(buf), "%s: negative step", who);
        return set_err(SW_E_ARG, buf);
    }
    const long n = (long)rows * cols;
    (void)n;
    const long tiles = (long)((rows + FLUX_TY - 1) / FLUX_TY) * ((cols + FLUX_TX - 1) / FLUX_TX);
    const unsigned tg = (unsigned)(((tiles + 7) / 8) * 8);
    if (planes) {
        hipLaunchKernelGGL(k_sun_step<true>, dim3(tg), dim3(256), 0, S(stream), (const float4 *)hpack,
                           walls, aspect_rad, cls, buildings, steps, model, (const int4 *)sun_tab, stats, (uint2 *)src,
                           step, rows, cols, planes);
        return hip_check(hipGetLastError(), "k_sun_step<planes>");
    }
    hipLaunchKernelGGL(k_sun_step<false>, dim3(tg), dim3(256), 0, S(stream), (const float4 *)hpack, walls,
                       aspect_rad, cls, buildings, steps, model, (const int4 *)sun_tab, stats, (uint2 *)src, step,
                       rows, cols, (float *)nullptr);
    return hip_check(hipGetLastError(), "k_sun_step");
}

int sw_sun_step(const void *hpack, const float *walls, const float *aspect_rad, const uint8_t *cls,
                const float *buildings, const sw_step_t *steps, const sw_model_t *model, const int32_t *sun_tab,
                const uint32_t *stats, uint32_t *src, int step, int rows, int cols, void *stream) {
    return sun_step(hpack, walls, aspect_rad, cls, buildings, steps, model, sun_tab, stats, src, nullptr, step, rows,
                    cols, stream, "sw_sun_step");
}

int sw_sun_step_planes(const void *hpack, const float *walls, const float *aspect_rad, const uint8_t *cls,
                       const float *buildings, const sw_step_t *steps, const sw_model_t *model,
                       const int32_t *sun_tab, const uint32_t *stats, uint32_t *src, float *planes, int step,
                       int rows, int cols, void *stream) {
    if (!planes) return set_err(SW_E_ARG, "sw_sun_step_planes: NULL planes");
    return sun_step(hpack, walls, aspect_rad, cls, buildings, steps, model, sun_tab, stats, src, planes, step, rows,
                    cols, stream, "sw_sun_step_planes");
}

int sw_flux_step(const uint32_t *masks, const uint32_t *src, const float *svfbuveg, const float *tan_asvf,
                 const float *tan_svfalfa, const uint32_t *walkface, const float *buildings, const uint8_t *cls,
                 const sw_step_t *steps, const sw_model_t *model, float *state, float *out, int out_mask,
                 int step, int walk_halo, int rows, int cols, void *stream) {
    int rc = check_shape(rows, cols, "sw_flux_step");
    if (rc) return rc;
    if (!masks || !src || !svfbuveg || !tan_asvf || !tan_svfalfa || !walkface || !buildings || !cls || !steps ||
        !model || !state || !out)
        return set_err(SW_E_ARG, "sw_flux_step: NULL pointer");
    if (walk_halo < -1 || walk_halo > 48) return set_err(SW_E_ARG, "sw_flux_step: walk halo out of range");
    const int tiles = ((cols + FLUX_TX - 1) / FLUX_TX) * ((rows + FLUX_TY - 1) / FLUX_TY);
    const int grid = ((tiles + 7) / 8) * 8;
    const size_t lds = flux_lds_bytes(walk_halo);
    if (lds > 65536) {
        rc = hip_check(hipFuncSetAttribute((const void *)k_flux_step<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds), "hipFuncSetAttribute");
        if (rc) return rc;
    }
    if (walk_halo < 0) {
        hipLaunchKernelGGL(k_flux_step<false>, dim3(grid), dim3(256), 0, S(stream), masks, (const uint2 *)src,
                           svfbuveg, tan_asvf, tan_svfalfa, walkface, buildings, cls, steps, model, state, out,
                           out_mask, step, walk_halo, rows, cols);
        return hip_check(hipGetLastError(), "k_flux_step<night>");
    }
    hipLaunchKernelGGL(k_flux_step<true>, dim3(grid), dim3(256), lds, S(stream), masks, (const uint2 *)src,
                       svfbuveg, tan_asvf, tan_svfalfa, walkface, buildings, cls, steps, model, state, out,
                       out_mask, step, walk_halo, rows, cols);
    return hip_check(hipGetLastError(), "k_flux_step");
}

// the sun march runs on the same 16 x 16 tiles as the flux step (8 x 8 pixels per wave)
static int item_blocks_sun(const sw_tile_t &t) {
    return ((t.cols + FLUX_TX - 1) / FLUX_TX) * ((t.rows + FLUX_TY - 1) / FLUX_TY);
}
static int item_blocks_flux(const sw_tile_t &t) {
    return ((t.cols + FLUX_TX - 1) / FLUX_TX) * ((t.rows + FLUX_TY - 1) / FLUX_TY);
}

int sw_batch_layout(sw_batch_item_t *items, int nitems, int32_t *sun_blocks, int32_t *flux_blocks) {
    if (!items || nitems <= 0 || !sun_blocks || !flux_blocks)
        return set_err(SW_E_ARG, "sw_batch_layout: NULL pointer or empty batch");
    long sb = 0, fb = 0;
    for (int k = 0; k < nitems; ++k) {
        int rc = check_shape(items[k].tile.rows, items[k].tile.cols, "sw_batch_layout");
        if (rc) return rc;
        items[k].sun_block0 = (int32_t)sb;
        items[k].flux_block0 = (int32_t)fb;
        sb += item_blocks_sun(items[k].tile);
        fb += item_blocks_flux(items[k].tile);
        if (sb > (1L << 30) || fb > (1L << 30)) return set_err(SW_E_ARG, "sw_batch_layout: batch too large");
    }
    *sun_blocks = (int32_t)sb;
    *flux_blocks = (int32_t)fb;
    return SW_OK;
}

// Host checks of a batch; returns the grid totals of items_host (laid out by sw_batch_layout).
static int check_batch(const sw_batch_item_t *items, int nitems, long *sun_total, long *flux_total) {
    long sb = 0, fb = 0;
    for (int k = 0; k < nitems; ++k) {
        const sw_tile_t &t = items[k].tile;
        int rc = check_shape(t.rows, t.cols, "sw_solve_batch");
        if (rc) return rc;
        if (!t.hpack || !t.walls || !t.aspect_rad || !t.buildings || !t.cls || !t.stats || !t.masks || !t.svfbuveg ||
            !t.tan_asvf || !t.tan_svfalfa || !t.walkface || !t.src || !t.state || !items[k].out || !items[k].steps ||
            !items[k].sun_tab)
            return set_err(SW_E_ARG, "sw_solve_batch: NULL pointer in an item");
        if (items[k].sun_block0 != sb || items[k].flux_block0 != fb)
            return set_err(SW_E_ARG, "sw_solve_batch: items not laid out by sw_batch_layout");
        sb += item_blocks_sun(t);
        fb += item_blocks_flux(t);
    }
    *sun_total = sb;
    *flux_total = fb;
    return SW_OK;
}

int sw_solve_batch(const sw_batch_item_t *items_dev, const sw_batch_item_t *items_host, int nitems,
                   const sw_model_t *model, const int32_t *is_day, int t0, int t1, int out_mask, int walk_halo,
                   int reset_state, void *stream) {
    if (!items_dev || !items_host || nitems <= 0 || !model || !is_day)
        return set_err(SW_E_ARG, "sw_solve_batch: NULL pointer or empty batch");
    if (t0 < 0 || t1 < t0) return set_err(SW_E_ARG, "sw_solve_batch: invalid step range");
    if (out_mask <= 0 || out_mask > 255) return set_err(SW_E_ARG, "sw_solve_batch: out_mask selects nothing");
    if (walk_halo < 0 || walk_halo > 48) return set_err(SW_E_ARG, "sw_solve_batch: walk halo out of range");
    long sun_total = 0, flux_total = 0;
    int rc = check_batch(items_host, nitems, &sun_total, &flux_total);
    if (rc) return rc;
    if (reset_state) {
        for (int k = 0; k < nitems; ++k) {
            const sw_tile_t &t = items_host[k].tile;
            rc = sw_fill_f32(t.state, 0.f, 6 * (int64_t)t.rows * t.cols, stream);
            if (rc) return rc;
        }
    }
    const int sun_grid = (int)(((sun_total + 7) / 8) * 8);
    const int flux_grid = (int)(((flux_total + 7) / 8) * 8);
    const size_t lds = flux_lds_bytes(walk_halo);
    if (lds > 65536) {
        rc = hip_check(hipFuncSetAttribute((const void *)k_flux_batch<true>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                       "hipFuncSetAttribute");
        if (rc) return rc;
    }
    for (int i = t0; i < t1; ++i) {
        if (is_day[i - t0]) {
            hipLaunchKernelGGL(k_sun_batch, dim3(sun_grid), dim3(256), 0, S(stream), (const BatchItemDev *)items_dev, nitems,
                               (int)sun_total, model, i);
            rc = hip_check(hipGetLastError(), "k_sun_batch");
            if (rc) return rc;
            hipLaunchKernelGGL(k_flux_batch<true>, dim3(flux_grid), dim3(256), lds, S(stream),
                               (const BatchItemDev *)items_dev, nitems,
                               (int)flux_total, model, out_mask, i - t0, i, walk_halo);
            rc = hip_check(hipGetLastError(), "k_flux_batch");
        } else {
            hipLaunchKernelGGL(k_flux_batch<false>, dim3(flux_grid), dim3(256), 0, S(stream),
                               (const BatchItemDev *)items_dev, nitems,
                               (int)flux_total, model, out_mask, i - t0, i, -1);
            rc = hip_check(hipGetLastError(), "k_flux_batch<night>");
        }
        if (rc) return rc;
    }
    return SW_OK;
}

int sw_utci(const float *ta, const float *rh, const float *tmrt, const float *va, float *out, int64_t n,
            void *stream) {
    if (n < 0) return set_err(SW_E_ARG, "sw_utci: negative n");
    if (n == 0) return SW_OK;
    hipLaunchKernelGGL(k_utci, dim3(grid_for(n, 256)), dim3(256), 0, S(stream), ta, rh, tmrt, va, out, (long)n);
    return hip_check(hipGetLastError(), "k_utci");
}

}  // extern "C"
