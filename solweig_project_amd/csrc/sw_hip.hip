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

int swi_fmt_err(int code, const char *fmt, const char *who) {
    char msg[256];
    snprintf(msg, sizeof(msg), fmt, who);
    return swi_set_err(code, msg);
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

#include "sw_day_body.h"


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
    // After the first step the state of the wave's 64 lanes lives in three lane masks (SGPRs): SH (sh),
    // VS (vegsh), VBB (vbshvegsh sum > 0), plus A (lanes still marching). A step's compares are ballots
    // and its update is five scalar bit operations for the whole wave:
    //   SH |= c1; VS = (VS | (c2 & ~c3)) & ~SH; VBB |= VS.
    // The update may run on every lane: a lane that reads no cell this chunk (it skips the chunk or has
    // stopped) has c1 = c2 = c3 = 0 and gets VS &= ~SH, VBB |= VS, which is exactly a skipped chunk's
    // effect, and a stopped lane is unchanged by it (after any update VS implies !SH and VBB covers VS;
    // a lane stopped at its first step has VS = 0). Lanes that read no cell point their loads past the
    // buffer (the range check returns zeros without a memory access); on this path heights and dz are
    // >= 0 and so is the target's height, so a zero source compares false.
    constexpr uint32_t kNoRead = 0xF0000000u;
    SvfLane st;
    st.sh = false; st.vs = false; st.vbb = false;
    bool active = valid && count > 0;
    if (active) {
        const int4 e = T[0];
        const H3 h = ld3(rs, tpb + (uint32_t)e.w);
        svf_first(st, at, veg2t, h.x, h.y, h.z, __int_as_float(e.z));
        active = !(st.sh & !st.vs);
    }
    uint64_t SH = __builtin_amdgcn_ballot_w64(st.sh), VS = __builtin_amdgcn_ballot_w64(st.vs), VBB = 0;
    uint64_t A = __builtin_amdgcn_ballot_w64(active);
    int s = 1;
    for (; s + CHUNK <= count; s += CHUNK) {
        if (A == 0) break;
        const int4 *__restrict__ Tc = T + s;
        int4 e[CHUNK];
#pragma unroll
        for (int k = 0; k < CHUNK; ++k) e[k] = Tc[k];
        // the chunk's offset bounding box and smallest dz, the same for every lane (the tables move
        // monotonically in x and y with non-decreasing dz, so the first and last steps bound them)
        const int xl = min(e[0].x, e[CHUNK - 1].x), xh = max(e[0].x, e[CHUNK - 1].x);
        const int yl = min(e[0].y, e[CHUNK - 1].y), yh = max(e[0].y, e[CHUNK - 1].y);
        const float dzl = __int_as_float(e[0].z), dl = __int_as_float(e[CHUNK - 1].z);
        bool skip = false;
        const int R0 = (pr + xl) >> 3, R1 = (pr + xh) >> 3, C0 = (pc + yl) >> 3, C1 = (pc + yh) >> 3;
        if (__builtin_amdgcn_inverse_ballot_w64(A) && R1 - R0 <= 1 && C1 - C0 <= 1) {
            // coarse cells past the grid read 0 (range-checked loads), which the steps they cover
            // would read as well (row offsets by 24-bit products: |R| and Wc stay far below 2^23)
            const int q0 = __mul24(R0, Wc), q1 = __mul24(R1, Wc);
            const uint32_t o00 = (uint32_t)(q0 + C0) * 4u, o01 = (uint32_t)(q0 + C1) * 4u;
            const uint32_t o10 = (uint32_t)(q1 + C0) * 4u, o11 = (uint32_t)(q1 + C1) * 4u;
            const bool inb = R0 >= 0 && C0 >= 0 && R1 < Hc && C1 < Wc;
            const float m = fmaxf(fmaxf(ldc(crs, o00), ldc(crs, o01)), fmaxf(ldc(crs, o10), ldc(crs, o11)));
            skip = inb && (m - dzl <= at);
        }
        const uint64_t N = A & ~__builtin_amdgcn_ballot_w64(skip);
        if (N != 0) {
            // lanes that skip or have stopped read past the buffer
            const uint32_t base = __builtin_amdgcn_inverse_ballot_w64(N) ? tpb : kNoRead;
            H3 h[CHUNK];
#pragma unroll
            for (int k = 0; k < CHUNK; ++k) h[k] = ld3(rs, base + (uint32_t)e[k].w);
#pragma unroll
            for (int k = 0; k < CHUNK; ++k) {
                const float dz = __int_as_float(e[k].z);
                const uint64_t c1 = __builtin_amdgcn_ballot_w64(h[k].x - dz > at);
                const uint64_t c2 = __builtin_amdgcn_ballot_w64(h[k].y - dz > at);
                const uint64_t c3 = __builtin_amdgcn_ballot_w64(h[k].z - dz > at);
                SH |= c1;
                VS = (VS | (c2 & ~c3)) & ~SH;
                VBB |= VS;
            }
        } else {
            VS &= ~SH;
            VBB |= VS;
        }
        // exits on the chunk's last step (within the pad, an out-of-grid cell is one past the edges)
        const int lr = pr + e[CHUNK - 1].x - SW_SVF_PAD, lc = pc + e[CHUNK - 1].y - SW_SVF_PAD;
        const uint64_t out = __builtin_amdgcn_ballot_w64(((unsigned)lr >= (unsigned)rows) |
                                                         ((unsigned)lc >= (unsigned)cols));
        const uint64_t low = __builtin_amdgcn_ballot_w64(hmax - dl <= at);
        A &= ~((SH & ~VS) | out | low);
    }
    for (; s < count; ++s) {
        if (A == 0) break;
        const int4 e = T[s];
        const uint32_t base = __builtin_amdgcn_inverse_ballot_w64(A) ? tpb : kNoRead;
        const H3 h = ld3(rs, base + (uint32_t)e.w);
        const float dz = __int_as_float(e.z);
        const uint64_t c1 = __builtin_amdgcn_ballot_w64(h.x - dz > at);
        const uint64_t c2 = __builtin_amdgcn_ballot_w64(h.y - dz > at);
        const uint64_t c3 = __builtin_amdgcn_ballot_w64(h.z - dz > at);
        SH |= c1;
        VS = (VS | (c2 & ~c3)) & ~SH;
        VBB |= VS;
        const uint64_t out = __builtin_amdgcn_ballot_w64(outside(h));
        const uint64_t low = __builtin_amdgcn_ballot_w64(hmax - dz <= at);
        A &= ~((SH & ~VS) | out | low);
    }
    st.sh = __builtin_amdgcn_inverse_ballot_w64(SH);
    st.vs = __builtin_amdgcn_inverse_ballot_w64(VS);
    st.vbb = __builtin_amdgcn_inverse_ballot_w64(VBB);
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
    // Block -> (tile block bx, word w, part sub). With split > 1 the parts of one tile's words are
    // the blocks L, L + 8, ..., L + 8 (split - 1) of the linear dispatch order L = x + gx y: the same
    // XCD (workgroups go round-robin over the 8 XCDs) and dispatched together, so the partial writes
    // of each mask line merge in that XCD's L2 before it is written back once. gx is a multiple of 8,
    // so the map is a bijection and keeps bx % 8 == L % 8 (the XCD xcd_tile assumes).
    int bx = blockIdx.x, w, sub;
    if (split == 1) {
        w = blockIdx.y;
        sub = 0;
    } else {
        const int L = blockIdx.x + gridDim.x * blockIdx.y, grp = L >> 3;
        const int q = (grp / split) * 8 + (L & 7);
        sub = grp - (grp / split) * split;
        w = q / (int)gridDim.x;
        bx = q - w * (int)gridDim.x;
    }
    int tile = xcd_tile(bx, gridDim.x);
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


__global__ void k_utci(const float *__restrict__ ta, const float *__restrict__ rh,
                       const float *__restrict__ tmrt, const float *__restrict__ va,
                       float *__restrict__ out, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = utci_point(ta[i], rh[i], tmrt[i], va[i]);
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
int swi_check_batch(const sw_batch_item_t *items, int nitems, long *sun_total, long *flux_total, const char *who) {
    long sb = 0, fb = 0;
    for (int k = 0; k < nitems; ++k) {
        const sw_tile_t &t = items[k].tile;
        int rc = check_shape(t.rows, t.cols, who);
        if (rc) return rc;
        if (!t.hpack || !t.walls || !t.aspect_rad || !t.buildings || !t.cls || !t.stats || !t.masks || !t.svfbuveg ||
            !t.tan_asvf || !t.tan_svfalfa || !t.walkface || !t.src || !t.state || !items[k].out || !items[k].steps ||
            !items[k].sun_tab)
            return swi_fmt_err(SW_E_ARG, "%s: NULL pointer in an item", who);
        if (items[k].sun_block0 != sb || items[k].flux_block0 != fb)
            return swi_fmt_err(SW_E_ARG, "%s: items not laid out by sw_batch_layout", who);
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
    int rc = swi_check_batch(items_host, nitems, &sun_total, &flux_total, "sw_solve_batch");
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
