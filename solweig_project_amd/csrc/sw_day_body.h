// Device code of the day step, shared by libsolweig_hip's translation units:
// math helpers, the sun march with the wall tail (sun_step_body), the batch
// item mirror, UTCI/WBGT and the fused flux step (flux_body).
//
// Step records and the model are read through pointers in the constant
// address space (SW_CONST): the compiler then always reads their fields with
// scalar loads, also inside a loop whose stores it cannot tell apart from
// them, and can re-load a field instead of spilling it.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "sw_internal.h"

// (the host pass of hipcc also parses the device functions: there the
// qualifier is dropped, since the host has no such address space)
#if defined(__HIP_DEVICE_COMPILE__)
#define SW_CONST __attribute__((address_space(4)))
#else
#define SW_CONST
#endif

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

// threadIdx.x behind an empty asm: the day-step bodies read the lane index
// through it, so that a loop around a body (the persistent day kernel) cannot
// hoist the lane-derived indices out of the loop and keep them live across it.
__device__ __forceinline__ int sw_tid() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

__device__ __forceinline__ bool wave_any(bool v) {
    return __ballot(v) != 0ull;
}

__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float funkey(uint32_t k) {
    return (k & 0x80000000u) ? __uint_as_float(k & 0x7FFFFFFFu) : __uint_as_float(~k);
}

typedef unsigned int sw_u3 __attribute__((ext_vector_type(3)));

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

// The same march with the wave's vegsh and marching lanes in lane masks (SGPRs), for waves whose
// every lane has a >= 0 and vegetation >= 0 and steps with dz >= 0 (then every lane takes the exits,
// and an out-of-grid source is the lane's last step). f, the vegetation running maximum and vb stay
// per lane. The running maxima and compares run on every lane: a lane that reads no cell (stopped,
// invalid or past the grid) points its load past the buffer and gets sources 0 - dz <= 0, which
// leave f >= a >= 0 and the vegetation maximum >= 0 unchanged and compare false, exactly as the
// reference's zero source does for a lane at its out-of-grid step; vb counts only marching lanes.
// The vegsh update VS = (VS | V2) & ~SH leaves a stopped lane unchanged (after every step vegsh
// implies !sh).
template <bool TRACK>
__device__ __forceinline__ SunOut march_sun_masks(__amdgpu_buffer_rsrc_t hrs, uint32_t tbytes, int r, int c,
                                                  int rows, int cols, float at, float vegt, bool valid, bool wall,
                                                  const int4 *__restrict__ tab, int count, float amax, float hmax) {
    constexpr uint32_t kNoRead = 0xF0000000u;
    float f = at, svv = vegt;
    int vb = 0, stop = count;
    uint64_t VS = 0, A = __builtin_amdgcn_ballot_w64(valid);
    const uint64_t WALL = __builtin_amdgcn_ballot_w64(wall);
    for (int s = 0; s < count; ++s) {
        if (A == 0) break;
        const int4 e = tab[s];
        const float dzs = __int_as_float(e.z), dzp = __int_as_float(e.w);
        const uint64_t INB = __builtin_amdgcn_ballot_w64((unsigned)(r + e.x) < (unsigned)rows) &
                             __builtin_amdgcn_ballot_w64((unsigned)(c + e.y) < (unsigned)cols);
        const uint32_t off = tbytes + (uint32_t)(e.x * cols + e.y) * 16u;
        const uint32_t o = __builtin_amdgcn_inverse_ballot_w64(A & INB) ? off : kNoRead;
        const sw_u3 v = __builtin_amdgcn_raw_buffer_load_b96(hrs, o, 0, 0);
        const float hx = __uint_as_float(v.x), hy = __uint_as_float(v.y), hz = __uint_as_float(v.z);
        const float tmp = hx - dzs, tv = hy - dzs, tv2 = hz - dzs, lf = hy - dzp, lg = hz - dzp;
        f = fmaxf(f, tmp);
        svv = fmaxf(svv, tv);
        const uint64_t SH = __builtin_amdgcn_ballot_w64(f > at);
        const uint64_t c1 = __builtin_amdgcn_ballot_w64(tv > at), c2 = __builtin_amdgcn_ballot_w64(tv2 > at);
        const uint64_t c3 = __builtin_amdgcn_ballot_w64(lf > at), c4 = __builtin_amdgcn_ballot_w64(lg > at);
        // v2 = 1 when one to three of the four conditions hold (four count as none)
        const uint64_t V2 = (c1 | c2 | c3 | c4) & ~(c1 & c2 & c3 & c4);
        VS = (VS | V2) & ~SH;
        vb += __builtin_amdgcn_inverse_ballot_w64(VS & A) ? 1 : 0;
        const uint64_t low = __builtin_amdgcn_ballot_w64(hmax - dzp <= at);
        const uint64_t stopped = A & (~INB | (~WALL & ((SH & ~VS) | low)));
        if (TRACK) stop = __builtin_amdgcn_inverse_ballot_w64(stopped) ? s : stop;
        A &= ~stopped;
        if (!(amax >= dzs)) break;
    }
    SunOut out;
    out.sh = (f > at) ? 1 : 0;
    out.vs = __builtin_amdgcn_inverse_ballot_w64(VS) ? 1 : 0;
    out.vb = vb; out.shvo = f - at; out.shvoveg = svv; out.stop = stop;
    return out;
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
// walked, the step's diffuse-sky terms of each pair of patches (16 float2 per pair, one per (sh, vegsh)
// state of both) take its place.
#define SW_NPAIR ((SW_NPATCH + 1) / 2)
#define FLUX_TX 16
#define FLUX_TY 16

static inline size_t flux_lds_bytes(int walk_halo) {
    if (walk_halo < 0) return 0;
    const size_t tile = (size_t)(FLUX_TX + 2 * walk_halo) * (FLUX_TY + 2 * walk_halo);   // one flag byte per cell
    const size_t table = (size_t)SW_NPAIR * 16 * sizeof(float2);   // overlays the tile after the walk
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
    const int tid = sw_tid();
    const long n = (long)rows * cols;
    const float amax = funkey(stats[0]);
    const float hmax = funkey(stats[1]);
    long t;
    bool valid;
    int r, c;
    if (TILED) {
        // a 16 x 16 tile of the flux grid, 8 x 8 pixels per wave
        const int tiles_x = (cols + FLUX_TX - 1) / FLUX_TX;
        const int wv = tid >> 6, ln = tid & 63;
        const int rr = (tile / tiles_x) * FLUX_TY + (wv >> 1) * 8 + (ln >> 3);
        const int cc = (tile % tiles_x) * FLUX_TX + (wv & 1) * 8 + (ln & 7);
        valid = rr < rows && cc < cols;
        r = valid ? rr : 0;
        c = valid ? cc : 0;
        t = (long)r * cols + c;
    } else {
        t = (long)tile * 256 + tid;
        valid = (t < n);
        r = valid ? (int)(t / cols) : 0;
        c = valid ? (int)(t - (long)r * cols) : 0;
    }
    const SW_CONST sw_step_t *sp = (const SW_CONST sw_step_t *)(steps + step);
    const SW_CONST sw_model_t *mdc = (const SW_CONST sw_model_t *)model;
    float at = 0.f, vegt = 0.f, w = 0.f;
    if (valid) {
        const float4 h = hp[t];
        at = h.x;
        vegt = h.y;
        w = walls[t];
    }
    // waves whose lanes all take the exits (no negative height, dz >= 0) march in lane masks
    const int4 *__restrict__ stab = suntab + sp->sun_off;
    const int scount = sp->sun_count;
    const bool masks = (scount == 0 || __int_as_float(stab[scount - 1].z) >= 0.f) &&
                       __builtin_amdgcn_ballot_w64(valid && !(at >= 0.f && vegt >= 0.f)) == 0;
    SunOut o;
    if (masks) {
        const __amdgpu_buffer_rsrc_t hrs = __builtin_amdgcn_make_buffer_rsrc((void *)hp, (short)0,
                                                                             (int)(n * 16), 0x00020000);
        o = march_sun_masks<PLANES>(hrs, valid ? (uint32_t)t * 16u : 0u, r, c, rows, cols, at, vegt, valid,
                                    w > 0.f, stab, scount, amax, hmax);
    } else {
        o = march_sun_pixel<true, PLANES>(hp, rows, cols, r, c, at, vegt, valid, w > 0.f, stab, scount, amax,
                                          hmax);
    }
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
    const float se = mdc->sbc * mdc->emis_class[k];
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
// Kside_veg_v2022a), each product rounded in the reference's order, summed over the patch pair
// (2j, 2j + 1) for each of the 16 states of the pair: entry 16 j + i, bits of i {sh_2j, sh_2j+1,
// vegsh_2j, vegsh_2j+1}. One LDS read per pair replaces two; the reference adds the patches one at
// a time, so the pair sum reassociates its sum (float32 rounding only).
__device__ __forceinline__ float2 diffuse_term(const SW_CONST sw_step_t *__restrict__ sp, int p, int st) {
    const float d00 = 0.f - (1.f - 0.f) * 0.97f, d10 = 1.f - (1.f - 0.f) * 0.97f;  // diffsh for vegsh == 0
    const float diffsh = (st & 2) ? ((st & 1) ? 1.f : 0.f) : ((st & 1) ? d10 : d00);
    const sw_patch_k_t q = sp->pk[p];
    return make_float2(diffsh * q.lv, diffsh * q.lumchi * q.angl * q.ster);
}

__device__ __forceinline__ void diffuse_table(const SW_CONST sw_step_t *__restrict__ sp, float2 *__restrict__ tab) {
    const int tid = sw_tid();
    for (int i = tid; i < SW_NPAIR * 16; i += blockDim.x) {
        const int j = i >> 4, b = i & 15, p = 2 * j;
        float2 v = diffuse_term(sp, p, (b & 1) | ((b >> 1) & 2));
        if (p + 1 < SW_NPATCH) {   // the last pair has one patch: its entries ignore the second's bits
            const float2 c = diffuse_term(sp, p + 1, ((b >> 1) & 1) | ((b >> 2) & 2));
            v = make_float2(v.x + c.x, v.y + c.y);
        }
        tab[i] = v;
    }
}

// The 153 patches are 8 altitude bands of consecutive patches (31, 30, 28, 24, 19, 13, 7 and 1
// patches at 6, 18, ..., 90 degrees; the reference's patch_option 2 vault). Steradian, cos_alt and
// sin_alt are the same for every patch of a band, so every per-patch constant of a flag-only sum
// (lsky_*, lv_*, k_sh, lh_*, refl * ster * cos | sin) is the same within a band too: a sum
// sum_p flag_p * c_p is sum_b n_b * c_b, where n_b counts the set flags of band b (an exact small
// integer). That is 8 fma per sum instead of 153. The reference adds the patches one at a time
// (solweig.py:1514-1611 and 1615-1625), so the two differ by float32 rounding of the sum only.
#define SW_NBAND 8
__device__ constexpr int kBandStart[SW_NBAND + 1] = {0, 31, 61, 89, 113, 132, 145, 152, SW_NPATCH};

// set flags of band b in a 153-bit patch flag (bit p % 32 of word p / 32; bits past patch 152 ignored)
__device__ __forceinline__ float band_count(const uint32_t *f, int b) {
    uint32_t c = 0u;
#pragma unroll
    for (int w = 0; w < SW_MASK_WORDS; ++w) {
        const int lo = max(kBandStart[b] - 32 * w, 0), hi = min(kBandStart[b + 1] - 32 * w, 32);
        if (lo < hi) {
            const uint32_t m = (hi - lo == 32) ? 0xffffffffu : (((1u << (hi - lo)) - 1u) << lo);
            c += (uint32_t)__builtin_popcount(f[w] & m);
        }
    }
    return (float)c;
}

template <bool DAY, bool ROLLED>
__device__ __forceinline__ void patch_pass_a(const SW_CONST sw_step_t *__restrict__ sp, const uint32_t *msh,
                                             const uint32_t *mvs, const uint32_t *mv1, const uint32_t *mv2,
                                             float hsvf, const float2 *__restrict__ s_diff, PassA &o) {
    float aniLum = 0.f, KsideD = 0.f, Kref_sun = 0.f, Kref_sh = 0.f, Kref_veg = 0.f;
    float Ldown_sky = 0.f, Ldown_veg = 0.f, Ldown_sun = 0.f, Ldown_sh = 0.f;
    float Lside_sky = 0.f, Lside_veg = 0.f, Lside_sun = 0.f, Lside_sh = 0.f;
    // per-patch booleans as 153-bit flags: tvs = vegsh == 0 or vbsh == 0; tsky = sh and vegsh;
    // tsh = (1 - sh) * vbsh == 1
    uint32_t ftvs[SW_MASK_WORDS], ftsky[SW_MASK_WORDS], ftsh[SW_MASK_WORDS];
#pragma unroll
    for (int w = 0; w < SW_MASK_WORDS; ++w) {
        ftvs[w] = ~mvs[w] | ~(mv1[w] | mv2[w]);
        ftsky[w] = msh[w] & mvs[w];
        ftsh[w] = ~msh[w] & mv1[w];
    }
    // flag-only sums by band counts, in band order
#pragma unroll
    for (int b = 0; b < SW_NBAND; ++b) {
        const SW_CONST sw_patch_k_t &q = sp->pk[kBandStart[b]];
        const float nsky = band_count(ftsky, b), nvs = band_count(ftvs, b);
        Ldown_sky = __builtin_fmaf(nsky, q.lsky_down, Ldown_sky);
        Lside_sky = __builtin_fmaf(nsky, q.lsky_side, Lside_sky);
        Lside_veg = __builtin_fmaf(nvs, q.lv_side, Lside_veg);
        Ldown_veg = __builtin_fmaf(nvs, q.lv_down, Ldown_veg);
        if (DAY) Kref_veg = __builtin_fmaf(nvs, q.k_sh, Kref_veg);
        if (!DAY) {
            // night steps have no sun-facing window (l_sun == 0 and sunlit_l == 0 for every patch,
            // solweig.py:1514-1611 with altitude <= 0): every shaded wall patch counts as shaded
            const SW_CONST sw_patch_l_t &u = sp->pl[kBandStart[b]];
            const float nsh = band_count(ftsh, b);
            Lside_sh = __builtin_fmaf(nsh, u.lh_side, Lside_sh);
            Ldown_sh = __builtin_fmaf(nsh, u.lh_down, Ldown_sh);
        }
    }
    if (DAY) {
#pragma unroll
        for (int w = 0; w < SW_MASK_WORDS; ++w) {
            const uint32_t ash = w == 0 ? msh[0] : w == 1 ? msh[1] : w == 2 ? msh[2] : w == 3 ? msh[3] : msh[4];
            const uint32_t avs = w == 0 ? mvs[0] : w == 1 ? mvs[1] : w == 2 ? mvs[2] : w == 3 ? mvs[3] : mvs[4];
            const uint32_t btsh = w == 0 ? ftsh[0] : w == 1 ? ftsh[1] : w == 2 ? ftsh[2] : w == 3 ? ftsh[3] : ftsh[4];
            const int pend = min(w * 32 + 32, SW_NPATCH);
            // the diffuse sky (per-patch Perez luminance) by patch pairs, the wall patches
            // (per-patch sun thresholds) per patch in the reference's patch order
            auto wall = [&](int p, uint32_t k) {
                const bool tsh = __builtin_amdgcn_ubfe(btsh, k, 1u) != 0u;
                // wall-patch terms need tsh; a wave whose lanes all have tsh == 0 adds
                // only zeros there and skips them
                if (wave_any(tsh)) {
                    const sw_patch_k_t q = sp->pk[p];
                    const sw_patch_l_t u = sp->pl[p];
                    const float x = hsvf + q.yi_k;
                    // sum + c * b with b in {0, 1}: fma(b, c, sum) rounds once, like the reference
                    Kref_sun = __builtin_fmaf((tsh && x < q.sun_lt) ? 1.f : 0.f, q.k_sun, Kref_sun);
                    Kref_sh = __builtin_fmaf((tsh && x > q.sun_gt) ? 1.f : 0.f, q.k_sh, Kref_sh);
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
            auto pair = [&](int p) {   // p even: patches p and p + 1 share a word
                const uint32_t k = (uint32_t)(p - w * 32);
                const uint32_t st = __builtin_amdgcn_ubfe(ash, k, 2u) | (__builtin_amdgcn_ubfe(avs, k, 2u) << 2);
                const float2 dv = s_diff[8 * p + (int)st];   // 16 (p / 2) + st
                aniLum = aniLum + dv.x;
                KsideD = KsideD + dv.y;
                wall(p, k);
                if (p + 1 < pend) wall(p + 1, k + 1u);
            };
            // ROLLED (the walking kernel, which holds both passes): no unrolling. Unrolled, its
            // iterations keep several patches' scalar constants live at once and the register
            // allocator spills.
            if (ROLLED) {
#pragma unroll 1
                for (int p = w * 32; p < pend; p += 2) pair(p);
            } else {
#pragma unroll 2
                for (int p = w * 32; p < pend; p += 2) pair(p);
            }
        }
    }
    o.aniLum = aniLum; o.KsideD = KsideD; o.Kref_sun = Kref_sun; o.Kref_sh = Kref_sh; o.Kref_veg = Kref_veg;
    o.Ldown_sky = Ldown_sky; o.Ldown_veg = Ldown_veg; o.Ldown_sun = Ldown_sun; o.Ldown_sh = Ldown_sh;
    o.Lside_sky = Lside_sky; o.Lside_veg = Lside_veg; o.Lside_sun = Lside_sun; o.Lside_sh = Lside_sh;
}

// Ground-view walk of one azimuth (sunonsurface_2018a, solweig.py:136-288).
// Every boolean of the walk is a float 0 or 1, as in the reference: the
// running minimum of building flags is a product f, the "wall seen" flag is
// a maximum (of 0/1 values: max(bub, tw * f) = min(1, bub + tw * f), one fma
// whose result the median clamps to [0, 1]), and "wall or building seen" (bubw) equals 1 - f because f never
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
                                             const SW_CONST sw_model_t *__restrict__ md, int j, int nsec, int nfirst,
                                             int lbase, int lw, int r, int c, int rows, int cols, float fb,
                                             WalkAcc &acc, WalkAcc &snap) {
    float f = fb, bub = 0.f;
    float tl = 0.f, ta = 0.f, tn = 0.f, tb = 0.f, tw = 0.f;
    bool live = true;
    acc.kb = 0.f; acc.fs = 0.f; acc.wlupsh = 0.f; acc.walbsh = 0.f; acc.walbnosh = 0.f;
    const SW_CONST int32_t *wdx = md->walk_dx + j * SW_WALK_MAX, *wdy = md->walk_dy + j * SW_WALK_MAX;
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
        bub = __builtin_amdgcn_fmed3f(__builtin_fmaf(tw, f, bub), 0.f, 1.f);                 \
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
        bub = __builtin_amdgcn_fmed3f(__builtin_fmaf(D.w, f, bub), 0.f, 1.f);                \
        acc.kb += bub;                                                                       \
        acc.fs += f;                                                                         \
    }
#define SW_WALK_GROUP(S)                                                                     \
    if (f != 0.f) {                                                                          \
        const uint32_t e0 = s_src[lbase + wdx[S] * lw + wdy[S]];                             \
        const uint32_t e1 = s_src[lbase + wdx[S + 1] * lw + wdy[S + 1]];                     \
        const uint32_t e2 = s_src[lbase + wdx[S + 2] * lw + wdy[S + 2]];                     \
        const uint32_t e3 = s_src[lbase + wdx[S + 3] * lw + wdy[S + 3]];                     \
        const float4 d0 = s_dec[e0], d1 = s_dec[e1], d2 = s_dec[e2], d3 = s_dec[e3];          \
        SW_WALK_ACC(e0, d0) SW_WALK_ACC(e1, d1) SW_WALK_ACC(e2, d2) SW_WALK_ACC(e3, d3)      \
    } else {                                                                                 \
        acc.kb = __builtin_fmaf(bub, 4.f, acc.kb);                                           \
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
    const int tid = sw_tid();
    extern __shared__ uint2 s_src[];
    __shared__ float s_alb[SW_NCLASS];
    __shared__ float4 s_dec[128];                 // walk source decode by flag bits 0-6: {alb * shadow, alb, ground, sunwall}
    __shared__ float s_tlw[SW_WALK_MAX + 1];      // k-fold sequential sums of lwall
    __shared__ float s_tal[SW_WALK_MAX + 1];      // k-fold sequential sums of albedo_b
    const SW_CONST sw_step_t *sp = (const SW_CONST sw_step_t *)(steps + step);
    const SW_CONST sw_model_t *md = (const SW_CONST sw_model_t *)model;
    const int tiles_x = (cols + FLUX_TX - 1) / FLUX_TX;
    const int tiles_y = (rows + FLUX_TY - 1) / FLUX_TY;
    if (tile >= tiles_x * tiles_y) return;  // whole block idle (uniform)
    const int r0 = (tile / tiles_x) * FLUX_TY, c0 = (tile % tiles_x) * FLUX_TX;
    // each wave covers an 8 x 8 square of the 16 x 16 tile
    const int wv = tid >> 6, ln = tid & 63;
    const int ty = (wv >> 1) * 8 + (ln >> 3), tx = (wv & 1) * 8 + (ln & 7);
    const int r = r0 + ty, c = c0 + tx;
    const bool valid = (r < rows) && (c < cols);
    const long n = (long)rows * cols;
    const long t = valid ? (long)r * cols + c : 0;
    const bool day = WALK && sp->is_day != 0;
    const int lw = FLUX_TX + 2 * halo, lh = FLUX_TY + 2 * halo;
    const float omp = sp->one_m_psi;
    if (tid < SW_NCLASS) s_alb[tid] = md->alb_class[tid];
    // halo < walk_halo (incl. the tile-less launch, halo = -1) refuses a day step below
    const bool staged = day && halo >= md->walk_halo;
    if (staged) {
        // stage the walk sources of the tile + halo (out-of-grid cells are
        // never read: the walk tracks liveness on global coordinates)
        for (int i = tid; i < lw * lh; i += blockDim.x) {
            const int ly = i / lw, lx = i - ly * lw;
            const int gr = r0 - halo + ly, gc = c0 - halo + lx;
            uint32_t v = 0u;
            if ((unsigned)gr < (unsigned)rows && (unsigned)gc < (unsigned)cols)
                v = ((const uint32_t *)src)[2 * ((long)gr * cols + gc) + 1];
            ((uint8_t *)s_src)[i] = (uint8_t)(v >> 4);
        }
        if (tid < 128) {
            // decode of a walk source's flag byte {sh, vegsh, sunwall, ground, class}: its Lup source
            // (the sun step's expression, so the same bits), albedo x shadow, albedo signed by the
            // ground flag (the sign bit carries it, also for albedo 0), sunwall
            const int shb = tid & 1, vsb = (tid >> 1) & 1, k = (tid >> 4) & 7;
            const float al = md->alb_class[k];
            const float shd = (float)shb - (1.f - (float)vsb) * omp;
            const float se = md->sbc * md->emis_class[k];
            const float xx = sp->tg_class[k] * shd + sp->ta + 273.15f;
            const float tl = se * pow4(xx) - se * sp->tak4;
            s_dec[tid] = make_float4(tl, al * shd, ((tid >> 3) & 1) ? al : -al,
                                             (float)((tid >> 2) & 1));
        }
        if (tid == 128) {
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
    // that starts at +0 and only grows); the flag-only sums go by band
    // counts, the diffuse-sky and wall-patch sums per patch ----
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
    // reflected term uses patches where any of sh, vegsh, vbsh is 0; per band, the reference's
    // per-patch product (refl * ster) * cos | sin, added once per flagged patch (band counts)
    uint32_t ftr[SW_MASK_WORDS];
#pragma unroll
    for (int w = 0; w < SW_MASK_WORDS; ++w) ftr[w] = ~(msh[w] & mvs[w] & (mv1[w] | mv2[w]));
#pragma unroll
    for (int b = 0; b < SW_NBAND; ++b) {
        const SW_CONST sw_patch_k_t &q = sp->pk[kBandStart[b]];
        const float ntr = band_count(ftr, b);
        const float rs = refl * q.ster;
        Lside_ref = __builtin_fmaf(ntr, rs * q.cos_alt, Lside_ref);
        Ldown_ref = __builtin_fmaf(ntr, rs * q.sin_alt, Ldown_ref);
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
