/*
 * SOLWEIG CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's torch-CPU algorithm for the
 * full-solve hot path, used to check the HIP kernels (tests, smoke check)
 * and as the timed CPU baseline of bench.py. Nothing in the product links
 * or loads this file.
 *
 * Every routine keeps the reference's evaluation order and float32
 * rounding points (one rounding per torch elementwise op), and its
 * plane-sweep loop order where the reference sweeps planes:
 *   or_march_svf      shadow()                       solweig_gpu/shadow.py:166-322
 *   or_svf_fold       svf_calculator fold            solweig_gpu/shadow.py:545-613
 *   or_march_sun      shadowingfunction_wallheight_23 solweig_gpu/solweig.py:1079-1224
 *   or_gvf            gvf_2018a / sunonsurface_2018a solweig_gpu/solweig.py:75-402
 *   or_flux           Solweig_2022a_calc flux part   solweig_gpu/solweig.py:2101-2299
 *                     (TsWaveDelay :451-488, cylindric_wedge :404-449,
 *                      Kup_veg :490-508, Kside_veg :564-832,
 *                      shaded_or_sunlit :524-562, define_patch :1439-1648)
 *   or_utci           utci_calculator                calculate_utci.py:259-330
 *
 * Transcendentals are evaluated in double and rounded to float (a
 * correctly rounded float32 function in all but vanishingly rare cases);
 * the reference's torch lanes are within 1 ulp of that, which the float
 * tolerances absorb. Comparisons, masks and table-driven marches involve
 * no transcendental and are bit-exact.
 *
 * Parallelism: OpenMP over rows inside each plane operation.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "utci_terms.inc"

#define NP 153

static inline float tmaxf(float x, float y) { return (x != x || x > y) ? x : y; }
static inline float tminf(float x, float y) { return (x != x || x < y) ? x : y; }
static inline float f_sin(float x) { return (float)sin((double)x); }
static inline float f_cos(float x) { return (float)cos((double)x); }
static inline float f_tan(float x) { return (float)tan((double)x); }
static inline float f_atan(float x) { return (float)atan((double)x); }
static inline float f_acos(float x) { return (float)acos((double)x); }
static inline float f_asin(float x) { return (float)asin((double)x); }
static inline float f_exp(float x) { return (float)exp((double)x); }
static inline float f_log(float x) { return (float)log((double)x); }
static inline float f_pow(float x, float e) { return (float)pow((double)x, (double)e); }
static inline float f_sqrt(float x) { return sqrtf(x); }

/* torch.pow with a scalar exponent: 2 and 3 are products, others pow. */
static inline float t_powi(float x, int e) {
    if (e == 0) return 1.0f;
    if (e == 1) return x;
    if (e == 2) return x * x;
    if (e == 3) return x * x * x;
    return f_pow(x, (float)e);
}

/* ------------------------------------------------------------------------ */
/* SVF patch march (shadow(), bush-free)                                     */
/* ------------------------------------------------------------------------ */
void or_march_svf(const float *a, const float *veg, const float *veg2, int rows, int cols,
                  const int *dxs, const int *dys, const float *dzs, int count,
                  float *sh, float *vegsh, float *vbsh, float *fwork)
{
    const long n = (long)rows * cols;
    for (long k = 0; k < n; ++k) { fwork[k] = a[k]; sh[k] = 0.f; vegsh[k] = 0.f; vbsh[k] = 0.f; }
    for (int s = 0; s < count; ++s) {
        const int dx = dxs[s], dy = dys[s];
        const float dz = dzs[s];
        const int first = (s == 0);
        #pragma omp parallel for schedule(static)
        for (int r = 0; r < rows; ++r) {
            const int sr = r + dx;
            const int rin = (sr >= 0 && sr < rows);
            for (int c = 0; c < cols; ++c) {
                const long t = (long)r * cols + c;
                const int sc = c + dy;
                float tmp = 0.f, tv = 0.f, tv2 = 0.f;
                if (rin && sc >= 0 && sc < cols) {
                    const long q = (long)sr * cols + sc;
                    tv = veg[q] - dz;
                    tv2 = veg2[q] - dz;
                    tmp = a[q] - dz;
                }
                const float at = a[t];
                float f = tmaxf(fwork[t], tmp);
                fwork[t] = f;
                float shv = sh[t];
                if (f > at) shv = 1.f;
                if (f <= at) shv = 0.f;
                sh[t] = shv;
                const float v2 = (float)(tv > at) - (float)(tv2 > at);
                float vs = tmaxf(vegsh[t], v2);
                if (vs * shv > 0.f) vs = 0.f;
                float vb = vs + vbsh[t];
                if (first) {
                    float fv = tv - tmp;
                    if (fv <= 0.f) fv = 1000.f;
                    if (fv < dz) vs = 1.f;
                    vs = vs * (float)(veg2[t] > at);
                    vb = 0.f;
                }
                vegsh[t] = vs;
                vbsh[t] = vb;
            }
        }
    }
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < n; ++k) {
        sh[k] = 1.f - sh[k];
        float vb = vbsh[k];
        if (vb > 0.f) vb = 1.f;
        vb = vb - vegsh[k];
        float vs = vegsh[k];
        if (vs > 0.f) vs = 1.f;
        vegsh[k] = 1.f - vs;
        vbsh[k] = 1.f - vb;
    }
}

/* ------------------------------------------------------------------------ */
/* SVF fold: per patch p, per annulus k of its ring, in the reference order. */
/*   w_iso/w_aniso: (8,12) ring weights; ring[p]; nann[8]; dirm[p] bits     */
/*   E=1 S=2 W=4 N=8. Masks are (NP, n) float cubes (patch-major).          */
/*   out: 15 fields in order svf,E,S,W,N, veg,vegE,S,W,N, aveg,E,S,W,N + tot */
/* ------------------------------------------------------------------------ */
void or_svf_fold(const float *shm, const float *vegm, const float *vbm, long n,
                 const float *w_iso, const float *w_aniso, const int *ring, const int *nann,
                 const int *dirm, const float *veg2, float *out)
{
    #pragma omp parallel for schedule(static)
    for (long t = 0; t < n; ++t) {
        float s[5] = {0}, v[5] = {0}, b[5] = {0};
        for (int p = 0; p < NP; ++p) {
            const float sh = shm[(long)p * n + t], vs = vegm[(long)p * n + t], vb = vbm[(long)p * n + t];
            const int i = ring[p], dm = dirm[p];
            for (int k = 0; k < nann[i]; ++k) {
                const float wi = w_iso[i * 12 + k], wa = w_aniso[i * 12 + k];
                s[0] = s[0] + wi * sh;
                const float wsh = wa * sh;
                if (dm & 1) s[1] = s[1] + wsh;
                if (dm & 2) s[2] = s[2] + wsh;
                if (dm & 4) s[3] = s[3] + wsh;
                if (dm & 8) s[4] = s[4] + wsh;
                v[0] = v[0] + wi * vs;
                b[0] = b[0] + wi * vb;
                if (dm & 1) { v[1] = v[1] + wa * vs; b[1] = b[1] + wa * vb; }
                if (dm & 2) { v[2] = v[2] + wa * vs; b[2] = b[2] + wa * vb; }
                if (dm & 4) { v[3] = v[3] + wa * vs; b[3] = b[3] + wa * vb; }
                if (dm & 8) { v[4] = v[4] + wa * vs; b[4] = b[4] + wa * vb; }
            }
        }
        const float c = 3.0459e-004f;
        s[2] = s[2] + c;
        s[3] = s[3] + c;
        for (int q = 0; q < 5; ++q) if (s[q] > 1.f) s[q] = 1.f;
        const float last = (veg2[t] == 0.f) ? c : 0.f;
        v[2] = v[2] + last; v[3] = v[3] + last; b[2] = b[2] + last; b[3] = b[3] + last;
        for (int q = 0; q < 5; ++q) { if (v[q] > 1.f) v[q] = 1.f; if (b[q] > 1.f) b[q] = 1.f; }
        for (int q = 0; q < 5; ++q) {
            out[(long)q * n + t] = s[q];
            out[(long)(5 + q) * n + t] = v[q];
            out[(long)(10 + q) * n + t] = b[q];
        }
        const float trans = 0.03f;
        out[15L * n + t] = s[0] - (1.f - v[0]) * (1.f - trans);
    }
}

/* ------------------------------------------------------------------------ */
/* Sun march with dzprev + wall tail (shadowingfunction_wallheight_23)      */
/*   out planes: vegsh, sh, vbsh, wallsh, wallsun, wallshve, facesh, facesun */
/* ------------------------------------------------------------------------ */
void or_march_sun(const float *a, const float *veg, const float *veg2, int rows, int cols,
                  const int *dxs, const int *dys, const float *dzs, const float *dzp, int count,
                  const float *walls, const float *aspect_rad, int branch, float azilow, float azihigh,
                  float *out, float *work)
{
    const long n = (long)rows * cols;
    float *o_vegsh = out, *o_sh = out + n, *o_vbsh = out + 2 * n, *o_wallsh = out + 3 * n;
    float *o_wallsun = out + 4 * n, *o_wallshve = out + 5 * n, *o_facesh = out + 6 * n, *o_facesun = out + 7 * n;
    float *f = work, *shvoveg = work + n;
    for (long k = 0; k < n; ++k) {
        f[k] = a[k]; shvoveg[k] = veg[k]; o_sh[k] = 0.f; o_vbsh[k] = 0.f; o_vegsh[k] = 0.f;
    }
    for (int s = 0; s < count; ++s) {
        const int dx = dxs[s], dy = dys[s];
        const float dz = dzs[s], dzprev = dzp[s];
        #pragma omp parallel for schedule(static)
        for (int r = 0; r < rows; ++r) {
            const int sr = r + dx;
            const int rin = (sr >= 0 && sr < rows);
            for (int c = 0; c < cols; ++c) {
                const long t = (long)r * cols + c;
                const int sc = c + dy;
                float tmp = 0.f, tv = 0.f, tv2 = 0.f, lf = 0.f, lg = 0.f;
                if (rin && sc >= 0 && sc < cols) {
                    const long q = (long)sr * cols + sc;
                    tv = veg[q] - dz;
                    tv2 = veg2[q] - dz;
                    tmp = a[q] - dz;
                    lf = veg[q] - dzprev;
                    lg = veg2[q] - dzprev;
                }
                const float at = a[t];
                const float fv = tmaxf(f[t], tmp);
                f[t] = fv;
                shvoveg[t] = tmaxf(shvoveg[t], tv);
                const float shv = (fv > at) ? 1.f : 0.f;
                o_sh[t] = shv;
                float v2 = (float)(tv > at) + (float)(tv2 > at);
                v2 = v2 + (float)(lf > at);
                v2 = v2 + (float)(lg > at);
                if (v2 == 4.f) v2 = 0.f;
                if (v2 > 0.f) v2 = 1.f;
                float vs = tmaxf(o_vegsh[t], v2);
                if (vs * shv > 0.f) vs = 0.f;
                o_vegsh[t] = vs;
                o_vbsh[t] = o_vbsh[t] + vs;
            }
        }
    }
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < n; ++k) {
        const float at = a[k];
        const float shv = 1.f - o_sh[k];
        float vb = o_vbsh[k];
        if (vb > 0.f) vb = 1.f;
        vb = vb - o_vegsh[k];
        float vs = o_vegsh[k];
        if (vs > 0.f) vs = 1.f;
        const float svv = (shvoveg[k] - at) * vs;
        vs = 1.f - vs;
        vb = 1.f - vb;
        const float shvo = f[k] - at;
        const float w = walls[k];
        const float wallbol = (w > 0.f) ? 1.f : 0.f;
        const float asp = aspect_rad[k];
        float facesh;
        if (branch == 0) facesh = (float)((asp < azilow) || (asp >= azihigh)) - wallbol + 1.f;
        else facesh = (float)((asp > azilow) || (asp <= azihigh)) * -1.f + 1.f;
        const float facesun = ((facesh + wallbol == 1.f) && (w > 0.f)) ? 1.f : 0.f;
        float wallsun = w - shvo;
        if (wallsun < 0.f) wallsun = 0.f;
        if (facesh == 1.f) wallsun = 0.f;
        const float wallsh = w - wallsun;
        float wallshve = svv * wallbol;
        wallshve = wallshve - wallsh;
        if (wallshve < 0.f) wallshve = 0.f;
        wallsun = wallsun - wallshve;
        if (wallsun < 0.f) wallsun = 0.f;
        if (wallshve > w) wallshve = w;
        o_vegsh[k] = vs; o_sh[k] = shv; o_vbsh[k] = vb;
        o_wallsh[k] = wallsh; o_wallsun[k] = wallsun; o_wallshve[k] = wallshve;
        o_facesh[k] = facesh; o_facesun[k] = facesun;
    }
}

/* ------------------------------------------------------------------------ */
/* Ground view factors (gvf_2018a over 18 x sunonsurface_2018a).            */
/* in planes: buildings, shadow, Lup(src), alb, sunwall, walls, aspect_rad  */
/* tables: wdx/wdy (18, nsec); wbranch/wlo/whi (18); wdir (18)              */
/* out (15 planes): Lup, alb, albnosh, LupE, albE, albnoshE, LupS.., W.., N.. */
/*   (sums over azimuths; the caller normalises)                             */
/* ------------------------------------------------------------------------ */
void or_gvf(const float *bu, const float *shd, const float *lup, const float *alb,
            const float *sunwall, const float *walls, const float *aspect_rad, int rows, int cols,
            const int *wdx, const int *wdy, int nsec, const int *wbranch, const float *wlo,
            const float *whi, const int *wdir, float first, float second, float lwall,
            float albedo_b, float *out, float *work)
{
    const long n = (long)rows * cols;
    for (long k = 0; k < 15 * n; ++k) out[k] = 0.f;
    /* work planes */
    float *tb = work, *ts = work + n, *tl = work + 2 * n, *ta = work + 3 * n, *tan_ = work + 4 * n,
          *tw = work + 5 * n, *f = work + 6 * n, *bub = work + 7 * n, *bubw = work + 8 * n,
          *wsh = work + 9 * n, *wwall = work + 10 * n, *wlupsh = work + 11 * n, *wlwall = work + 12 * n,
          *walbsh = work + 13 * n, *walbwall = work + 14 * n, *walbnosh = work + 15 * n,
          *walbwallnosh = work + 16 * n;
    float *f_wwall = work + 17 * n, *f_wsh = work + 18 * n, *f_wlwall = work + 19 * n,
          *f_wlupsh = work + 20 * n, *f_walbwall = work + 21 * n, *f_walbsh = work + 22 * n,
          *f_walbwallnosh = work + 23 * n, *f_walbnosh = work + 24 * n;
    for (int j = 0; j < 18; ++j) {
        for (long k = 0; k < n; ++k) {
            tb[k] = ts[k] = tl[k] = ta[k] = tan_[k] = tw[k] = 0.f;
            f[k] = bu[k];
            bub[k] = bubw[k] = 0.f;
            wsh[k] = wwall[k] = wlupsh[k] = wlwall[k] = walbsh[k] = walbwall[k] = walbnosh[k] = walbwallnosh[k] = 0.f;
        }
        for (int s = 0; s < nsec; ++s) {
            const int dx = wdx[j * nsec + s], dy = wdy[j * nsec + s];
            const int snap = ((float)(s + 1) <= first);
            #pragma omp parallel for schedule(static)
            for (int r = 0; r < rows; ++r) {
                const int sr = r + dx;
                const int rin = (sr >= 0 && sr < rows);
                for (int c = 0; c < cols; ++c) {
                    const long t = (long)r * cols + c;
                    const int sc = c + dy;
                    if (rin && sc >= 0 && sc < cols) {
                        const long q = (long)sr * cols + sc;
                        tb[t] = bu[q]; ts[t] = shd[q]; tl[t] = lup[q];
                        ta[t] = alb[q] * shd[q]; tan_[t] = alb[q]; tw[t] = sunwall[q];
                    }
                    const float fv = tminf(f[t], tb[t]);
                    f[t] = fv;
                    wsh[t] = wsh[t] + ts[t] * fv;
                    wlupsh[t] = wlupsh[t] + tl[t] * fv;
                    walbsh[t] = walbsh[t] + ta[t] * fv;
                    walbnosh[t] = walbnosh[t] + tan_[t] * fv;
                    const float tempb = tw[t] * fv;
                    const float tempbwall = fv * -1.f + 1.f;
                    bub[t] = ((tempb + bub[t]) > 0.f) ? 1.f : 0.f;
                    bubw[t] = ((tempbwall + bubw[t]) > 0.f) ? 1.f : 0.f;
                    wlwall[t] = wlwall[t] + bub[t] * lwall;
                    walbwall[t] = walbwall[t] + bub[t] * albedo_b;
                    wwall[t] = wwall[t] + bub[t];
                    walbwallnosh[t] = walbwallnosh[t] + bubw[t] * albedo_b;
                    if (snap) {
                        f_wwall[t] = wwall[t]; f_wsh[t] = wsh[t]; f_wlwall[t] = wlwall[t];
                        f_wlupsh[t] = wlupsh[t]; f_walbwall[t] = walbwall[t]; f_walbsh[t] = walbsh[t];
                        f_walbwallnosh[t] = walbwallnosh[t]; f_walbnosh[t] = walbnosh[t];
                    }
                }
            }
        }
        const int dm = wdir[j];
        const int br = wbranch[j];
        const float lo = wlo[j], hi = whi[j];
        const float fp1 = first + 1.f, sp1 = second + 1.f;
        #pragma omp parallel for schedule(static)
        for (long t = 0; t < n; ++t) {
            const float wallbol = (walls[t] > 0.f) ? 1.f : 0.f;
            const float asp = aspect_rad[t];
            float facesh;
            if (br == 0) facesh = (float)((asp < lo) || (asp >= hi)) - wallbol + 1.f;
            else facesh = (float)((asp > lo) || (asp <= hi)) * -1.f + 1.f;
            float keep = (float)(wwall[t] == second) - facesh;
            if (keep == -1.f) keep = 0.f;
            const float wsi1 = (f_wwall[t] > 0.f) ? 1.f : 0.f;
            const float wi1 = (f_walbwallnosh[t] > 0.f) ? 1.f : 0.f;
            const float wsi2 = (wwall[t] > 0.f) ? 1.f : 0.f;
            const float wi2 = (walbwallnosh[t] > 0.f) ? 1.f : 0.f;
            float lw2 = wlwall[t], aw2 = walbwall[t];
            if (keep == 1.f) { lw2 = 0.f; aw2 = 0.f; }
            const float gl1 = ((f_wlwall[t] + f_wlupsh[t]) / fp1) * wsi1 + (f_wlupsh[t]) / first * (wsi1 * -1.f + 1.f);
            const float gl2 = ((lw2 + wlupsh[t]) / sp1) * wsi2 + (wlupsh[t]) / second * (wsi2 * -1.f + 1.f);
            const float ga1 = ((f_walbwall[t] + f_walbsh[t]) / fp1) * wsi1 + (f_walbsh[t]) / first * (wsi1 * -1.f + 1.f);
            const float ga2 = ((aw2 + walbsh[t]) / sp1) * wsi2 + (walbsh[t]) / second * (wsi2 * -1.f + 1.f);
            const float gn1 = ((f_walbwallnosh[t] + f_walbnosh[t]) / fp1) * wi1 + (f_walbnosh[t]) / first * (wi1 * -1.f + 1.f);
            const float gn2 = ((walbwallnosh[t] + walbnosh[t]) / second) * wi2 + (walbnosh[t]) / second * (wi2 * -1.f + 1.f);
            const float nb = bu[t] * -1.f + 1.f;
            float gLup = (gl1 * 0.5f + gl2 * 0.4f) / 0.9f;
            gLup = gLup + lup[t] * nb;
            float galb = (ga1 * 0.5f + ga2 * 0.4f) / 0.9f;
            galb = galb + alb[t] * nb * shd[t];
            float gnosh = (gn1 * 0.5f + gn2 * 0.4f) / 0.9f;
            gnosh = gnosh * bu[t] + alb[t] * nb;
            out[t] += gLup; out[n + t] += galb; out[2 * n + t] += gnosh;
            if (dm & 1) { out[3 * n + t] += gLup; out[4 * n + t] += galb; out[5 * n + t] += gnosh; }
            if (dm & 2) { out[6 * n + t] += gLup; out[7 * n + t] += galb; out[8 * n + t] += gnosh; }
            if (dm & 4) { out[9 * n + t] += gLup; out[10 * n + t] += galb; out[11 * n + t] += gnosh; }
            if (dm & 8) { out[12 * n + t] += gLup; out[13 * n + t] += galb; out[14 * n + t] += gnosh; }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* UTCI (utci_calculator + utci_polynomial)                                 */
/* ------------------------------------------------------------------------ */
static float utci_point(float ta, float rh, float tmrt, float va)
{
    if (ta <= -999.f || rh <= -999.f || va <= -999.f || tmrt <= -999.f) return -999.f;
    static const float g[8] = {-2.8365744E3f, -6.028076559E3f, 1.954263612E1f, -2.737830188E-2f,
                               1.6261698E-5f, 7.0229056E-10f, -1.8680009E-13f, 2.7150305f};
    const float tk = ta + 273.15f;
    float es = g[7] * f_log(tk);
    for (int i = 0; i < 7; ++i) {
        const int e = i + 1 - 3;
        float p;
        if (e == -2) p = 1.f / (tk * tk);
        else if (e == -1) p = 1.f / tk;
        else p = t_powi(tk, e);
        es = es + g[i] * p;
    }
    es = f_exp(es) * 0.01f;
    const float ehPa = es * rh / 100.f;
    const float D = tmrt - ta;
    const float Pa = ehPa / 10.0f;
    float pw[4][7];
    const float base[4] = {ta, va, D, Pa};
    for (int v = 0; v < 4; ++v)
        for (int e = 0; e <= 6; ++e) pw[v][e] = t_powi(base[v], e);
    float u = ta;
    for (int k = 0; k < OR_UTCI_NTERMS; ++k) {
        float term = OR_UTCI_COEF[k];
        for (int v = 0; v < 4; ++v) {
            const int e = OR_UTCI_EXP[k][v];
            if (e) term = term * pw[v][e];
        }
        u = u + term;
    }
    return u;
}

void or_utci(const float *ta, const float *rh, const float *tmrt, const float *va, long n, float *out)
{
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) out[i] = utci_point(ta[i], rh[i], tmrt[i], va[i]);
}

/* ------------------------------------------------------------------------ */
/* Flux assembly -> Tmrt -> UTCI for one time step.                         */
/* ------------------------------------------------------------------------ */
typedef struct {
    int is_day;
    float ta, ta_mat, rh, va;
    float radI, radD, radG, sinalt, cosalt_k, tan_zen, one_m_psi;
    float tak4;            /* f32((Ta+273.15)^4) */
    int tw_first, tw_update;
    float tw_w, tw_omw;
    float sunlit_k, shaded_k, sunlit_l, shaded_l, veg_l;
    float absK, absL, Fside, Fup, Fcyl, albedo_b, ewall, sbc;
} or_step_t;

typedef struct {
    const float *lv, *lumchi, *yi_k, *yi_l, *lsky_down, *lsky_side;
    const int *l_sun;
    const float *alt, *ster, *cos_alt, *sin_alt, *angl;
} or_patch_t;

/*
 * Values a caller may supply in place of the oracle's own derivations. The reference's flat
 * radiation-step bundle (solweig_core/numba_cpu/radiation.py:1266-1818, its pinned form in
 * native/cuda/data/rad_pins.json) carries them precomputed:
 *   fsh        F_sh plane (cylindric_wedge, solweig.py:404-449) instead of tan_svfalfa;
 *   sun_pb/shd_pb  the shaded_or_sunlit decisions (solweig.py:524-562), `stride` bytes per
 *              pixel, bit i of byte i / 8: Kside patch p at bit p, a guarded define_patch
 *              patch p at bit kside_n + rank[p];
 *   night_lup  night Lup plane instead of SBC emis (Ta + 273.15)^4.
 * NULL members (or a NULL struct) keep the oracle's derivations.
 */
typedef struct {
    const float *fsh;
    const uint8_t *sun_pb, *shd_pb;
    int stride, kside_n;
    const int *rank;
    const float *night_lup;
} or_override_t;

static inline float pbitf(const uint8_t *pb, long t, int stride, int i)
{
    return ((pb[t * stride + (i >> 3)] >> (i & 7)) & 1) ? 1.f : 0.f;
}

/*
 * Static per-pixel inputs: svfbuveg, tan_asvf, tan_svfalfa, buildings, emis, alb, tg (per-pixel
 * Tg of this step), masks sh/veg/vbsh cubes (NP, n).
 * day inputs: vegsh_sun/sh_sun (march), gvf (15 summed planes from or_gvf).
 * state (in/out): tgmap (6 planes: Lup, E, S, W, N, TgOut).
 * outputs: tmrt, utci, kdown, kup, ldown, lup, shadow; o_extra (NULL or 10 planes): Lside,
 * Kside, KsideI, KsideD, dRad, TgOut, Keast, Ksouth, Kwest, Knorth (zero K terms and TgOut = Ta
 * at night, as solweig.py:2183-2227 leaves them).
 */
void or_flux(const or_step_t *sp, const or_patch_t *pt, long n,
             const float *svfbuveg, const float *tan_asvf, const float *tan_svfalfa,
             const float *bu, const float *emis, const float *alb, const float *tg,
             const float *shm, const float *vegm, const float *vbm,
             const float *sh_sun, const float *vegsh_sun, const float *gvf,
             float *tgmap, float *o_tmrt, float *o_utci, float *o_kdown, float *o_kup,
             float *o_ldown, float *o_lup, float *o_shadow, const or_override_t *ov, float *o_extra)
{
    const int given_fsh = ov && ov->fsh;
    const int given_bits = ov && ov->sun_pb && ov->shd_pb;
    const int given_lup = ov && ov->night_lup;
    const float rad2deg = (float)(180.0 / 3.14159265358979323846);
    const float pi_f = (float)3.14159265358979323846;
    const float two_pi = (float)(2.0 * 3.14159265358979323846);
    #pragma omp parallel for schedule(static)
    for (long t = 0; t < n; ++t) {
        float Kdown = 0.f, Kup = 0.f, KupE = 0.f, KupS = 0.f, KupW = 0.f, KupN = 0.f, Kside = 0.f;
        float Lup, LupE, LupS, LupW, LupN, shadow = 0.f;
        float KsideI = 0.f, KsideD = 0.f, dRad = 0.f, TgOut = sp->ta;
        const float sbc_emis = sp->sbc * emis[t];
        if (sp->is_day) {
            shadow = sh_sun[t] - (1.f - vegsh_sun[t]) * sp->one_m_psi;
            /* ground view factor normalisation */
            const float lup_air = sbc_emis * sp->tak4;
            float g[15];
            for (int q = 0; q < 15; ++q) g[q] = gvf[(long)q * n + t];
            const float gLup = g[0] / 18.f + lup_air;
            const float galb = g[1] / 18.f, gnosh = g[2] / 18.f;
            const float gLupE = g[3] / 9.f + lup_air, galbE = g[4] / 9.f, gnoshE = g[5] / 9.f;
            const float gLupS = g[6] / 9.f + lup_air, galbS = g[7] / 9.f, gnoshS = g[8] / 9.f;
            const float gLupW = g[9] / 9.f + lup_air, galbW = g[10] / 9.f, gnoshW = g[11] / 9.f;
            const float gLupN = g[12] / 9.f + lup_air, galbN = g[13] / 9.f, gnoshN = g[14] / 9.f;
            /* thermal delay (TsWaveDelay_2015a x 6, same weights) */
            const float tg0[6] = {gLup, gLupE, gLupS, gLupW, gLupN, tg[t] * shadow + sp->ta};
            float L[6];
            for (int q = 0; q < 6; ++q) {
                float T1 = sp->tw_first ? tg0[q] : tgmap[(long)q * n + t];
                if (sp->tw_update) {
                    T1 = tg0[q] * sp->tw_omw + T1 * sp->tw_w;
                    L[q] = T1;
                } else {
                    L[q] = tg0[q] * sp->tw_omw + T1 * sp->tw_w;
                }
                tgmap[(long)q * n + t] = T1;
            }
            Lup = L[0]; LupE = L[1]; LupS = L[2]; LupW = L[3]; LupN = L[4];
            TgOut = L[5];
            /* cylinder wedge F_sh */
            float F_sh;
            if (given_fsh) {
                F_sh = ov->fsh[t];
            } else {
                const float ta_ = tan_svfalfa[t], tb = sp->tan_zen;
                const float xa = 1.f - 2.f / (ta_ * tb);
                const float ha = 2.f / (ta_ * tb);
                const float ba = 1.f / ta_;
                const float hkil = 2.f * ba * ha;
                float ukil = 0.f;
                if (xa < 0.f) {
                    const float qa = tb / 2.f;
                    const float Za = f_sqrt(ba * ba - (qa * qa) / 4.f);
                    const float phi = f_atan(Za / qa);
                    const float A = (f_sin(phi) - phi * f_cos(phi)) / (1.f - f_cos(phi));
                    ukil = 2.f * ba * xa * A;
                }
                const float Ssurf = hkil + ukil;
                F_sh = (two_pi * ba - Ssurf) / (two_pi * ba);
                if (F_sh != F_sh) F_sh = 0.5f;
            }
            /* diffuse sky + Kside patch loop */
            float aniLum = 0.f, Kref_sun = 0.f, Kref_sh = 0.f, Kref_veg = 0.f;
            const float hsvf = tan_asvf[t];
            for (int p = 0; p < NP; ++p) {
                const float sh = shm[(long)p * n + t], vs = vegm[(long)p * n + t], vb = vbm[(long)p * n + t];
                const float diffsh = sh - (1.f - vs) * 0.97f;
                aniLum = aniLum + diffsh * pt->lv[p];
                KsideD = KsideD + diffsh * pt->lumchi[p] * pt->angl[p] * pt->ster[p];
                const float tvs = (vs == 0.f || vb == 0.f) ? 1.f : 0.f;
                Kref_veg = Kref_veg + sp->shaded_k * tvs * pt->ster[p] * pt->cos_alt[p];
                const float tsh = ((1.f - sh) * vb == 1.f) ? 1.f : 0.f;
                float sunlit, shaded;
                if (given_bits) {
                    sunlit = pbitf(ov->sun_pb, t, ov->stride, p);
                    shaded = pbitf(ov->shd_pb, t, ov->stride, p);
                } else {
                    const float sd = f_atan(hsvf + pt->yi_k[p]) * rad2deg;
                    sunlit = (sd < pt->alt[p]) ? 1.f : 0.f;
                    shaded = (sd > pt->alt[p]) ? 1.f : 0.f;
                }
                Kref_sun = Kref_sun + sp->sunlit_k * sunlit * tsh * pt->ster[p] * pt->cos_alt[p];
                Kref_sh = Kref_sh + sp->shaded_k * shaded * tsh * pt->ster[p] * pt->cos_alt[p];
            }
            dRad = aniLum * sp->radD;
            const float sb = 1.f - svfbuveg[t];
            const float mix = sp->radG * (1.f - F_sh) + sp->radD * F_sh;
            Kdown = sp->radI * shadow * sp->sinalt + dRad + sp->albedo_b * sb * mix;
            const float kdiff = sp->radD * svfbuveg[t] + sp->albedo_b * sb * mix;
            Kup = galb * sp->radI * sp->sinalt + kdiff * gnosh;
            KupE = galbE * sp->radI * sp->sinalt + kdiff * gnoshE;
            KupS = galbS * sp->radI * sp->sinalt + kdiff * gnoshS;
            KupW = galbW * sp->radI * sp->sinalt + kdiff * gnoshW;
            KupN = galbN * sp->radI * sp->sinalt + kdiff * gnoshN;
            KsideI = shadow * sp->radI * sp->cosalt_k;
            Kside = KsideI + KsideD + Kref_sun + Kref_sh + Kref_veg;
        } else {
            const float x = sp->ta + 273.15f;
            Lup = given_lup ? ov->night_lup[t] : sbc_emis * t_powi(x, 4);
            LupE = LupS = LupW = LupN = Lup;
        }
        /* longwave: define_patch_characteristics, two passes */
        float Ldown_sky = 0.f, Ldown_veg = 0.f, Ldown_sun = 0.f, Ldown_sh = 0.f, Ldown_ref = 0.f;
        float Lside_sky = 0.f, Lside_veg = 0.f, Lside_sun = 0.f, Lside_sh = 0.f, Lside_ref = 0.f;
        const float hsvf = tan_asvf[t];
        for (int p = 0; p < NP; ++p) {
            const float sh = shm[(long)p * n + t], vs = vegm[(long)p * n + t], vb = vbm[(long)p * n + t];
            const float tsky = (sh == 1.f && vs == 1.f) ? 1.f : 0.f;
            Ldown_sky = Ldown_sky + tsky * pt->lsky_down[p];
            Lside_sky = Lside_sky + tsky * pt->lsky_side[p];
            const float tvs = (vs == 0.f || vb == 0.f) ? 1.f : 0.f;
            Lside_veg = Lside_veg + sp->veg_l * pt->ster[p] * pt->cos_alt[p] * tvs;
            Ldown_veg = Ldown_veg + sp->veg_l * pt->ster[p] * pt->sin_alt[p] * tvs;
            const float tsh = ((1.f - sh) * vb == 1.f) ? 1.f : 0.f;
            if (pt->l_sun[p]) {
                float sunlit, shaded;
                if (given_bits) {
                    sunlit = pbitf(ov->sun_pb, t, ov->stride, ov->kside_n + ov->rank[p]);
                    shaded = pbitf(ov->shd_pb, t, ov->stride, ov->kside_n + ov->rank[p]);
                } else {
                    const float sd = f_atan(hsvf + pt->yi_l[p]) * rad2deg;
                    sunlit = (sd < pt->alt[p]) ? 1.f : 0.f;
                    shaded = (sd > pt->alt[p]) ? 1.f : 0.f;
                }
                Lside_sun = Lside_sun + sp->sunlit_l * sunlit * pt->ster[p] * pt->cos_alt[p] * tsh;
                Lside_sh = Lside_sh + sp->shaded_l * shaded * pt->ster[p] * pt->cos_alt[p] * tsh;
                Ldown_sun = Ldown_sun + sp->sunlit_l * sunlit * pt->ster[p] * pt->sin_alt[p] * tsh;
                Ldown_sh = Ldown_sh + sp->shaded_l * shaded * pt->ster[p] * pt->sin_alt[p] * tsh;
            } else {
                Lside_sh = Lside_sh + sp->shaded_l * pt->ster[p] * pt->cos_alt[p] * tsh;
                Ldown_sh = Ldown_sh + sp->shaded_l * pt->ster[p] * pt->sin_alt[p] * tsh;
            }
        }
        const float refl = ((Ldown_sky + Lup) * (1.f - sp->ewall) * 0.5f) / pi_f;
        for (int p = 0; p < NP; ++p) {
            const float sh = shm[(long)p * n + t], vs = vegm[(long)p * n + t], vb = vbm[(long)p * n + t];
            const float tr = (sh == 0.f || vs == 0.f || vb == 0.f) ? 1.f : 0.f;
            Lside_ref = Lside_ref + refl * pt->ster[p] * pt->cos_alt[p] * tr;
            Ldown_ref = Ldown_ref + refl * pt->ster[p] * pt->sin_alt[p] * tr;
        }
        const float Lside = Lside_sky + Lside_veg + Lside_sh + Lside_sun + Lside_ref;
        const float Ldown = Ldown_sky + Ldown_veg + Ldown_sh + Ldown_sun + Ldown_ref;
        /* Tmrt (cylinder, anisotropic sky) */
        const float Keast = KupE * 0.5f, Kwest = KupW * 0.5f, Knorth = KupN * 0.5f, Ksouth = KupS * 0.5f;
        const float Least = LupE * 0.5f, Lwest = LupW * 0.5f, Lnorth = LupN * 0.5f, Lsouth = LupS * 0.5f;
        const float Sstr = sp->absK * (Kside * sp->Fcyl + (Kdown + Kup) * sp->Fup + (Knorth + Keast + Ksouth + Kwest) * sp->Fside)
                         + sp->absL * ((Ldown + Lup) * sp->Fup + Lside * sp->Fcyl + (Lnorth + Least + Lsouth + Lwest) * sp->Fside);
        const float tmrt = f_sqrt(f_sqrt(Sstr / (sp->absL * sp->sbc))) - 273.2f;
        o_tmrt[t] = tmrt;
        o_kdown[t] = Kdown; o_kup[t] = Kup; o_ldown[t] = Ldown; o_lup[t] = Lup; o_shadow[t] = shadow;
        if (o_extra) {
            const float ex[10] = {Lside, Kside, KsideI, KsideD, dRad, TgOut, Keast, Ksouth, Kwest, Knorth};
            for (int q = 0; q < 10; ++q) o_extra[(long)q * n + t] = ex[q];
        }
        if (bu[t] == 1.f) o_utci[t] = utci_point(sp->ta_mat, sp->rh, tmrt, sp->va);
        else o_utci[t] = NAN;
    }
}

/* per-pixel static derivations of the day loop (utci_process.py:867-897) */
void or_static(const float *svf, const float *svfveg, long n, float *svfbuveg, float *tan_asvf,
               float *tan_svfalfa)
{
    #pragma omp parallel for schedule(static)
    for (long t = 0; t < n; ++t) {
        svfbuveg[t] = svf[t] - (1.0f - svfveg[t]) * 0.97f;
        const float asvf = f_acos(f_sqrt(svf[t]));
        tan_asvf[t] = f_tan(asvf);
        float tmp = svf[t] + svfveg[t] - 1.0f;
        if (tmp < 0.f) tmp = 0.f;
        const float sa = f_asin(f_exp(f_log(1.0f - tmp) / 2.0f));
        tan_svfalfa[t] = f_tan(sa);
    }
}

/* GVF walk source planes of one day step (sunonsurface_2018a prologue,
 * solweig.py:121-134, and gvf_2018a sunwall, solweig.py:342):
 * shadow = sh - (1 - vegsh) * (1 - psi); Lup = SBC emis (Tg shadow + Ta + 273.15)^4 - SBC emis (Ta+273.15)^4;
 * sunwall = (wallsun / walls * buildings == 1). */
void or_walk_sources(const float *sh_sun, const float *vegsh_sun, const float *wallsun,
                     const float *walls, const float *bu, const float *emis, const float *tg, long n,
                     float one_m_psi, float ta, float tak4, float sbc, float *shadow, float *lup,
                     float *sunwall)
{
    #pragma omp parallel for schedule(static)
    for (long t = 0; t < n; ++t) {
        const float s = sh_sun[t] - (1.f - vegsh_sun[t]) * one_m_psi;
        shadow[t] = s;
        const float se = sbc * emis[t];
        const float x = tg[t] * s + ta + 273.15f;
        lup[t] = se * t_powi(x, 4) - se * tak4;
        const float q = wallsun[t] / walls[t] * bu[t];
        sunwall[t] = (q == 1.f) ? 1.f : 0.f;
    }
}

/* Threads an OpenMP region of this library runs on (what the timed CPU baseline used). */
int or_threads(void) { return omp_get_max_threads(); }
