// Wall heights and wall aspect of a DSM tile on the device
// (solweig_gpu/walls_aspect.py:27-166).
//
//   k_find_walls   4-neighbour maximum minus the cell, limit, zero border   (:27-58)
//   k_wall_aspect  rotating line filter over 180 rotations per wall cell,
//                  building-side decision from float32 DSM sums, gradient
//                  fallback for cells left at 0                             (:78-166)
//
// One thread per cell; a wave is 64 consecutive cells of one row. The
// rotated filters are a small table built by the host with the same
// scipy.ndimage.rotate calls as the reference (an exact table beats a
// restated interpolation: the rounded masks decide ties at 0.5). The table
// index h is uniform over the wave, so its counts and offsets are scalar
// loads. Every window access is clamped to the grid, so a malformed table
// cannot read outside the planes.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "sw_internal.h"

#define set_err swi_set_err
#define hip_check swi_hip_check
#define S swi_stream

namespace {

// float32 division rounded once: the double quotient of two floats rounds to
// the correctly rounded float quotient (53 >= 2 * 24 + 2).
__device__ __forceinline__ float fdiv_rn(float a, float b) { return (float)((double)a / (double)b); }

__device__ __forceinline__ bool wave_any(bool v) { return __ballot(v) != 0ull; }

struct Window {
    const float *__restrict__ p;
    int i, j, rows, cols;
    __device__ __forceinline__ float at(const int8_t *o) const {
        const int y = min(max(i + (int)o[0], 0), rows - 1);
        const int x = min(max(j + (int)o[1], 0), cols - 1);
        return p[(long)y * cols + x];
    }
};

// np.sum of a float32 selection (walls_aspect.py:149): the identity +0.0
// plus the pairwise sum, which is one running sum from -0.0 below 8 values
// and, from 8 to 128 values, eight running sums over blocks of 8 combined
// pairwise, then the remainder in order.
__device__ float np_sum_f32(const Window &w, const int8_t (*off)[2], int n) {
    if (n < 8) {
        float r = -0.0f;
        for (int k = 0; k < n; ++k) r += w.at(off[k]);
        return 0.0f + r;
    }
    float r0 = w.at(off[0]), r1 = w.at(off[1]), r2 = w.at(off[2]), r3 = w.at(off[3]);
    float r4 = w.at(off[4]), r5 = w.at(off[5]), r6 = w.at(off[6]), r7 = w.at(off[7]);
    int k = 8;
    const int lim = n - (n & 7);
    for (; k < lim; k += 8) {
        r0 += w.at(off[k + 0]);
        r1 += w.at(off[k + 1]);
        r2 += w.at(off[k + 2]);
        r3 += w.at(off[k + 3]);
        r4 += w.at(off[k + 4]);
        r5 += w.at(off[k + 5]);
        r6 += w.at(off[k + 6]);
        r7 += w.at(off[k + 7]);
    }
    float res = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
    for (; k < n; ++k) res += w.at(off[k]);
    return 0.0f + res;
}

__global__ __launch_bounds__(256) void k_find_walls(const float *__restrict__ a, float *__restrict__ w, int rows,
                                                    int cols, double lim) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= rows || j >= cols) return;
    const long p = (long)i * cols + j;
    float out = 0.f;
    if (i > 0 && j > 0 && i < rows - 1 && j < cols - 1) {
        const float n = a[p - cols], s = a[p + cols], e = a[p + 1], o = a[p - 1];
        // np.max over the 4 neighbours: NaN if any is NaN
        const float m = (isnan(n) || isnan(s) || isnan(e) || isnan(o)) ? NAN : fmaxf(fmaxf(n, o), fmaxf(e, s));
        double d = (double)m - (double)a[p];
        if (d < lim) d = 0.0;
        out = (float)d;
    }
    w[p] = out;
}

__global__ __launch_bounds__(256) void k_wall_aspect(const float *__restrict__ walls, const float *__restrict__ a,
                                                     const sw_wall_filters_t *__restrict__ F, float *__restrict__ out,
                                                     int rows, int cols, int fs, float d1, float d2) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    const bool in = i < rows && j < cols;
    const long p = (long)i * cols + j;
    const int r = fs >> 1;
    const bool table_ok = F->size == fs;
    const bool wall = in && walls[p] > 0.f;
    // the reference's loop bounds: ceil(fs/2)-1 <= i < rows-ceil(fs/2)-1 (walls_aspect.py:144-145)
    const bool act = wall && i >= r && i < rows - r - 2 && j >= r && j < cols - r - 2;
    int y = 0, x = 0;
    if (table_ok && wave_any(act)) {
        const Window ww{walls, i, j, rows, cols};
        const Window wa{a, i, j, rows, cols};
        int z = 0;
        for (int h = 0; h < SW_WALL_ROTATIONS; ++h) {
            const int nl = min(F->count[h][0], SW_WALL_LIST_CAP);
            if (act) {
                int c = 0;
                for (int k = 0; k < nl; ++k) c += ww.at(F->off[h][0][k]) > 0.f;
                if (c > z) {               // strictly larger: first maximum kept
                    z = c;
                    const float s1 = np_sum_f32(wa, F->off[h][1], min(F->count[h][1], SW_WALL_LIST_CAP));
                    const float s2 = np_sum_f32(wa, F->off[h][2], min(F->count[h][2], SW_WALL_LIST_CAP));
                    x = s1 > s2 ? 1 : 2;
                    y = 270 - h;
                }
            }
        }
    }
    if (!in) return;
    if (x == 1) y -= 180;
    if (y < 0) y += 360;
    // np.gradient of the float32 DSM with spacing 1/scale: central
    // differences inside, one-sided at the edges (walls_aspect.py:92)
    const float c = a[p];
    const float fy = i == 0 ? fdiv_rn(a[p + cols] - c, d1)
                   : i == rows - 1 ? fdiv_rn(c - a[p - cols], d1) : fdiv_rn(a[p + cols] - a[p - cols], d2);
    const float fx = j == 0 ? fdiv_rn(a[p + 1] - c, d1)
                   : j == cols - 1 ? fdiv_rn(c - a[p - 1], d1) : fdiv_rn(a[p + 1] - a[p - 1], d2);
    float res;
    if (!table_ok) {
        res = NAN;
    } else if (wall && y == 0) {
        // cart2pol(fy, fx): arctan2(fx, fy), negated, wrapped by float32(2 pi),
        // divided by float32(pi / 180) (walls_aspect.py:93-96, 163)
        float asp = (float)atan2((double)fx, (double)fy);
        asp = -asp;
        if (asp < 0.f) asp = asp + 6.28318548202514648438f;
        res = fdiv_rn(asp, 0.01745329238474369049f);
    } else {
        // y + 0 * aspect: NaN where the gradient is NaN
        res = (isnan(fx) || isnan(fy)) ? NAN : (float)y;
    }
    out[p] = res;
}

}  // namespace

extern "C" {

int sw_find_walls(const float *dsm, float *walls, int rows, int cols, double walllimit, void *stream) {
    int rc = swi_check_shape(rows, cols, "sw_find_walls");
    if (rc) return rc;
    if (!dsm || !walls) return set_err(SW_E_ARG, "sw_find_walls: NULL pointer");
    const dim3 blk(64, 4), grd((cols + 63) / 64, (rows + 3) / 4);
    if (grd.y > 65535) return set_err(SW_E_ARG, "sw_find_walls: more than 262140 rows");
    hipLaunchKernelGGL(k_find_walls, grd, blk, 0, S(stream), dsm, walls, rows, cols, walllimit);
    return hip_check(hipGetLastError(), "k_find_walls");
}

int sw_wall_filter_size(double scale) {
    const double f = floor((scale + 1e-10) * 9);
    if (!(f < 1e6)) return -1;  // NaN or absurd scale
    int fs = (int)f;
    if (fs <= 2)
        fs = 3;
    else if (fs != 9 && fs % 2 == 0)
        fs += 1;
    return fs;
}

int sw_wall_aspect(const float *walls, const float *dsm, const sw_wall_filters_t *filters, float *aspect, int rows,
                   int cols, double scale, void *stream) {
    int rc = swi_check_shape(rows, cols, "sw_wall_aspect");
    if (rc) return rc;
    if (!walls || !dsm || !filters || !aspect) return set_err(SW_E_ARG, "sw_wall_aspect: NULL pointer");
    if (rows < 2 || cols < 2) return set_err(SW_E_ARG, "sw_wall_aspect: rows and cols must be >= 2 (gradient)");
    if (!(scale > 0.0) || !isfinite(scale)) return set_err(SW_E_ARG, "sw_wall_aspect: scale must be finite and > 0");
    const int fs = sw_wall_filter_size(scale);
    if (fs < 3 || fs > 255) return set_err(SW_E_ARG, "sw_wall_aspect: filter size above 255 (scale too large)");
    const double dx = 1.0 / scale;
    const dim3 blk(64, 4), grd((cols + 63) / 64, (rows + 3) / 4);
    if (grd.y > 65535) return set_err(SW_E_ARG, "sw_wall_aspect: more than 262140 rows");
    hipLaunchKernelGGL(k_wall_aspect, grd, blk, 0, S(stream), walls, dsm, filters, aspect, rows, cols, fs,
                       (float)dx, (float)(2.0 * dx));
    return hip_check(hipGetLastError(), "k_wall_aspect");
}

}  // extern "C"
