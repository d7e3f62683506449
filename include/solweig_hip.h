/*
 * solweig_hip.h -- C ABI of libsolweig_hip.so (MI355X / gfx950).
 *
 * The SOLWEIG full-solve hot path behind a plain C boundary: sky-view-factor
 * sweep (153-patch ray march + annulus fold), sun march with the wall tail,
 * ground-view-factor walk and flux assembly into Tmrt with a fused UTCI
 * epilogue. All buffers are device pointers owned by the caller; the
 * library allocates nothing on the solve path. Every entry point takes the
 * HIP stream to enqueue on (NULL = legacy default stream), returns 0 on
 * success or a nonzero code, and records a message retrievable with
 * sw_last_error() (thread-local). Entry points enqueue work and do not
 * synchronise; call sw_sync() or synchronise the stream before reading.
 *
 * Conventions kept from the reference's native C ABI
 * (native/cuda/src/sw_api.cu:24-29, 64-85, 209-344): int return codes plus
 * a last-error string, "pointers first, scalars last", version and
 * build-flag queries.
 */
#ifndef SOLWEIG_HIP_H
#define SOLWEIG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SW_ABI_VERSION 1
#define SW_NPATCH 153          /* sky patches (patch_option 2) */
#define SW_MASK_WORDS 5        /* ceil(153 / 32) */
#define SW_MASK_PLANES 4       /* sh, vegsh, vbsh==1, vbsh==2 */
#define SW_NFIELDS 16          /* 15 SVF fields + SVFtotal */
#define SW_WALK_AZ 18          /* ground-view-factor search azimuths */
#define SW_WALK_MAX 64         /* max walk steps per azimuth */
#define SW_NCLASS 8            /* land-cover class codes 0..7 */
#define SW_SVF_PAD 16          /* border cells of the march's padded height buffer */

/* Error codes */
#define SW_OK 0
#define SW_E_ARG 1             /* invalid argument (shape, pointer, range) */
#define SW_E_HIP 2             /* HIP runtime error */
#define SW_E_UNSUPPORTED 3     /* input outside the supported domain */

/* Mask cube layout: masks[(plane * SW_MASK_WORDS + word) * n + pixel], bit
 * (patch % 32) of word (patch / 32). Planes: 0 sh (1 = sky not blocked by
 * buildings), 1 vegsh (1 = not blocked by vegetation), 2 vbsh == 1,
 * 3 vbsh == 2 (vbsh == 0 when neither bit is set). */

/* Per-patch constants of one step, laid out patch-major so the flux kernel
 * fetches a patch's record with one wide scalar load. Products are float32
 * in the reference's operation order; the kernel adds them where the
 * patch's boolean factors are 1. */
typedef struct sw_patch_k {
    float lv, lumchi, angl, ster;   /* Perez luminance, lumChi, incidence cosine, steradian */
    float k_sh;                      /* (shaded_k * ster) * cos_alt: Kref_veg, Kref_sh */
    float k_sun;                     /* (sunlit_k * ster) * cos_alt: Kref_sun */
    float yi_k;                      /* shortwave shaded_or_sunlit offset */
    float sun_lt;                    /* x < sun_lt  <=> atan(x)*180/pi < alt  (sunlit) */
    float sun_gt;                    /* x > sun_gt  <=> atan(x)*180/pi > alt  (shaded) */
    float lsky_down, lsky_side;      /* sky longwave per patch */
    float lv_side, lv_down;          /* (veg_l * ster) * cos_alt | sin_alt */
    float cos_alt, sin_alt;
    int32_t l_sun;                   /* longwave sun-facing window active */
} sw_patch_k_t;

typedef struct sw_patch_l {
    float yi_l;                      /* longwave shaded_or_sunlit offset */
    float ls_side, ls_down;          /* (sunlit_l * ster) * cos_alt | sin_alt */
    float lh_side, lh_down;          /* (shaded_l * ster) * cos_alt | sin_alt */
    float pad[3];
} sw_patch_l_t;

/* Per-time-step constants (host-computed scalar chain), one per step in
 * device memory so a captured day replays without host work. */
typedef struct sw_step {
    int32_t is_day;
    int32_t sun_off;          /* offset of this step's sun march table */
    int32_t sun_count;        /* steps in the sun march table (day) */
    int32_t facesh_branch;    /* 0, 1, 2: face-shading window form */
    float azilow, azihigh;    /* face-shading window (radians, float32) */
    float ta, ta_mat, rh, va, tak4;
    float radI, radD, radG, sinalt, cosalt_k, tan_zen, one_m_psi;
    int32_t tw_first, tw_update;
    float tw_w, tw_omw;
    float sunlit_k, shaded_k, sunlit_l, shaded_l, veg_l;
    float tgwall, lwall;
    float tg_class[SW_NCLASS];
    float wbgt_w;             /* f32(0.7 * isobaric wet-bulb temperature, degC) */
    float wbgt_a;             /* black globe: h_cg / (emissivity * sigma), h_cg = 13.455 va^0.6 */
    int32_t pad0[1];
    sw_patch_k_t pk[SW_NPATCH];
    sw_patch_l_t pl[SW_NPATCH];
} sw_step_t;

/* Step-invariant model constants and walk tables. */
typedef struct sw_model {
    float absK, absL, Fside, Fup, Fcyl, albedo_b, ewall, sbc;
    float first, second;      /* walk snapshot / length (rounded, scaled) */
    int32_t nsec;             /* walk steps per azimuth (<= SW_WALK_MAX) */
    int32_t walk_halo;        /* max |walk offset| over all azimuths (<= 48) */
    float alb_class[SW_NCLASS], emis_class[SW_NCLASS];
    int32_t walk_dx[SW_WALK_AZ * SW_WALK_MAX], walk_dy[SW_WALK_AZ * SW_WALK_MAX];
    int32_t walk_branch[SW_WALK_AZ], walk_dir[SW_WALK_AZ];
    float walk_lo[SW_WALK_AZ], walk_hi[SW_WALK_AZ];
} sw_model_t;

/* Device buffers of one tile's day loop (host struct of device pointers;
 * see the entry points below for each buffer's layout). */
typedef struct sw_tile {
    int32_t rows, cols;
    const void *hpack;
    const float *walls, *aspect_rad, *buildings;
    const uint8_t *cls;
    const uint32_t *stats, *masks;
    const float *svfbuveg, *tan_asvf, *tan_svfalfa;
    const uint32_t *walkface;
    uint32_t *src;            /* 2 x uint32 per pixel, rewritten each day step */
    float *state;             /* 6 planes, carried across steps */
} sw_tile_t;

/* One item of a batched day loop: a tile (or an edit scenario of a tile),
 * its day plan and its output buffer. Items of one batch share the model
 * constants and the day/night pattern (same site, day, forcing and pixel
 * size); step records and sun tables are per item because the sun march
 * ends at the tile's edge. sun_block0 / flux_block0 are filled by
 * sw_batch_layout. */
typedef struct sw_batch_item {
    sw_tile_t tile;
    const sw_step_t *steps;   /* the item's step records */
    const int32_t *sun_tab;   /* the item's sun-march tables */
    float *out;               /* step i at out + (i - t0) * popcount(out_mask) * rows * cols */
    int32_t sun_block0;       /* first block of the item in the batched sun-march grid */
    int32_t flux_block0;      /* first block of the item in the batched flux grid */
} sw_batch_item_t;

/* ---- runtime ------------------------------------------------------------
 * Mirrors the reference native ABI's runtime surface (sw_api.cu:24-171,
 * sw_residency.cu:53-331): queries, counted device allocation, pinned host
 * memory, sync/async/2D copies, streams (default priority), events and
 * stream capture into executable graphs. */
int sw_abi_version(void);
const char *sw_last_error(void);
const char *sw_build_flags(void);
int sw_device_count(int *count);
int sw_device_name(int device, char *buf, int len);
int sw_mem_info(int64_t *free_bytes, int64_t *total_bytes);
int sw_sync(void *stream);
int sw_alloc(void **ptr, int64_t bytes);
int sw_free(void *ptr);
int sw_alloc_stats(int64_t *live_bytes, int64_t *n_alloc, int64_t *n_free);
int sw_host_alloc(void **ptr, int64_t bytes);
int sw_host_free(void *ptr);
int sw_copy(void *dst, const void *src, int64_t bytes);
int sw_copy_async(void *dst, const void *src, int64_t bytes, void *stream);
int sw_copy2d_async(void *dst, int64_t dpitch, const void *src, int64_t spitch, int64_t width_bytes,
                    int64_t height, void *stream);
/* Device -> pinned host copy on the DMA engines, using no compute units (hipMemcpyAsync with
 * hipMemcpyDeviceToDeviceNoCU: pinned host memory is device-addressable), so output downloads
 * overlap solves on other streams at full PCIe rate. dst must be pinned host memory
 * (sw_host_alloc). No counterpart in the reference (its residency ABI copies with
 * cudaMemcpyAsync, sw_residency.cu:53-331). */
int sw_copy_to_host(void *dst, const void *src, int64_t bytes, void *stream);
/* Upload from pinned host memory by a copy kernel (the compute units read the host buffer), so an
 * upload never queues behind downloads on the DMA engines. Both pointers 16-byte aligned; a source
 * that is not pinned host memory is refused (SW_E_ARG). No counterpart in the reference (its
 * uploads are torch/cudaMemcpy copies). */
int sw_copy_to_device_cu(void *dst, const void *src, int64_t bytes, void *stream);
int sw_fill_f32(float *dst, float value, int64_t n, void *stream);
int sw_stream_create(void **stream);
int sw_stream_destroy(void *stream);
int sw_event_create(void **event);
int sw_event_destroy(void *event);
int sw_event_record(void *event, void *stream);
int sw_event_sync(void *event);
int sw_event_elapsed(void *start, void *end, float *ms);
int sw_stream_wait_event(void *stream, void *event);
int sw_graph_begin(void *stream);
int sw_graph_end(void *stream, void **graph_exec);
int sw_graph_launch(void *graph_exec, void *stream);
int sw_graph_destroy(void *graph_exec);

/* ---- SVF sweep ----------------------------------------------------------
 * Replaces shadow.svf_calculator (solweig_gpu/shadow.py:409-625) and the
 * per-patch shadow() march (shadow.py:166-322).
 * hpack: float4 per pixel {dsm, vegdsm, vegdsm2, 0} (sw_pack_heights).
 * stats[1] (hmax, max over the three planes) enables the exact early exits.
 * masks: SW_MASK_PLANES*SW_MASK_WORDS*n uint32. */
int sw_pack_heights(const float *a, const float *vegdsm, const float *vegdsm2, void *hpack,
                    int rows, int cols, void *stream);
/* Scene compose on the device (utci_process.py:667-679 and the land-cover
 * cleanup :738-744): hpack {dsm, vegdsm, vegdsm2}, vegdsm2, buildings
 * (ground = 1), class codes, aspect in radians, and tile statistics
 * stats[4] = {key(amax), key(hmax), flags, 0} where key() is the
 * order-preserving uint32 image of a float. flags: 1 non-finite height,
 * 4 bush layer present, 8 non-integer land-cover code (inputs outside the
 * supported domain: the caller must refuse the solve). landcover may be
 * NULL (no land cover). */
int sw_compose(const float *dsm, const float *dem, const float *cdsm, const float *aspect_deg,
               const float *landcover, void *hpack, float *vegdsm2, float *buildings, uint8_t *cls,
               float *aspect_rad, uint32_t *stats, int rows, int cols, void *stream);
/* Patch tables: tab[offs[p] .. offs[p+1]) holds int4 {dx, dy, dz bits,
 * 12 * (dx * (cols + 2 * SW_SVF_PAD) + dy)} (byte offset in the padded
 * buffer) for every geometrically valid step of
 * patch p (consecutive steps move at most one cell); the library applies
 * the reference's amax stop test with amax from stats, writing the executed
 * step count of each patch to counts[153]. hpad is caller-owned workspace of
 * sw_svf_workspace_floats(rows, cols) floats: the padded heights,
 * (rows + 2 * SW_SVF_PAD) * (cols + 2 * SW_SVF_PAD) * 3 floats (below 2 GiB), followed
 * by the maxima of max(building, canopy) over SW_SVF_COARSE x SW_SVF_COARSE cells of
 * the padded grid, which let the march skip chunks of steps that cannot change a
 * mask (no counterpart in the reference; outputs unchanged); stats[3] is
 * rewritten (1 when some height is negative, which selects the
 * bounds-checked march). */
#define SW_SVF_COARSE 8
int64_t sw_svf_workspace_floats(int rows, int cols);
int sw_svf_march(const void *hpack, const int32_t *offs, const int32_t *tab, int32_t *counts, void *hpad,
                 uint32_t *masks, uint32_t *stats, int rows, int cols, void *stream);
/* March launch configuration (no counterpart in the reference; chosen per tile shape by the
 * autotuner, solweig_project_amd.solver.MarchTuner). None of it changes any output:
 *   chunk  steps per gather chunk of the fast path: 4, 8 (default) or 12;
 *   shape  a wave's pixels in the 16 x 16 tile: 0 an 8 x 8 square (default), 1 a 4 x 16 row
 *          strip, 2 a 16 x 4 column strip;
 *   split  blocks per 32-patch mask word: 1 (default), 2 or 4. */
typedef struct sw_march_config {
    int32_t chunk, shape, split, pad;
} sw_march_config_t;
int sw_svf_march_cfg(const void *hpack, const int32_t *offs, const int32_t *tab, int32_t *counts, void *hpad,
                     uint32_t *masks, uint32_t *stats, int rows, int cols, const sw_march_config_t *cfg,
                     void *stream);
/* Fold the masks into the 16 fields [svf, E, S, W, N, veg, vegE.., aveg, avegE.., total]
 * (shadow.py:545-613), in the reference's per-annulus summation order. */
int sw_svf_fold(const uint32_t *masks, const float *vegdsm2, const float *w_iso,
                const float *w_aniso, const int32_t *ring, const int32_t *nann,
                const int32_t *dirm, float *fields, int rows, int cols, void *stream);
/* Expand the masks to the reference's float cubes (rows, cols, 153):
 * shmat, vegshmat, vbshvegshmat (shadow.py:531-554). Any output may be NULL. */
int sw_unpack_masks(const uint32_t *masks, float *shmat, float *vegshmat, float *vbshmat,
                    int rows, int cols, void *stream);

/* Pack the reference's float cubes shmat, vegshmat, vbshvegshmat
 * (rows, cols, 153; the cubes of svf_calculator's return tuple,
 * shadow.py:625) into the mask planes. flags[0] gets bit 1 when a value is
 * outside {0, 1} ({0, 1, 2} for vbsh); the caller refuses such a bundle. */
int sw_pack_masks(const float *shmat, const float *vegshmat, const float *vbshmat, uint32_t *masks,
                  uint32_t *flags, int rows, int cols, void *stream);
/* Tile statistics for marches on caller-packed heights (the svf_calculator
 * entry, where the caller supplies amaxvalue): stats = {key(amax),
 * key(max of the three planes), flags (1: non-finite height), 0}. */
int sw_height_stats(const void *hpack, float amax, uint32_t *stats, int rows, int cols, void *stream);
/* Replace stats[0] by key(amax): a caller-given domain maximum
 * (run_utci_window's amaxvalue, utci_process.py:675-676). */
int sw_set_amax(uint32_t *stats, float amax, void *stream);

/* Single-patch marches with float outputs, the drop-in form of the
 * reference native ABI (sw_api.cu:209-228). */
int sw_run_march_svf_shadow(const float *a, const float *vegdem, const float *vegdem2,
                            const int32_t *dxs, const int32_t *dys, const float *dzs,
                            float *sh, float *vegsh, float *vbsh, int rows, int cols, int count,
                            void *stream);
int sw_run_march_wallheight23(const float *a, const float *vegdem, const float *vegdem2,
                              const int32_t *dxs, const int32_t *dys, const float *dzs,
                              const float *dzprevs, float *sh, float *vegsh, float *vbsh,
                              int rows, int cols, int count, void *stream);

/* ---- time loop ------------------------------------------------------------
 * Static per-pixel fields of the day loop (utci_process.py:867-897):
 * svfbuveg, tan(asvf), tan(svfalfa), and walkface: bit j set where the
 * face-shading term of ground-view azimuth j is 0 (sunonsurface_2018a,
 * solweig.py:236-252). */
int sw_tile_static(const float *fields, const float *walls, const float *aspect_rad,
                   const sw_model_t *model, float *svfbuveg, float *tan_asvf, float *tan_svfalfa,
                   uint32_t *walkface, int rows, int cols, void *stream);

/* Sun march + wall tail + ground-view-factor walk sources for one day step
 * (shadowingfunction_wallheight_23, solweig.py:1079-1224; sunonsurface
 * prologue, solweig.py:121-134). src: 2 x uint32 per pixel {Lup bits, flags}
 * with flag bits 4..10 = {sh, vegsh, sunwall, ground, land-cover class x3}.
 * sun_tab: int4 {dx, dy, dz bits, dzprev bits} per march step. */
int sw_sun_step(const void *hpack, const float *walls, const float *aspect_rad,
                const uint8_t *cls, const float *buildings, const sw_step_t *steps,
                const sw_model_t *model, const int32_t *sun_tab, const uint32_t *stats,
                uint32_t *src, int step, int rows, int cols, void *stream);
/* sw_sun_step that also writes the reference's eight output planes of the
 * march and wall tail, the return tuple of shadowingfunction_wallheight_23
 * (solweig.py:1187-1224): planes[q * n + pixel] for q = 0 vegsh, 1 sh,
 * 2 vbshvegsh, 3 wallsh, 4 wallsun, 5 wallshve, 6 facesh, 7 facesun
 * (float32, 8 * n floats). The march, its exits and the tail are the ones
 * the day loop runs (same kernel body); src is written as by sw_sun_step.
 * Replaces the native lane's sw_run_march_wallheight23 + wall tail
 * (native/cuda/src/sw_api.cu:219-228). */
int sw_sun_step_planes(const void *hpack, const float *walls, const float *aspect_rad,
                       const uint8_t *cls, const float *buildings, const sw_step_t *steps,
                       const sw_model_t *model, const int32_t *sun_tab, const uint32_t *stats,
                       uint32_t *src, float *planes, int step, int rows, int cols, void *stream);

/* Flux assembly -> Tmrt -> UTCI for one step (Solweig_2022a_calc,
 * solweig.py:2101-2311; utci_calculator, calculate_utci.py:259-330).
 * state: 6 planes (Tgmap1, E, S, W, N, TgOut1), read and updated.
 * out: planes selected by out_mask bits: 1 tmrt, 2 utci, 4 shadow,
 * 8 kdown, 16 kup, 32 ldown, 64 lup, 128 wbgt (utci_process.py:1021-1037:
 * black globe from Tmrt, the step's air temperature and wbgt_a, blended
 * with wbgt_w by shadow < 0.1), written at out + k * n for the k-th
 * selected plane in that bit order. walk_halo sizes the shared-memory tile
 * of walk sources and must be at least model->walk_halo for a day step;
 * -1 launches without the tile (night steps: less shared memory, more
 * waves). A day step launched with a smaller halo writes NaN outputs. */
int sw_flux_step(const uint32_t *masks, const uint32_t *src, const float *svfbuveg,
                 const float *tan_asvf, const float *tan_svfalfa, const uint32_t *walkface,
                 const float *buildings, const uint8_t *cls, const sw_step_t *steps,
                 const sw_model_t *model, float *state, float *out, int out_mask, int step,
                 int walk_halo, int rows, int cols, void *stream);

/* ---- whole stages ---------------------------------------------------------
 * SVF sweep = sw_svf_march + sw_svf_fold (svf_calculator, shadow.py:409-625). */
int sw_svf_sweep(const void *hpack, uint32_t *stats, const int32_t *offs, const int32_t *tab,
                 const float *vegdsm2, const float *w_iso, const float *w_aniso, const int32_t *ring,
                 const int32_t *nann, const int32_t *dirm, int32_t *counts, void *hpad, uint32_t *masks,
                 float *fields, int rows, int cols, void *stream);
/* sw_svf_sweep with a march configuration (NULL: the defaults) */
int sw_svf_sweep_cfg(const void *hpack, uint32_t *stats, const int32_t *offs, const int32_t *tab,
                     const float *vegdsm2, const float *w_iso, const float *w_aniso, const int32_t *ring,
                     const int32_t *nann, const int32_t *dirm, int32_t *counts, void *hpad, uint32_t *masks,
                     float *fields, int rows, int cols, const sw_march_config_t *cfg, void *stream);
/* Scratch-free SVF sweep. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops and harnesses that forbid hipMalloc). sw_svf_sweep_cfg without the
 * counts and hpad workspace arguments: both live in a static device array of
 * SW_SVF_SCRATCH_BYTES that each GPU the library runs on holds until the
 * process ends. Limit: sw_svf_workspace_floats(rows, cols) * 4 + 1024
 * <= SW_SVF_SCRATCH_BYTES (tiles up to about 2300 x 2300); larger tiles
 * are refused with SW_E_ARG before anything is enqueued. Calls on one GPU
 * share its array: calls that overlap on one GPU (several streams or host
 * threads) race; serialise them on one stream. Allocates nothing (no
 * hipMalloc, hipFree or allocating library routine). Outputs equal
 * sw_svf_sweep_cfg's. */
#define SW_SVF_SCRATCH_BYTES (64LL << 20)
int sw_svf_sweep_noalloc(const void *hpack, uint32_t *stats, const int32_t *offs, const int32_t *tab,
                         const float *vegdsm2, const float *w_iso, const float *w_aniso, const int32_t *ring,
                         const int32_t *nann, const int32_t *dirm, uint32_t *masks, float *fields, int rows,
                         int cols, const sw_march_config_t *cfg, void *stream);
/* Day loop over steps [t0, t1) (the time loop of run_utci_window,
 * utci_process.py:917-1049): per step the sun march when is_day[i - t0]
 * (host array) and the flux step; outputs of step i at
 * out + (i - t0) * popcount(out_mask) * n. reset_state clears the 6 state
 * planes first (fill kernel, capture-safe). No host synchronisation. */
int sw_solve_day(const sw_tile_t *tile, const sw_step_t *steps, const sw_model_t *model,
                 const int32_t *sun_tab, const int32_t *is_day, int t0, int t1, float *out, int out_mask,
                 int walk_halo, int reset_state, void *stream);
/* sw_solve_day recording events (created with sw_event_create) around the
 * launches of step i: events[3 * (i - t0) + 0] before the sun march (or the
 * flux step of a night), + 1 between sun march and flux step, + 2 after the
 * flux step; for per-kernel timing without host work between launches. */
int sw_solve_day_timed(const sw_tile_t *tile, const sw_step_t *steps, const sw_model_t *model,
                       const int32_t *sun_tab, const int32_t *is_day, void **events, int t0, int t1, float *out,
                       int out_mask, int walk_halo, int reset_state, void *stream);

/* Sparse recomputation over listed 16 x 16 tiles (tile index = tile_row *
 * ceil(cols / 16) + tile_col; device arrays). Replaces the reference's
 * sparse recompute of edited regions (solweig_gpu/incremental/sparse_work.py,
 * solver.py:595-957). sw_svf_sweep_tiles: step counts and padded heights of
 * the whole tile, then the march and the fold only on the listed tiles (the
 * others keep their masks and fields). sw_solve_day_tiles: sw_solve_day with
 * the sun march on sun_tiles and the flux step on flux_tiles only; the caller
 * lists every tile whose outputs can change in flux_tiles and every tile
 * within the walk reach of those in sun_tiles. Outputs of the other tiles
 * are left untouched. */
int sw_svf_sweep_tiles(const void *hpack, uint32_t *stats, const int32_t *offs, const int32_t *tab,
                       const float *vegdsm2, const float *w_iso, const float *w_aniso, const int32_t *ring,
                       const int32_t *nann, const int32_t *dirm, int32_t *counts, void *hpad, uint32_t *masks,
                       float *fields, const int32_t *tiles, int ntiles, int rows, int cols, void *stream);
int sw_solve_day_tiles(const sw_tile_t *tile, const sw_step_t *steps, const sw_model_t *model,
                       const int32_t *sun_tab, const int32_t *is_day, int t0, int t1, float *out, int out_mask,
                       int walk_halo, int reset_state, const int32_t *sun_tiles, int nsun,
                       const int32_t *flux_tiles, int nflux, void *stream);

/* Batched day loop (the reference solves tiles one after another,
 * utci_process.py:1199-1663; edit scenarios one request at a time,
 * solweig_gpu/incremental/worker.py:1588-1595). sw_batch_layout assigns each
 * host item its block ranges and returns the grid totals. sw_solve_batch
 * then runs steps [t0, t1) of every item with one sun-march launch per day
 * step and one flux launch per step over all items: items_dev is a device
 * copy of items_host (the same nitems items, after sw_batch_layout). Every
 * item's outputs equal those of sw_solve_day on that item alone. */
int sw_batch_layout(sw_batch_item_t *items, int nitems, int32_t *sun_blocks, int32_t *flux_blocks);
int sw_solve_batch(const sw_batch_item_t *items_dev, const sw_batch_item_t *items_host, int nitems,
                   const sw_model_t *model, const int32_t *is_day, int t0, int t1, int out_mask, int walk_halo,
                   int reset_state, void *stream);

/* The same batched day loop in one persistent launch (per 64 steps): every
 * block stays resident and runs the sun-march and flux phases of every step,
 * drawing the batch's tiles from per-XCD counters, with grid-wide barriers
 * between the phases (the reference has no counterpart; it replaces the
 * per-step launches of sw_solve_batch, whose outputs it equals bit for bit).
 * The grid is as many blocks as the device holds at once (cooperative launch;
 * a plain launch of the same grid under stream capture), at most one per
 * tile. `sync` is a device workspace of SW_DAY_SYNC_WORDS uint32 (barrier and
 * tile counters, cleared on the stream before each launch); after the stream
 * completes, sync[SW_DAY_SYNC_ERR] != 0 means a barrier wait exceeded
 * timeout_s and the launch stopped early (its outputs are incomplete). Two
 * such launches must not run at once on one GPU. */
#define SW_DAY_SYNC_WORDS 256
#define SW_DAY_SYNC_ERR 9
int sw_solve_batch_persistent(const sw_batch_item_t *items_dev, const sw_batch_item_t *items_host, int nitems,
                              const sw_model_t *model, const int32_t *is_day, int t0, int t1, int out_mask,
                              int walk_halo, int reset_state, uint32_t *sync, double timeout_s, void *stream);

/* UTCI of elementwise inputs (utci_calculator); -999 inputs give -999. */
int sw_utci(const float *ta, const float *rh, const float *tmrt, const float *va, float *out,
            int64_t n, void *stream);

/* ---- walls and wall aspect (tile inputs) -----------------------------------
 * Wall heights: replaces findwalls(dem_array, walllimit)
 * (solweig_gpu/walls_aspect.py:27-58). Per cell: maximum of the 4
 * neighbours (NaN if any is NaN) minus the cell, formed in double; values
 * below walllimit become 0; the one-cell border is 0. Written as float32,
 * the dtype of the wall raster the reference writes (walls_aspect.py:214-228). */
int sw_find_walls(const float *dsm, float *walls, int rows, int cols, double walllimit, void *stream);

/* Rotated line/building filters of the wall-aspect method, one set per
 * rotation h = 0..179 (walls_aspect.py:121-142, built with
 * scipy.ndimage.rotate by the host). For each h: count[h][0] offsets of the
 * rotated line filter (cells equal to 1), count[h][1] / count[h][2] offsets of
 * the building filter cells equal to 1 / 2, in row-major window order;
 * off[h][l][k] = (dy, dx) from the window centre. size is the filter size
 * the table was built for; counts are at most SW_WALL_LIST_CAP. */
#define SW_WALL_ROTATIONS 180
#define SW_WALL_LIST_CAP 128
typedef struct {
    int32_t size;
    int32_t count[SW_WALL_ROTATIONS][3];
    int8_t off[SW_WALL_ROTATIONS][3][SW_WALL_LIST_CAP][2];
} sw_wall_filters_t;

/* Filter size for a scale in pixels per metre (walls_aspect.py:114-118):
 * floor((scale + 1e-10) * 9), 3 if that is <= 2, made odd by +1 unless 9. */
int sw_wall_filter_size(double scale);

/* Wall aspect in degrees: replaces filter1Goodwin_as_aspect_v3(walls, scale, a)
 * (walls_aspect.py:99-166). filters is a device copy of the table for
 * sw_wall_filter_size(scale); a table of another size gives NaN everywhere.
 * Wall cells (walls > 0) of the window interior take the rotation with the
 * largest line count (first maximum kept), 270 - h, minus 180 when the
 * building half 1 has the larger float32 DSM sum (np.sum order); cells
 * still at 0 take the DSM gradient aspect (get_ders, :78-97). rows and
 * cols must be at least 2 (np.gradient's minimum). */
int sw_wall_aspect(const float *walls, const float *dsm, const sw_wall_filters_t *filters, float *aspect,
                   int rows, int cols, double scale, void *stream);

/* ---- SVF surrogate network (no counterpart in the reference: a learned
 * replacement of the SVF sweep, gated by and falling back to the exact
 * kernel) ---------------------------------------------------------------------
 * Activations are NHWC bf16 (uint16_t storage) with channel counts that are
 * multiples of 16; weights are bf16 [cout][ks * ks][cin]; biases, gradients
 * and optimiser state are f32. A network pass is a list of ops run in order
 * on one stream (capturable); sw_cnn_fwd and sw_cnn_bwd run forward and
 * backward lists. */
#define SW_CNN_CONV 1             /* out = act(conv(in0 [up] ++ in1) + bias), split at o1 */
#define SW_CNN_CONV_BWD_DATA 2    /* d(input) = conv(dz = in0, flipped weight); ws holds the flip */
#define SW_CNN_CONV_BWD_WEIGHT 3  /* gw, gb from dz and the forward inputs; ws: chunk partials */
#define SW_CNN_POOL2 4            /* 2x2 average pool, in0 (n, h, w, c1) -> out0 (n, h/2, w/2, c1) */
#define SW_CNN_POOL2_BWD 5        /* out0 = in1 (optional addend) + 0.25 * in0 broadcast */
#define SW_CNN_UP2_BWD 6          /* out0 (half) = 2x2 sums of in0 (n, h, w, c1) */
#define SW_CNN_RELU_BWD 7         /* out0 = in0 where in1 > 0 (c1 channels) */
#define SW_CNN_HEAD_LOSS 8        /* sigmoid-MSE over the core window; ws: per-block losses */
typedef struct sw_cnn_op {
    int32_t kind;
    int32_t n, h, w;              /* batch and full-resolution size of the op */
    int32_t c1, c2, up, o1;       /* input channels (in0 half resolution when up), output split */
    int32_t cout, ks, act, nchunk;/* act: 0 none, 1 ReLU, 2 sigmoid; nchunk: weight-gradient split */
    int32_t core, c0, nl, pad;    /* head loss: core window [c0, c0 + core), nl labelled channels;
                                   * forward conv: core > 0 computes only the square [c0, c0 + core)^2
                                   * of each image (inference; the rest of the output is untouched) */
    const void *in0, *in1;
    void *out0, *out1;
    float *outf;
    const void *weight;
    const float *bias;
    const void *dz;
    float *gw, *gb;
    void *ws;
    /* forward conv epilogues (3x3, 16 output channels only): pool_out receives the 2x2 average
     * pool of the stored output (bit-identical to an SW_CNN_POOL2 op after it); head_w (bf16
     * [16][16]) and head_b (f32 [16]) run a 1x1 sigmoid head on the stored output into outf */
    void *pool_out;
    const void *head_w;
    const float *head_b;
    /* forward 3x3 conv with c1 = 16: when site is set, in0 is unused and the input is the window at
     * site_origins[img] (int32 row, col) of the site's composed heights (float4 {a, vegdsm, vegdsm2,
     * -} per cell, site_rows x site_cols): channels 0-2 = height * 0.02 in bf16, the rest and cells
     * outside the site zero (the values sw_cnn_pack_input writes) */
    const void *site;
    const int32_t *site_origins;
    int32_t site_rows, site_cols;
} sw_cnn_op_t;

int sw_cnn_fwd(const sw_cnn_op_t *ops, int nops, void *stream);
int sw_cnn_bwd(const sw_cnn_op_t *ops, int nops, void *stream);
/* workspace bytes an op needs (data gradient: the flipped kernel; weight gradient: chunk partials) */
int sw_cnn_bwd_scratch_bytes(const sw_cnn_op_t *op, int64_t *bytes);
/* Adam step on f32 master weights (bias-corrected with the step count); p_bf16 receives the bf16 copy */
int sw_adam(float *p, float *m, float *v, const float *g, uint16_t *p_bf16, int64_t n, float lr, float b1, float b2,
            float eps, int step, void *stream);
/* network input from composed heights (hpack of sw_compose) at window origins (int32 pairs row, col) */
int sw_cnn_pack_input(const void *hpack, const int32_t *origins, uint16_t *x, int n, int hw, int rows, int cols,
                      void *stream);
/* labels [n][core][core][nl] from SVF field planes at the window cores (origin + c0) */
int sw_cnn_gather_labels(const float *fields, const int32_t *origins, const int32_t *field_map, float *labels, int n,
                         int core, int c0, int nl, int rows, int cols, void *stream);
/* surrogate outputs (f32 [n][hw][hw][cs]) of the window cores into SVF field planes */
int sw_cnn_scatter_cores(const float *out, const int32_t *origins, const int32_t *field_map, float *fields, int n,
                         int hw, int cs, int core, int c0, int nl, int rows, int cols, void *stream);
/* Surrogate gate (no counterpart in the reference). sw_gate_error: per window w, the mean over its
 * sampled pixels pix[seg[w] .. seg[w + 1]) (flat indices inside window w's core) of the mean
 * absolute difference between the network output out [w][hw][hw][cs] and the exact field planes
 * fields[fm[k]] over k < nl; fixed-order reduction (repeatable). sw_gate_write: the output cores
 * [c0, c0 + core) of every window into fields[fm[k]] and SVFtotal (plane 15 = svf - (1 - svfveg) *
 * 0.97, from output channels 0 and 5) except at pixels with sampled[p] != 0. */
int sw_gate_error(const float *out, const float *fields, const int32_t *org, const int32_t *pix, const int32_t *seg,
                  const int32_t *fm, float *werr, int nwin, int hw, int cs, int nl, int rows, int cols, void *stream);
int sw_gate_write(const float *out, const int32_t *org, const int32_t *fm, const uint8_t *sampled, float *fields,
                  int nwin, int hw, int cs, int core, int c0, int nl, int rows, int cols, void *stream);
/* sw_gate_max_error: as sw_gate_error, but werr[w] is the largest absolute difference of any of the nl
 * fields at any sampled pixel of window w (no counterpart in the reference). */
int sw_gate_max_error(const float *out, const float *fields, const int32_t *org, const int32_t *pix,
                      const int32_t *seg, const int32_t *fm, float *werr, int nwin, int hw, int cs, int nl, int rows,
                      int cols, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* SOLWEIG_HIP_H */
