/* updc_hip.h -- C ABI of the MI355X (gfx950) UPDC v17 reconstruction library.
 *
 * Plain C types only (pointers, sizes, ints, doubles): the library can be
 * bound from ctypes, cgo, JNI or a C program.  Every entry point returns a
 * status code (UPDC_OK on success) instead of calling exit(); the message of
 * the last failure on the calling thread is available from updc_last_error().
 * Host pointers are read during the call only (the library copies what it
 * keeps).  One context is driven by one host thread at a time.
 *
 * Interfaces replaced (reference files cited as path:line):
 *   updc_all            <- UPDC_ALL(...)           updc/c_core/updc_core.c:1289-1293
 *                          and the updc_jwst_c argv  updc/c_core/updc_core.c:1566-1573
 *   updc_self_test      <- self_test()             updc/c_core/updc_core.c:1545-1564
 *   updc_polygon_area   <- updc_polygon_area()     updc/c_core/overlap_geometry.h:15
 *   updc_overlap_area   <- updc_overlap_area()     updc/c_core/overlap_geometry.h:16
 *   updc_ctx_* / updc_add_exposure* / updc_prepare / updc_iterate
 *                       <- the in-process stages of UPDC_ALL: manifest rows
 *                          (:426-461), LinkMap build/open (:530-663), weights
 *                          (:378-424), PSF spectra (:777-876), initial drizzle
 *                          (:878-983), denominator (:985-1086) and the
 *                          iteration loop (:1410-1479).
 */
#ifndef UPDC_HIP_H
#define UPDC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define UPDC_OK 0
#define UPDC_ERR_ARG 1      /* invalid argument or dimensions                  */
#define UPDC_ERR_HIP 2      /* HIP runtime / device failure                    */
#define UPDC_ERR_IO 3       /* file missing, short, or with trailing bytes     */
#define UPDC_ERR_NUMERIC 4  /* a numeric guard of the reference fired          */
#define UPDC_ERR_STATE 5    /* call out of order (e.g. iterate before prepare) */
#define UPDC_ERR_NOMEM 6    /* device or host allocation failed                */
#define UPDC_ERR_FORMAT 7   /* invalid / obsolete LinkMap or cache file        */

/* ---- precision tiers ----------------------------------------------------- */
#define UPDC_PREC_F32 0   /* the reference's float32 arithmetic (docs/PRECISION.md) */
#define UPDC_PREC_F64 1   /* plain FP64 form                                        */
#define UPDC_PREC_MIXED 2 /* FP32 FFTs, PSF and S_k products; FP64 spectrum sum over
                             exposures, denominator, raw_mu test and ratio         */

/* ---- context flags ------------------------------------------------------- */
#define UPDC_FLAG_GRAPH 1u      /* capture each iteration as a hipGraph and replay */
#define UPDC_FLAG_QUIET 2u      /* no informational stdout                          */
#define UPDC_FLAG_PER_EXPOSURE 4u /* one launch per exposure and stage (no batching) */
#define UPDC_FLAG_AUTOTUNE 8u   /* time the candidate FFT shapes on the GPU at prepare
                                   (when the context was created with a 0 x 0 shape) */
#define UPDC_FLAG_FUSED_ADJOINT 16u /* batched form: the adjoint gather fused into the row
                                       R2C (the S_k^T wr_k images are never stored), instead
                                       of the default: the gather as its own launch for all
                                       exposures, then a plain row R2C of the images */
#define UPDC_FLAG_COLUMN_SPECTRA 32u /* batched form: the spectra a column pass writes for a row
                                        C2R (blurred spectra, correction) column-major, instead
                                        of the default where the passes allow it: column pairs
                                        interleaved element by element (64-byte row-pass reads);
                                        results are identical bit for bit */

/* ---- LinkMap ABI (updc/cache_abi.py:9-12, updc_core.c:44-54) -------------- */
#define UPDC_LINKMAP_MAGIC 0x324D4C55u
#define UPDC_LINKMAP_VERSION 2u
#define UPDC_LINKMAP_ABS32 1u
#define UPDC_LINKMAP_DELTA16 2u
#define UPDC_LINKMAP_HEADER_BYTES 64u

/* ---- drop-in solver -------------------------------------------------------- */
/* Same 18 parameters, file contract and caches as UPDC_ALL
 * (updc_core.c:1289-1543); returns a status instead of exit(EXIT_FAILURE).
 * The worker-count parameters are accepted and validated as the reference
 * does (>= 1); the GPU path has no thread pools.  plan_measure, autotune,
 * retune_fft and wisdom_path steer the GPU FFT shape autotuner and its cache. */
int updc_all(const char *manifest, int width, int height, int input_width, int input_height, int exposures,
             int transpose_psf, int iterations, const char *iter_prefix, const char *weight_output,
             int fft_worker_count, int exposure_worker_count, int preprocess_worker_count, int plan_measure,
             int autotune, int retune_fft, const char *wisdom_path, int save_every);

/* Message of the last failure on this thread ("" if none). */
const char *updc_last_error(void);

/* updc_core.c:1545-1564: unit-square area, half-square overlap, LinkMap header
 * size.  Returns 0 when every check passes (and prints the reference's OK line
 * when verbose != 0). */
int updc_self_test(int verbose);

/* Host-side geometry with the reference's precision policy (f64 clipping,
 * f32 local-origin shoelace with fmaf): overlap.c:72-116.  xy holds count
 * (x, y) pairs; quad_xy holds 4 pairs. */
float updc_polygon_area(const double *xy, int count);
float updc_overlap_area(const double *quad_xy, int x, int y);

/* ---- in-memory context API --------------------------------------------- */
typedef struct updc_ctx updc_ctx;

/* Per-(iteration, exposure) FreezeMu record: one mufreeze CSV row
 * (updc_core.c:1425-1431). */
typedef struct {
    int32_t iteration;
    int32_t exposure; /* 1-based, as in the CSV */
    uint64_t n_new_frozen;
    uint64_t n_total_frozen;
    double min_raw_mu;
} updc_freeze_row;

/* Per-iteration summary (the reference's per-iteration stdout line,
 * updc_core.c:1473-1478), with GPU stage times in milliseconds. */
typedef struct {
    int32_t iteration;
    int32_t pad;
    double signed_sum;
    uint64_t n_negative;
    uint64_t n_new_frozen;
    uint64_t n_total_frozen;
    double min_raw_mu;
    double ms_total;
} updc_iter_info;

/* Set-up stage times of a context in seconds (the setup rows of the
 * reference's timing CSV, updc_core.c:1487-1495): the FFT shape autotune and
 * the transform tables (fft_autotune, fft_plan), the PSF spectra including
 * their cache (psf_cache), LinkMap build or validation and upload over all
 * updc_add_exposure* calls (polygon_links), weight scale and normalisation
 * (weight_scale), the transposed LinkMaps (no reference counterpart: the
 * reference scatters), the initial drizzle and the denominator
 * (initial_drizzle, sensitivity). */
typedef struct {
    double fft_autotune, fft_plan, psf_spectra, linkmaps, weight_scale, csc_transpose, initial_drizzle, sensitivity;
} updc_setup_times;

/* One iteration's timed stages in seconds, named after the reference's
 * per-iteration timing rows (updc_core.c:1496-1505).  eager = 1: measured
 * with HIP events around the launches; the X^ column transform is fused into
 * the forward stage, the adjoint-image gather counts as detector time, and
 * with the update applied inside the correction's inverse transform that
 * time is in adjoint_reduce_ifft and update is 0.  The per-exposure form
 * times only the fused exposure wall.  eager = 0 (hipGraph replay): only total
 * is measured (the call's average); the stage fields are NaN. */
typedef struct {
    int32_t iteration;
    int32_t eager;
    double estimate_fft, fused_exposure_wall, forward_ifft_critical, fused_detector_critical, scatter_fft_critical,
        adjoint_reduce_ifft, update, total;
} updc_iter_times;

/* fft_width/fft_height of 0 pick the shape at prepare: the smallest even
 * 7-smooth width and 7-smooth height >= target + max PSF - 1
 * (updc_core.c:117-126), or with UPDC_FLAG_AUTOTUNE the fastest candidate
 * timed on the GPU (updc_core.c:242-291).  precision is one of UPDC_PREC_*;
 * device is the HIP device ordinal. */
int updc_ctx_create(updc_ctx **out, int width, int height, int fft_width, int fft_height, int precision,
                    unsigned flags, int device);
int updc_ctx_destroy(updc_ctx *ctx);
/* Replace the UPDC_FLAG_* set (e.g. switch hipGraph replay on or off between
 * updc_iterate calls; per-kernel times are recorded for eager launches only). */
int updc_ctx_set_flags(updc_ctx *ctx, unsigned flags);

/* Drizzle worker topology of the initial estimate (updc_core.c:917-973):
 * 1 = the serial summation order, n > 1 = per-worker partial sums combined in
 * worker order.  Default 1. */
int updc_ctx_set_preprocess_workers(updc_ctx *ctx, int workers);

/* FFT shape autotuner (replaces select_fft_shape and its cache,
 * updc_core.c:170-291).  Candidates: the first five admissible sizes per axis
 * from the minimum (even 7-smooth widths, 7-smooth heights), the diagonal
 * pairs, then (w2,h0), (w0,h2), (w4,h0), (w0,h4); up to 9 shapes.  Each is
 * timed, with the edge-stage and the LDS-only column kernel, on the
 * transforms of one batched iteration: X^, the broadcast column pass with the
 * PSF products over all exposures, their row inverses, the exposures' row
 * R2Cs, the summed adjoint column pass and the final row C2R (the reference
 * times one R2C + C2R round trip instead). */
int updc_fft_shape_candidates(int min_width, int min_height, int *widths, int *heights, int capacity,
                              int *count);
/* Cache directory of the autotune result (the reference keeps
 * fft_autotune_v1_*.txt in the cache root, updc_core.c:181-223); "" or NULL
 * disables it.  retune != 0 ignores a cached result (updc_core.c:198). */
int updc_ctx_set_fft_cache(updc_ctx *ctx, const char *dir, int retune);
/* Persistent set-up caches of the reference (updc_core.c:145-158, 790-843,
 * 905-976, 1024-1083) in directory `dir` ("" or NULL: none), for the F32
 * tier (the reference's arithmetic; the other tiers never read or write
 * them): psf_spectrum_%02d_%dx%d_t%d.c64 (FH x (FW/2+1) complex64, row-major,
 * FFTW's R2C layout), initial_estimate.f32 and initial_coverage.f32 (H x W),
 * denominator_%dx%d.f32 (H x W).  A file is used when its size is exact,
 * else recomputed and written (temporary file + rename).  Set before
 * updc_prepare. */
int updc_ctx_set_setup_cache(updc_ctx *ctx, const char *dir);
/* Caches the last updc_prepare read: bit 0 PSF spectra, bit 1 initial
 * estimate and coverage, bit 2 denominator. */
int updc_get_setup_cache_hits(updc_ctx *ctx, int *hits);
/* Layout of the spectra the batched iteration's column passes write, under
 * the context's current flags (prepared context): 0 column-major, 1 column
 * pairs interleaved (see UPDC_FLAG_COLUMN_SPECTRA).  No reference counterpart: the reference's
 * spectra are FFTW's row-major half spectra, internal to each transform. */
int updc_get_spectrum_layout(updc_ctx *ctx, int *layout);

/* Column kernel of the 2-D FFT: -1 automatic (edge-stage where the column
 * length allows), 0 LDS-only, 1 edge-stage where the length allows.  Set
 * before updc_prepare; no reference counterpart (a tile choice). */
int updc_ctx_set_column_kernel(updc_ctx *ctx, int kind);
/* The last autotune: count rows of (width, height, column_edge, ms); ms < 0
 * marks an unsupported shape.  cached = 1 when the result came from the cache
 * file (one row). */
int updc_get_autotune(updc_ctx *ctx, int *widths, int *heights, int *column_edge, double *ms, int capacity,
                      int *count, int *cached);

/* One exposure (one manifest row, updc_core.c:426-461).  sci/wht: ih*iw raw
 * float32 planes (not yet normalised).  psf: ph*pw float32 raster (stored
 * transposed when psf_transposed, updc_core.c:814-820).  The LinkMap is
 * either a complete ULM2 image (linkmap/linkmap_bytes, validated as
 * updc_core.c:485-496 does) or built on the GPU from a vertex lattice of
 * (ih+1)*(iw+1) (x,y) float64 pairs in target pixels (updc_core.c:530-609).
 * Exposures may differ in iw/ih and PSF size. */
int updc_add_exposure(updc_ctx *ctx, const float *sci, const float *wht, int input_width, int input_height,
                      const void *linkmap, size_t linkmap_bytes, const float *psf, int psf_width, int psf_height,
                      int psf_transposed, double background, double pedestal);
int updc_add_exposure_vertices(updc_ctx *ctx, const float *sci, const float *wht, int input_width,
                               int input_height, const double *vertices, const float *psf, int psf_width,
                               int psf_height, int psf_transposed, double background, double pedestal);

/* Bytes of exposure k's LinkMap (valid after add) and a copy of its ULM2
 * image (for the persistent .links cache, updc_core.c:574-601). */
int updc_linkmap_size(updc_ctx *ctx, int k, size_t *bytes);
int updc_get_linkmap(updc_ctx *ctx, int k, void *out, size_t capacity);

/* Setup: weight scale and normalisation, CSC transpose, PSF spectra, initial
 * drizzle, denominator and its maximum (updc_core.c:1367-1398). */
int updc_prepare(updc_ctx *ctx);
/* Run n iterations.  rows (optional) receives n*N freeze records, info
 * (optional) n per-iteration summaries.  When a numeric guard of the
 * reference fires (UPDC_ERR_NUMERIC: non-finite forward mean or ratio,
 * updc_core.c:1160-1169; non-finite or substantially negative correction,
 * non-finite estimate, :1441-1453), the iterations after the failing one do
 * not run, the message names the first failure the reference's serial loops
 * meet (kind, iteration, exposure and pixel or target), and rows/info hold
 * only what the reference would have written before stopping: info for the
 * completed iterations, freeze rows for those plus, for a failure in the
 * update, the failing iteration's rows (the reference writes them before its
 * update).  Entries after that are zero; updc_get_iterate_progress gives the
 * counts.  After a detector-stage failure the estimate is the last completed
 * iteration's; after an update failure it holds the failing update. */
int updc_iterate(updc_ctx *ctx, int n, updc_freeze_row *rows, updc_iter_info *info);
/* Of the last updc_iterate call: completed = iterations whose update ran,
 * freeze_rows = iterations whose freeze rows are valid (the UPDC_ALL loop's
 * per-iteration CSV and stdout lines, updc_core.c:1410-1479).  Both equal n
 * after a successful call.  Does not reset updc_last_error(). */
int updc_get_iterate_progress(updc_ctx *ctx, int *completed, int *freeze_rows);
/* Set-up times of the context (valid after updc_prepare). */
int updc_get_setup_times(updc_ctx *ctx, updc_setup_times *out);
/* Per-iteration stage times of the last updc_iterate call: *count entries
 * (its n), the first min(count, capacity) copied to out. */
int updc_get_iteration_times(updc_ctx *ctx, updc_iter_times *out, int capacity, int *count);

/* Results and setup products (host copies; f32 outputs are what the reference
 * writes, f64 variants return the working precision). */
int updc_get_estimate(updc_ctx *ctx, float *out);
int updc_get_estimate_f64(updc_ctx *ctx, double *out);
int updc_set_estimate_f64(updc_ctx *ctx, const double *in);
int updc_get_wht(updc_ctx *ctx, float *out);                /* coverage * weight_scale */
int updc_get_denominator_f64(updc_ctx *ctx, double *out);   /* H x W                    */
int updc_get_initial_f64(updc_ctx *ctx, double *estimate, double *coverage);
int updc_get_psf_spectrum_f64(updc_ctx *ctx, int k, double *out); /* FH x (FW/2+1) complex, interleaved */
int updc_get_scalars(updc_ctx *ctx, double *weight_scale, double *max_sensitivity, int *fft_width,
                     int *fft_height, uint64_t *masked_pixels);
int updc_get_frozen(updc_ctx *ctx, int k, uint8_t *out);
/* Sizes of exposure k: detector pixels, LinkMap links and encoding, and the
 * links of its CSC transpose (pixels with w > 0 only; valid after prepare). */
int updc_get_exposure_info(updc_ctx *ctx, int k, uint64_t *nin, uint64_t *nlinks, uint32_t *encoding,
                           uint64_t *ncsc);

/* The transposed LinkMap of exposure k that the iteration's adjoint gather
 * reads: compact = 1 for the 6-byte-per-link form (every 64-target tile has
 * at most 255 links per target and spans at most 65536 source pixels), 0 for
 * the 8-byte form; bytes = its size.  No reference counterpart (the
 * reference scatters, updc_core.c:1171-1176). */
int updc_get_csc_layout(updc_ctx *ctx, int k, int *compact, uint64_t *bytes);

/* Stage entry points (parity tests, autograd).  All host buffers, f64 at the
 * boundary, computed in the context's precision on the GPU. */
/* blur_k = crop(C2R(X^ . P^_k)) / nf for every exposure: out is N*H*W. */
int updc_stage_forward(updc_ctx *ctx, const double *x, double *blur_out);
/* Noise-free model H_k x = S_k P_k x for every pixel of every exposure
 * (flattened, sum of nin), the forward gather of updc_core.c:1154-1158. */
int updc_stage_model(updc_ctx *ctx, const double *x, double *pred_out);
/* Detector stage of exposure k given blur_k (H*W): FreezeMu update of the
 * context's masks, wr (nin) and the transposed-gather scatter (H*W).
 * stats = {n_new, n_total, min_raw_mu}. */
int updc_stage_detector(updc_ctx *ctx, int k, const double *blur, double *wr_out, double *scatter_out,
                        double *stats);
/* Adjoint half: C = crop(C2R(sum_k conj(P^_k) . R2C(pad(s_k)))) / nf where
 * s_k = S_k^T v_k, v is the flattened per-pixel vector (sum of nin). */
int updc_stage_adjoint(updc_ctx *ctx, const double *v, double *out);
/* S_k^T v for one exposure (H*W), via the transposed gather. */
int updc_stage_scatter(updc_ctx *ctx, int k, const double *v, double *out);
/* S_k^T v_k for every exposure at once (out: N*H*W) through the iteration's
 * own adjoint-image launch: the scatter half of the detector loop,
 * updc_core.c:1171-1176 (v is the flattened per-pixel vector, sum of nin;
 * pixels with w <= 0 contribute nothing). */
int updc_stage_adjoint_images(updc_ctx *ctx, const double *v, double *out);
/* Zero every FreezeMu mask. */
int updc_reset_frozen(updc_ctx *ctx);

/* ---- PSF projector ----------------------------------------------------------
 * Polygon push-forward of a PSF onto the target grid (psf_projector.c:82-172):
 * input pixel (x, y) of the width x height PSF maps to the quad of vertices
 * (y, x), (y+1, x), (y+1, x+1), (y, x+1) of the (height+1) x (width+1) x 2
 * lattice (target-pixel offsets from the PSF centre); every overlap area
 * > 1e-8 adds value * area / full_area to the output pixel, in ascending input
 * order (fmaf).  Output: out_width x out_height = (2 hx + 1) x (2 hy + 1),
 * hx = ceil(max|x| - 0.5 - 1e-7) (0 if negative), row-major, on the GPU of
 * `device`.  out = NULL queries the shape.  Errors as the reference's:
 * negative/non-finite PSF, zero input sum, non-finite vertex, |output_sum /
 * input_sum - 1| > 5e-5 (UPDC_ERR_NUMERIC; the output and sums are still
 * written).  overlaps = the number of links. */
int updc_psf_project(int device, const float *psf, int width, int height, const double *vertices, float *out,
                     size_t out_capacity, int *out_width, int *out_height, double *input_sum, double *output_sum,
                     uint64_t *overlaps);

/* ---- device-pointer operators (the autograd wrapper) ------------------------
 * On a prepared context, with DEVICE pointers of the context's element type
 * (float for UPDC_PREC_F32 and UPDC_PREC_MIXED, double for UPDC_PREC_F64),
 * enqueued on `stream` (a hipStream_t; NULL is HIP's null stream, as
 * everywhere in HIP -- torch's default stream) without synchronising.  Spectra are column-major half spectra: complex
 * element (row r, column v) at v * pitch + r, spectrum_columns = FW/2 + 1
 * columns of spectrum_pitch >= FH rows; rows >= FH are padding and never
 * read.  Calls on one context share a workspace: issue them on one stream.
 * No reference counterpart as an API: they expose the iteration's own stages
 * (ws_forward_target / ws_inverse_crop, updc_core.c:728-776; the LinkMap
 * gather and scatter, updc_core.c:1154-1176). */
int updc_op_layout(updc_ctx *ctx, int *fft_width, int *fft_height, int *spectrum_columns, int *spectrum_pitch,
                   int *element_bytes);
/* spectrum = R2C of the rows x cols image (row pitch `pitch` elements), zero
 * padded to FH x FW; unnormalised. */
int updc_op_r2c(updc_ctx *ctx, const void *image, int rows, int cols, int pitch, void *spectrum, void *stream);
/* image = scale * C2R(spectrum), cropped to rows x cols (row pitch `pitch`). */
int updc_op_c2r(updc_ctx *ctx, const void *spectrum, void *image, int rows, int cols, int pitch, double scale,
                void *stream);
/* Allocation-free variant of updc_op_c2r, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  Same arguments, kernels and results (bit for
 * bit) as updc_op_c2r; its column-inverse workspace is a static device array of
 * UPDC_NOALLOC_WORKSPACE_BYTES (128 MiB) that every GPU loading the library
 * holds until the process ends, instead of a buffer the context allocates on
 * its first updc_op_c2r call.  The spectrum's workspace, (FW/2 + 1) * pitch
 * elements of 8 (F32, MIXED) or 16 (F64) bytes, must fit: an FP32 5600 x 5600
 * transform does (125.5 MB), an FP64 one does not (UPDC_ERR_ARG, nothing
 * launched).  An FFT axis too long for one workgroup (four-step) also works
 * in the scratch updc_prepare reserved, so that path allocates nothing either.
 * Every call on one GPU shares that array, whatever the context:
 * calls that overlap in time on one GPU (other streams, other threads) race. */
#define UPDC_NOALLOC_WORKSPACE_BYTES 134217728u
int updc_op_c2r_noalloc(updc_ctx *ctx, const void *spectrum, void *image, int rows, int cols, int pitch,
                        double scale, void *stream);
/* out[p] = sum_z frac_z image[t_z] for the nin_k pixels of exposure k (S_k,
 * the fmaf chain of updc_core.c:1154-1158); image is H x W. */
int updc_op_gather(updc_ctx *ctx, int k, const void *image, void *out, void *stream);
/* image[t] = sum_p frac v[p] over the pixels with w > 0 (S_k^T restricted to
 * the weighted pixels, the transposed gather of updc_core.c:1171-1176);
 * v has nin_k elements, image is H x W and fully written. */
int updc_op_scatter(updc_ctx *ctx, int k, const void *v, void *image, void *stream);

/* 2-D FFT self-check (the GPU transforms on their own): image FH x FW real ->
 * half spectrum FH x (FW/2+1) interleaved complex, and back (unnormalised). */
int updc_fft_r2c_f64(updc_ctx *ctx, const double *image, double *spectrum);
int updc_fft_c2r_f64(updc_ctx *ctx, const double *spectrum, double *image);

/* Average device time (ms) per launch of each kernel class over the last
 * updc_iterate call, measured with HIP events on the library's stream:
 * names[i] / ms[i] / launches[i] for i < *count (capacity in, count out). */
int updc_get_stage_times(updc_ctx *ctx, char names[][32], double *ms, int *launches, int *count);

/* Measurement aid with no counterpart in the reference: copies `bytes` bytes
 * between two device buffers `reps` times with a streaming kernel
 * (k_copy_ref) and returns the average time per copy.  It is the known-size
 * workload against which HBM byte counters (FETCH_SIZE / WRITE_SIZE) are
 * scaled. */
int updc_copy_calibration(size_t bytes, int reps, int device, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* UPDC_HIP_H */
