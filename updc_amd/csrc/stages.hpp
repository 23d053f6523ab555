// Launchers of the non-FFT kernels (geometry, setup, detector, update).
#pragma once
#include "kernels.hpp"

namespace updc {

// ---- geometry (device restatement of overlap.c, bit-exact) ----------------
// Pass 1 of the LinkMap build (updc_core.c:549-568): per-pixel link count,
// first target, and flags bit 0 when some stencil spans more than 65535.
// PSF projector link lists (psf_projector.c:82-146), see stages.hip.
hipError_t pp_count(const double *vert, const float *val, int iw, int ih, int hx, int hy, uint32_t *cnt,
                    hipStream_t st);
hipError_t pp_fill(const double *vert, const float *val, int iw, int ih, int hx, int hy, const uint32_t *off,
                   uint32_t *t32, float *frac, unsigned int *flags, hipStream_t st);
hipError_t lm_count(const double *vert, int iw, int ih, int W, int H, uint32_t *cnt, uint32_t *first,
                    unsigned int *flags, hipStream_t st);
// Pass 2 (updc_core.c:586-597): fill delta16/target32 and frac at off[p].
hipError_t lm_fill(const double *vert, int iw, int ih, int W, int H, const uint32_t *off, const uint32_t *base,
                   uint32_t encoding, uint16_t *d16, uint32_t *t32, float *frac, unsigned int *flags,
                   hipStream_t st);

// Validation of an externally supplied LinkMap: offsets non-decreasing and
// within nlinks, every target < nout, fractions finite.  flags bit 1 offsets,
// bit 2 targets, bit 4 fractions.
hipError_t lm_validate(const uint32_t *off, const uint32_t *base, const uint16_t *d16, const uint32_t *t32,
                       const float *frac, uint32_t encoding, uint32_t nin, uint64_t nlinks, uint64_t nout,
                       unsigned int *flags, hipStream_t st);

// ---- scans ---------------------------------------------------------------
// out[0..n] = exclusive prefix sum of in[0..n), out[n] = total (mod 2^32);
// *total64 (device) receives the exact 64-bit total.  scratch: >= scan_scratch(n) bytes.
size_t scan_scratch_bytes(size_t n);
hipError_t scan_u32(const uint32_t *in, uint32_t *out, size_t n, void *scratch,
                    unsigned long long *total64, hipStream_t st);

// ---- CSC transpose of one exposure's LinkMap ------------------------------
template <class TW>
hipError_t csc_count(const ExpDev *e_dev, uint32_t nin, const TW *w, uint32_t *tcnt, hipStream_t st);
template <class TW>
hipError_t csc_fill(const ExpDev *e_dev, uint32_t nin, const TW *w, uint32_t *cursor, uint2 *tent, hipStream_t st);
hipError_t csc_sort(const uint32_t *toff, uint2 *tent, size_t nout, hipStream_t st);
// Compact copy of a sorted CSC (ExpDev::c_*): c_toff[ntiles+1], c_pb[ntiles],
// c_cnt[nout], c_dp/c_fr[ncsc], ntiles = ceil(nout/64).  flags (device, zeroed
// by the caller): bit 0 a target with more than 255 entries, bit 1 a tile
// spanning more than 65535 source pixels (then the copy is unusable).
hipError_t csc16_build(const uint32_t *toff, const uint2 *tent, size_t nout, uint32_t *c_toff, uint32_t *c_pb,
                       uint8_t *c_cnt, uint16_t *c_dp, float *c_fr, unsigned int *flags, hipStream_t st);

// ---- weights (updc_core.c:364-373) --------------------------------------
// F32 semantics (TW = float) or FP64 (TW = double).  err bits UPDC_E_WEIGHT /
// UPDC_E_YEFF; masked counts pixels set to zero.
template <class TW>
hipError_t normalize_weights(const float *sci_raw, const float *w_raw, size_t n, float scale, double scale_d,
                             float ped_f, double ped_d, TW *sci, TW *w, unsigned long long *masked,
                             unsigned int *err, hipStream_t st);

// ---- initial drizzle (updc_core.c:905-973) ---------------------------------
template <class TF>
hipError_t drizzle(const ExpDev *ex, int nexp, const TF *sci, const TF *w, int nworkers, size_t nout, TF *estimate,
                   TF *coverage, unsigned int *err, hipStream_t st);

// ---- detector stage (updc_core.c:1152-1177, forward half) -------------------
// blur: per exposure H*W at stride blur_bstride.  Writes wr (0 for inactive
// pixels), updates frozen, accumulates diag[k] and err.
template <class TF, class TA>
hipError_t gather_freeze(const ExpDev *ex, int kbeg, int kend, uint32_t max_nin, const TF *blur,
                         size_t blur_bstride, const TF *sci, const TF *w, uint8_t *frozen, TF *wr, ExpDiag *diag,
                         IterDiag *idiag, hipStream_t st);

// Forward gather only: pred[p] = sum_z fma(frac, blur[t], .) for every pixel
// with links (weights not consulted), all exposures in one launch.
template <class TF>
hipError_t predict(const ExpDev *ex, int nexp, uint32_t max_nin, const TF *blur, size_t blur_bstride, TF *pred,
                   hipStream_t st);

// ---- multiplicative update (updc_core.c:1437-1457) --------------------------
// C and D in the accumulation type TA (the mixed tier keeps both in FP64).
template <class TF, class TA>
hipError_t update(TF *x, const TA *C, const TA *D, size_t nout, TA max_sens, IterDiag *idiag, double *partials,
                  unsigned long long *negs, unsigned int *ticket, hipStream_t st);
int update_blocks();

// ---- misc -----------------------------------------------------------------
template <class T> hipError_t fill(T *p, T v, size_t n, hipStream_t st);
template <class TD> hipError_t max_key(const TD *p, size_t n, unsigned long long *key, hipStream_t st);
template <class TF>
hipError_t psf_wrap(const float *psf, int pw, int ph, int transposed, int FW, int FH, TF inv_sum, TF *image,
                    hipStream_t st);
template <class TS, class TD> hipError_t convert(const TS *s, TD *d, size_t n, hipStream_t st);
hipError_t copy_ref(const void *src, void *dst, size_t bytes, hipStream_t st);
hipError_t init_diag(ExpDiag *d, int n, IterDiag *it, int nit, hipStream_t st);
// Iteration bookkeeping that stays inside a captured graph: copy the current
// diagnostics to history slot *ctr, advance *ctr and clear the current slot.
hipError_t commit_diag(ExpDiag *cur, int nexp, IterDiag *icur, ExpDiag *hist, IterDiag *ihist, int cap,
                       unsigned int *ctr, hipStream_t st);
// out[t] = (S_k^T v)[t] for one exposure through the CSC gather (H*W targets).
template <class TF> hipError_t csc_apply(const ExpDev *e_dev, const TF *v, size_t nout, TF *out, hipStream_t st);
// All exposures of the batch: out[k * nout + t] = (S_k^T v_k)[t] over the CSC,
// the compact copy where an exposure has one (any_c16: some do; any_wide:
// some do not).
template <class TF>
hipError_t csc_image(const ExpDev *ex, int nexp, const TF *v, size_t nout, TF *out, bool any_c16, bool any_wide,
                     hipStream_t st);

}  // namespace updc
