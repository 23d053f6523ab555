// Host interface of the batched 2-D real FFT used by the iteration.
//
// Layout in HBM.  A real image is row-major with its own pitch.  A half
// spectrum (FH rows x NC = FW/2+1 columns) is stored COLUMN-major: element
// (row r, column v) at v * RP + r, with the row pitch RP a multiple of 16
// elements (rp() below).  The column passes then read and write each column
// as one contiguous run (full 128-byte requests), and the row passes, which
// handle RB rows per workgroup, touch RB consecutive elements per column.
// Rows beyond the zero-padded input height are never transformed or stored,
// and the inverse only produces the cropped [0,H) x [0,W) window, which is all
// the reference ever reads back (updc_core.c:728-744).
#pragma once
#include "kernels.hpp"
#include <vector>
#include <algorithm>
#include <string>

namespace updc {

enum ColMode { COL_FWD = 0, COL_INV = 1, COL_BCAST = 2, COL_ADJ = 3, COL_ACC = 4 };
enum RowMode { ROW_PLAIN = 0, ROW_GATHER = 1 };

// Spectrum layouts.  SL_COL: column-major, (r, v) at v * RP + r.  SL_PAIR:
// the columns in pairs (2p, 2p+1) interleaved element by element, (r, v) at
// ((v >> 1) * RP + r) * 2 + (v & 1).  A row C2R tile of RB rows then reads
// RB * 2 consecutive elements per pair (64 bytes for FP32 RB = 4 and FP64
// RB = 2) instead of RB per column (32 bytes), while the column pass that
// writes the spectrum stores every other element of a 2 RP run; it runs the
// two columns of a pair on one XCD, back to back, so its L2 merges their
// halves of every line.  Measured for 8 spectra of 2801 columns x 4096 rows
// (FP32, 734 MB): a 4-row tile pass reads in 0.22 ms instead of 0.30, the
// column pass writes in 0.15 ms instead of 0.13.  The batched iteration
// uses SL_PAIR for the spectra a column pass writes and a row C2R reads (the
// blurred spectra, the correction) where both kernels support it
// (pair_ok).  The spectra a row R2C writes stay SL_COL: the same layout made
// the adjoint images' row R2C and the summed-adjoint column pass slower
// together, by 0.06 ms at that size.  PSF spectra, the per-exposure form and
// the operators are SL_COL throughout.
enum SpecLayout { SL_COL = 0, SL_PAIR = 1 };

struct FftShape {
    int FW = 0, FH = 0, M = 0, NC = 0;
};

// Column pitch (elements) of a spectrum holding `rows` rows.
inline int rp(int rows) { return (rows + 15) / 16 * 16; }

// Lines too long for one workgroup's LDS (rows of more than M = 4096 points,
// columns of more than 8192) run as a four-step transform: L = L1 x L2, index
// n = n1 + L1 n2 in, k = k2 + L2 k1 out.  Pass A: for each n1, the L2-point
// transform over n2 of the strided inputs, times exp(-+2 pi i n1 k2 / L),
// stored to a line-major scratch at k2 L1 + n1.  Pass B: for each k2, the
// L1-point transform over n1, stored at k2 + L2 k1.  Both passes transform
// B adjacent sub-lines per workgroup (interleaved in LDS), so every global
// access moves B consecutive elements.
enum LongLoad { LL_CX = 0, LL_CX_MUL = 1, LL_REAL = 2, LL_GATHER = 3 };
enum LongStore { LS_CX = 0, LS_REAL = 1 };
struct LineIO {
    const void *p = nullptr;  // source or destination base
    size_t bstride = 0;       // elements between batches (lpb lines each)
    size_t pitch = 0;         // elements between the lines of one batch
    int lpb = 1;              // lines per batch
    int nvalid = 0;           // loads: elements read (LL_REAL: real columns, LL_GATHER: W), zero beyond;
                              // stores: elements kept (LS_REAL: real columns)
    const void *psf = nullptr;  // LL_CX_MUL: line l of batch b times psf[b pbstride + l ppitch + n]
    size_t pbstride = 0, ppitch = 0;
    const ExpDev *ex = nullptr;  // LL_GATHER: exposure b; wr per source pixel
    const void *wr = nullptr;
    double scale = 1.0;  // LS_REAL
};
template <class T> struct LongPlan {
    int L = 0, L1 = 0, L2 = 0, lgb = 0;  // B = 1 << lgb sub-lines per workgroup
    FftLen p1{}, p2{};                    // L1- and L2-point plans
    Cx<T> *tw1 = nullptr, *tw2 = nullptr, *roots = nullptr;  // roots[m] = exp(-2 pi i m / L)
};

template <class T> struct Fft2 {
    FftShape s;
    FftLen row{}, col{};
    Cx<T> *tw_row = nullptr, *tw_col = nullptr, *post = nullptr;  // device tables
    int col_re = 0;          // edge radix of the column plan (0: LDS-only column kernel)
    bool col_pp = false;     // edge column kernel with two LDS transform buffers (out-of-place stages)
    bool allow_pp = true;    // set before init(): false keeps the edge column kernel in place
    bool allow_ten = true;   // set before init(): false keeps radix-10 edges out of FP32 column plans
    // FP32 only: the column plan with FP64 twiddles (the accumulate pass's FP64
    // inverse of the summed spectrum runs it)
    FftLen col_wide{};
    void *tw_col_wide = nullptr;
    bool allow_edge = true;  // set before init(): false forces the LDS-only column kernel
    bool row_persistent = true;  // set before init(): persistent pipelined row passes where they fit;
                                 // after init(): the persistent row C2R passes run
    bool r2c_persistent = false;  // after init(): the persistent row R2C runs (it fits longer rows than the C2R)
    bool c2r_long = false;        // after init(): FP32 rows too long for the 4-row persistent C2R take a 2-row one
                                  // (row_c2r only; row_c2r_update needs row_persistent)
    bool row_odd_first = true;   // set before init(): an odd radix first in the row plan (LDS banks)
    int row_grid = 0;            // workgroups of the persistent row passes (one per CU)
    int row_ntw = 0;             // entries of the row plan's stage twiddle tables
    // Four-step rows / columns where one workgroup cannot hold a line, with
    // their device scratch (a, b: complex T line sets; c: one summed spectrum
    // in the accumulation type).
    bool long_row = false, long_col = false;
    LongPlan<T> lrow, lcol;
    Cx<T> *lt_a = nullptr, *lt_b = nullptr;
    void *lt_c = nullptr;
    size_t lt_a_n = 0, lt_b_n = 0, lt_c_bytes = 0;

    // Returns "" on success or an error message.
    std::string init(int FW, int FH);
    void release();
    // Four-step scratch for transforms of up to `nexp` images (or spectra)
    // at the full FFT size, allocated now rather than at the first call (a
    // call under graph capture cannot allocate).  No-op when every line fits.
    hipError_t long_reserve(int nexp);
    // Whether the four-step scratch already covers a row C2R of `rows` rows
    // and a column pass of one spectrum (true when no axis is four-step).
    bool long_covers(int rows) const {
        const size_t r = long_row ? (size_t)rows * s.M : 0, c = long_col ? (size_t)s.NC * s.FH : 0;
        return std::max(r, c) <= lt_a_n && r <= lt_b_n;
    }
    // elements of one spectrum buffer in either layout (SL_PAIR stores a
    // phantom partner of the last column when NC is odd)
    size_t spec_elems(int rows) const { return (size_t)(s.NC + (s.NC & 1)) * rp(rows); }
    // Whether the batched passes accept SL_PAIR spectra: the row C2R passes
    // (row_c2r, row_c2r_update) where they are persistent (row_c2r also in
    // its 2-row form) and no row is four-step, the column passes' outputs
    // (colpass, colacc) where the edge column kernel runs.
    bool row_pair_ok() const;
    bool col_pair_ok() const { return col_re > 0 && !long_col; }
    bool pair_ok() const { return row_pair_ok() && col_pair_ok(); }

    // Row R2C of rows [0, rows) of `batch` images -> column-major spectra
    // out[b * out_bstride + v * RP + r], RP = rp(rp_rows).
    //   ROW_PLAIN : in[b*in_bstride + r*in_pitch + c], c < win (zero beyond)
    //   ROW_GATHER: value at target (r, c<W) = CSC gather of wr for exposure b
    hipError_t row_r2c(int mode, const T *in, size_t in_bstride, int in_pitch, int win, const ExpDev *ex,
                       const T *wr, int W, Cx<T> *out, size_t out_bstride, int rp_rows, int rows, int batch,
                       hipStream_t st);
    // Row C2R of rows [0, rows) of `batch` column-major spectra (pitch
    // rp(rp_rows); lin: their layout) -> out[b*out_bstride + r*out_pitch + c]
    // for c < wout, times scale.
    hipError_t row_c2r(const Cx<T> *in, size_t in_bstride, int rp_rows, T *out, size_t out_bstride, int out_pitch,
                       int wout, T scale, int rows, int batch, hipStream_t st, int lin = SL_COL);
    // Row C2R of one spectrum (rows [0, H), pitch rp(rp_rows)) fused with the
    // multiplicative update of x (W x H) by C = scale * C2R against D, the
    // signed sum and negative count going to idiag (partials/negs: one slot
    // per workgroup, ticket: zero between launches).  hipErrorNotSupported
    // when the persistent row passes do not apply: run row_c2r and update.
    template <class TF>
    hipError_t row_c2r_update(const Cx<T> *in, int rp_rows, TF *x, const T *D, int W, int H, T scale, T max_sens,
                              IterDiag *idiag, double *partials, unsigned long long *negs, unsigned int *ticket,
                              hipStream_t st, int lin = SL_COL);
    // Column passes (complex, length FH) over column-major spectra.  Rows >=
    // hin read as zero; rows [0, hout) are stored.  Input pitch rp(in_rows),
    // output pitch rp(out_rows); PSF spectra have pitch rp(FH).
    //   COL_FWD / COL_INV : in[b] -> out[b], b < batch
    //   COL_BCAST: X = FFT(in); out[k] = IFFT(X * psf[k]) for k < nexp
    //   COL_ADJ  : out = IFFT(sum_k conj(psf[k]) * FFT(in[k]))   (sum in TA)
    // in and the PSF spectra are SL_COL; lout: the layout of out (SL_PAIR
    // needs col_pair_ok() and a batched mode: COL_BCAST, COL_ADJ, COL_ACC).
    template <class TA>
    hipError_t colpass(int mode, const Cx<T> *in, size_t in_bstride, int in_rows, Cx<T> *out, size_t out_bstride,
                       int out_rows, const Cx<T> *psf, size_t psf_bstride, int hin, int hout, int batch,
                       hipStream_t st, int lout = SL_COL);
    // COL_ACC: A = sum_k conj(psf[k]) * FFT(in[k]) kept in TA (TA wider than
    // T).  fa == nullptr: out = A, all FH rows, pitch rp(FH).  Else out =
    // IFFT(A) computed in TA inside the same pass with plan fa (the TA plan of
    // the same height, accepted by acc_inverse_ok), rows [0, hout), pitch
    // rp(out_rows).
    template <class TA>
    hipError_t colacc(const Cx<T> *in, size_t in_bstride, int in_rows, Cx<TA> *out, int out_rows, int hout,
                      const Cx<T> *psf, size_t psf_bstride, int hin, int batch, const FftLen *fa, hipStream_t st,
                      int lout = SL_COL);
    // Whether colacc can run the TA inverse with plan fa (same radices as the
    // edge kernel's plan, registers and LDS within the column configuration).
    bool acc_inverse_ok(const FftLen &fa) const;
    // out = IFFT(acc) in TA with plan fa, the inverse colacc runs, for a sum
    // already in memory (FH rows, pitch rp(FH)); rows [0, hout), pitch rp(out_rows).
    template <class TA>
    hipError_t acc_inverse(const Cx<TA> *acc, Cx<TA> *out, int out_rows, int hout, const FftLen &fa, hipStream_t st);
    template <class TA>
    hipError_t colpass_impl(int mode, const Cx<T> *in, size_t in_bstride, int in_rows, Cx<T> *out,
                            size_t out_bstride, int out_rows, const Cx<T> *psf, size_t psf_bstride, int hin,
                            int hout, int batch, FftLen fa, size_t lds, hipStream_t st, int lout = SL_COL);
    // Plain (per-exposure) form helpers over full column-major spectra (pitch rp(FH)):
    //   out = a * psf_k ;  acc += conj(psf_k) * a  (acc in TA) ;  out = (T) acc
    hipError_t spec_mul(const Cx<T> *a, const Cx<T> *psf_k, Cx<T> *out, hipStream_t st);
    template <class TA> hipError_t spec_acc(const Cx<T> *a, const Cx<T> *psf_k, Cx<TA> *acc, hipStream_t st);
    template <class TA> hipError_t spec_cast(const Cx<TA> *acc, Cx<T> *out, hipStream_t st);

    // four-step pieces (fft.hip)
    hipError_t long_space(size_t a, size_t b, size_t cbytes, hipStream_t st);
    template <class TS>
    hipError_t long_a(const LongPlan<T> &lp, bool inv, int ld, const LineIO &src, size_t lines, Cx<T> *tmp,
                      hipStream_t st);
    hipError_t long_b(const LongPlan<T> &lp, bool inv, int store, const Cx<T> *tmp, const LineIO &dst, size_t lines,
                      hipStream_t st);
    hipError_t row_r2c_long(int mode, const T *in, size_t in_bstride, int in_pitch, int win, const ExpDev *ex,
                            const T *wr, int W, Cx<T> *out, size_t out_bstride, int rp_rows, int rows, int batch,
                            hipStream_t st);
    hipError_t row_c2r_long(const Cx<T> *in, size_t in_bstride, int rp_rows, T *out, size_t out_bstride,
                            int out_pitch, int wout, T scale, int rows, int batch, hipStream_t st);
    template <class TA>
    hipError_t colpass_long(int mode, const Cx<T> *in, size_t in_bstride, int in_rows, Cx<T> *out,
                            size_t out_bstride, int out_rows, const Cx<T> *psf, size_t psf_bstride, int hin,
                            int hout, int batch, hipStream_t st);
};

}  // namespace updc
