// Host interface of the batched 2-D real FFT used by the iteration.
//
// Layout in HBM: a real image is row-major with its own pitch; a half
// spectrum is FH rows x NC = FW/2+1 columns of interleaved complex, row pitch
// NCP (NC rounded up to 16 elements so every row starts 128-byte aligned).
// The 2-D R2C is a row pass (half-length complex FFT + split) followed by a
// column pass; the C2R is the column pass followed by the row pass.  Rows
// beyond the zero-padded input height are never transformed or read, and the
// inverse only produces the cropped [0,H) x [0,W) window, which is all the
// reference ever reads back (updc_core.c:728-744).
#pragma once
#include "kernels.hpp"
#include <vector>
#include <string>

namespace updc {

enum ColMode { COL_FWD = 0, COL_INV = 1, COL_BCAST = 2, COL_ADJ = 3 };
enum RowMode { ROW_PLAIN = 0, ROW_GATHER = 1 };

struct FftShape {
    int FW = 0, FH = 0, M = 0, NC = 0, NCP = 0;
};

template <class T> struct Fft2 {
    FftShape s;
    FftLen row{}, col{};
    Cx<T> *tw_row = nullptr, *tw_col = nullptr, *post = nullptr;  // device tables
    int row_nt = 256, col_nt = 1024, col_lines = 1, col_lg = 0, row_vmax = 24, col_vmax = 24;
    int lds_limit = 0;

    // Returns "" on success or an error message.
    std::string init(int FW, int FH);
    void release();

    // Row R2C: rows [0, rows) of `batch` images.
    //   ROW_PLAIN : in[b*in_bstride + r*in_pitch + c], c < win (zero beyond)
    //   ROW_GATHER: value at target (r, c<W) = CSC gather of wr for exposure b
    hipError_t row_r2c(int mode, const T *in, size_t in_bstride, int in_pitch, int win, const ExpDev *ex,
                       const T *wr, int W, Cx<T> *out, size_t out_bstride, int rows, int batch, hipStream_t st);
    // Row C2R: rows [0, rows) of `batch` spectra -> out[b*out_bstride + r*out_pitch + c] for c < wout, times scale.
    hipError_t row_c2r(const Cx<T> *in, size_t in_bstride, T *out, size_t out_bstride, int out_pitch, int wout,
                       T scale, int rows, int batch, hipStream_t st);
    // Column passes (complex, length FH).  Rows >= hin read as zero; rows
    // [0, hout) are stored.
    //   COL_FWD / COL_INV : in[b] -> out[b], b < batch
    //   COL_BCAST: X = FFT(in); out[k] = IFFT(X * psf[k]) for k < nexp
    //   COL_ADJ  : out = IFFT(sum_k conj(psf[k]) * FFT(in[k]))   (sum in TA)
    //   (psf tile-major, psf_bstride = tile_stride())
    // PSF spectra for COL_BCAST / COL_ADJ are stored tile-major: for column
    // tile t (columns t*lines .. t*lines+lines-1), rows 0..FH-1, `lines`
    // interleaved columns, so one workgroup reads its whole tile contiguously.
    int ntiles() const { return (s.NC + col_lines - 1) / col_lines; }
    size_t tile_stride() const { return (size_t)ntiles() * col_lines * s.FH; }  // elements per spectrum
    hipError_t to_tiles(const Cx<T> *src, Cx<T> *dst, hipStream_t st);   // [FH][NCP] -> tile-major
    hipError_t from_tiles(const Cx<T> *src, Cx<T> *dst, hipStream_t st); // tile-major -> [FH][NCP]
    // Plain (per-exposure) form helpers over full FH x NCP spectra:
    //   out = a * psf_k                         (psf tile-major)
    //   acc += conj(psf_k) * a  (acc in TA)     (psf tile-major)
    //   out = (T) acc
    hipError_t spec_mul(const Cx<T> *a, const Cx<T> *psf_k, Cx<T> *out, hipStream_t st);
    template <class TA> hipError_t spec_acc(const Cx<T> *a, const Cx<T> *psf_k, Cx<TA> *acc, hipStream_t st);
    template <class TA> hipError_t spec_cast(const Cx<TA> *acc, Cx<T> *out, hipStream_t st);
    template <class TA>
    hipError_t colpass(int mode, const Cx<T> *in, size_t in_bstride, Cx<T> *out, size_t out_bstride, const Cx<T> *psf,
                   size_t psf_bstride, int hin, int hout, int batch, hipStream_t st);
};

}  // namespace updc
