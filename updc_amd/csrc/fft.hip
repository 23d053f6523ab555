// Batched 2-D real FFT for gfx950: LDS-resident Stockham passes.
//
// Replaces the FFTW3f R2C/C2R calls of the reference iteration
// (updc/c_core/updc_core.c:733-744, 761-776, 988-995) with three kernels:
//   k_row_r2c : per row, a half-length complex FFT in LDS plus the real split
//               (optionally fed by the atomic-free CSC gather of S_k^T);
//   k_row_c2r : per row, the inverse split plus half-length inverse FFT,
//               writing only the cropped columns, times 1/nf;
//   k_col     : per tile of adjacent columns, the complex column FFT in LDS;
//               the BCAST and ADJ forms keep one spectrum column tile in
//               registers and loop over all exposures inside the workgroup,
//               so X^ and the summed adjoint spectrum never touch HBM.
#include "fft.hpp"
#include <cmath>
#include <algorithm>

namespace updc {

// ------------------------------------------------------------------ kernels
template <class T> __device__ __forceinline__ T csc_gather(const ExpDev &e, const T *__restrict__ w, uint32_t t) {
    // scatter[t] = sum over source pixels p (ascending) of fma(frac, w[p], .)
    // from 0 -- the same fmaf chain as updc_core.c:1173-1176 / :1064, since the
    // reference visits pixels in ascending p and each (p,t) pair once.
    const uint32_t a = e.toff[t];
    const uint32_t n = e.toff[t + 1] - a;
    // entries and their w loads issued together, then the chain in order
    constexpr int U = 6;
    uint2 en[U];
    T wv[U];
#pragma unroll
    for (int j = 0; j < U; ++j) en[j] = (uint32_t)j < n ? e.tent[a + j] : make_uint2(0u, 0u);
#pragma unroll
    for (int j = 0; j < U; ++j) wv[j] = (uint32_t)j < n ? w[en[j].x] : (T)0;
    T acc = (T)0;
#pragma unroll
    for (int j = 0; j < U; ++j)
        if ((uint32_t)j < n) acc = fmat((T)__uint_as_float(en[j].y), wv[j], acc);
    for (uint32_t j = U; j < n; ++j) {
        const uint2 x = e.tent[a + j];
        acc = fmat((T)__uint_as_float(x.y), w[x.x], acc);
    }
    return acc;
}

template <class T, int NT, int VMAX, int MODE>
__global__ __launch_bounds__(NT) void k_row_r2c(FftLen f, const Cx<T> *__restrict__ post, const T *__restrict__ in,
                                                size_t in_bstride, int in_pitch, int win,
                                                const ExpDev *__restrict__ ex, const T *__restrict__ wr, int W,
                                                Cx<T> *__restrict__ out, size_t out_bstride, int out_pitch) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cx<T> *z = reinterpret_cast<Cx<T> *>(smem_raw);
    const int r = blockIdx.x, b = blockIdx.y, M = f.L;
    if (MODE == ROW_PLAIN) {
        const T *row = in + (size_t)b * in_bstride + (size_t)r * in_pitch;
        for (int n = threadIdx.x; n < M; n += NT) {
            const int c = 2 * n;
            const T x0 = c < win ? row[c] : (T)0;
            const T x1 = c + 1 < win ? row[c + 1] : (T)0;
            z[sw(n)] = cx<T>(x0, x1);
        }
    } else {
        const ExpDev &e = ex[b];
        const T *w = wr + e.pix;
        const uint32_t t0 = (uint32_t)r * (uint32_t)W;
        for (int n = threadIdx.x; n < M; n += NT) {
            const int c = 2 * n;
            const T x0 = c < W ? csc_gather<T>(e, w, t0 + c) : (T)0;
            const T x1 = c + 1 < W ? csc_gather<T>(e, w, t0 + c + 1) : (T)0;
            z[sw(n)] = cx<T>(x0, x1);
        }
    }
    __syncthreads();
    fft_lds<false, T, NT, VMAX>(z, f, 0);
    // Split: X[v] = (Z[v] + Z*[M-v])/2 - i/2 exp(-2 pi i v/FW) (Z[v] - Z*[M-v]), v in [0, M].
    Cx<T> *o = out + (size_t)b * out_bstride + (size_t)r * out_pitch;
    const T half = (T)0.5;
    for (int v = threadIdx.x; v <= M; v += NT) {
        const Cx<T> zv = z[sw(v == M ? 0 : v)];
        const Cx<T> zm = cconj(z[sw(v == 0 ? 0 : M - v)]);
        const Cx<T> fe = cscale(zv + zm, half);
        const Cx<T> d = zv - zm;
        const Cx<T> fo = cx<T>(d.im * half, -d.re * half);  // (zv - zm) / (2i)
        o[v] = fe + cmul(post[v], fo);
    }
}

template <class T, int NT, int VMAX>
__global__ __launch_bounds__(NT) void k_row_c2r(FftLen f, const Cx<T> *__restrict__ post, const Cx<T> *__restrict__ in,
                                                size_t in_bstride, int in_pitch, T *__restrict__ out,
                                                size_t out_bstride, int out_pitch, int wout, T scale) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cx<T> *z = reinterpret_cast<Cx<T> *>(smem_raw);
    const int r = blockIdx.x, b = blockIdx.y, M = f.L;
    const Cx<T> *y = in + (size_t)b * in_bstride + (size_t)r * in_pitch;
    // Merge: C[v] = E[v] + i O[v], E = Y[v] + Y*[M-v], O = (Y[v] - Y*[M-v]) exp(+2 pi i v/FW).
    for (int v = threadIdx.x; v < M; v += NT) {
        const Cx<T> yv = y[v];
        const Cx<T> ym = cconj(y[M - v]);
        const Cx<T> e = yv + ym;
        const Cx<T> o = cmul(yv - ym, cconj(post[v]));
        z[sw(v)] = cx<T>(e.re - o.im, e.im + o.re);
    }
    __syncthreads();
    fft_lds<true, T, NT, VMAX>(z, f, 0);
    T *orow = out + (size_t)b * out_bstride + (size_t)r * out_pitch;
    for (int n = threadIdx.x; n < M; n += NT) {
        const int c = 2 * n;
        const Cx<T> v = z[sw(n)];
        if (c < wout) orow[c] = v.re * scale;
        if (c + 1 < wout) orow[c + 1] = v.im * scale;
    }
}

template <class T, class TA, int NT, int VMAX, int XMAX, int MODE>
__global__ __launch_bounds__(NT) void k_col(FftLen f, int nc, int lg, int pitch, int hin, int hout,
                                            const Cx<T> *__restrict__ in, size_t in_bs, Cx<T> *__restrict__ out,
                                            size_t out_bs, const Cx<T> *__restrict__ psf, size_t psf_bs, int nexp,
                                            int ntiles) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cx<T> *buf = reinterpret_cast<Cx<T> *>(smem_raw);
    const int tile = xcd_tile(blockIdx.x, ntiles);
    if (tile < 0) return;
    const int lines = 1 << lg, lm = lines - 1;
    const int c0 = tile << lg;
    const int FH = f.L;
    const int tot = FH << lg;
    const int tid = threadIdx.x;

    auto load = [&](const Cx<T> *src) {
        for (int e = tid; e < tot; e += NT) {
            const int n = e >> lg, col = c0 + (e & lm);
            buf[sw(e)] = (n < hin && col < nc) ? src[(size_t)n * pitch + col] : cx<T>((T)0, (T)0);
        }
        __syncthreads();
    };
    auto store = [&](Cx<T> *dst) {
        const int lim = hout << lg;
        for (int e = tid; e < lim; e += NT) {
            const int n = e >> lg, col = c0 + (e & lm);
            if (col < nc) dst[(size_t)n * pitch + col] = buf[sw(e)];
        }
    };

    if (MODE == COL_FWD || MODE == COL_INV) {
        const int b = blockIdx.y;
        load(in + (size_t)b * in_bs);
        fft_lds<MODE == COL_INV, T, NT, VMAX>(buf, f, lg);
        store(out + (size_t)b * out_bs);
    } else if (MODE == COL_BCAST) {
        // X^ column tile: forward FFT once, kept in registers across exposures.
        load(in);
        fft_lds<false, T, NT, VMAX>(buf, f, lg);
        Cx<T> xr[XMAX];
#pragma unroll
        for (int j = 0; j < XMAX; ++j) {
            const int e = tid + j * NT;
            xr[j] = e < tot ? buf[sw(e)] : cx<T>((T)0, (T)0);
        }
        for (int k = 0; k < nexp; ++k) {
            __syncthreads();
            const Cx<T> *pk = psf + (size_t)k * psf_bs + (size_t)tile * tot;  // tile-major, contiguous
#pragma unroll
            for (int j = 0; j < XMAX; ++j) {
                const int e = tid + j * NT;
                if (e < tot) {
                    // (a+ib)(c+id) as a*c-b*d, a*d+b*c (updc_core.c:756-758, adjoint=0)
                    buf[sw(e)] = cmul(xr[j], pk[e]);
                }
            }
            __syncthreads();
            fft_lds<true, T, NT, VMAX>(buf, f, lg);
            store(out + (size_t)k * out_bs);
        }
    } else {  // COL_ADJ
        Cx<TA> acc[XMAX];
#pragma unroll
        for (int j = 0; j < XMAX; ++j) acc[j] = cx<TA>((TA)0, (TA)0);
        for (int k = 0; k < nexp; ++k) {
            __syncthreads();
            load(in + (size_t)k * in_bs);
            fft_lds<false, T, NT, VMAX>(buf, f, lg);
            const Cx<T> *pk = psf + (size_t)k * psf_bs + (size_t)tile * tot;  // tile-major, contiguous
#pragma unroll
            for (int j = 0; j < XMAX; ++j) {
                const int e = tid + j * NT;
                if (e < tot) {
                    const Cx<T> p = pk[e];
                    const Cx<T> s = buf[sw(e)];
                    // re += a*c + b*d ; im += b*c - a*d  (updc_core.c:991-993)
                    const T pre = s.re * p.re + s.im * p.im;
                    const T pim = s.im * p.re - s.re * p.im;
                    acc[j].re += (TA)pre;
                    acc[j].im += (TA)pim;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < XMAX; ++j) {
            const int e = tid + j * NT;
            if (e < tot) buf[sw(e)] = cx<T>((T)acc[j].re, (T)acc[j].im);
        }
        __syncthreads();
        fft_lds<true, T, NT, VMAX>(buf, f, lg);
        store(out);
    }
}

template <class T>
__global__ void k_to_tiles(const Cx<T> *__restrict__ src, int pitch, int nc, int FH, int lg, size_t total,
                           Cx<T> *__restrict__ dst, bool back) {
    const size_t tot = (size_t)FH << lg;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const size_t tile = i / tot, e = i % tot;
        const int n = (int)(e >> lg), col = (int)(tile << lg) + (int)(e & ((1u << lg) - 1));
        if (!back) dst[i] = col < nc ? src[(size_t)n * pitch + col] : cx<T>((T)0, (T)0);
        else if (col < nc) dst[(size_t)n * pitch + col] = src[i];
    }
}

// Elementwise spectrum kernels of the per-exposure form (the reference's
// ws_multiply_spectrum / accumulate_adjoint_spectrum, updc_core.c:753-760,
// 988-995), over [FH][NCP] spectra with tile-major PSF spectra.
template <class T, class TA, int OP>
__global__ void k_spec(const Cx<T> *__restrict__ a, const Cx<T> *__restrict__ psf, Cx<T> *__restrict__ out,
                       Cx<TA> *__restrict__ acc, int pitch, int nc, int FH, int lg) {
    const size_t n = (size_t)FH * nc;
    const size_t tot = (size_t)FH << lg;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int r = (int)(i / nc), col = (int)(i % nc);
        const size_t g = (size_t)r * pitch + col;
        if (OP == 2) {
            out[g] = cx<T>((T)acc[g].re, (T)acc[g].im);
            continue;
        }
        const size_t tix = (size_t)(col >> lg) * tot + ((size_t)r << lg) + (col & ((1 << lg) - 1));
        const Cx<T> p = psf[tix], v = a[g];
        if (OP == 0) {
            out[g] = cmul(v, p);
        } else {
            const T pre = v.re * p.re + v.im * p.im;
            const T pim = v.im * p.re - v.re * p.im;
            acc[g].re += (TA)pre;
            acc[g].im += (TA)pim;
        }
    }
}

// ------------------------------------------------------------------ host side
// Thread counts and per-thread register budgets.  One column workgroup owns
// most of a CU's LDS, so it runs 8 waves (2 per SIMD, at most 256 VGPRs);
// the budgets below keep the column kernels free of spills in both precisions.
static constexpr int ROW_NT = 512, ROW_VMAX = 12;
template <class T> struct ColCfg;
template <> struct ColCfg<float> { static constexpr int NT = 512, VMAX = 16, XMAX = 16; };
template <> struct ColCfg<double> { static constexpr int NT = 512, VMAX = 16, XMAX = 12; };
static constexpr size_t LDS_BUDGET = 150 * 1024;
// LDS elements for n FFT values: sw() permutes inside 16-element blocks, so
// buffers are rounded up to whole blocks.
static inline size_t lds_elems(size_t n) { return (n + 15) & ~(size_t)15; }

static bool factor_radices(int L, FftLen &f) {
    int n = L, k = 0;
    auto push = [&](int r) { if (k < UPDC_MAX_STAGES) f.rad[k++] = r; };
    while (n % 8 == 0) { push(8); n /= 8; }
    if (n % 4 == 0) { push(4); n /= 4; }
    if (n % 2 == 0) { push(2); n /= 2; }
    while (n % 7 == 0) { push(7); n /= 7; }
    while (n % 5 == 0) { push(5); n /= 5; }
    while (n % 3 == 0) { push(3); n /= 3; }
    f.L = L;
    f.nst = k;
    long ns = 1;
    for (int i = 0; i < k; ++i) {
        f.ns[i] = (int)ns;
        f.mg[i] = (unsigned int)(((1ull << 32) + (unsigned long long)ns - 1) / (unsigned long long)ns);
        ns *= f.rad[i];
    }
    return n == 1 && k <= UPDC_MAX_STAGES && L < 65536;
}

// Every stage must fit: R * ceil(lines*L/R / NT) <= VMAX.
static bool fits(const FftLen &f, int lines, int nt, int vmax) {
    for (int s = 0; s < f.nst; ++s) {
        const int R = f.rad[s];
        const long nb = (long)lines * (f.L / R);
        const long per = (nb + nt - 1) / nt;
        if (per > vmax / R) return false;
    }
    return true;
}

template <class T> static Cx<T> *upload_roots(int L, int count, int sign, std::string &err) {
    // table[m] = exp(sign * 2 pi i m / L), m < count, evaluated in long double
    std::vector<Cx<T>> h(count);
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int m = 0; m < count; ++m) {
        const long double a = twopi * (long double)m / (long double)L;
        h[m].re = (T)cosl(a);
        h[m].im = (T)(sign * sinl(a));
    }
    Cx<T> *d = nullptr;
    if (hipMalloc(&d, sizeof(Cx<T>) * count) != hipSuccess) { err = "hipMalloc twiddles failed"; return nullptr; }
    if (hipMemcpy(d, h.data(), sizeof(Cx<T>) * count, hipMemcpyHostToDevice) != hipSuccess) {
        err = "twiddle upload failed";
        hipFree(d);
        return nullptr;
    }
    return d;
}

template <class T, int NT, int VMAX, int MODE> static void *row_r2c_fn() { return (void *)k_row_r2c<T, NT, VMAX, MODE>; }
template <class T, int NT, int VMAX> static void *row_c2r_fn() { return (void *)k_row_c2r<T, NT, VMAX>; }
template <class T, class TA, int MODE> static void *col_fn() {
    return (void *)k_col<T, TA, ColCfg<T>::NT, ColCfg<T>::VMAX, ColCfg<T>::XMAX, MODE>;
}

template <class T> std::string Fft2<T>::init(int FW, int FH) {
    if (FW < 2 || FH < 1 || (FW & 1)) return "FFT width must be even and >= 2";
    s.FW = FW; s.FH = FH; s.M = FW / 2; s.NC = FW / 2 + 1; s.NCP = (s.NC + 15) / 16 * 16;
    if (!factor_radices(s.M, row)) return "FFT width/2 is not 7-smooth";
    if (!factor_radices(FH, col)) return "FFT height is not 7-smooth";
    if (lds_elems(s.M) * sizeof(Cx<T>) > LDS_BUDGET || !fits(row, 1, ROW_NT, ROW_VMAX))
        return "FFT width too large for the LDS-resident row pass";
    constexpr int CNT = ColCfg<T>::NT, CVMAX = ColCfg<T>::VMAX, CXMAX = ColCfg<T>::XMAX;
    row_nt = ROW_NT; row_vmax = ROW_VMAX; col_nt = CNT; col_vmax = CVMAX;
    col_lines = 0;
    for (int t = 16; t >= 1; t >>= 1) {
        if (lds_elems((size_t)t * FH) * sizeof(Cx<T>) > LDS_BUDGET) continue;
        if ((long)t * FH > (long)CNT * CXMAX) continue;
        if (!fits(col, t, CNT, CVMAX)) continue;
        col_lines = t;
        col_lg = 0;
        while ((1 << col_lg) < t) ++col_lg;
        break;
    }
    if (!col_lines) return "FFT height too large for the LDS-resident column pass";
    std::string err;
    tw_row = upload_roots<T>(s.M, s.M, -1, err);
    if (!tw_row) return err;
    tw_col = upload_roots<T>(FH, FH, -1, err);
    if (!tw_col) return err;
    post = upload_roots<T>(FW, s.M + 1, -1, err);
    if (!post) return err;
    row.tw = tw_row;
    col.tw = tw_col;
    const int row_lds = (int)(lds_elems(s.M) * sizeof(Cx<T>));
    const int col_lds = (int)(lds_elems((size_t)col_lines * FH) * sizeof(Cx<T>));
    void *fns_row[] = {row_r2c_fn<T, ROW_NT, ROW_VMAX, ROW_PLAIN>(), row_r2c_fn<T, ROW_NT, ROW_VMAX, ROW_GATHER>(),
                       row_c2r_fn<T, ROW_NT, ROW_VMAX>()};
    for (void *fn : fns_row)
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, row_lds) != hipSuccess)
            return "cannot raise LDS limit for the row pass";
    void *fns_col[] = {col_fn<T, T, COL_FWD>(), col_fn<T, T, COL_INV>(), col_fn<T, T, COL_BCAST>(),
                       col_fn<T, float, COL_ADJ>(), col_fn<T, double, COL_ADJ>()};
    for (void *fn : fns_col)
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, col_lds) != hipSuccess)
            return "cannot raise LDS limit for the column pass";
    lds_limit = col_lds;
    return "";
}

template <class T> hipError_t Fft2<T>::to_tiles(const Cx<T> *src, Cx<T> *dst, hipStream_t st) {
    const size_t total = tile_stride();
    hipLaunchKernelGGL((k_to_tiles<T>), dim3((unsigned)std::min<size_t>((total + 255) / 256, 16384)), dim3(256), 0, st,
                       src, s.NCP, s.NC, s.FH, col_lg, total, dst, false);
    return hipGetLastError();
}
template <class T> hipError_t Fft2<T>::from_tiles(const Cx<T> *src, Cx<T> *dst, hipStream_t st) {
    const size_t total = tile_stride();
    hipLaunchKernelGGL((k_to_tiles<T>), dim3((unsigned)std::min<size_t>((total + 255) / 256, 16384)), dim3(256), 0, st,
                       src, s.NCP, s.NC, s.FH, col_lg, total, dst, true);
    return hipGetLastError();
}

template <class T> hipError_t Fft2<T>::spec_mul(const Cx<T> *a, const Cx<T> *psf_k, Cx<T> *out, hipStream_t st) {
    const unsigned g = (unsigned)std::min<size_t>(((size_t)s.FH * s.NC + 255) / 256, 16384);
    hipLaunchKernelGGL((k_spec<T, T, 0>), dim3(g), dim3(256), 0, st, a, psf_k, out, (Cx<T> *)nullptr, s.NCP, s.NC,
                       s.FH, col_lg);
    return hipGetLastError();
}
template <class T>
template <class TA>
hipError_t Fft2<T>::spec_acc(const Cx<T> *a, const Cx<T> *psf_k, Cx<TA> *acc, hipStream_t st) {
    const unsigned g = (unsigned)std::min<size_t>(((size_t)s.FH * s.NC + 255) / 256, 16384);
    hipLaunchKernelGGL((k_spec<T, TA, 1>), dim3(g), dim3(256), 0, st, a, psf_k, (Cx<T> *)nullptr, acc, s.NCP, s.NC,
                       s.FH, col_lg);
    return hipGetLastError();
}
template <class T>
template <class TA>
hipError_t Fft2<T>::spec_cast(const Cx<TA> *acc, Cx<T> *out, hipStream_t st) {
    const unsigned g = (unsigned)std::min<size_t>(((size_t)s.FH * s.NC + 255) / 256, 16384);
    hipLaunchKernelGGL((k_spec<T, TA, 2>), dim3(g), dim3(256), 0, st, (const Cx<T> *)nullptr, (const Cx<T> *)nullptr,
                       out, (Cx<TA> *)acc, s.NCP, s.NC, s.FH, col_lg);
    return hipGetLastError();
}

template <class T> void Fft2<T>::release() {
    if (tw_row) hipFree(tw_row);
    if (tw_col) hipFree(tw_col);
    if (post) hipFree(post);
    tw_row = tw_col = post = nullptr;
}

template <class T>
hipError_t Fft2<T>::row_r2c(int mode, const T *in, size_t in_bstride, int in_pitch, int win, const ExpDev *ex,
                            const T *wr, int W, Cx<T> *out, size_t out_bstride, int rows, int batch, hipStream_t st) {
    if (rows <= 0 || batch <= 0) return hipSuccess;
    const size_t lds = lds_elems(s.M) * sizeof(Cx<T>);
    dim3 grid(rows, batch);
    if (mode == ROW_PLAIN)
        hipLaunchKernelGGL((k_row_r2c<T, ROW_NT, ROW_VMAX, ROW_PLAIN>), grid, dim3(ROW_NT), lds, st, row, post, in,
                           in_bstride, in_pitch, win, ex, wr, W, out, out_bstride, s.NCP);
    else
        hipLaunchKernelGGL((k_row_r2c<T, ROW_NT, ROW_VMAX, ROW_GATHER>), grid, dim3(ROW_NT), lds, st, row, post, in,
                           in_bstride, in_pitch, win, ex, wr, W, out, out_bstride, s.NCP);
    return hipGetLastError();
}

template <class T>
hipError_t Fft2<T>::row_c2r(const Cx<T> *in, size_t in_bstride, T *out, size_t out_bstride, int out_pitch, int wout,
                            T scale, int rows, int batch, hipStream_t st) {
    if (rows <= 0 || batch <= 0) return hipSuccess;
    const size_t lds = lds_elems(s.M) * sizeof(Cx<T>);
    hipLaunchKernelGGL((k_row_c2r<T, ROW_NT, ROW_VMAX>), dim3(rows, batch), dim3(ROW_NT), lds, st, row, post, in,
                       in_bstride, s.NCP, out, out_bstride, out_pitch, wout, scale);
    return hipGetLastError();
}

template <class T>
template <class TA>
hipError_t Fft2<T>::colpass(int mode, const Cx<T> *in, size_t in_bstride, Cx<T> *out, size_t out_bstride, const Cx<T> *psf,
                        size_t psf_bstride, int hin, int hout, int batch, hipStream_t st) {
    constexpr int CNT = ColCfg<T>::NT, CVMAX = ColCfg<T>::VMAX, CXMAX = ColCfg<T>::XMAX;
    const int ntiles = (s.NC + col_lines - 1) / col_lines;
    const int gx = (ntiles + 7) / 8 * 8;
    const size_t lds = lds_elems((size_t)col_lines * s.FH) * sizeof(Cx<T>);
    hin = std::min(hin, s.FH);
    hout = std::min(hout, s.FH);
    switch (mode) {
        case COL_FWD:
            hipLaunchKernelGGL((k_col<T, T, CNT, CVMAX, CXMAX, COL_FWD>), dim3(gx, batch), dim3(CNT), lds, st,
                               col, s.NC, col_lg, s.NCP, hin, hout, in, in_bstride, out, out_bstride, psf,
                               psf_bstride, batch, ntiles);
            break;
        case COL_INV:
            hipLaunchKernelGGL((k_col<T, T, CNT, CVMAX, CXMAX, COL_INV>), dim3(gx, batch), dim3(CNT), lds, st,
                               col, s.NC, col_lg, s.NCP, hin, hout, in, in_bstride, out, out_bstride, psf,
                               psf_bstride, batch, ntiles);
            break;
        case COL_BCAST:
            hipLaunchKernelGGL((k_col<T, T, CNT, CVMAX, CXMAX, COL_BCAST>), dim3(gx, 1), dim3(CNT), lds, st,
                               col, s.NC, col_lg, s.NCP, hin, hout, in, in_bstride, out, out_bstride, psf,
                               psf_bstride, batch, ntiles);
            break;
        default:
            hipLaunchKernelGGL((k_col<T, TA, CNT, CVMAX, CXMAX, COL_ADJ>), dim3(gx, 1), dim3(CNT), lds, st,
                               col, s.NC, col_lg, s.NCP, hin, hout, in, in_bstride, out, out_bstride, psf,
                               psf_bstride, batch, ntiles);
            break;
    }
    return hipGetLastError();
}

template struct Fft2<float>;
template struct Fft2<double>;
template hipError_t Fft2<float>::spec_acc<float>(const Cx<float> *, const Cx<float> *, Cx<float> *, hipStream_t);
template hipError_t Fft2<float>::spec_acc<double>(const Cx<float> *, const Cx<float> *, Cx<double> *, hipStream_t);
template hipError_t Fft2<double>::spec_acc<double>(const Cx<double> *, const Cx<double> *, Cx<double> *, hipStream_t);
template hipError_t Fft2<float>::spec_cast<float>(const Cx<float> *, Cx<float> *, hipStream_t);
template hipError_t Fft2<float>::spec_cast<double>(const Cx<double> *, Cx<float> *, hipStream_t);
template hipError_t Fft2<double>::spec_cast<double>(const Cx<double> *, Cx<double> *, hipStream_t);
template hipError_t Fft2<float>::colpass<float>(int, const Cx<float> *, size_t, Cx<float> *, size_t, const Cx<float> *,
                                            size_t, int, int, int, hipStream_t);
template hipError_t Fft2<float>::colpass<double>(int, const Cx<float> *, size_t, Cx<float> *, size_t, const Cx<float> *,
                                             size_t, int, int, int, hipStream_t);
template hipError_t Fft2<double>::colpass<double>(int, const Cx<double> *, size_t, Cx<double> *, size_t,
                                              const Cx<double> *, size_t, int, int, int, hipStream_t);
template hipError_t Fft2<double>::colpass<float>(int, const Cx<double> *, size_t, Cx<double> *, size_t,
                                             const Cx<double> *, size_t, int, int, int, hipStream_t);

}  // namespace updc
