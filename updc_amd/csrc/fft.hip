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
#include <type_traits>

namespace updc {

// ------------------------------------------------------------------ kernels
// Row gather batching: RG_GB targets per lane per batch, the first RG_GU CSC
// entries of each loaded together, further entries one at a time.  A target
// pixel overlaps at most 4 pixels of a detector twice as coarse, so at that
// scale no entry takes the one-at-a-time path; 6 x 4 is the widest batch that
// fits the 128 VGPRs of a 1024-thread workgroup without spills.
static constexpr int RG_GB = 6, RG_GU = 4;
template <class T> constexpr int row_gather_gb() { return sizeof(T) == 4 ? RG_GB : 4; }  // FP64: the wv values take twice the VGPRs

// A real row's (x[2n], x[2n+1]) pair as one 8-byte (FP32) or 16-byte (FP64)
// access when `vec` (the host checked an even pitch and batch stride and an
// aligned base, so every pair is aligned), else two scalar accesses.
template <class T> struct Pair;
template <> struct Pair<float> { typedef float2 type; };
template <> struct Pair<double> { typedef double2 type; };
template <class T> __device__ __forceinline__ Cx<T> load_pair(const T *row, int c, int lim, bool vec) {
    if (!vec) return cx<T>(c < lim ? row[c] : (T)0, c + 1 < lim ? row[c + 1] : (T)0);
    if (c + 1 < lim) {
        const typename Pair<T>::type v = *reinterpret_cast<const typename Pair<T>::type *>(row + c);
        return cx<T>(v.x, v.y);
    }
    return cx<T>(c < lim ? row[c] : (T)0, (T)0);
}
template <class T> __device__ __forceinline__ void store_pair(T *row, int c, int lim, T a, T b, bool vec) {
    if (!vec) {
        if (c < lim) row[c] = a;
        if (c + 1 < lim) row[c + 1] = b;
        return;
    }
    if (c + 1 < lim) {
        typename Pair<T>::type v;
        v.x = a;
        v.y = b;
        *reinterpret_cast<typename Pair<T>::type *>(row + c) = v;
    } else if (c < lim) {
        row[c] = a;
    }
}

// SL_PAIR loads of the row C2R (SpecLayout in fft.hpp): a lane reads 16
// bytes, both columns of a pair (FP32) or one of them (FP64, two lanes per
// pair and row), so a tile's lanes cover whole pair runs.
template <class T> constexpr int pair_cpl() { return sizeof(T) == 4 ? 2 : 1; }  // complex values per lane
template <class T> constexpr int pair_lpp() { return 2 / pair_cpl<T>(); }       // lanes per pair and row
template <class T> __device__ __forceinline__ void load_cpl(const Cx<T> *p, Cx<T> *v) {
    if constexpr (sizeof(T) == 4) {
        const float4 q = *reinterpret_cast<const float4 *>(p);
        v[0] = cx<float>(q.x, q.y);
        v[1] = cx<float>(q.z, q.w);
    } else {
        v[0] = p[0];
    }
}

// Real split of RB transformed rows in LDS (row j at slot (n << LG) + j) into
// the half spectrum X[v] = (Z[v] + Z*[M-v])/2 - i/2 exp(-2 pi i v/FW)
// (Z[v] - Z*[M-v]), v in [0, M], stored column-major at rows r0 + j < rows.
template <class T, int NT, int RB>
__device__ __forceinline__ void split_store(const Cx<T> *z, const Cx<T> *__restrict__ post, Cx<T> *__restrict__ o,
                                            int M, int RP, int r0, int rows) {
    constexpr int LG = RB == 1 ? 0 : RB == 2 ? 1 : RB == 4 ? 2 : 3;
    const T half = (T)0.5;
    for (int idx = threadIdx.x; idx < ((M + 1) << LG); idx += NT) {
        const int v = idx >> LG, j = idx & (RB - 1);
        if (r0 + j >= rows) continue;
        const Cx<T> zv = z[((v == M ? 0 : v) << LG) + j];
        const Cx<T> zm = cconj(z[((v == 0 ? 0 : M - v) << LG) + j]);
        const Cx<T> fe = cscale(zv + zm, half);
        const Cx<T> d = zv - zm;
        const Cx<T> fo = cx<T>(d.im * half, -d.re * half);  // (zv - zm) / (2i)
        o[(size_t)v * RP + r0 + j] = fe + cmul(post[v], fo);
    }
}

// Minimum waves per SIMD requested of the row kernels.  Their 1024-thread
// workgroups (one per CU: 89.6 KB of LDS at M = 2800) already hold them to
// four waves per SIMD and 128 VGPRs; the request adds no further limit.
template <class T, int NT> constexpr int row_min_waves() { return sizeof(T) == 4 ? 2048 / NT : 1024 / NT; }

// Row R2C: RB consecutive rows per workgroup (interleaved in LDS, row j at
// slot (n << LG) + j), one half-length complex FFT per row, then the real
// split written column-major, so RB rows of one column land contiguously.
template <class T, int NT, int VMAX, int RB, int MODE>
__global__ __launch_bounds__(NT, (row_min_waves<T, NT>())) void k_row_r2c(FftLen f, const Cx<T> *__restrict__ post, const T *__restrict__ in,
                                                size_t in_bstride, int in_pitch, int win,
                                                const ExpDev *__restrict__ ex, const T *__restrict__ wr, int W,
                                                Cx<T> *__restrict__ out, size_t out_bstride, int RP, int rows, int vec) {
    constexpr int LG = RB == 1 ? 0 : RB == 2 ? 1 : RB == 4 ? 2 : 3;
    static_assert((1 << LG) == RB, "RB must be 1, 2, 4 or 8");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cx<T> *z = reinterpret_cast<Cx<T> *>(smem_raw);
    // XCD-aware: the row blocks sharing a spectrum column's 128-byte lines run
    // on one XCD, back to back, so its L2 serves them all
    const int rblk = xcd_tile(blockIdx.x, (rows + RB - 1) / RB);
    if (rblk < 0) return;
    const int r0 = rblk * RB, b = blockIdx.y, M = f.L;
    if (MODE == ROW_PLAIN) {
#pragma unroll
        for (int j = 0; j < RB; ++j) {
            const int r = r0 + j;
            const T *row = in + (size_t)b * in_bstride + (size_t)r * in_pitch;
            for (int n = threadIdx.x; n < M; n += NT) {
                const int c = 2 * n;
                z[(n << LG) + j] = r < rows ? load_pair(row, c, win, vec != 0) : cx<T>((T)0, (T)0);
            }
        }
    } else {
        // Target column c of row j is the real (c even) or imaginary part of
        // complex slot ((c/2) << LG) + j.  Lanes take consecutive targets; a
        // batch of GB targets per lane issues its offset loads, then its
        // entry loads, then its w loads together (three round trips per batch).
        constexpr int GB = row_gather_gb<T>();
        const ExpDev &e = ex[b];
        const T *w = wr + e.pix;
        T *zf = reinterpret_cast<T *>(z);
        const int wc = min(W, 2 * M);
        for (int j = 0; j < RB; ++j)
            for (int c = wc + threadIdx.x; c < 2 * M; c += NT) zf[2 * (((c >> 1) << LG) + j) + (c & 1)] = (T)0;
        const int per_row = (wc + NT - 1) / NT, total = RB * per_row;
        for (int u0 = 0; u0 < total; u0 += GB) {
            uint32_t a[GB], cnt[GB];
            int slot[GB];
#pragma unroll
            for (int g = 0; g < GB; ++g) {
                const int u = u0 + g, j = u / per_row, c = threadIdx.x + (u - j * per_row) * NT, r = r0 + j;
                slot[g] = (u < total && c < wc) ? 2 * (((c >> 1) << LG) + j) + (c & 1) : -1;
                const bool ok = slot[g] >= 0 && r < rows;
                const uint32_t t = (uint32_t)r * (uint32_t)W + (uint32_t)c;
                a[g] = ok ? e.toff[t] : 0u;
                cnt[g] = ok ? e.toff[t + 1] : 0u;
            }
            uint2 en[GB][RG_GU];
            T wv[GB][RG_GU];
#pragma unroll
            for (int g = 0; g < GB; ++g) {
                cnt[g] -= a[g];
#pragma unroll
                for (int k = 0; k < RG_GU; ++k) en[g][k] = (uint32_t)k < cnt[g] ? e.tent[a[g] + k] : make_uint2(0u, 0u);
            }
#pragma unroll
            for (int g = 0; g < GB; ++g)
#pragma unroll
                for (int k = 0; k < RG_GU; ++k) wv[g][k] = (uint32_t)k < cnt[g] ? w[en[g][k].x] : (T)0;
#pragma unroll
            for (int g = 0; g < GB; ++g) {
                // scatter[t] = sum over source pixels p (ascending) of fma(frac, w[p], .)
                // from 0 -- the fmaf chain of updc_core.c:1173-1176 / :1064, since
                // the reference visits pixels in ascending p and each (p,t) pair once
                T acc = (T)0;
#pragma unroll
                for (int k = 0; k < RG_GU; ++k)
                    if ((uint32_t)k < cnt[g]) acc = fmat((T)__uint_as_float(en[g][k].y), wv[g][k], acc);
                for (uint32_t k = RG_GU; k < cnt[g]; ++k) {
                    const uint2 x = e.tent[a[g] + k];
                    acc = fmat((T)__uint_as_float(x.y), w[x.x], acc);
                }
                if (slot[g] >= 0) zf[slot[g]] = acc;
            }
        }
    }
    __syncthreads();
    fft_lds<false, T, NT, VMAX>(z, f, LG);  // the RB rows together
    split_store<T, NT, RB>(z, post, out + (size_t)b * out_bstride, M, RP, r0, rows);
}

// Row C2R: RB rows per workgroup.  The RB-element column runs of the
// spectrum (v = 0..M) are staged in LDS once, merged in place into
// half-length complex rows (one thread forms both members of a (v, M-v) pair,
// so the merge needs no second buffer), transformed together, and written
// cropped and scaled, row-major.
template <class T, int NT, int VMAX, int RB>
__global__ __launch_bounds__(NT, (row_min_waves<T, NT>())) void k_row_c2r(FftLen f, const Cx<T> *__restrict__ post, const Cx<T> *__restrict__ in,
                                                size_t in_bstride, int RP, T *__restrict__ out, size_t out_bstride,
                                                int out_pitch, int wout, T scale, int rows, int vec) {
    constexpr int LG = RB == 1 ? 0 : RB == 2 ? 1 : RB == 4 ? 2 : 3;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cx<T> *z = reinterpret_cast<Cx<T> *>(smem_raw);  // (M + 1) * RB slots
    // XCD-aware: the row blocks sharing a spectrum column's 128-byte lines run
    // on one XCD, back to back, so its L2 serves them all
    const int rblk = xcd_tile(blockIdx.x, (rows + RB - 1) / RB);
    if (rblk < 0) return;
    const int r0 = rblk * RB, b = blockIdx.y, M = f.L;
    const Cx<T> *y = in + (size_t)b * in_bstride;
    for (int idx = threadIdx.x; idx < ((M + 1) << LG); idx += NT) {
        const int v = idx >> LG, j = idx & (RB - 1);
        z[idx] = r0 + j < rows ? y[(size_t)v * RP + r0 + j] : cx<T>((T)0, (T)0);
    }
    __syncthreads();
    // Merge: C[v] = E[v] + i O[v], E = Y[v] + Y*[M-v], O = (Y[v] - Y*[M-v]) exp(+2 pi i v/FW).
    for (int idx = threadIdx.x; idx < ((M / 2 + 1) << LG); idx += NT) {
        const int v = idx >> LG, j = idx & (RB - 1);
        const int u = M - v;  // partner
        const Cx<T> yv = z[(v << LG) + j], yu = z[(u << LG) + j];
        const Cx<T> e1 = yv + cconj(yu), o1 = cmul(yv - cconj(yu), cconj(post[v]));
        const Cx<T> cv = cx<T>(e1.re - o1.im, e1.im + o1.re);
        if (v != 0 && u != v) {
            const Cx<T> e2 = yu + cconj(yv), o2 = cmul(yu - cconj(yv), cconj(post[u]));
            z[(u << LG) + j] = cx<T>(e2.re - o2.im, e2.im + o2.re);
        }
        z[(v << LG) + j] = cv;
    }
    __syncthreads();
    fft_lds<true, T, NT, VMAX>(z, f, LG);  // the RB rows together
#pragma unroll
    for (int j = 0; j < RB; ++j) {
        if (r0 + j >= rows) continue;
        T *orow = out + (size_t)b * out_bstride + (size_t)(r0 + j) * out_pitch;
        for (int n = threadIdx.x; n < M; n += NT) {
            const int c = 2 * n;
            const Cx<T> v = z[(n << LG) + j];
            store_pair(orow, c, wout, v.re * scale, v.im * scale, vec != 0);
        }
    }
}

// ---- persistent, software-pipelined row passes ------------------------------
// One workgroup per CU (the LDS holds one tile of RB rows) walks a range of
// tiles; tile t = (exposure t / nblk, row block t % nblk).  XCD x owns the
// contiguous tiles [x * tper, (x + 1) * tper), taken round robin by its
// gridDim.x / 8 workgroups, so the row blocks sharing a spectrum column's
// 128-byte lines run on one XCD at the same time.  The next tile's input is
// loaded into registers while the current tile is transformed, which hides
// the load latency that a lone workgroup per CU cannot otherwise cover.
__device__ __forceinline__ void row_tiles(int ntiles, int &t, int &t_end, int &step) {
    const int xcd = blockIdx.x & 7, l = blockIdx.x >> 3;
    step = gridDim.x >> 3;
    const int tper = (ntiles + 7) >> 3;
    t = xcd * tper + l;
    t_end = min(ntiles, (xcd + 1) * tper);
}

// The persistent C2R workgroups copy the stage twiddles, and the split roots
// where PL, into LDS behind the tile, once.  Read from global memory, each
// table load would make its wave wait for the tile prefetch in flight as well
// (vector memory loads retire in order), and the prefetch would overlap
// little (0.71 -> 0.64 ms for 8 exposures of 4096 rows on a 5600-point grid).
// The persistent R2C keeps its tables in global memory: there the extra LDS
// reads cost more than the overlap gains (0.46 -> 0.64 ms measured).  FP64
// rows leave the roots in global memory (no LDS room).
template <class T, int NT, bool PL>
__device__ __forceinline__ void row_tables_lds(const FftLen &f, const Cx<T> *__restrict__ post, int ntw,
                                               Cx<T> *lds, const Cx<T> *&twl, const Cx<T> *&pl) {
    const Cx<T> *tw = (const Cx<T> *)f.tw;
    for (int i = threadIdx.x; i < ntw; i += NT) lds[i] = tw[i];
    twl = lds;
    if constexpr (PL) {
        Cx<T> *pp = lds + ntw;
        for (int i = threadIdx.x; i <= f.L; i += NT) pp[i] = post[i];
        pl = pp;
    } else {
        pl = post;
    }
}
template <class T> struct RowPL;
template <> struct RowPL<float> { static constexpr bool value = true; };
template <> struct RowPL<double> { static constexpr bool value = false; };

// C2R tiles of a persistent workgroup (as k_row_c2r).  PFM complex values per
// thread cover (M + 1) * RB slots (SL_COL) or (M + 2) / 2 pairs of RB rows
// (CP: SL_PAIR spectra).  emit(b, r0, z) consumes each transformed
// tile from LDS (z[(n << LG) + j] = columns 2n, 2n+1 of row r0 + j, unscaled).
template <class T, int NT, int VMAX, int RB, int PFM, bool CP, class Emit>
__device__ __forceinline__ void row_c2r_tiles(const FftLen &f, const Cx<T> *twl, const Cx<T> *__restrict__ post,
                                              const Cx<T> *__restrict__ in, size_t in_bstride, int RP, int rows,
                                              int batch, Cx<T> *z, Emit &emit) {
    constexpr int LG = RB == 1 ? 0 : RB == 2 ? 1 : RB == 4 ? 2 : 3;
    // CP (SL_PAIR spectra): a slot is CPL complex values of one pair and row
    constexpr int CPL = CP ? pair_cpl<T>() : 1, LPP = CP ? pair_lpp<T>() : 1, NS = PFM / CPL;
    const int M = f.L, nblk = (rows + RB - 1) / RB, nsl = (M + 1) << LG;
    int t, t_end, step;
    row_tiles(nblk * batch, t, t_end, step);
    Cx<T> pre[PFM];
    // !CP: slot idx = tid + m NT holds column v = idx / RB, row j = idx % RB; NT
    // is a multiple of RB, so j is fixed per thread and v advances by NT / RB
    // per slot.  CP: slot idx holds lane part h = idx % LPP of pair
    // p = idx / (LPP RB), row j = (idx / LPP) % RB; h and j are fixed per
    // thread and p advances by NT / (LPP RB) per slot.
    const int jt = CP ? (threadIdx.x / LPP) & (RB - 1) : threadIdx.x & (RB - 1);
    const int ht = threadIdx.x % LPP, ut = threadIdx.x / (LPP * RB);
    constexpr int USTEP = NT / (LPP * RB);
    const int np = (M + 2) >> 1;  // CP: pairs per row
    const size_t vstep = (size_t)(NT >> LG) * RP, pstep = (size_t)USTEP * RP * 2;
    // prefetch slots [m0, m1) of tile tt
    auto load = [&](int tt, int m0, int m1) {
        const int b = tt / nblk, r0 = (tt - b * nblk) * RB;
        const bool rok = r0 + jt < rows;
        if constexpr (CP) {
            const Cx<T> *p0 = in + (size_t)b * in_bstride + (size_t)ut * RP * 2 + (size_t)(r0 + jt) * 2 + ht * CPL;
#pragma unroll
            for (int m = 0; m < NS; ++m) {
                if (m < m0 || m >= m1) continue;
                if (ut + m * USTEP < np && rok) {
                    load_cpl(p0 + m * pstep, pre + m * CPL);
                } else {
#pragma unroll
                    for (int c = 0; c < CPL; ++c) pre[m * CPL + c] = cx<T>((T)0, (T)0);
                }
            }
        } else {
            const Cx<T> *p0 = in + (size_t)b * in_bstride + (size_t)(threadIdx.x >> LG) * RP + r0 + jt;
#pragma unroll
            for (int m = 0; m < PFM; ++m) {
                if (m < m0 || m >= m1) continue;
                const int idx = threadIdx.x + m * NT;
                pre[m] = (idx < nsl && rok) ? p0[m * vstep] : cx<T>((T)0, (T)0);
            }
        }
    };
    // The next tile's prefetch is issued in thirds: before the merge, after
    // it, and after the first transform stage.  Issued at once, the loads
    // (32-byte pieces of 2801 column runs per tile) back up the wave's memory
    // queue and stall it at the issue: 0.65 -> 0.58 ms for 8 x 4096 rows at
    // M = 2800, against 0.62 in halves and 0.63-0.67 in four or six parts.
    // FP64 tiles issue theirs at once: split, the FP64 update kernel spills.
    constexpr int P3 = sizeof(T) == 4 ? (NS + 2) / 3 : NS;
    if (t < t_end) load(t, 0, NS);
    for (; t < t_end; t += step) {
        const int b = t / nblk, r0 = (t - b * nblk) * RB;
        const bool more = t + step < t_end;
        __syncthreads();  // the previous tile's LDS reads are done
        // The prefetched tile has landed (vmcnt(0)) on every path.  The
        // compiler's own waits below sit in lane-masked branches that a wave
        // may skip, after which it would wait again before reusing a
        // register: later, with the next tile's first third in flight.
        __builtin_amdgcn_s_waitcnt(0x0F70);
        if constexpr (CP) {
#pragma unroll
            for (int m = 0; m < NS; ++m) {
                const int u = ut + m * USTEP;
                if (u >= np) continue;
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const int v = 2 * u + ht * CPL + c;
                    if (v <= M) z[(v << LG) + jt] = pre[m * CPL + c];
                }
            }
        } else {
#pragma unroll
            for (int m = 0; m < PFM; ++m) {
                const int idx = threadIdx.x + m * NT;
                if (idx < nsl) z[idx] = pre[m];
            }
        }
        if (more) load(t + step, 0, P3);
        __syncthreads();
        // Merge: C[v] = E[v] + i O[v], E = Y[v] + Y*[M-v], O = (Y[v] - Y*[M-v]) exp(+2 pi i v/FW).
        for (int idx = threadIdx.x; idx < ((M / 2 + 1) << LG); idx += NT) {
            const int v = idx >> LG, j = idx & (RB - 1);
            const int u = M - v;
            const Cx<T> yv = z[(v << LG) + j], yu = z[(u << LG) + j];
            const Cx<T> e1 = yv + cconj(yu), o1 = cmul(yv - cconj(yu), cconj(post[v]));
            const Cx<T> cv = cx<T>(e1.re - o1.im, e1.im + o1.re);
            if (v != 0 && u != v) {
                const Cx<T> e2 = yu + cconj(yv), o2 = cmul(yu - cconj(yv), cconj(post[u]));
                z[(u << LG) + j] = cx<T>(e2.re - o2.im, e2.im + o2.re);
            }
            z[(v << LG) + j] = cv;
        }
        __syncthreads();
        if (more) load(t + step, P3, 2 * P3);
        auto third = [&](int s) __attribute__((always_inline)) {
            if (s == 0 && more) load(t + step, 2 * P3, NS);
        };
        fft_lds_hook<true, T, NT, VMAX>(z, f, twl, LG, -1, third);
        emit(b, r0, z);
    }
}

template <class T, int NT, int VMAX, int RB, int PFM, bool CP>
__global__ __launch_bounds__(NT, (row_min_waves<T, NT>())) void k_row_c2r_p(
    FftLen f, const Cx<T> *__restrict__ post, int ntw, const Cx<T> *__restrict__ in, size_t in_bstride, int RP,
    T *__restrict__ out, size_t out_bstride, int out_pitch, int wout, T scale, int rows, int batch, int vec) {
    constexpr int LG = RB == 1 ? 0 : RB == 2 ? 1 : RB == 4 ? 2 : 3;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cx<T> *z = reinterpret_cast<Cx<T> *>(smem_raw);
    const int M = f.L;
    const Cx<T> *twl, *pl;
    row_tables_lds<T, NT, RowPL<T>::value>(f, post, ntw, z + ((M + 1) << LG), twl, pl);
    auto emit = [&](int b, int r0, const Cx<T> *zz) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < RB; ++j) {
            if (r0 + j >= rows) continue;
            T *orow = out + (size_t)b * out_bstride + (size_t)(r0 + j) * out_pitch;
            for (int n = threadIdx.x; n < M; n += NT) {
                const int c = 2 * n;
                const Cx<T> v = zz[(n << LG) + j];
                store_pair(orow, c, wout, v.re * scale, v.im * scale, vec != 0);
            }
        }
    };
    row_c2r_tiles<T, NT, VMAX, RB, PFM, CP>(f, twl, pl, in, in_bstride, RP, rows, batch, z, emit);
}

// Persistent C2R of the summed adjoint spectrum (batch 1) with the update
// x <- x * max(C/D, 0) (as k_update in stages.hip: the same guards and the
// same operations on each element, so x is bit-identical) applied where C
// leaves the transform, instead of storing C and reading it back.  The
// signed sum and the negative count: per-thread sums over the thread's
// elements, then per workgroup, then the last workgroup to finish (ticket)
// adds the workgroup partials in workgroup order.
template <class T, class TF, int NT, int VMAX, int RB, int PFM, bool CP>
__global__ __launch_bounds__(NT, (row_min_waves<T, NT>())) void k_row_c2r_upd_p(
    FftLen f, const Cx<T> *__restrict__ post, int ntw, const Cx<T> *__restrict__ in, int RP, TF *__restrict__ x,
    const T *__restrict__ D, int W, T scale, T max_sens, int rows, IterDiag *idiag, double *partials,
    unsigned long long *negs, unsigned int *ticket) {
    constexpr int LG = RB == 1 ? 0 : RB == 2 ? 1 : RB == 4 ? 2 : 3;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cx<T> *z = reinterpret_cast<Cx<T> *>(smem_raw);
    const int M = f.L;
    const Cx<T> *twl, *pl;
    row_tables_lds<T, NT, RowPL<T>::value>(f, post, ntw, z + ((M + 1) << LG), twl, pl);
    const T thr = sizeof(T) == 4 ? (T)1e-8f : (T)1e-8;
    const T neg = sizeof(T) == 4 ? (T)(-1.0e-6f) : (T)(-1.0e-6);
    double sum = 0.0;
    unsigned long long nneg = 0, badkey = ~0ull;
    unsigned int bad = 0;
    // the detector stage of this iteration, or an earlier iteration, failed:
    // the reference never reaches this update, so x stays as it is
    const bool hold = idiag->halt || (idiag->err & (UPDC_E_MU | UPDC_E_RATIO));
    auto upd = [&](size_t t, T c, T d, TF xv) __attribute__((always_inline)) {
        unsigned int b = 0;
        if (d > thr) {
            T fct = c / d;
            if (!isfinite(fct)) b = b ? b : UPDC_E_FACTOR;
            if (fct < (T)0) {
                if (c < neg * max_sens) b = b ? b : UPDC_E_NEGCORR;
                fct = (T)0;
            }
            xv = (TF)((T)xv * fct);
        } else {
            xv = (TF)0;
        }
        if (!isfinite(xv)) b = b ? b : UPDC_E_ESTIMATE;
        if (b) {
            bad |= b;
            const unsigned long long key =
                err_key_update(t, b == UPDC_E_FACTOR ? 0u : b == UPDC_E_NEGCORR ? 1u : 2u);
            badkey = badkey < key ? badkey : key;
        }
        x[t] = xv;
        sum += (double)xv;
        nneg += xv < (TF)0 ? 1 : 0;
    };
    // The D and x values of a column step are loaded before its updates: one
    // update's store would otherwise hold up the next element's loads (a
    // wave's vector memory counter drains in issue order).
    auto emit = [&](int, int r0, const Cx<T> *zz) __attribute__((always_inline)) {
        if (hold) return;
#pragma unroll
        for (int j = 0; j < RB; ++j) {
            if (r0 + j >= rows) continue;
            const size_t rowt = (size_t)(r0 + j) * W;
            for (int n0 = threadIdx.x; n0 < M; n0 += 2 * NT) {
                T dv[2][2];
                TF xv[2][2];
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        // clamped index: every load in bounds, unused values discarded
                        const size_t tt = rowt + min(2 * (n0 + u * NT) + h, W - 1);
                        dv[u][h] = D[tt];
                        xv[u][h] = x[tt];
                    }
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int n = n0 + u * NT, c = 2 * n;
                    if (n >= M) break;
                    const Cx<T> v = zz[(n << LG) + j];
                    if (c < W) upd(rowt + c, v.re * scale, dv[u][0], xv[u][0]);
                    if (c + 1 < W) upd(rowt + c + 1, v.im * scale, dv[u][1], xv[u][1]);
                }
            }
        }
    };
    row_c2r_tiles<T, NT, VMAX, RB, PFM, CP>(f, twl, pl, in, 0, RP, rows, 1, z, emit);

    if (bad) {
        atomicOr(&idiag->err, bad);
        atomicMin(&idiag->err_key, badkey);
    }
    // workgroup reduction in LDS (the transform is done with z)
    __syncthreads();
    double *sh = reinterpret_cast<double *>(smem_raw);
    unsigned long long *shn = reinterpret_cast<unsigned long long *>(sh + NT);
    __shared__ bool last;
    sh[threadIdx.x] = sum;
    shn[threadIdx.x] = nneg;
    __syncthreads();
    for (int o = NT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sh[threadIdx.x] += sh[threadIdx.x + o];
            shn[threadIdx.x] += shn[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = sh[0];
        negs[blockIdx.x] = shn[0];
        __threadfence();
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (last) {
        __threadfence();
        double s2 = 0.0;
        unsigned long long n2 = 0;
        for (int i = threadIdx.x; i < (int)gridDim.x; i += NT) {
            s2 += __builtin_nontemporal_load(&partials[i]);
            n2 += __builtin_nontemporal_load(&negs[i]);
        }
        sh[threadIdx.x] = s2;
        shn[threadIdx.x] = n2;
        __syncthreads();
        for (int o = NT / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) {
                sh[threadIdx.x] += sh[threadIdx.x + o];
                shn[threadIdx.x] += shn[threadIdx.x + o];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            idiag->signed_sum = sh[0];
            idiag->n_negative = shn[0];
            *ticket = 0u;
        }
    }
}

// R2C of real rows (as k_row_r2c<..., ROW_PLAIN>).  PFR (x[2n], x[2n+1]) pairs
// per thread and row cover M pairs.  VEC: rows whose pairs are 8-byte
// aligned (pairs_aligned), read with one load per pair; a kernel that also
// held the unaligned form's code spilled (FP32, 12 bytes).
template <class T, int NT, int VMAX, int RB, int PFR, bool VEC>
__global__ __launch_bounds__(NT, (row_min_waves<T, NT>())) void k_row_r2c_p(
    FftLen f, const Cx<T> *__restrict__ post, const T *__restrict__ in, size_t in_bstride, int in_pitch, int win,
    Cx<T> *__restrict__ out, size_t out_bstride, int RP, int rows, int batch) {
    constexpr int LG = RB == 1 ? 0 : RB == 2 ? 1 : RB == 4 ? 2 : 3;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cx<T> *z = reinterpret_cast<Cx<T> *>(smem_raw);
    const int M = f.L, nblk = (rows + RB - 1) / RB;
    int t, t_end, step;
    row_tiles(nblk * batch, t, t_end, step);
    Cx<T> pre[RB][PFR];
    auto load = [&](int tt) {
        const int b = tt / nblk, r0 = (tt - b * nblk) * RB;
#pragma unroll
        for (int j = 0; j < RB; ++j) {
            const int r = r0 + j;
            const T *row = in + (size_t)b * in_bstride + (size_t)r * in_pitch;
#pragma unroll
            for (int m = 0; m < PFR; ++m) {
                const int n = threadIdx.x + m * NT, c = 2 * n;
                const bool ok = n < M && r < rows;
                pre[j][m] = ok ? load_pair(row, c, win, VEC) : cx<T>((T)0, (T)0);
            }
        }
    };
    if (t < t_end) load(t);
    for (; t < t_end; t += step) {
        const int b = t / nblk, r0 = (t - b * nblk) * RB;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RB; ++j)
#pragma unroll
            for (int m = 0; m < PFR; ++m) {
                const int n = threadIdx.x + m * NT;
                if (n < M) z[(n << LG) + j] = pre[j][m];
            }
        if (t + step < t_end) load(t + step);
        __syncthreads();
        fft_lds<false, T, NT, VMAX>(z, f, LG);
        split_store<T, NT, RB>(z, post, out + (size_t)b * out_bstride, M, RP, r0, rows);
    }
}

// Column pass: one spectrum column per workgroup, read and written as one
// contiguous column-major run.  BCAST keeps the X^ column in registers and
// loops over all exposures; ADJ keeps the running sum over exposures in
// registers; neither X^ nor the summed adjoint spectrum touches HBM.
template <class T, class TA, int NT, int VMAX, int XMAX, int MODE>
__global__ __launch_bounds__(NT) void k_col(FftLen f, int nc, const Cx<T> *__restrict__ in, size_t in_bs, int in_rp,
                                            Cx<T> *__restrict__ out, size_t out_bs, int out_rp,
                                            const Cx<T> *__restrict__ psf, size_t psf_bs, int psf_rp, int hin,
                                            int hout, int nexp, FftLen fa) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cx<T> *buf = reinterpret_cast<Cx<T> *>(smem_raw);
    const int col = xcd_tile(blockIdx.x, nc);
    if (col < 0) return;
    const int FH = f.L;
    const int tid = threadIdx.x;

    auto load = [&](const Cx<T> *src) {
        const Cx<T> *c = src + (size_t)col * in_rp;
        for (int n = tid; n < FH; n += NT) buf[n] = n < hin ? c[n] : cx<T>((T)0, (T)0);
        __syncthreads();
    };
    auto store = [&](Cx<T> *dst) {
        Cx<T> *c = dst + (size_t)col * out_rp;
        for (int n = tid; n < hout; n += NT) c[n] = buf[n];
    };

    if (MODE == COL_FWD || MODE == COL_INV) {
        const int b = blockIdx.y;
        load(in + (size_t)b * in_bs);
        fft_lds<MODE == COL_INV, T, NT, VMAX>(buf, f, 0);
        store(out + (size_t)b * out_bs);
    } else if (MODE == COL_BCAST) {
        load(in);
        fft_lds<false, T, NT, VMAX>(buf, f, 0);
        Cx<T> xr[XMAX];
#pragma unroll
        for (int j = 0; j < XMAX; ++j) {
            const int e = tid + j * NT;
            xr[j] = e < FH ? buf[e] : cx<T>((T)0, (T)0);
        }
        for (int k = 0; k < nexp; ++k) {
            __syncthreads();
            const Cx<T> *pk = psf + (size_t)k * psf_bs + (size_t)col * psf_rp;
#pragma unroll
            for (int j = 0; j < XMAX; ++j) {
                const int e = tid + j * NT;
                // (a+ib)(c+id) as a*c-b*d, a*d+b*c (updc_core.c:756-758, adjoint=0)
                if (e < FH) buf[e] = cmul(xr[j], pk[e]);
            }
            __syncthreads();
            fft_lds<true, T, NT, VMAX>(buf, f, 0);
            store(out + (size_t)k * out_bs);
        }
    } else {  // COL_ADJ, COL_ACC
        Cx<TA> acc[XMAX];
        // (COL_ACC with nexp < 0: the summed spectrum from memory, TA at pitch in_rp)
        const Cx<TA> *ia = reinterpret_cast<const Cx<TA> *>(in) + (size_t)col * in_rp;
#pragma unroll
        for (int j = 0; j < XMAX; ++j) {
            const int e = tid + j * NT;
            acc[j] = (MODE == COL_ACC && nexp < 0 && e < hin) ? ia[e] : cx<TA>((TA)0, (TA)0);
        }
        for (int k = 0; k < nexp; ++k) {
            __syncthreads();
            load(in + (size_t)k * in_bs);
            fft_lds<false, T, NT, VMAX>(buf, f, 0);
            const Cx<T> *pk = psf + (size_t)k * psf_bs + (size_t)col * psf_rp;
#pragma unroll
            for (int j = 0; j < XMAX; ++j) {
                const int e = tid + j * NT;
                if (e < FH) {
                    const Cx<T> p = pk[e];
                    const Cx<T> s = buf[e];
                    // re += a*c + b*d ; im += b*c - a*d  (updc_core.c:991-993)
                    const T pre = s.re * p.re + s.im * p.im;
                    const T pim = s.im * p.re - s.re * p.im;
                    acc[j].re += (TA)pre;
                    acc[j].im += (TA)pim;
                }
            }
        }
        if (MODE == COL_ACC) {
            Cx<TA> *o = reinterpret_cast<Cx<TA> *>(out) + (size_t)col * out_rp;
            if (fa.L > 0) {
                // the inverse of the sum in TA (plan fa, TA twiddles) over the
                // whole LDS allocation, FH TA values
                Cx<TA> *bufa = reinterpret_cast<Cx<TA> *>(smem_raw);
                __syncthreads();
#pragma unroll
                for (int j = 0; j < XMAX; ++j) {
                    const int e = tid + j * NT;
                    if (e < FH) bufa[e] = acc[j];
                }
                __syncthreads();
                fft_lds<true, TA, NT, VMAX>(bufa, fa, 0);
                for (int n = tid; n < hout; n += NT) o[n] = bufa[n];
                return;
            }
            // the sum itself, in TA, for an inverse transform in TA
#pragma unroll
            for (int j = 0; j < XMAX; ++j) {
                const int e = tid + j * NT;
                if (e < hout) o[e] = acc[j];
            }
            return;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < XMAX; ++j) {
            const int e = tid + j * NT;
            if (e < FH) buf[e] = cx<T>((T)acc[j].re, (T)acc[j].im);
        }
        __syncthreads();
        fft_lds<true, T, NT, VMAX>(buf, f, 0);
        store(out);
    }
}

// Column pass with register-resident edge stages, for column plans whose
// first and last radix are both RE (LR = FH / RE).  Element i + q LR (i < LR)
// of a column is value q of butterfly i: the positions the first Stockham
// stage reads and the last one writes.  Thread tid holds butterflies
// tid + r NT.  The column loads therefore feed the first stage directly, the
// last stage leaves the natural-order spectrum in registers, where the PSF
// product and the stores use the same coalesced pattern, and an inverse
// transform that follows starts from those registers -- two LDS passes and
// their barriers fewer per transform than the LDS-only kernel.  The
// butterflies, twiddles and their order are those of the LDS-only kernel.
// PF: the next PSF column (BCAST) or the current one (ADJ) is loaded into
// registers before the transform that precedes its use.  The accumulate pass
// (COL_ACC) loads its PSF column after the transform: held through it, the
// column's registers and the FP64 sum made the pass spill once the
// transforms' products became fused multiply-adds (0.66 -> 0.83 ms at 5600
// rows; issued between the middle stages and the last edge stage, it still
// spills), and without them it runs faster than with the prefetch before
// (0.625 ms).
// EV: complex values per thread per register set (the edge butterflies of
// one thread, EV / RE of them).
// PP: the LDS holds two transform buffers (2 FH values).  The middle stages
// then run out of place, one barrier each instead of two, and the last edge
// stage reads one buffer while the next transform's first edge stage writes
// the other, so it needs no barrier of its own.  (The stage twiddles stay in
// global memory: read from LDS, with a butterfly's R-1 twiddles adjacent,
// they add bank-conflicted LDS traffic to passes that LDS bounds already, and
// the broadcast and accumulate passes slowed by 10 %.)
template <class T, class TA, int NT, int VMAX, int RE, int MODE, bool PF, int EV, bool PP, bool CP>
__global__ __launch_bounds__(NT) void k_col_e(FftLen f, int nc, const Cx<T> *__restrict__ in, size_t in_bs, int in_rp,
                                              Cx<T> *__restrict__ out, size_t out_bs, int out_rp,
                                              const Cx<T> *__restrict__ psf, size_t psf_bs, int psf_rp, int hin,
                                              int hout, int nexp, FftLen fa) {
    constexpr int NBE = EV / RE;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cx<T> *buf = reinterpret_cast<Cx<T> *>(smem_raw);
    // CP (output in SL_PAIR): the two columns of a pair run on one XCD, back
    // to back, so its L2 merges their halves of every line they write
    const int col = CP ? xcd_tile_even(blockIdx.x, nc) : xcd_tile(blockIdx.x, nc);
    if (col < 0) return;
    // column col of an output spectrum with pitch rpitch: its first element;
    // el(c, n): its element n
    auto cbase = [&](auto *p, int rpitch) __attribute__((always_inline)) {
        return CP ? p + (size_t)(col >> 1) * rpitch * 2 + (col & 1) : p + (size_t)col * rpitch;
    };
    auto el = [](auto *c, int n) __attribute__((always_inline)) -> auto & { return CP ? c[(size_t)n * 2] : c[n]; };
    const int LR = f.L / RE, ns = f.nst, tid = threadIdx.x;
    const Cx<T> *twl = (const Cx<T> *)f.tw + f.to[ns - 1];
    Cx<T> rb[NBE][RE];
    // ping-pong state: the buffer the next first edge stage writes
    Cx<T> *bx = buf, *by = buf + f.L;
    const Cx<T> *twb = (const Cx<T> *)f.tw;

    auto load = [&](Cx<T>(&d)[NBE][RE], const Cx<T> *c, int h) {
#pragma unroll
        for (int r = 0; r < NBE; ++r)
#pragma unroll
            for (int q = 0; q < RE; ++q) {
                const int i = tid + r * NT, n = i + q * LR;
                d[r][q] = (i < LR && n < h) ? c[n] : cx<T>((T)0, (T)0);
            }
    };
    auto store = [&](Cx<T> *c) {
#pragma unroll
        for (int r = 0; r < NBE; ++r)
#pragma unroll
            for (int q = 0; q < RE; ++q) {
                const int i = tid + r * NT, n = i + q * LR;
                if (i < LR && n < hout) el(c, n) = rb[r][q];
            }
    };
    auto fwd = [&]() {
        if constexpr (PP) {
            edge_first<RE, false, T, NT, NBE>(rb, bx, LR);
            fft_lds_range_pp<false, T, NT, VMAX>(bx, by, f, twb, 1, ns - 1);
            edge_last<RE, false, T, NT, NBE, false>(rb, bx, LR, twl);
            Cx<T> *t = bx;  // the next first edge stage writes the other buffer
            bx = by;
            by = t;
        } else {
            edge_first<RE, false, T, NT, NBE>(rb, buf, LR);
            fft_lds_range<false, T, NT, VMAX>(buf, f, 1, ns - 1);
            edge_last<RE, false, T, NT, NBE>(rb, buf, LR, twl);
        }
    };
    auto inv = [&]() {
        if constexpr (PP) {
            edge_first<RE, true, T, NT, NBE>(rb, bx, LR);
            fft_lds_range_pp<true, T, NT, VMAX>(bx, by, f, twb, 1, ns - 1);
            edge_last<RE, true, T, NT, NBE, false>(rb, bx, LR, twl);
            Cx<T> *t = bx;
            bx = by;
            by = t;
        } else {
            edge_first<RE, true, T, NT, NBE>(rb, buf, LR);
            fft_lds_range<true, T, NT, VMAX>(buf, f, 1, ns - 1);
            edge_last<RE, true, T, NT, NBE>(rb, buf, LR, twl);
        }
    };

    if (MODE == COL_FWD || MODE == COL_INV) {
        const int b = blockIdx.y;
        load(rb, in + (size_t)b * in_bs + (size_t)col * in_rp, hin);
        if (MODE == COL_FWD) fwd();
        else inv();
        store(cbase(out + (size_t)b * out_bs, out_rp));
    } else if (MODE == COL_BCAST) {
        load(rb, in + (size_t)col * in_rp, hin);
        Cx<T> pf[NBE][RE];
        if (PF && nexp > 0) load(pf, psf + (size_t)col * psf_rp, f.L);
        fwd();
        Cx<T> xr[NBE][RE];
#pragma unroll
        for (int r = 0; r < NBE; ++r)
#pragma unroll
            for (int q = 0; q < RE; ++q) xr[r][q] = rb[r][q];
        for (int k = 0; k < nexp; ++k) {
            if (!PF) load(pf, psf + (size_t)k * psf_bs + (size_t)col * psf_rp, f.L);
            // (a+ib)(c+id) as a*c-b*d, a*d+b*c (updc_core.c:756-758, adjoint=0)
#pragma unroll
            for (int r = 0; r < NBE; ++r)
#pragma unroll
                for (int q = 0; q < RE; ++q) rb[r][q] = cmul(xr[r][q], pf[r][q]);
            if (PF && k + 1 < nexp) load(pf, psf + (size_t)(k + 1) * psf_bs + (size_t)col * psf_rp, f.L);
            inv();
            store(cbase(out + (size_t)k * out_bs, out_rp));
        }
    } else {  // COL_ADJ, COL_ACC
        Cx<TA> acc[NBE][RE];
        if (MODE == COL_ACC && nexp < 0) {
            // the summed spectrum from memory (TA, pitch in_rp): only its inverse runs
            const Cx<TA> *ia = reinterpret_cast<const Cx<TA> *>(in) + (size_t)col * in_rp;
#pragma unroll
            for (int r = 0; r < NBE; ++r)
#pragma unroll
                for (int q = 0; q < RE; ++q) {
                    const int i = tid + r * NT, n = i + q * LR;
                    acc[r][q] = (i < LR && n < hin) ? ia[n] : cx<TA>((TA)0, (TA)0);
                }
        } else {
#pragma unroll
            for (int r = 0; r < NBE; ++r)
#pragma unroll
                for (int q = 0; q < RE; ++q) acc[r][q] = cx<TA>((TA)0, (TA)0);
        }
        for (int k = 0; k < nexp; ++k) {
            load(rb, in + (size_t)k * in_bs + (size_t)col * in_rp, hin);
            Cx<T> pf[NBE][RE];
            if (PF) load(pf, psf + (size_t)k * psf_bs + (size_t)col * psf_rp, f.L);
            fwd();
            if (!PF) load(pf, psf + (size_t)k * psf_bs + (size_t)col * psf_rp, f.L);
#pragma unroll
            for (int r = 0; r < NBE; ++r)
#pragma unroll
                for (int q = 0; q < RE; ++q) {
                    const Cx<T> p = pf[r][q], sv = rb[r][q];
                    // re += a*c + b*d ; im += b*c - a*d  (updc_core.c:991-993)
                    const T pre = sv.re * p.re + sv.im * p.im;
                    const T pim = sv.im * p.re - sv.re * p.im;
                    acc[r][q].re += (TA)pre;
                    acc[r][q].im += (TA)pim;
                }
        }
        if (MODE == COL_ACC) {
            Cx<TA> *o = cbase(reinterpret_cast<Cx<TA> *>(out), out_rp);
            if (fa.L > 0) {
                // the inverse of the sum in TA (plan fa: the same radices, TA
                // twiddles) over the whole LDS allocation, FH TA values; the
                // last forward transform's edge stage ended with a barrier,
                // except in ping-pong mode
                if (PP) __syncthreads();
                Cx<TA> *bufa = reinterpret_cast<Cx<TA> *>(smem_raw);
                const Cx<TA> *twla = (const Cx<TA> *)fa.tw + fa.to[ns - 1];
                edge_first<RE, true, TA, NT, NBE>(acc, bufa, LR);
                fft_lds_range<true, TA, NT, VMAX>(bufa, fa, 1, ns - 1);
                edge_last<RE, true, TA, NT, NBE>(acc, bufa, LR, twla);
            }
            // the sum (or its inverse), in TA
#pragma unroll
            for (int r = 0; r < NBE; ++r)
#pragma unroll
                for (int q = 0; q < RE; ++q) {
                    const int i = tid + r * NT, n = i + q * LR;
                    if (i < LR && n < hout) el(o, n) = acc[r][q];
                }
            return;
        }
#pragma unroll
        for (int r = 0; r < NBE; ++r)
#pragma unroll
            for (int q = 0; q < RE; ++q) rb[r][q] = cx<T>((T)acc[r][q].re, (T)acc[r][q].im);
        inv();
        store(cbase(out, out_rp));
    }
}

// Elementwise spectrum kernels of the per-exposure form (the reference's
// ws_multiply_spectrum / accumulate_adjoint_spectrum, updc_core.c:753-760,
// 988-995), over full column-major spectra of pitch RP.
template <class T, class TA, int OP>
__global__ void k_spec(const Cx<T> *__restrict__ a, const Cx<T> *__restrict__ psf, Cx<T> *__restrict__ out,
                       Cx<TA> *__restrict__ acc, int RP, int nc, int FH) {
    const size_t n = (size_t)nc * RP;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        if ((int)(i % RP) >= FH) continue;
        if (OP == 2) {
            out[i] = cx<T>((T)acc[i].re, (T)acc[i].im);
            continue;
        }
        const Cx<T> p = psf[i], v = a[i];
        if (OP == 0) {
            out[i] = cmul(v, p);
        } else {
            const T pre = v.re * p.re + v.im * p.im;
            const T pim = v.im * p.re - v.re * p.im;
            acc[i].re += (TA)pre;
            acc[i].im += (TA)pim;
        }
    }
}

// ------------------------------------------------------------------ four-step lines
// (the decomposition is described with LongPlan in fft.hpp).  Line o of a
// LineIO is line o % lpb of batch o / lpb.
template <class T> __device__ __forceinline__ size_t line_base(const LineIO &io, size_t o) {
    const size_t b = o / (size_t)io.lpb;
    return b * io.bstride + (o - b * (size_t)io.lpb) * io.pitch;
}
// target t of exposure e: S_k^T wr at t, the fmaf chain over the transposed
// map's entries in ascending source pixel (as the row gather above)
template <class T> __device__ __forceinline__ T gather_target(const ExpDev &e, const T *__restrict__ w, uint32_t t) {
    T acc = (T)0;
    for (uint32_t a = e.toff[t], z = e.toff[t + 1]; a < z; ++a) {
        const uint2 x = e.tent[a];
        acc = fmat((T)__uint_as_float(x.y), w[x.x], acc);
    }
    return acc;
}
template <class T, class TS, int LD>
__device__ __forceinline__ Cx<T> long_load(const LineIO &io, size_t o, int n) {
    if (LD == LL_CX) {
        const Cx<TS> v = n < io.nvalid ? ((const Cx<TS> *)io.p)[line_base<T>(io, o) + n] : cx<TS>((TS)0, (TS)0);
        return cx<T>((T)v.re, (T)v.im);
    } else if (LD == LL_CX_MUL) {
        const size_t b = o / (size_t)io.lpb, l = o - b * (size_t)io.lpb;
        const Cx<T> x = ((const Cx<T> *)io.p)[l * io.pitch + n];
        const Cx<T> q = ((const Cx<T> *)io.psf)[b * io.pbstride + l * io.ppitch + n];
        return cmul(x, q);
    } else if (LD == LL_REAL) {
        const T *row = (const T *)io.p + line_base<T>(io, o);
        const int c = 2 * n;
        return cx<T>(c < io.nvalid ? row[c] : (T)0, c + 1 < io.nvalid ? row[c + 1] : (T)0);
    } else {  // LL_GATHER: line = target row r of exposure b, columns c < W
        const size_t b = o / (size_t)io.lpb, r = o - b * (size_t)io.lpb;
        const ExpDev &e = io.ex[b];
        const T *w = (const T *)io.wr + e.pix;
        const int c = 2 * n, W = io.nvalid;
        const uint32_t t = (uint32_t)(r * (size_t)W) + (uint32_t)c;
        return cx<T>(c < W ? gather_target<T>(e, w, t) : (T)0, c + 1 < W ? gather_target<T>(e, w, t + 1) : (T)0);
    }
}

// Pass A: workgroup (line o, block of B sub-lines n1b..n1b+B-1).
template <class T, class TS, bool INV, int LD, int NT, int VMAX>
__global__ __launch_bounds__(NT) void k_long_a(FftLen p2, const Cx<T> *__restrict__ roots, int L1, int lgb, int nblk,
                                               size_t line0, LineIO src, Cx<T> *__restrict__ tmp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cx<T> *buf = reinterpret_cast<Cx<T> *>(smem_raw);
    const size_t o = line0 + blockIdx.x / (unsigned)nblk;
    const int n1b = (int)(blockIdx.x % (unsigned)nblk) << lgb;
    const int L2 = p2.L, B = 1 << lgb, nsl = L2 << lgb;
    for (int idx = threadIdx.x; idx < nsl; idx += NT) {
        const int j = idx & (B - 1), n2 = idx >> lgb, n1 = n1b + j;
        buf[idx] = n1 < L1 ? long_load<T, TS, LD>(src, o, n1 + L1 * n2) : cx<T>((T)0, (T)0);
    }
    __syncthreads();
    fft_lds<INV, T, NT, VMAX>(buf, p2, lgb);
    Cx<T> *d = tmp + o * (size_t)L1 * L2;
    for (int idx = threadIdx.x; idx < nsl; idx += NT) {
        const int j = idx & (B - 1), k2 = idx >> lgb, n1 = n1b + j;
        if (n1 >= L1) continue;
        Cx<T> w = roots[n1 * k2];  // n1 k2 < L1 L2: no reduction needed
        if (INV) w.im = -w.im;
        d[(size_t)k2 * L1 + n1] = cmul(buf[idx], w);
    }
}

// Pass B: workgroup (line o, block of B sub-lines k2b..k2b+B-1).
template <class T, bool INV, int ST, int NT, int VMAX>
__global__ __launch_bounds__(NT) void k_long_b(FftLen p1, int L2, int lgb, int nblk, size_t line0,
                                               const Cx<T> *__restrict__ tmp, LineIO dst) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Cx<T> *buf = reinterpret_cast<Cx<T> *>(smem_raw);
    const size_t o = line0 + blockIdx.x / (unsigned)nblk;
    const int k2b = (int)(blockIdx.x % (unsigned)nblk) << lgb;
    const int L1 = p1.L, B = 1 << lgb, nsl = L1 << lgb;
    const Cx<T> *srcl = tmp + o * (size_t)L1 * L2;
    // sub-line j is the contiguous run srcl[(k2b + j) L1 ...]: lanes walk it in order
    for (int idx = threadIdx.x; idx < nsl; idx += NT) {
        const int j = idx / L1, n1 = idx - j * L1, k2 = k2b + j;
        buf[(n1 << lgb) + j] = k2 < L2 ? srcl[(size_t)k2 * L1 + n1] : cx<T>((T)0, (T)0);
    }
    __syncthreads();
    fft_lds<INV, T, NT, VMAX>(buf, p1, lgb);
    const size_t base = line_base<T>(dst, o);
    for (int idx = threadIdx.x; idx < nsl; idx += NT) {
        const int j = idx & (B - 1), k1 = idx >> lgb, k2 = k2b + j;
        if (k2 >= L2) continue;
        const int k = k2 + L2 * k1;
        const Cx<T> v = buf[idx];
        if (ST == LS_CX) {
            if (k < dst.nvalid) ((Cx<T> *)dst.p)[base + k] = v;
        } else {
            T *row = (T *)dst.p + base;
            const T sc = (T)dst.scale;
            const int c = 2 * k;
            if (c < dst.nvalid) row[c] = v.re * sc;
            if (c + 1 < dst.nvalid) row[c + 1] = v.im * sc;
        }
    }
}

// Real split of four-step row transforms: Z (row-major lines b * rows + r of
// M points) -> column-major half spectra out[b][v RP + r], v in [0, M], the
// same split as k_row_r2c.  A 16-row x 64-column tile passes through LDS so
// both sides are read and written in runs.
template <class T>
__global__ __launch_bounds__(256) void k_long_split(const Cx<T> *__restrict__ Z, int M, int rows,
                                                    const Cx<T> *__restrict__ post, Cx<T> *__restrict__ out,
                                                    size_t out_bstride, int RP) {
    constexpr int TR = 16, TV = 64;
    __shared__ Cx<T> tile[TR][TV + 1];
    const int v0 = blockIdx.x * TV, r0 = blockIdx.y * TR, b = blockIdx.z;
    const T half = (T)0.5;
    for (int i = threadIdx.x; i < TR * TV; i += 256) {
        const int rr = i / TV, vv = i - rr * TV, r = r0 + rr, v = v0 + vv;
        if (r >= rows || v > M) continue;
        const Cx<T> *zl = Z + ((size_t)b * rows + r) * M;
        const Cx<T> zv = zl[v == M ? 0 : v];
        const Cx<T> zm = cconj(zl[v == 0 ? 0 : M - v]);
        const Cx<T> fe = cscale(zv + zm, half);
        const Cx<T> d = zv - zm;
        const Cx<T> fo = cx<T>(d.im * half, -d.re * half);
        tile[rr][vv] = fe + cmul(post[v], fo);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < TR * TV; i += 256) {
        const int vv = i / TR, rr = i - vv * TR, r = r0 + rr, v = v0 + vv;
        if (r < rows && v <= M) out[(size_t)b * out_bstride + (size_t)v * RP + r] = tile[rr][vv];
    }
}

// Merge of four-step row inverses: column-major half spectra -> row-major
// half-length complex lines Z[b * rows + r][v], v < M, the same merge as
// k_row_c2r.
template <class T>
__global__ __launch_bounds__(256) void k_long_merge(const Cx<T> *__restrict__ in, size_t in_bstride, int RP, int M,
                                                    int rows, const Cx<T> *__restrict__ post, Cx<T> *__restrict__ Z) {
    constexpr int TR = 16, TV = 64;
    __shared__ Cx<T> tile[TR][TV + 1];
    const int v0 = blockIdx.x * TV, r0 = blockIdx.y * TR, b = blockIdx.z;
    const Cx<T> *y = in + (size_t)b * in_bstride;
    for (int i = threadIdx.x; i < TR * TV; i += 256) {
        const int vv = i / TR, rr = i - vv * TR, r = r0 + rr, v = v0 + vv;
        if (r >= rows || v >= M) continue;
        const Cx<T> yv = y[(size_t)v * RP + r], yu = y[(size_t)(M - v) * RP + r];
        const Cx<T> e1 = yv + cconj(yu), o1 = cmul(yv - cconj(yu), cconj(post[v]));
        tile[rr][vv] = cx<T>(e1.re - o1.im, e1.im + o1.re);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < TR * TV; i += 256) {
        const int rr = i / TV, vv = i - rr * TV, r = r0 + rr, v = v0 + vv;
        if (r < rows && v < M) Z[((size_t)b * rows + r) * M + v] = tile[rr][vv];
    }
}

// Summed adjoint spectrum of four-step columns: out[v opitch + r] =
// sum_k conj(psf_k) X_k (k ascending, products in T, sum in TA, as k_col's
// COL_ADJ / COL_ACC), X_k line-major (k NC + v) FH + r.
template <class T, class TA>
__global__ __launch_bounds__(256) void k_long_adjsum(const Cx<T> *__restrict__ X, int NC, int FH, int nexp,
                                                     const Cx<T> *__restrict__ psf, size_t pbs, int prp,
                                                     Cx<TA> *__restrict__ out, int opitch) {
    const size_t n = (size_t)NC * FH;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const size_t v = i / FH, r = i - v * FH;
        Cx<TA> acc = cx<TA>((TA)0, (TA)0);
        for (int k = 0; k < nexp; ++k) {
            const Cx<T> p = psf[(size_t)k * pbs + v * prp + r];
            const Cx<T> s = X[(size_t)k * n + i];
            const T pre = s.re * p.re + s.im * p.im;
            const T pim = s.im * p.re - s.re * p.im;
            acc.re += (TA)pre;
            acc.im += (TA)pim;
        }
        out[v * opitch + r] = acc;
    }
}

// ------------------------------------------------------------------ host side
// Thread counts and per-thread register budgets.  A column workgroup holds
// one FH-point column (44.8 KB f32 / 89.6 KB f64 at FH = 5600) and runs 8
// waves; the budgets keep the column kernels free of spills.  Row
// workgroups hold RB rows (RB x M points) to write RB-element column runs.
template <class T> struct RowCfg;
// PFM / PFR: prefetch registers of the persistent row kernels; shapes beyond
// them (FP32: M > 3071) use the one-tile-per-workgroup kernels.
template <> struct RowCfg<float> { static constexpr int NT = 1024, VMAX = 16, RB = 4, PFM = 12, PFR = 4; };
template <> struct RowCfg<double> { static constexpr int NT = 1024, VMAX = 10, RB = 2, PFM = 6, PFR = 3; };
// FP32 rows too long for the 4-row persistent C2R (M > 3071) take 2-row
// tiles: 8 prefetch registers cover M <= 4095, and the tile with the LDS
// tables fits (123 KB at M = 3840).
struct RowCfgL { static constexpr int NT = 1024, VMAX = 16, RB = 2, PFM = 8; };
template <class T> struct ColCfg;
// PF: PSF prefetch in the edge column kernel (FP64 would spill with it).
// EV: complex values per thread of the edge kernel's register sets.
template <> struct ColCfg<float> {
    static constexpr int NT = 1024, VMAX = 10, XMAX = 8, EV = 10;
    static constexpr bool PF = true;
};
template <> struct ColCfg<double> {
    static constexpr int NT = 512, VMAX = 16, XMAX = 12, EV = 16;
    static constexpr bool PF = false;
};
static constexpr size_t LDS_BUDGET = 150 * 1024;

// Radix order: 8s, a 4 or 2, then 7s, 5s, 3s.  With `edge`, a radix that
// occurs twice (preferring 8, 5, 4, 7, 3) is moved to both ends, so the
// first and last stages can run in registers; returns that radix in *re
// (0 when none repeats).  With `odd_first` (and no edge), the largest odd
// radix goes first: the first stage writes each butterfly's outputs R apart,
// and for the interleaved row lines an even R (8: 256-byte strides) sends a
// wave's writes to one LDS bank window, while an odd R spreads them.
static bool factor_radices(int L, FftLen &f, bool edge = false, int *re = nullptr, bool odd_first = false,
                           bool sixteen = false, bool ten = false) {
    int n = L, k = 0;
    int rad[UPDC_MAX_STAGES + 8];
    auto push = [&](int r) { if (k < UPDC_MAX_STAGES) rad[k++] = r; };
    // ten (with edge): radix-10 edge stages where L has two factors 2 and two
    // factors 5 (5600 = 10 8 7 10: four stages instead of 5 8 4 7 5)
    if (edge && ten && L % 100 == 0) {
        push(10);
        n /= 100;
    }
    // sixteen: radix 16 where it saves a stage on the power of two (2^4 = 16
    // against 8 2, 2^7 = 16 8 against 8 8 2); the caller checks it fits
    int p2 = 0;
    for (int m = n; m % 2 == 0; m /= 2) ++p2;
    if (sixteen && !edge && p2 % 3 != 0 && p2 >= 4) { push(16); n /= 16; }
    while (n % 8 == 0) { push(8); n /= 8; }
    if (n % 4 == 0) { push(4); n /= 4; }
    if (n % 2 == 0) { push(2); n /= 2; }
    while (n % 7 == 0) { push(7); n /= 7; }
    while (n % 5 == 0) { push(5); n /= 5; }
    while (n % 3 == 0) { push(3); n /= 3; }
    if (re) *re = 0;
    if (k >= 1 && rad[0] == 10) {
        // the second radix-10 edge goes last
        push(10);
        if (re) *re = 10;
    } else if (edge && k >= 2) {
        for (int cand : {8, 5, 4, 7, 3}) {
            int cnt = 0;
            for (int i = 0; i < k; ++i) cnt += rad[i] == cand;
            if (cnt < 2) continue;
            int mid[UPDC_MAX_STAGES], m = 0, skip = 2;
            for (int i = 0; i < k; ++i) {
                if (rad[i] == cand && skip > 0) { --skip; continue; }
                mid[m++] = rad[i];
            }
            rad[0] = cand;
            for (int i = 0; i < m; ++i) rad[1 + i] = mid[i];
            rad[k - 1] = cand;
            if (re) *re = cand;
            break;
        }
    }
    if (odd_first && !(edge && re && *re)) {
        for (int cand : {7, 5, 3}) {
            int at = -1;
            for (int i = 0; i < k; ++i)
                if (rad[i] == cand) { at = i; break; }
            if (at < 0) continue;
            for (int i = at; i > 0; --i) rad[i] = rad[i - 1];
            rad[0] = cand;
            break;
        }
    }
    for (int i = 0; i < k; ++i) f.rad[i] = rad[i];
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

// `lines` interleaved transforms per stage must fit: R * ceil(lines*(L/R) / NT) <= VMAX.
static bool fits(const FftLen &f, int lines, int nt, int vmax) {
    for (int s = 0; s < f.nst; ++s) {
        const int R = f.rad[s];
        const long per = ((long)lines * (f.L / R) + nt - 1) / nt;
        if (per > vmax / R) return false;
    }
    return true;
}

// Per-stage twiddle tables (layout in common.hpp's FftLen), long double
// arithmetic rounded once.
template <class T> static Cx<T> *upload_stage_twiddles(FftLen &f, std::string &err) {
    std::vector<Cx<T>> h;
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int s = 0; s < f.nst; ++s) {
        const int R = f.rad[s], Ns = f.ns[s];
        f.to[s] = (int)h.size();
        for (int k = 0; k < Ns; ++k)
            for (int q = 1; q < R; ++q) {
                const long double a = twopi * (long double)(q * k) / (long double)(Ns * R);
                Cx<T> w;
                w.re = (T)cosl(a);
                w.im = (T)(-sinl(a));
                h.push_back(w);
            }
    }
    if (h.empty()) h.push_back(Cx<T>{(T)1, (T)0});
    Cx<T> *d = nullptr;
    if (hipMalloc(&d, sizeof(Cx<T>) * h.size()) != hipSuccess) { err = "hipMalloc twiddles failed"; return nullptr; }
    if (hipMemcpy(d, h.data(), sizeof(Cx<T>) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        err = "twiddle upload failed";
        (void)hipFree(d);
        return nullptr;
    }
    return d;
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
        (void)hipFree(d);
        return nullptr;
    }
    return d;
}

template <class T, int MODE> static void *row_r2c_fn() {
    return (void *)k_row_r2c<T, RowCfg<T>::NT, RowCfg<T>::VMAX, RowCfg<T>::RB, MODE>;
}
// Pair accesses (load_pair / store_pair) need every (2n, 2n+1) pair of every
// row and batch image aligned to 2 elements.
template <class T> static int pairs_aligned(const void *p, int pitch, size_t bstride) {
    return p && ((uintptr_t)p % (2 * sizeof(T))) == 0 && pitch % 2 == 0 && bstride % 2 == 0;
}
template <class T, bool CP> static void *row_c2r_p_fn() {
    return (void *)k_row_c2r_p<T, RowCfg<T>::NT, RowCfg<T>::VMAX, RowCfg<T>::RB, RowCfg<T>::PFM, CP>;
}
template <bool CP> static void *row_c2r_pl_fn() {
    return (void *)k_row_c2r_p<float, RowCfgL::NT, RowCfgL::VMAX, RowCfgL::RB, RowCfgL::PFM, CP>;
}
template <class T, class TF, bool CP> static void *row_c2r_upd_fn() {
    return (void *)k_row_c2r_upd_p<T, TF, RowCfg<T>::NT, RowCfg<T>::VMAX, RowCfg<T>::RB, RowCfg<T>::PFM, CP>;
}
template <class T, bool VEC> static void *row_r2c_p_fn() {
    return (void *)k_row_r2c_p<T, RowCfg<T>::NT, RowCfg<T>::VMAX, RowCfg<T>::RB, RowCfg<T>::PFR, VEC>;
}
template <class T> static void *row_c2r_fn() {
    return (void *)k_row_c2r<T, RowCfg<T>::NT, RowCfg<T>::VMAX, RowCfg<T>::RB>;
}
// LDS of a row workgroup: RB interleaved rows of M points, plus the Nyquist
// column the C2R merge stages.
template <class T> static size_t row_lds(int M) { return (size_t)RowCfg<T>::RB * (M + 1) * sizeof(Cx<T>); }
// ... and of a persistent C2R workgroup: the tile, the stage twiddles and
// (FP32) the split roots
template <class T> static size_t row_lds_p(int M, int ntw) {
    return row_lds<T>(M) + ((size_t)ntw + (RowPL<T>::value ? (size_t)M + 1 : 0)) * sizeof(Cx<T>);
}
// ... and of the FP32 2-row persistent C2R
static size_t row_lds_pl(int M, int ntw) {
    return (size_t)RowCfgL::RB * (M + 1) * sizeof(Cx<float>) + ((size_t)ntw + M + 1) * sizeof(Cx<float>);
}
// entries of a plan's stage twiddle tables (as upload_stage_twiddles lays them out)
static int stage_twiddle_count(const FftLen &f) {
    long n = 0;
    for (int s = 0; s < f.nst; ++s) n += (long)f.ns[s] * (f.rad[s] - 1);
    return (int)std::max<long>(n, 1);
}
template <class T, class TA, int MODE> static void *col_fn() {
    return (void *)k_col<T, TA, ColCfg<T>::NT, ColCfg<T>::VMAX, ColCfg<T>::XMAX, MODE>;
}

template <class T, class TA, int MODE, bool PF, bool PP, bool CP> static void *col_e_fn_pf(int re) {
    constexpr int NT = ColCfg<T>::NT, V = ColCfg<T>::VMAX, EV = ColCfg<T>::EV;
    switch (re) {
        case 3: return (void *)k_col_e<T, TA, NT, V, 3, MODE, PF, EV, PP, CP>;
        case 4: return (void *)k_col_e<T, TA, NT, V, 4, MODE, PF, EV, PP, CP>;
        case 5: return (void *)k_col_e<T, TA, NT, V, 5, MODE, PF, EV, PP, CP>;
        case 7: return (void *)k_col_e<T, TA, NT, V, 7, MODE, PF, EV, PP, CP>;
        case 10:
            if constexpr (EV >= 10) return (void *)k_col_e<T, TA, NT, V, 10, MODE, PF, EV, PP, CP>;
            return nullptr;
        case 8: return (void *)k_col_e<T, TA, NT, V, 8, MODE, PF, EV, PP, CP>;
        default: return nullptr;
    }
}
// pp: the two-buffer kernel (FP32 only; FP64 columns fill the LDS with one).
// cp: the output in SL_PAIR (the batched modes only; nullptr for COL_FWD/COL_INV)
template <class T, class TA, int MODE> static void *col_e_fn(int re, bool pp, bool cp = false) {
    constexpr bool PF = ColCfg<T>::PF && (MODE == COL_BCAST || MODE == COL_ADJ);  // not COL_ACC (k_col_e)
    if (cp) {
        if constexpr (MODE == COL_FWD || MODE == COL_INV) return nullptr;
        if constexpr (sizeof(T) == 4)
            if (pp) return col_e_fn_pf<T, TA, MODE, PF, true, true>(re);
        return col_e_fn_pf<T, TA, MODE, PF, false, true>(re);
    }
    if constexpr (sizeof(T) == 4)
        if (pp) return col_e_fn_pf<T, TA, MODE, PF, true, false>(re);
    return col_e_fn_pf<T, TA, MODE, PF, false, false>(re);
}

// COL_ACC kernels exist only for a sum wider than the transform type.
template <class T, class TA> static void *col_acc_e_fn(int re, bool pp, bool cp = false) {
    if constexpr (sizeof(TA) > sizeof(T)) return col_e_fn<T, TA, COL_ACC>(re, pp, cp);
    return nullptr;
}
// LDS of the two-buffer column kernel: two lines
template <class T> static size_t col_pp_lds(int FH, const FftLen &) { return 2 * (size_t)FH * sizeof(Cx<T>); }

// ------------------------------------------------------------------ four-step host side
// 256 threads, 16 complex values per thread and stage; B = 16 (FP32) or 8
// (FP64) sub-lines make 128-byte runs.
template <class T> struct LongCfg;
template <> struct LongCfg<float> { static constexpr int NT = 256, VMAX = 16, LGB = 4; };
template <> struct LongCfg<double> { static constexpr int NT = 256, VMAX = 16, LGB = 3; };
static constexpr size_t LONG_LDS = 64 * 1024;  // per workgroup: several resident per CU

static bool seven_smooth_len(long n) {
    for (int p : {2, 3, 5, 7})
        while (n > 1 && n % p == 0) n /= p;
    return n == 1;
}

// L = L1 x L2 with both sub-plans fitting B interleaved sub-lines in one
// workgroup's registers and LDS: the widest B first (the longest global
// runs), then the most even split.
template <class T> static std::string long_plan(int L, LongPlan<T> &lp) {
    constexpr int NT = LongCfg<T>::NT, VMAX = LongCfg<T>::VMAX;
    int best1 = 0, bestlg = -1;
    long bestd = 0;
    for (int L1 = 2; L1 <= L / 2; ++L1) {
        if (L % L1) continue;
        const int L2 = L / L1;
        FftLen a{}, b{};
        if (!factor_radices(L1, a, false, nullptr, true) || !factor_radices(L2, b, false, nullptr, true)) continue;
        for (int lg = LongCfg<T>::LGB; lg >= 0; --lg) {
            const int B = 1 << lg;
            if (!fits(a, B, NT, VMAX) || !fits(b, B, NT, VMAX) ||
                (size_t)B * std::max(L1, L2) * sizeof(Cx<T>) > LONG_LDS)
                continue;
            const long d = std::labs((long)L1 - (long)L2);
            if (lg > bestlg || (lg == bestlg && d < bestd)) {
                best1 = L1;
                bestlg = lg;
                bestd = d;
            }
            break;
        }
    }
    if (bestlg < 0) return "FFT length " + std::to_string(L) + " has no split for the four-step pass";
    lp.L = L;
    lp.L1 = best1;
    lp.L2 = L / best1;
    lp.lgb = bestlg;
    factor_radices(lp.L1, lp.p1, false, nullptr, true);
    factor_radices(lp.L2, lp.p2, false, nullptr, true);
    std::string err;
    lp.tw1 = upload_stage_twiddles<T>(lp.p1, err);
    if (!lp.tw1) return err;
    lp.tw2 = upload_stage_twiddles<T>(lp.p2, err);
    if (!lp.tw2) return err;
    lp.roots = upload_roots<T>(L, L, -1, err);
    if (!lp.roots) return err;
    lp.p1.tw = lp.tw1;
    lp.p2.tw = lp.tw2;
    return "";
}
template <class T> static void long_free(LongPlan<T> &lp) {
    for (Cx<T> *p : {lp.tw1, lp.tw2, lp.roots})
        if (p) (void)hipFree(p);
    lp = LongPlan<T>();
}

template <class T> std::string Fft2<T>::init(int FW, int FH) {
    if (FW < 2 || FH < 1 || (FW & 1)) return "FFT width must be even and >= 2";
    s.FW = FW; s.FH = FH; s.M = FW / 2; s.NC = FW / 2 + 1;
    if (!seven_smooth_len(s.M)) return "FFT width/2 is not 7-smooth";
    if (!seven_smooth_len(FH)) return "FFT height is not 7-smooth";
    if ((size_t)s.NC * (size_t)FH > ((size_t)1 << 40)) return "FFT size too large";
    // rows of RB lines in one workgroup's LDS and registers, else four-step rows
    long_row = !factor_radices(s.M, row, false, nullptr, row_odd_first) || row_lds<T>(s.M) > LDS_BUDGET ||
               !fits(row, RowCfg<T>::RB, RowCfg<T>::NT, RowCfg<T>::VMAX);
    if (!long_row) {
        // a radix-16 stage where it saves one and fits the row tiles' registers
        FftLen r16{};
        if (factor_radices(s.M, r16, false, nullptr, row_odd_first, true) && r16.nst < row.nst &&
            fits(r16, RowCfg<T>::RB, RowCfg<T>::NT, RowCfg<T>::VMAX))
            row = r16;
    }
    constexpr int CNT = ColCfg<T>::NT, CVMAX = ColCfg<T>::VMAX, CXMAX = ColCfg<T>::XMAX;
    // one column per workgroup, else four-step columns.  The edge kernel holds
    // EV / RE butterflies of each thread in registers; the LDS-only column
    // kernel keeps XMAX values per thread in registers
    // radix-10 edges (FP32 plans, where the edge kernel holds them)
    long_col = !factor_radices(FH, col, allow_edge, &col_re, false, false, allow_ten && sizeof(T) == 4);
    if (!long_col && col_re == 10 && FH / 10 > (ColCfg<T>::EV / 10) * ColCfg<T>::NT)
        long_col = !factor_radices(FH, col, allow_edge, &col_re);
    if (!long_col && col_re && FH / col_re > (ColCfg<T>::EV / col_re) * ColCfg<T>::NT) col_re = 0;
    // radix 10 runs only as an edge stage (the LDS stage loops have no radix-10 case)
    for (int i = 0; i < col.nst && !long_col; ++i)
        if (col.rad[i] == 10 && (col_re != 10 || (i != 0 && i != col.nst - 1)))
            return "internal: radix 10 outside the column plan's edge stages";
    long_col = long_col || (size_t)FH * sizeof(Cx<T>) > LDS_BUDGET || (col_re == 0 && FH > CNT * CXMAX) ||
               !fits(col, 1, CNT, CVMAX);
    if (long_col) col_re = 0;
    std::string err;
    if (long_row) {
        if (!(err = long_plan<T>(s.M, lrow)).empty()) return err;
    } else {
        tw_row = upload_stage_twiddles<T>(row, err);
        if (!tw_row) return err;
        row.tw = tw_row;
        row_ntw = stage_twiddle_count(row);
    }
    if (long_col) {
        if (!(err = long_plan<T>(FH, lcol)).empty()) return err;
    } else {
        tw_col = upload_stage_twiddles<T>(col, err);
        if (!tw_col) return err;
        col.tw = tw_col;
        if constexpr (sizeof(T) < sizeof(double)) {
            // the same plan with FP64 twiddles: the accumulate pass's FP64 inverse
            col_wide = col;
            tw_col_wide = upload_stage_twiddles<double>(col_wide, err);
            if (!tw_col_wide) return err;
            col_wide.tw = tw_col_wide;
        }
    }
    post = upload_roots<T>(FW, s.M + 1, -1, err);
    if (!post) return err;
    const int rlds = (int)row_lds<T>(s.M), rlds_p = (int)row_lds_p<T>(s.M, row_ntw);
    // two transform buffers where they fit the LDS budget (the FP32 edge
    // kernel's out-of-place stages)
    col_pp = allow_pp && col_re && sizeof(T) == 4 && col_pp_lds<T>(FH, col) <= LDS_BUDGET;
    const int col_lds = (int)(col_pp ? col_pp_lds<T>(FH, col) : (size_t)FH * sizeof(Cx<T>));
    void *fns_row[] = {row_r2c_fn<T, ROW_PLAIN>(), row_r2c_fn<T, ROW_GATHER>(), row_c2r_fn<T>(),
                       row_r2c_p_fn<T, false>(), row_r2c_p_fn<T, true>()};
    void *fns_row_p[] = {row_c2r_p_fn<T, false>(), row_c2r_upd_fn<T, float, false>(), row_c2r_upd_fn<T, T, false>(),
                         row_c2r_p_fn<T, true>(), row_c2r_upd_fn<T, float, true>(), row_c2r_upd_fn<T, T, true>()};
    // persistent row passes: one workgroup per CU, prefetch registers must cover a tile
    {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
            return "cannot query the device";
        row_grid = std::max(8, prop.multiProcessorCount / 8 * 8);
        constexpr int RB = RowCfg<T>::RB, NT = RowCfg<T>::NT;
        // the R2C keeps its twiddle tables in global memory, so it fits rows
        // the C2R (tables in LDS, a larger prefetch) does not
        r2c_persistent = row_persistent && !long_row && s.M <= RowCfg<T>::PFR * NT && (size_t)rlds <= LDS_BUDGET;
        const bool allow = row_persistent;
        row_persistent = row_persistent && !long_row && (s.M + 1) * RB <= RowCfg<T>::PFM * NT &&
                         s.M <= RowCfg<T>::PFR * NT && (size_t)rlds_p <= LDS_BUDGET;
        if constexpr (sizeof(T) == 4) {
            c2r_long = allow && !long_row && !row_persistent && (s.M + 1) * RowCfgL::RB <= RowCfgL::PFM * NT &&
                       row_lds_pl(s.M, row_ntw) <= LDS_BUDGET;
            if (c2r_long)
                for (void *fn : {row_c2r_pl_fn<false>(), row_c2r_pl_fn<true>()})
                    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)row_lds_pl(s.M, row_ntw)) != hipSuccess)
                        return "cannot raise LDS limit for the row pass";
        }
    }
    if (!long_row)
        for (void *fn : fns_row)
            if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, rlds) != hipSuccess)
                return "cannot raise LDS limit for the row pass";
    if (row_persistent)
        for (void *fn : fns_row_p)
            if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, rlds_p) != hipSuccess)
                return "cannot raise LDS limit for the row pass";
    if (long_col) return "";
    void *fns_col[] = {col_fn<T, T, COL_FWD>(), col_fn<T, T, COL_INV>(), col_fn<T, T, COL_BCAST>(),
                       col_fn<T, float, COL_ADJ>(), col_fn<T, double, COL_ADJ>()};
    for (void *fn : fns_col)
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, col_lds) != hipSuccess)
            return "cannot raise LDS limit for the column pass";
    if (col_re) {
        void *fns_e[] = {col_e_fn<T, T, COL_FWD>(col_re, col_pp), col_e_fn<T, T, COL_INV>(col_re, col_pp),
                         col_e_fn<T, T, COL_BCAST>(col_re, col_pp), col_e_fn<T, T, COL_ADJ>(col_re, col_pp),
                         col_e_fn<T, double, COL_ADJ>(col_re, col_pp),
                         col_e_fn<T, T, COL_BCAST>(col_re, col_pp, true), col_e_fn<T, T, COL_ADJ>(col_re, col_pp, true),
                         col_e_fn<T, double, COL_ADJ>(col_re, col_pp, true)};
        for (void *fn : fns_e)
            if (!fn || hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, col_lds) != hipSuccess)
                return "cannot raise LDS limit for the column pass";
    }
    if constexpr (sizeof(T) < sizeof(double)) {
        // the accumulate pass may run its FP64 inverse over FH double values
        void *fn_acc = col_re ? col_acc_e_fn<T, double>(col_re, col_pp) : col_fn<T, double, COL_ACC>();
        const int acc_lds = (int)std::min<size_t>((size_t)FH * sizeof(Cx<double>), LDS_BUDGET);
        if (!fn_acc || hipFuncSetAttribute(fn_acc, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           std::max(col_lds, acc_lds)) != hipSuccess)
            return "cannot raise LDS limit for the column pass";
        void *fn_acc_cp = col_re ? col_acc_e_fn<T, double>(col_re, col_pp, true) : nullptr;
        if (fn_acc_cp && hipFuncSetAttribute(fn_acc_cp, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             std::max(col_lds, acc_lds)) != hipSuccess)
            return "cannot raise LDS limit for the column pass";
    }
    return "";
}

template <class T> void Fft2<T>::release() {
    if (tw_row) (void)hipFree(tw_row);
    if (tw_col) (void)hipFree(tw_col);
    if (tw_col_wide) (void)hipFree(tw_col_wide);
    tw_col_wide = nullptr;
    col_wide = FftLen{};
    if (post) (void)hipFree(post);
    tw_row = tw_col = post = nullptr;
    long_free(lrow);
    long_free(lcol);
    long_row = long_col = false;
    if (lt_a) (void)hipFree(lt_a);
    if (lt_b) (void)hipFree(lt_b);
    if (lt_c) (void)hipFree(lt_c);
    lt_a = lt_b = nullptr;
    lt_c = nullptr;
    lt_a_n = lt_b_n = lt_c_bytes = 0;
}

template <class T>
hipError_t Fft2<T>::row_r2c(int mode, const T *in, size_t in_bstride, int in_pitch, int win, const ExpDev *ex,
                            const T *wr, int W, Cx<T> *out, size_t out_bstride, int rp_rows, int rows, int batch,
                            hipStream_t st) {
    if (rows <= 0 || batch <= 0) return hipSuccess;
    if (long_row)
        return row_r2c_long(mode, in, in_bstride, in_pitch, win, ex, wr, W, out, out_bstride, rp_rows, rows, batch,
                            st);
    constexpr int RB = RowCfg<T>::RB, NT = RowCfg<T>::NT, VMAX = RowCfg<T>::VMAX;
    const size_t lds = row_lds<T>(s.M);
    dim3 grid(((rows + RB - 1) / RB + 7) / 8 * 8, batch);  // a multiple of 8 for xcd_tile
    const int vec = pairs_aligned<T>(in, in_pitch, in_bstride);
    if (mode == ROW_PLAIN && r2c_persistent) {
        if (vec)
            hipLaunchKernelGGL((k_row_r2c_p<T, NT, VMAX, RB, RowCfg<T>::PFR, true>), dim3(row_grid), dim3(NT), lds, st,
                               row, post, in, in_bstride, in_pitch, win, out, out_bstride, rp(rp_rows), rows, batch);
        else
            hipLaunchKernelGGL((k_row_r2c_p<T, NT, VMAX, RB, RowCfg<T>::PFR, false>), dim3(row_grid), dim3(NT), lds,
                               st, row, post, in, in_bstride, in_pitch, win, out, out_bstride, rp(rp_rows), rows, batch);
        return hipGetLastError();
    }
    if (mode == ROW_PLAIN)
        hipLaunchKernelGGL((k_row_r2c<T, NT, VMAX, RB, ROW_PLAIN>), grid, dim3(NT), lds, st, row, post, in, in_bstride,
                           in_pitch, win, ex, wr, W, out, out_bstride, rp(rp_rows), rows, vec);
    else
        hipLaunchKernelGGL((k_row_r2c<T, NT, VMAX, RB, ROW_GATHER>), grid, dim3(NT), lds, st, row, post, in, in_bstride,
                           in_pitch, win, ex, wr, W, out, out_bstride, rp(rp_rows), rows, 0);
    return hipGetLastError();
}

template <class T>
hipError_t Fft2<T>::row_c2r(const Cx<T> *in, size_t in_bstride, int rp_rows, T *out, size_t out_bstride,
                            int out_pitch, int wout, T scale, int rows, int batch, hipStream_t st, int lin) {
    if (lin != SL_COL && (lin != SL_PAIR || !row_pair_ok())) return hipErrorNotSupported;
    if (rows <= 0 || batch <= 0) return hipSuccess;
    if (long_row) return row_c2r_long(in, in_bstride, rp_rows, out, out_bstride, out_pitch, wout, scale, rows, batch, st);
    constexpr int RB = RowCfg<T>::RB, NT = RowCfg<T>::NT, VMAX = RowCfg<T>::VMAX;
    const size_t lds = row_lds<T>(s.M);
    const int vec = pairs_aligned<T>(out, out_pitch, out_bstride);
    if (row_persistent) {
        if (lin == SL_PAIR)
            hipLaunchKernelGGL((k_row_c2r_p<T, NT, VMAX, RB, RowCfg<T>::PFM, true>), dim3(row_grid), dim3(NT),
                               row_lds_p<T>(s.M, row_ntw), st, row, post, row_ntw, in, in_bstride, rp(rp_rows), out,
                               out_bstride, out_pitch, wout, scale, rows, batch, vec);
        else
            hipLaunchKernelGGL((k_row_c2r_p<T, NT, VMAX, RB, RowCfg<T>::PFM, false>), dim3(row_grid), dim3(NT),
                               row_lds_p<T>(s.M, row_ntw), st, row, post, row_ntw, in, in_bstride, rp(rp_rows), out,
                               out_bstride, out_pitch, wout, scale, rows, batch, vec);
        return hipGetLastError();
    }
    if constexpr (sizeof(T) == 4) {
        if (c2r_long) {
            using L = RowCfgL;
            if (lin == SL_PAIR)
                hipLaunchKernelGGL((k_row_c2r_p<T, L::NT, L::VMAX, L::RB, L::PFM, true>), dim3(row_grid), dim3(L::NT),
                                   row_lds_pl(s.M, row_ntw), st, row, post, row_ntw, in, in_bstride, rp(rp_rows), out,
                                   out_bstride, out_pitch, wout, scale, rows, batch, vec);
            else
                hipLaunchKernelGGL((k_row_c2r_p<T, L::NT, L::VMAX, L::RB, L::PFM, false>), dim3(row_grid),
                                   dim3(L::NT), row_lds_pl(s.M, row_ntw), st, row, post, row_ntw, in, in_bstride,
                                   rp(rp_rows), out, out_bstride, out_pitch, wout, scale, rows, batch, vec);
            return hipGetLastError();
        }
    }
    const dim3 grid(((rows + RB - 1) / RB + 7) / 8 * 8, batch);  // a multiple of 8 for xcd_tile
    hipLaunchKernelGGL((k_row_c2r<T, NT, VMAX, RB>), grid, dim3(NT), lds, st, row, post,
                       in, in_bstride, rp(rp_rows), out, out_bstride, out_pitch, wout, scale, rows, vec);
    return hipGetLastError();
}

template <class T>
template <class TF>
hipError_t Fft2<T>::row_c2r_update(const Cx<T> *in, int rp_rows, TF *x, const T *D, int W, int H, T scale,
                                   T max_sens, IterDiag *idiag, double *partials, unsigned long long *negs,
                                   unsigned int *ticket, hipStream_t st, int lin) {
    if (!row_persistent) return hipErrorNotSupported;
    if (lin != SL_COL && (lin != SL_PAIR || !row_pair_ok())) return hipErrorNotSupported;
    constexpr int RB = RowCfg<T>::RB, NT = RowCfg<T>::NT, VMAX = RowCfg<T>::VMAX;
    if (lin == SL_PAIR)
        hipLaunchKernelGGL((k_row_c2r_upd_p<T, TF, NT, VMAX, RB, RowCfg<T>::PFM, true>), dim3(row_grid), dim3(NT),
                           row_lds_p<T>(s.M, row_ntw), st, row, post, row_ntw, in, rp(rp_rows), x, D, W, scale,
                           max_sens, H, idiag, partials, negs, ticket);
    else
        hipLaunchKernelGGL((k_row_c2r_upd_p<T, TF, NT, VMAX, RB, RowCfg<T>::PFM, false>), dim3(row_grid), dim3(NT),
                           row_lds_p<T>(s.M, row_ntw), st, row, post, row_ntw, in, rp(rp_rows), x, D, W, scale,
                           max_sens, H, idiag, partials, negs, ticket);
    return hipGetLastError();
}

template <class T>
template <class TA>
hipError_t Fft2<T>::colpass(int mode, const Cx<T> *in, size_t in_bstride, int in_rows, Cx<T> *out,
                            size_t out_bstride, int out_rows, const Cx<T> *psf, size_t psf_bstride, int hin,
                            int hout, int batch, hipStream_t st, int lout) {
    FftLen none{};  // no TA inverse
    return colpass_impl<TA>(mode, in, in_bstride, in_rows, out, out_bstride, out_rows, psf, psf_bstride, hin, hout,
                            batch, none, (size_t)s.FH * sizeof(Cx<T>), st, lout);
}

template <class T>
template <class TA>
hipError_t Fft2<T>::colpass_impl(int mode, const Cx<T> *in, size_t in_bstride, int in_rows, Cx<T> *out,
                                 size_t out_bstride, int out_rows, const Cx<T> *psf, size_t psf_bstride, int hin,
                                 int hout, int batch, FftLen fa, size_t lds, hipStream_t st, int lout) {
    if (lout != SL_COL && lout != SL_PAIR) return hipErrorInvalidValue;
    const bool cp = lout == SL_PAIR;
    if (cp && (!col_pair_ok() || mode == COL_FWD || mode == COL_INV)) return hipErrorNotSupported;
    if (long_col) {
        if (fa.L > 0) return hipErrorInvalidValue;  // no fused TA inverse on four-step columns
        return colpass_long<TA>(mode, in, in_bstride, in_rows, out, out_bstride, out_rows, psf, psf_bstride, hin,
                                hout, batch, st);
    }
    constexpr int CNT = ColCfg<T>::NT, CVMAX = ColCfg<T>::VMAX, CXMAX = ColCfg<T>::XMAX;
    const int gx = (s.NC + 7) / 8 * 8;
    hin = std::min(hin, s.FH);
    hout = std::min(hout, s.FH);
    const int irp = rp(in_rows), orp = rp(out_rows), prp = rp(s.FH);
    if (col_re) {
        void *fn = mode == COL_FWD     ? col_e_fn<T, T, COL_FWD>(col_re, col_pp)
                   : mode == COL_INV   ? col_e_fn<T, T, COL_INV>(col_re, col_pp)
                   : mode == COL_BCAST ? col_e_fn<T, T, COL_BCAST>(col_re, col_pp, cp)
                   : mode == COL_ADJ   ? col_e_fn<T, TA, COL_ADJ>(col_re, col_pp, cp)
                                       : col_acc_e_fn<T, TA>(col_re, col_pp, cp);
        if (!fn) return hipErrorInvalidValue;
        const int nc = s.NC, nexp = batch;
        const Cx<T> *pin = in, *ppsf = psf;
        Cx<T> *pout = out;
        void *args[] = {(void *)&col, (void *)&nc, (void *)&pin, (void *)&in_bstride, (void *)&irp,
                        (void *)&pout, (void *)&out_bstride, (void *)&orp, (void *)&ppsf, (void *)&psf_bstride,
                        (void *)&prp, (void *)&hin, (void *)&hout, (void *)&nexp, (void *)&fa};
        // SL_PAIR: every XCD's column range starts at an even column (xcd_tile_even)
        const int gxe = cp ? (s.NC + 15) / 16 * 16 : gx;
        const dim3 grid(gxe, (mode == COL_FWD || mode == COL_INV) ? batch : 1);
        if (col_pp) lds = std::max(lds, col_pp_lds<T>(s.FH, col));
        return hipLaunchKernel(fn, grid, dim3(CNT), args, lds, st);
    }
    switch (mode) {
        case COL_FWD:
            hipLaunchKernelGGL((k_col<T, T, CNT, CVMAX, CXMAX, COL_FWD>), dim3(gx, batch), dim3(CNT), lds, st, col,
                               s.NC, in, in_bstride, irp, out, out_bstride, orp, psf, psf_bstride, prp, hin, hout,
                               batch, fa);
            break;
        case COL_INV:
            hipLaunchKernelGGL((k_col<T, T, CNT, CVMAX, CXMAX, COL_INV>), dim3(gx, batch), dim3(CNT), lds, st, col,
                               s.NC, in, in_bstride, irp, out, out_bstride, orp, psf, psf_bstride, prp, hin, hout,
                               batch, fa);
            break;
        case COL_BCAST:
            hipLaunchKernelGGL((k_col<T, T, CNT, CVMAX, CXMAX, COL_BCAST>), dim3(gx, 1), dim3(CNT), lds, st, col, s.NC,
                               in, in_bstride, irp, out, out_bstride, orp, psf, psf_bstride, prp, hin, hout, batch,
                               fa);
            break;
        case COL_ADJ:
            hipLaunchKernelGGL((k_col<T, TA, CNT, CVMAX, CXMAX, COL_ADJ>), dim3(gx, 1), dim3(CNT), lds, st, col, s.NC,
                               in, in_bstride, irp, out, out_bstride, orp, psf, psf_bstride, prp, hin, hout, batch,
                               fa);
            break;
        default:
            if constexpr (sizeof(TA) > sizeof(T)) {
                hipLaunchKernelGGL((k_col<T, TA, CNT, CVMAX, CXMAX, COL_ACC>), dim3(gx, 1), dim3(CNT), lds, st, col,
                                   s.NC, in, in_bstride, irp, out, out_bstride, orp, psf, psf_bstride, prp, hin, hout,
                                   batch, fa);
            } else {
                return hipErrorInvalidValue;
            }
            break;
    }
    return hipGetLastError();
}

template <class T>
template <class TA>
hipError_t Fft2<T>::acc_inverse(const Cx<TA> *acc, Cx<TA> *out, int out_rows, int hout, const FftLen &fa,
                                hipStream_t st) {
    if (!acc_inverse_ok(fa)) return hipErrorInvalidValue;
    // the accumulate pass with no exposures, its sum read from acc (FH rows)
    return colpass_impl<TA>(COL_ACC, reinterpret_cast<const Cx<T> *>(acc), 0, s.FH, reinterpret_cast<Cx<T> *>(out), 0,
                            out_rows, nullptr, 0, s.FH, hout, -1, fa, (size_t)s.FH * sizeof(Cx<TA>), st);
}

template <class T> bool Fft2<T>::row_pair_ok() const {
    constexpr int RB = RowCfg<T>::RB, NT = RowCfg<T>::NT;
    // the C2R prefetch slots must cover the tile's pairs: (M + 2) / 2 pairs of
    // RB rows, pair_lpp lanes each, pair_cpl complex values per lane
    const long slots = (long)(RowCfg<T>::PFM / pair_cpl<T>()) * NT;
    if (row_persistent) return !long_row && (long)((s.M + 2) / 2) * RB * pair_lpp<T>() <= slots;
    if constexpr (sizeof(T) == 4) {
        const long slots_l = (long)(RowCfgL::PFM / pair_cpl<T>()) * RowCfgL::NT;
        if (c2r_long) return (long)((s.M + 2) / 2) * RowCfgL::RB * pair_lpp<T>() <= slots_l;
    }
    return false;
}

template <class T> bool Fft2<T>::acc_inverse_ok(const FftLen &fa) const {
    constexpr int CNT = ColCfg<T>::NT, CVMAX = ColCfg<T>::VMAX;
    if (long_col || !fa.tw || fa.L != s.FH || (size_t)s.FH * sizeof(Cx<double>) > LDS_BUDGET || !fits(fa, 1, CNT, CVMAX)) return false;
    if (!col_re) return true;  // the LDS-only kernel runs plan fa as it is
    // the edge kernel keeps its own edge radix: plan fa must have the same radices
    if (fa.nst != col.nst) return false;
    for (int i = 0; i < col.nst; ++i)
        if (fa.rad[i] != col.rad[i]) return false;
    return true;
}

template <class T>
template <class TA>
hipError_t Fft2<T>::colacc(const Cx<T> *in, size_t in_bstride, int in_rows, Cx<TA> *out, int out_rows, int hout,
                           const Cx<T> *psf, size_t psf_bstride, int hin, int batch, const FftLen *fa,
                           hipStream_t st, int lout) {
    // the kernels write Cx<TA> elements at pitch rp(out_rows) through `out`
    if (fa && !acc_inverse_ok(*fa)) return hipErrorInvalidValue;
    FftLen none{};
    const size_t lds = (size_t)s.FH * (fa ? sizeof(Cx<TA>) : sizeof(Cx<T>));
    return colpass_impl<TA>(COL_ACC, in, in_bstride, in_rows, reinterpret_cast<Cx<T> *>(out), 0,
                            fa ? out_rows : s.FH, psf, psf_bstride, hin, fa ? hout : s.FH, batch, fa ? *fa : none,
                            lds, st, lout);
}

template <class T> hipError_t Fft2<T>::spec_mul(const Cx<T> *a, const Cx<T> *psf_k, Cx<T> *out, hipStream_t st) {
    const unsigned g = (unsigned)std::min<size_t>((spec_elems(s.FH) + 255) / 256, 16384);
    hipLaunchKernelGGL((k_spec<T, T, 0>), dim3(g), dim3(256), 0, st, a, psf_k, out, (Cx<T> *)nullptr, rp(s.FH), s.NC,
                       s.FH);
    return hipGetLastError();
}
template <class T>
template <class TA>
hipError_t Fft2<T>::spec_acc(const Cx<T> *a, const Cx<T> *psf_k, Cx<TA> *acc, hipStream_t st) {
    const unsigned g = (unsigned)std::min<size_t>((spec_elems(s.FH) + 255) / 256, 16384);
    hipLaunchKernelGGL((k_spec<T, TA, 1>), dim3(g), dim3(256), 0, st, a, psf_k, (Cx<T> *)nullptr, acc, rp(s.FH), s.NC,
                       s.FH);
    return hipGetLastError();
}
template <class T>
template <class TA>
hipError_t Fft2<T>::spec_cast(const Cx<TA> *acc, Cx<T> *out, hipStream_t st) {
    const unsigned g = (unsigned)std::min<size_t>((spec_elems(s.FH) + 255) / 256, 16384);
    hipLaunchKernelGGL((k_spec<T, TA, 2>), dim3(g), dim3(256), 0, st, (const Cx<T> *)nullptr, (const Cx<T> *)nullptr,
                       out, (Cx<TA> *)acc, rp(s.FH), s.NC, s.FH);
    return hipGetLastError();
}

template <class T> hipError_t Fft2<T>::long_space(size_t a, size_t b, size_t cbytes, hipStream_t st) {
    if (a <= lt_a_n && b <= lt_b_n && cbytes <= lt_c_bytes) return hipSuccess;
    // a captured launch sequence cannot allocate: long_reserve sizes the
    // scratch for the pipeline's calls beforehand
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
        return hipErrorStreamCaptureUnsupported;
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    auto grow = [](void **p, size_t &have, size_t need, size_t unit) -> hipError_t {
        if (need <= have) return hipSuccess;
        if (*p) (void)hipFree(*p);
        *p = nullptr;
        have = 0;
        const hipError_t r = hipMalloc(p, need * unit);
        if (r != hipSuccess) {
            *p = nullptr;
            return r;
        }
        have = need;
        return hipSuccess;
    };
    if ((e = grow((void **)&lt_a, lt_a_n, a, sizeof(Cx<T>))) != hipSuccess) return e;
    if ((e = grow((void **)&lt_b, lt_b_n, b, sizeof(Cx<T>))) != hipSuccess) return e;
    return grow(&lt_c, lt_c_bytes, cbytes, 1);
}

template <class T> hipError_t Fft2<T>::long_reserve(int nexp) {
    const size_t n = (size_t)std::max(nexp, 1);
    size_t a = 0, b = 0, cb = 0;
    if (long_row) a = b = n * (size_t)s.FH * (size_t)s.M;  // up to n images of FH rows
    if (long_col) {
        const size_t sp = n * (size_t)s.NC * (size_t)s.FH;  // n full spectra
        a = std::max(a, sp);
        b = std::max(b, sp);
        cb = (size_t)s.NC * s.FH * sizeof(Cx<double>);
    }
    return long_space(a, b, cb, nullptr);
}

template <class T, class TS> static void *long_a_fn(bool inv, int ld) {
    constexpr int NT = LongCfg<T>::NT, V = LongCfg<T>::VMAX;
    if constexpr (!std::is_same<T, TS>::value) {
        // a wider source: only the summed spectrum, read as T
        if (ld != LL_CX) return nullptr;
        return inv ? (void *)k_long_a<T, TS, true, LL_CX, NT, V> : (void *)k_long_a<T, TS, false, LL_CX, NT, V>;
    } else {
        switch (ld) {
            case LL_CX:
                return inv ? (void *)k_long_a<T, T, true, LL_CX, NT, V> : (void *)k_long_a<T, T, false, LL_CX, NT, V>;
            case LL_CX_MUL:
                return inv ? (void *)k_long_a<T, T, true, LL_CX_MUL, NT, V>
                           : (void *)k_long_a<T, T, false, LL_CX_MUL, NT, V>;
            case LL_REAL:
                return inv ? (void *)k_long_a<T, T, true, LL_REAL, NT, V>
                           : (void *)k_long_a<T, T, false, LL_REAL, NT, V>;
            case LL_GATHER:
                return inv ? (void *)k_long_a<T, T, true, LL_GATHER, NT, V>
                           : (void *)k_long_a<T, T, false, LL_GATHER, NT, V>;
            default: return nullptr;
        }
    }
}
template <class T> static void *long_b_fn(bool inv, int store) {
    constexpr int NT = LongCfg<T>::NT, V = LongCfg<T>::VMAX;
    if (store == LS_CX) return inv ? (void *)k_long_b<T, true, LS_CX, NT, V> : (void *)k_long_b<T, false, LS_CX, NT, V>;
    return inv ? (void *)k_long_b<T, true, LS_REAL, NT, V> : (void *)k_long_b<T, false, LS_REAL, NT, V>;
}

// Workgroups per launch stay below 2^22; longer line sets take several launches.
static constexpr size_t LONG_MAX_WG = (size_t)1 << 22;

template <class T>
template <class TS>
hipError_t Fft2<T>::long_a(const LongPlan<T> &lp, bool inv, int ld, const LineIO &src, size_t lines, Cx<T> *tmp,
                           hipStream_t st) {
    void *fn = long_a_fn<T, TS>(inv, ld);
    if (!fn || lp.L == 0) return hipErrorInvalidValue;
    FftLen p2 = lp.p2;
    const Cx<T> *roots = lp.roots;
    int L1 = lp.L1, lgb = lp.lgb, nblk = (lp.L1 + (1 << lp.lgb) - 1) >> lp.lgb;
    LineIO io = src;
    Cx<T> *t = tmp;
    const size_t lds = ((size_t)lp.L2 << lp.lgb) * sizeof(Cx<T>);
    const size_t per = std::max<size_t>(1, LONG_MAX_WG / (size_t)nblk);
    for (size_t l0 = 0; l0 < lines; l0 += per) {
        size_t line0 = l0;
        const size_t nl = std::min(per, lines - l0);
        void *args[] = {(void *)&p2, (void *)&roots, (void *)&L1, (void *)&lgb, (void *)&nblk, (void *)&line0,
                        (void *)&io, (void *)&t};
        const hipError_t e = hipLaunchKernel(fn, dim3((unsigned)(nl * nblk)), dim3(LongCfg<T>::NT), args, lds, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <class T>
hipError_t Fft2<T>::long_b(const LongPlan<T> &lp, bool inv, int store, const Cx<T> *tmp, const LineIO &dst,
                           size_t lines, hipStream_t st) {
    void *fn = long_b_fn<T>(inv, store);
    if (lp.L == 0) return hipErrorInvalidValue;
    FftLen p1 = lp.p1;
    int L2 = lp.L2, lgb = lp.lgb, nblk = (lp.L2 + (1 << lp.lgb) - 1) >> lp.lgb;
    const Cx<T> *t = tmp;
    LineIO io = dst;
    const size_t lds = ((size_t)lp.L1 << lp.lgb) * sizeof(Cx<T>);
    const size_t per = std::max<size_t>(1, LONG_MAX_WG / (size_t)nblk);
    for (size_t l0 = 0; l0 < lines; l0 += per) {
        size_t line0 = l0;
        const size_t nl = std::min(per, lines - l0);
        void *args[] = {(void *)&p1, (void *)&L2, (void *)&lgb, (void *)&nblk, (void *)&line0, (void *)&t,
                        (void *)&io};
        const hipError_t e = hipLaunchKernel(fn, dim3((unsigned)(nl * nblk)), dim3(LongCfg<T>::NT), args, lds, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Row R2C of rows longer than one workgroup: pass A reads the real rows as
// pairs (or gathers them), pass B leaves natural-order half-length
// transforms, the split kernel writes the column-major half spectra.
template <class T>
hipError_t Fft2<T>::row_r2c_long(int mode, const T *in, size_t in_bstride, int in_pitch, int win, const ExpDev *ex,
                                 const T *wr, int W, Cx<T> *out, size_t out_bstride, int rp_rows, int rows, int batch,
                                 hipStream_t st) {
    const size_t lines = (size_t)rows * batch, M = (size_t)s.M;
    if (lines > (size_t)INT32_MAX) return hipErrorInvalidValue;
    hipError_t e = long_space(lines * M, lines * M, 0, st);
    if (e != hipSuccess) return e;
    LineIO src;
    src.p = in;
    src.bstride = in_bstride;
    src.pitch = (size_t)in_pitch;
    src.lpb = rows;
    src.nvalid = mode == ROW_PLAIN ? win : W;
    src.ex = ex;
    src.wr = wr;
    if ((e = long_a<T>(lrow, false, mode == ROW_PLAIN ? LL_REAL : LL_GATHER, src, lines, lt_a, st)) != hipSuccess)
        return e;
    LineIO mid;
    mid.p = lt_b;
    mid.pitch = M;
    mid.lpb = (int)lines;
    mid.nvalid = s.M;
    if ((e = long_b(lrow, false, LS_CX, lt_a, mid, lines, st)) != hipSuccess) return e;
    const dim3 g((s.M + 1 + 63) / 64, (rows + 15) / 16, batch);
    hipLaunchKernelGGL(k_long_split<T>, g, dim3(256), 0, st, lt_b, s.M, rows, post, out, out_bstride, rp(rp_rows));
    return hipGetLastError();
}

template <class T>
hipError_t Fft2<T>::row_c2r_long(const Cx<T> *in, size_t in_bstride, int rp_rows, T *out, size_t out_bstride,
                                 int out_pitch, int wout, T scale, int rows, int batch, hipStream_t st) {
    const size_t lines = (size_t)rows * batch, M = (size_t)s.M;
    if (lines > (size_t)INT32_MAX) return hipErrorInvalidValue;
    hipError_t e = long_space(lines * M, lines * M, 0, st);
    if (e != hipSuccess) return e;
    const dim3 g((s.M + 63) / 64, (rows + 15) / 16, batch);
    hipLaunchKernelGGL(k_long_merge<T>, g, dim3(256), 0, st, in, in_bstride, rp(rp_rows), s.M, rows, post, lt_b);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    LineIO z;
    z.p = lt_b;
    z.pitch = M;
    z.lpb = (int)lines;
    z.nvalid = s.M;
    if ((e = long_a<T>(lrow, true, LL_CX, z, lines, lt_a, st)) != hipSuccess) return e;
    LineIO dst;
    dst.p = out;
    dst.bstride = out_bstride;
    dst.pitch = (size_t)out_pitch;
    dst.lpb = rows;
    dst.nvalid = wout;
    dst.scale = (double)scale;
    return long_b(lrow, true, LS_REAL, lt_a, dst, lines, st);
}

// Column passes over columns longer than one workgroup, composed from the
// two passes: BCAST as X^ = FFT(in) into scratch, then inverse passes whose
// loads multiply by each PSF spectrum; ADJ / ACC as forward passes of every
// exposure into scratch, the summed product with conj(P^_k), and (ADJ) its
// inverse.
template <class T>
template <class TA>
hipError_t Fft2<T>::colpass_long(int mode, const Cx<T> *in, size_t in_bstride, int in_rows, Cx<T> *out,
                                 size_t out_bstride, int out_rows, const Cx<T> *psf, size_t psf_bstride, int hin,
                                 int hout, int batch, hipStream_t st) {
    if (batch < 0) return hipErrorInvalidValue;
    const int NC = s.NC, FH = s.FH;
    const size_t L = (size_t)FH, nl = (size_t)NC, nb = (size_t)batch;
    if (nb * nl > (size_t)INT32_MAX) return hipErrorInvalidValue;
    LineIO src;
    src.p = in;
    src.bstride = in_bstride;
    src.pitch = (size_t)rp(in_rows);
    src.lpb = NC;
    src.nvalid = std::min(hin, FH);
    LineIO dst;
    dst.p = out;
    dst.bstride = out_bstride;
    dst.pitch = (size_t)rp(out_rows);
    dst.lpb = NC;
    dst.nvalid = std::min(hout, FH);
    LineIO full;  // line-major scratch, FH elements per line
    full.p = lt_b;
    full.pitch = L;
    full.nvalid = FH;
    hipError_t e;
    switch (mode) {
        case COL_FWD:
        case COL_INV: {
            if ((e = long_space(nb * nl * L, 0, 0, st)) != hipSuccess) return e;
            const bool inv = mode == COL_INV;
            if ((e = long_a<T>(lcol, inv, LL_CX, src, nb * nl, lt_a, st)) != hipSuccess) return e;
            return long_b(lcol, inv, LS_CX, lt_a, dst, nb * nl, st);
        }
        case COL_BCAST: {
            if (nb == 0) return hipSuccess;
            if ((e = long_space(nb * nl * L, nl * L, 0, st)) != hipSuccess) return e;
            src.bstride = 0;
            if ((e = long_a<T>(lcol, false, LL_CX, src, nl, lt_a, st)) != hipSuccess) return e;
            full.p = lt_b;
            full.lpb = NC;
            if ((e = long_b(lcol, false, LS_CX, lt_a, full, nl, st)) != hipSuccess) return e;
            LineIO mul = full;
            mul.psf = psf;
            mul.pbstride = psf_bstride;
            mul.ppitch = (size_t)rp(FH);
            if ((e = long_a<T>(lcol, true, LL_CX_MUL, mul, nb * nl, lt_a, st)) != hipSuccess) return e;
            return long_b(lcol, true, LS_CX, lt_a, dst, nb * nl, st);
        }
        case COL_ADJ:
        case COL_ACC: {
            if ((e = long_space(nb * nl * L, nb * nl * L, nl * L * sizeof(Cx<TA>), st)) != hipSuccess) return e;
            if ((e = long_a<T>(lcol, false, LL_CX, src, nb * nl, lt_a, st)) != hipSuccess) return e;
            full.p = lt_b;
            full.lpb = (int)std::max<size_t>(nb * nl, 1);
            if ((e = long_b(lcol, false, LS_CX, lt_a, full, nb * nl, st)) != hipSuccess) return e;
            Cx<TA> *acc = mode == COL_ACC ? reinterpret_cast<Cx<TA> *>(out) : reinterpret_cast<Cx<TA> *>(lt_c);
            const int opitch = mode == COL_ACC ? rp(out_rows) : FH;
            const unsigned g = (unsigned)std::min<size_t>((nl * L + 255) / 256, 16384);
            hipLaunchKernelGGL((k_long_adjsum<T, TA>), dim3(g), dim3(256), 0, st, lt_b, NC, FH, batch, psf,
                               psf_bstride, rp(FH), acc, opitch);
            if ((e = hipGetLastError()) != hipSuccess || mode == COL_ACC) return e;
            LineIO sum;
            sum.p = lt_c;
            sum.pitch = L;
            sum.lpb = NC;
            sum.nvalid = FH;
            if ((e = long_a<TA>(lcol, true, LL_CX, sum, nl, lt_a, st)) != hipSuccess) return e;
            dst.bstride = 0;
            return long_b(lcol, true, LS_CX, lt_a, dst, nl, st);
        }
        default: return hipErrorInvalidValue;
    }
}

template struct Fft2<float>;
template struct Fft2<double>;
template hipError_t Fft2<float>::row_c2r_update<float>(const Cx<float> *, int, float *, const float *, int, int,
                                                       float, float, IterDiag *, double *, unsigned long long *,
                                                       unsigned int *, hipStream_t, int);
template hipError_t Fft2<double>::row_c2r_update<float>(const Cx<double> *, int, float *, const double *, int, int,
                                                        double, double, IterDiag *, double *, unsigned long long *,
                                                        unsigned int *, hipStream_t, int);
template hipError_t Fft2<double>::row_c2r_update<double>(const Cx<double> *, int, double *, const double *, int,
                                                         int, double, double, IterDiag *, double *,
                                                         unsigned long long *, unsigned int *, hipStream_t, int);
template hipError_t Fft2<float>::spec_acc<float>(const Cx<float> *, const Cx<float> *, Cx<float> *, hipStream_t);
template hipError_t Fft2<float>::spec_acc<double>(const Cx<float> *, const Cx<float> *, Cx<double> *, hipStream_t);
template hipError_t Fft2<double>::spec_acc<double>(const Cx<double> *, const Cx<double> *, Cx<double> *, hipStream_t);
template hipError_t Fft2<float>::spec_cast<float>(const Cx<float> *, Cx<float> *, hipStream_t);
template hipError_t Fft2<float>::spec_cast<double>(const Cx<double> *, Cx<float> *, hipStream_t);
template hipError_t Fft2<double>::spec_cast<double>(const Cx<double> *, Cx<double> *, hipStream_t);
template hipError_t Fft2<float>::colpass<float>(int, const Cx<float> *, size_t, int, Cx<float> *, size_t, int,
                                                const Cx<float> *, size_t, int, int, int, hipStream_t, int);
template hipError_t Fft2<float>::colpass<double>(int, const Cx<float> *, size_t, int, Cx<float> *, size_t, int,
                                                 const Cx<float> *, size_t, int, int, int, hipStream_t, int);
template hipError_t Fft2<double>::colpass<double>(int, const Cx<double> *, size_t, int, Cx<double> *, size_t, int,
                                                  const Cx<double> *, size_t, int, int, int, hipStream_t, int);
template hipError_t Fft2<float>::colacc<double>(const Cx<float> *, size_t, int, Cx<double> *, int, int,
                                                const Cx<float> *, size_t, int, int, const FftLen *, hipStream_t, int);
template hipError_t Fft2<float>::acc_inverse<double>(const Cx<double> *, Cx<double> *, int, int, const FftLen &,
                                                     hipStream_t);

}  // namespace updc
