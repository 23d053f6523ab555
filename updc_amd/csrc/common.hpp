// Shared device/host helpers for the MI355X (gfx950) UPDC library.
//
// Everything here is written for CDNA4 directly: 64-lane wavefronts, LDS
// resident transforms, no CUDA shims.  The whole library is compiled with
// -ffp-contract=off so that only the explicit fma()/fmaf() calls fuse, which
// is what keeps the detector and geometry kernels bit-identical to the
// reference's float32 C core (built -std=c11, updc/c_core/Makefile:3).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define UPDC_HD __host__ __device__ __forceinline__

// ---------------------------------------------------------------- complex
template <class T> struct Cx { T re, im; };

template <class T> UPDC_HD Cx<T> cx(T r, T i) { Cx<T> c; c.re = r; c.im = i; return c; }
template <class T> UPDC_HD Cx<T> operator+(Cx<T> a, Cx<T> b) { return cx<T>(a.re + b.re, a.im + b.im); }
template <class T> UPDC_HD Cx<T> operator-(Cx<T> a, Cx<T> b) { return cx<T>(a.re - b.re, a.im - b.im); }
template <class T> UPDC_HD Cx<T> cmul(Cx<T> a, Cx<T> b) { return cx<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
// a * conj(b)
template <class T> UPDC_HD Cx<T> cmulc(Cx<T> a, Cx<T> b) { return cx<T>(a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im); }
template <class T> UPDC_HD Cx<T> cscale(Cx<T> a, T s) { return cx<T>(a.re * s, a.im * s); }
template <class T> UPDC_HD Cx<T> cconj(Cx<T> a) { return cx<T>(a.re, -a.im); }
// multiply by -i (forward) or +i (inverse)
template <bool INV, class T> UPDC_HD Cx<T> rot90(Cx<T> a) { return INV ? cx<T>(-a.im, a.re) : cx<T>(a.im, -a.re); }

// ---------------------------------------------------------------- math consts
// cos/sin(2*pi*m/P) for the odd prime radices, evaluated to 22 digits from
// their series (plain mathematics, no library internals).
template <int P> struct OddRoots;
template <> struct OddRoots<3> {
    static constexpr double c[2] = {1.0, -0.5};
    static constexpr double s[2] = {0.0, 0.8660254037844386467637};
};
template <> struct OddRoots<5> {
    static constexpr double c[3] = {1.0, 0.3090169943749474241022, -0.8090169943749474241023};
    static constexpr double s[3] = {0.0, 0.9510565162951535721164, 0.5877852522924731291687};
};
template <> struct OddRoots<7> {
    static constexpr double c[4] = {1.0, 0.6234898018587335305250, -0.2225209339563144042889, -0.9009688679024191262361};
    static constexpr double s[4] = {0.0, 0.7818314824680298087084, 0.9749279121818236070181, 0.4338837391175581204757};
};
#define UPDC_SQRT1_2 0.7071067811865475244008

// ---------------------------------------------------------------- FFT arithmetic
// The transforms' internal products use explicit fused multiply-adds (the
// build contracts nothing by itself): a twiddle product in two multiplies and
// two fmas instead of four multiplies and two adds, a butterfly's cosine and
// sine terms as fmas.  Every transform kernel inlines these same helpers, so
// the plain and batched forms still do identical arithmetic.  The spectrum
// products of the iteration (PSF multiply, adjoint accumulation) keep the
// reference's unfused form and do not use them.
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <class T> __device__ __forceinline__ Cx<T> cmul_f(Cx<T> a, Cx<T> w) {
    return cx<T>(fma_t(a.re, w.re, -(a.im * w.im)), fma_t(a.re, w.im, a.im * w.re));
}

// ---------------------------------------------------------------- butterflies
// In-place DFT of length R on v[0..R): y_j = sum_q v_q exp(-+2 pi i j q / R)
// (minus for forward, plus for inverse), unnormalised.
template <bool INV, class T> __device__ __forceinline__ void bfly2(Cx<T> *v) {
    Cx<T> a = v[0], b = v[1];
    v[0] = a + b; v[1] = a - b;
}
template <bool INV, class T> __device__ __forceinline__ void bfly4(Cx<T> *v) {
    Cx<T> a0 = v[0] + v[2], a1 = v[0] - v[2];
    Cx<T> b0 = v[1] + v[3], b1 = rot90<INV>(v[1] - v[3]);
    v[0] = a0 + b0; v[2] = a0 - b0;
    v[1] = a1 + b1; v[3] = a1 - b1;
}
template <bool INV, class T> __device__ __forceinline__ void bfly8(Cx<T> *v) {
    // two radix-4 on even/odd samples, then radix-2 with w8 twiddles
    Cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    Cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    bfly4<INV>(e);
    bfly4<INV>(o);
    const T h = (T)UPDC_SQRT1_2;
    // o1 *= w8^1, o2 *= w8^2 = -+i, o3 *= w8^3
    Cx<T> t1 = o[1], t3 = o[3];
    if (!INV) {
        o[1] = cx<T>((t1.re + t1.im) * h, (t1.im - t1.re) * h);
        o[3] = cx<T>((t3.im - t3.re) * h, -(t3.re + t3.im) * h);
    } else {
        o[1] = cx<T>((t1.re - t1.im) * h, (t1.im + t1.re) * h);
        o[3] = cx<T>(-(t3.re + t3.im) * h, (t3.re - t3.im) * h);
    }
    o[2] = rot90<INV>(o[2]);
    for (int j = 0; j < 4; ++j) { v[j] = e[j] + o[j]; v[j + 4] = e[j] - o[j]; }
}
template <int P, bool INV, class T> __device__ __forceinline__ void bfly_odd(Cx<T> *v) {
    constexpr int H = (P - 1) / 2;
    Cx<T> a[H + 1], b[H + 1];
    Cx<T> sum = v[0];
#pragma unroll
    for (int m = 1; m <= H; ++m) {
        a[m] = v[m] + v[P - m];
        b[m] = v[m] - v[P - m];
        sum = sum + a[m];
    }
    Cx<T> out[P];
    out[0] = sum;
#pragma unroll
    for (int j = 1; j <= H; ++j) {
        Cx<T> re = v[0], im = cx<T>((T)0, (T)0);
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            const int jm = (j * m) % P;
            const T c = (T)OddRoots<P>::c[jm <= H ? jm : P - jm];
            const T s = (jm <= H) ? (T)OddRoots<P>::s[jm] : -(T)OddRoots<P>::s[P - jm];
            re = cx<T>(fma_t(a[m].re, c, re.re), fma_t(a[m].im, c, re.im));
            im = cx<T>(fma_t(b[m].re, s, im.re), fma_t(b[m].im, s, im.im));
        }
        // forward: y_j = re - i*im, y_{P-j} = re + i*im
        Cx<T> iim = cx<T>(-im.im, im.re);
        if (!INV) { out[j] = re - iim; out[P - j] = re + iim; }
        else { out[j] = re + iim; out[P - j] = re - iim; }
    }
#pragma unroll
    for (int j = 0; j < P; ++j) v[j] = out[j];
}
// Radix 10 as a prime-factor (Good-Thomas) 2 x 5 DFT: input n = (5 n1 + 2 n2)
// mod 10 and output k = (5 k1 + 6 k2) mod 10 make exp(-+2 pi i n k / 10) the
// product of the 2- and 5-point kernels, so no twiddle sits between the two
// butterflies (no rounding beyond theirs).
template <bool INV, class T> __device__ __forceinline__ void bfly10(Cx<T> *v) {
    Cx<T> a[5] = {v[0], v[2], v[4], v[6], v[8]};
    Cx<T> b[5] = {v[5], v[7], v[9], v[1], v[3]};
    bfly_odd<5, INV>(a);
    bfly_odd<5, INV>(b);
    v[0] = a[0] + b[0]; v[5] = a[0] - b[0];
    v[6] = a[1] + b[1]; v[1] = a[1] - b[1];
    v[2] = a[2] + b[2]; v[7] = a[2] - b[2];
    v[8] = a[3] + b[3]; v[3] = a[3] - b[3];
    v[4] = a[4] + b[4]; v[9] = a[4] - b[4];
}
template <bool INV, class T> __device__ __forceinline__ void bfly16(Cx<T> *v) {
    // 16 = 4 x 4: X[k1 + 4 k2] = sum_n2 W4^(n2 k2) W16^(n2 k1) sum_n1 W4^(n1 k1) x[4 n1 + n2]
    constexpr double c[10] = {1.0, 0.9238795325112867561282, 0.7071067811865475244008, 0.3826834323650897717285,
                              0.0, -0.3826834323650897717285, -0.7071067811865475244008, -0.9238795325112867561282,
                              -1.0, -0.9238795325112867561282};
    constexpr double s[10] = {0.0, 0.3826834323650897717285, 0.7071067811865475244008, 0.9238795325112867561282,
                              1.0, 0.9238795325112867561282, 0.7071067811865475244008, 0.3826834323650897717285,
                              0.0, -0.3826834323650897717285};  // cos, sin (2 pi e / 16), e < 10
    Cx<T> y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        Cx<T> t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
        bfly4<INV>(t);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = t[k1];
    }
#pragma unroll
    for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
        for (int k1 = 1; k1 < 4; ++k1) {
            const int e = n2 * k1;
            y[n2][k1] = cmul_f(y[n2][k1], cx<T>((T)c[e], INV ? (T)s[e] : -(T)s[e]));
        }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        Cx<T> t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
        bfly4<INV>(t);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
}
template <int R, bool INV, class T> __device__ __forceinline__ void bfly(Cx<T> *v) {
    if constexpr (R == 2) bfly2<INV>(v);
    else if constexpr (R == 4) bfly4<INV>(v);
    else if constexpr (R == 8) bfly8<INV>(v);
    else if constexpr (R == 16) bfly16<INV>(v);
    else if constexpr (R == 10) bfly10<INV>(v);
    else bfly_odd<R, INV>(v);
}

// ---------------------------------------------------------------- FFT plan
#define UPDC_MAX_STAGES 16
// One transform length, split into Stockham radix stages.  tw points at the
// device twiddle tables of the FFT's real type, one per stage: stage s (radix
// R, Ns = product of the earlier radices) stores exp(-2 pi i q k / (Ns R)) at
// tw[to[s] + k (R-1) + q - 1] for k < Ns, 1 <= q < R, so the R-1 twiddles of a
// butterfly are adjacent and a wave's twiddle reads are contiguous.
// For stage s, ns[s] is the product of the earlier radices and mg[s] =
// ceil(2^32 / ns[s]), so i / ns[s] = umulhi(i, mg[s]) exactly for i, ns < 2^16
// (ns > 1; the first stage, ns = 1, divides trivially).
struct FftLen {
    int L;
    int nst;
    int rad[UPDC_MAX_STAGES];
    int ns[UPDC_MAX_STAGES];
    unsigned int mg[UPDC_MAX_STAGES];
    int to[UPDC_MAX_STAGES];
    const void *tw;
};

// One Stockham (decimation in time, self-sorting) radix-R stage over
// `1 << lg` interleaved transforms held in LDS: element n of line c lives at
// buf[(n << lg) + c].  Butterfly b handles line c = b & (lines-1), index
// i = b >> lg, reads the R values i + q L/R, applies the twiddles
// exp(-+2 pi i q k / (Ns R)) with k = i mod Ns, and writes the outputs to
// (i / Ns) Ns R + k + q Ns.  All reads finish before the barrier, so the stage
// runs in place in one LDS buffer.  VMAX bounds the complex values one thread
// holds per stage; the host checks R * ceil(nb / NT) <= VMAX.
template <int R, bool INV, class T, int NT, int VMAX>
__device__ __forceinline__ void fft_stage(Cx<T> *buf, int L, int lg, int cfix, int Ns, unsigned int mg,
                                          const Cx<T> *__restrict__ tws) {
    // cfix >= 0: transform only line cfix of the interleaved layout
    constexpr int NB = VMAX / R;
    static_assert(NB >= 1, "VMAX too small for radix");
    const bool one = cfix >= 0;
    const int lm = (1 << lg) - 1;
    // Redefine L/R opaquely (it stays in an SGPR): otherwise the compiler
    // hoists every radix case's per-thread LDS addresses out of the stage
    // loop and spills them to scratch.
    int LR = L / R;
    asm volatile("" : "+s"(LR));
    const int nb = one ? LR : (LR << lg);
    Cx<T> v[NB][R];
    const int tid = threadIdx.x;
#pragma unroll
    for (int r = 0; r < NB; ++r) {
        const int b = tid + r * NT;
        if (b < nb) {
            const int c = one ? cfix : (b & lm), i = one ? b : (b >> lg);
#pragma unroll
            for (int q = 0; q < R; ++q) v[r][q] = buf[((i + q * LR) << lg) + c];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NB; ++r) {
        const int b = tid + r * NT;
        if (b < nb) {
            const int c = one ? cfix : (b & lm), i = one ? b : (b >> lg);
            // Ns == 1 (first stage): mg = 2^32 does not fit, and i / 1 = i
            const int hi = Ns == 1 ? i : (int)__umulhi((unsigned int)i, mg);
            const int k = i - hi * Ns;
            if (Ns > 1) {
                const Cx<T> *wk = tws + k * (R - 1);
#pragma unroll
                for (int q = 1; q < R; ++q) {
                    Cx<T> w = wk[q - 1];
                    if (INV) w.im = -w.im;
                    v[r][q] = cmul_f(v[r][q], w);
                }
            }
            bfly<R, INV>(v[r]);
            const int j = hi * Ns * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) buf[((j + q * Ns) << lg) + c] = v[r][q];
        }
    }
    __syncthreads();
}

// Stages [s0, s1) over the LDS buffer (single line, lg = 0).
template <bool INV, class T, int NT, int VMAX>
__device__ __forceinline__ void fft_lds_range(Cx<T> *buf, const FftLen &p, int s0, int s1) {
    for (int s = s0; s < s1; ++s) {
        const int R = p.rad[s], Ns = p.ns[s];
        const unsigned int mg = p.mg[s];
        const Cx<T> *tw = (const Cx<T> *)p.tw + p.to[s];
        switch (R) {
            case 2: fft_stage<2, INV, T, NT, VMAX>(buf, p.L, 0, -1, Ns, mg, tw); break;
            case 3: fft_stage<3, INV, T, NT, VMAX>(buf, p.L, 0, -1, Ns, mg, tw); break;
            case 4: fft_stage<4, INV, T, NT, VMAX>(buf, p.L, 0, -1, Ns, mg, tw); break;
            case 5: fft_stage<5, INV, T, NT, VMAX>(buf, p.L, 0, -1, Ns, mg, tw); break;
            case 7: fft_stage<7, INV, T, NT, VMAX>(buf, p.L, 0, -1, Ns, mg, tw); break;
            case 16:
                if constexpr (VMAX >= 16) fft_stage<16, INV, T, NT, VMAX>(buf, p.L, 0, -1, Ns, mg, tw);
                break;
            default: fft_stage<8, INV, T, NT, VMAX>(buf, p.L, 0, -1, Ns, mg, tw); break;
        }
    }
}

// The same stage out of place: reads `in`, writes `out` (two LDS buffers), so
// one barrier (after the writes) separates it from the next stage instead of
// two.  Single line (lg = 0).
template <int R, bool INV, class T, int NT, int VMAX>
__device__ __forceinline__ void fft_stage_pp(const Cx<T> *__restrict__ in, Cx<T> *__restrict__ out, int L, int Ns,
                                             unsigned int mg, const Cx<T> *__restrict__ tws) {
    constexpr int NB = VMAX / R;
    static_assert(NB >= 1, "VMAX too small for radix");
    int LR = L / R;
    asm volatile("" : "+s"(LR));
    const int tid = threadIdx.x;
#pragma unroll
    for (int r = 0; r < NB; ++r) {
        const int i = tid + r * NT;
        if (i < LR) {
            Cx<T> v[R];
#pragma unroll
            for (int q = 0; q < R; ++q) v[q] = in[i + q * LR];
            const int hi = Ns == 1 ? i : (int)__umulhi((unsigned int)i, mg);
            const int k = i - hi * Ns;
            if (Ns > 1) {
                const Cx<T> *wk = tws + k * (R - 1);
#pragma unroll
                for (int q = 1; q < R; ++q) {
                    Cx<T> w = wk[q - 1];
                    if (INV) w.im = -w.im;
                    v[q] = cmul_f(v[q], w);
                }
            }
            bfly<R, INV>(v);
            const int j = hi * Ns * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) out[j + q * Ns] = v[q];
        }
    }
    __syncthreads();
}
// Stages [s0, s1) out of place, alternating between the buffers *cur and
// *oth; on return *cur holds the result.  twb: the plan's stage twiddle
// tables (p.tw or a copy of them, e.g. in LDS).
template <bool INV, class T, int NT, int VMAX>
__device__ __forceinline__ void fft_lds_range_pp(Cx<T> *&cur, Cx<T> *&oth, const FftLen &p, const Cx<T> *twb,
                                                 int s0, int s1) {
    for (int s = s0; s < s1; ++s) {
        const int R = p.rad[s], Ns = p.ns[s];
        const unsigned int mg = p.mg[s];
        const Cx<T> *tw = twb + p.to[s];
        switch (R) {
            case 2: fft_stage_pp<2, INV, T, NT, VMAX>(cur, oth, p.L, Ns, mg, tw); break;
            case 3: fft_stage_pp<3, INV, T, NT, VMAX>(cur, oth, p.L, Ns, mg, tw); break;
            case 4: fft_stage_pp<4, INV, T, NT, VMAX>(cur, oth, p.L, Ns, mg, tw); break;
            case 5: fft_stage_pp<5, INV, T, NT, VMAX>(cur, oth, p.L, Ns, mg, tw); break;
            case 7: fft_stage_pp<7, INV, T, NT, VMAX>(cur, oth, p.L, Ns, mg, tw); break;
            case 16:
                if constexpr (VMAX >= 16) fft_stage_pp<16, INV, T, NT, VMAX>(cur, oth, p.L, Ns, mg, tw);
                break;
            default: fft_stage_pp<8, INV, T, NT, VMAX>(cur, oth, p.L, Ns, mg, tw); break;
        }
        Cx<T> *t = cur;
        cur = oth;
        oth = t;
    }
}

// Register-resident edge stages of a transform whose first and last radix
// are both RE (LR = L / RE).  Thread tid owns butterflies i_r = tid + r NT
// (i_r < LR) and the RE values of each at element positions i_r + q LR, the
// positions the first stage reads and the last stage writes.
//   edge_first: stage 0 (Ns = 1, no twiddles) from registers into LDS
//               (outputs at i RE + q), then a barrier.
//   edge_last : the last stage (Ns = LR, k = i) from LDS back into registers,
//               then (SYNC) a barrier so the buffer may be rewritten.
template <int RE, bool INV, class T, int NT, int NBE>
__device__ __forceinline__ void edge_first(Cx<T> (&rb)[NBE][RE], Cx<T> *buf, int LR) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int r = 0; r < NBE; ++r) {
        const int i = tid + r * NT;
        if (i < LR) {
            Cx<T> v[RE];
#pragma unroll
            for (int q = 0; q < RE; ++q) v[q] = rb[r][q];
            bfly<RE, INV>(v);
#pragma unroll
            for (int q = 0; q < RE; ++q) buf[i * RE + q] = v[q];
        }
    }
    __syncthreads();
}
template <int RE, bool INV, class T, int NT, int NBE, bool SYNC = true>
__device__ __forceinline__ void edge_last(Cx<T> (&rb)[NBE][RE], const Cx<T> *buf, int LR,
                                          const Cx<T> *__restrict__ tws) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int r = 0; r < NBE; ++r) {
        const int i = tid + r * NT;
        if (i < LR) {
            Cx<T> v[RE];
#pragma unroll
            for (int q = 0; q < RE; ++q) v[q] = buf[i + q * LR];
            const Cx<T> *wk = tws + i * (RE - 1);
#pragma unroll
            for (int q = 1; q < RE; ++q) {
                Cx<T> w = wk[q - 1];
                if (INV) w.im = -w.im;
                v[q] = cmul_f(v[q], w);
            }
            bfly<RE, INV>(v);
#pragma unroll
            for (int q = 0; q < RE; ++q) rb[r][q] = v[q];
        } else {
            // a full definition, so the registers are dead between the edges
#pragma unroll
            for (int q = 0; q < RE; ++q) rb[r][q] = cx<T>((T)0, (T)0);
        }
    }
    if (SYNC) __syncthreads();
}

// Full transform over the LDS buffer (natural order in, natural order out):
// all `1 << lg` interleaved lines together (cfix < 0) or line cfix only.
// fft_lds_hook reads the stage twiddles from twb (the tables of p copied
// elsewhere, e.g. into LDS) instead of p.tw, and calls hook(s) after stage s.
template <bool INV, class T, int NT, int VMAX, class Hook>
__device__ __forceinline__ void fft_lds_hook(Cx<T> *buf, const FftLen &p, const Cx<T> *twb, int lg, int cfix,
                                             Hook &hook) {
    for (int s = 0; s < p.nst; ++s) {
        const int R = p.rad[s], Ns = p.ns[s];
        const unsigned int mg = p.mg[s];
        const Cx<T> *tw = twb + p.to[s];
        switch (R) {
            case 2: fft_stage<2, INV, T, NT, VMAX>(buf, p.L, lg, cfix, Ns, mg, tw); break;
            case 3: fft_stage<3, INV, T, NT, VMAX>(buf, p.L, lg, cfix, Ns, mg, tw); break;
            case 4: fft_stage<4, INV, T, NT, VMAX>(buf, p.L, lg, cfix, Ns, mg, tw); break;
            case 5: fft_stage<5, INV, T, NT, VMAX>(buf, p.L, lg, cfix, Ns, mg, tw); break;
            case 7: fft_stage<7, INV, T, NT, VMAX>(buf, p.L, lg, cfix, Ns, mg, tw); break;
            case 16:
                if constexpr (VMAX >= 16) fft_stage<16, INV, T, NT, VMAX>(buf, p.L, lg, cfix, Ns, mg, tw);
                break;
            default: fft_stage<8, INV, T, NT, VMAX>(buf, p.L, lg, cfix, Ns, mg, tw); break;
        }
        hook(s);
    }
}
template <bool INV, class T, int NT, int VMAX>
__device__ __forceinline__ void fft_lds_t(Cx<T> *buf, const FftLen &p, const Cx<T> *twb, int lg, int cfix = -1) {
    auto none = [](int) {};
    fft_lds_hook<INV, T, NT, VMAX>(buf, p, twb, lg, cfix, none);
}
template <bool INV, class T, int NT, int VMAX>
__device__ __forceinline__ void fft_lds(Cx<T> *buf, const FftLen &p, int lg, int cfix = -1) {
    fft_lds_t<INV, T, NT, VMAX>(buf, p, (const Cx<T> *)p.tw, lg, cfix);
}

// ---------------------------------------------------------------- XCD-aware tiles
// Workgroups are dealt round-robin to the 8 XCDs.  Remap so each XCD walks
// a contiguous range of tiles: neighbouring column tiles then share L2 lines
// of the same XCD.  Returns -1 for padding workgroups.
__device__ __forceinline__ int xcd_tile(int bid, int ntiles) {
    const int per = (ntiles + 7) / 8;
    const int t = (bid % 8) * per + bid / 8;
    return t < ntiles ? t : -1;
}
// As xcd_tile, with every XCD's range starting at an even tile, so that
// tiles 2p and 2p + 1 run on one XCD, back to back.  Needs
// 16 * ceil(ntiles / 16) workgroups.
__device__ __forceinline__ int xcd_tile_even(int bid, int ntiles) {
    const int per = (ntiles + 15) / 16 * 2;
    const int t = (bid % 8) * per + bid / 8;
    return t < ntiles ? t : -1;
}

// ---------------------------------------------------------------- ordered keys
// Monotone float/double -> unsigned maps for atomicMin of signed values.
__device__ __forceinline__ unsigned int fkey(float f) {
    unsigned int u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ unsigned long long dkey(double d) {
    unsigned long long u = (unsigned long long)__double_as_longlong(d);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
UPDC_HD float fkey_inv(unsigned int u) {
    u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
UPDC_HD double dkey_inv(unsigned long long u) {
    u = (u & 0x8000000000000000ull) ? (u & 0x7fffffffffffffffull) : ~u;
    double d;
    __builtin_memcpy(&d, &u, 8);
    return d;
}

// fused multiply-add in the storage type (fmaf for float, fma for double)
__device__ __forceinline__ float fmat(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmat(double a, double b, double c) { return __builtin_fma(a, b, c); }

// ---------------------------------------------------------------- global pointers
// A pointer the kernel reads out of device memory (the ExpDev records) is a
// generic pointer to the compiler, and loads through it become flat loads:
// 64-bit VGPR addresses, and every flat load also counts against the LDS
// counter, so an LDS read after it waits for the memory load as well.  gp()
// states that the pointer addresses global memory (all of these do).
// (The host pass of the same source keeps plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T> using gptr_t = __attribute__((address_space(1))) T *;
#else
template <class T> using gptr_t = T *;
#endif
template <class T> __device__ __forceinline__ gptr_t<const T> gp(const T *p) { return (gptr_t<const T>)p; }
// Read-only for the whole launch, and read at wave-uniform addresses: the
// constant address space lets the compiler use scalar loads that it waits for
// where the value is used (a vector load of a uniform value is moved to a
// scalar register right away, which waits for every load issued before it).
#if defined(__HIP_DEVICE_COMPILE__)
template <class T> using cptr_t = __attribute__((address_space(4))) T *;
#else
template <class T> using cptr_t = T *;
#endif
template <class T> __device__ __forceinline__ cptr_t<const T> cp(const T *p) { return (cptr_t<const T>)p; }
template <class T> __device__ __forceinline__ gptr_t<T> gp(T *p) { return (gptr_t<T>)p; }

// ---------------------------------------------------------------- wave scan
// Inclusive prefix sum over the 64 lanes of a wave with DPP row shifts and the
// row broadcasts of the GFX9 family (no LDS round trip).  Every lane of the
// wave must be active.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    int v = (int)x;
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
    return (uint32_t)v;
}
