// Non-FFT kernels of the MI355X UPDC library: LinkMap geometry, scans, the
// atomic-free CSC transpose, weight normalisation, initial drizzle, the fused
// gather + FreezeMu detector kernel and the multiplicative update.
//
// Compiled with -ffp-contract=off: every f64 clipping operation and every f32
// area/fraction operation rounds exactly as the reference's C11 build does,
// and only the explicit fma()/fmaf() calls fuse.
#include "stages.hpp"
#include <algorithm>

namespace updc {

// ===================================================================== geometry
struct DPt { double x, y; };

// Boundary tests and crossings of overlap.c:14-43 (0 left, 1 right, 2 bottom, 3 top).
__device__ __forceinline__ bool g_in(DPt p, int side, double v) {
    return side == 0 ? p.x >= v : side == 1 ? p.x <= v : side == 2 ? p.y >= v : p.y <= v;
}
__device__ __forceinline__ DPt g_cross(DPt a, DPt b, int side, double v) {
    DPt r = a;
    if (side < 2) {
        const double d = b.x - a.x;
        if (fabs(d) > 1.0e-15) {
            const double t = (v - a.x) / d;
            r.x = v;
            r.y = a.y + t * (b.y - a.y);
        }
    } else {
        const double d = b.y - a.y;
        if (fabs(d) > 1.0e-15) {
            const double t = (v - a.y) / d;
            r.x = a.x + t * (b.x - a.x);
            r.y = v;
        }
    }
    return r;
}
// Local-origin f32 shoelace (overlap.c:72-91).
__device__ __forceinline__ float g_area(const DPt *p, int n) {
    if (n < 3) return 0.0f;
    const double ox = p[0].x, oy = p[0].y;
    float acc = 0.0f;
    for (int i = 0; i < n; ++i) {
        const int j = (i + 1 == n) ? 0 : i + 1;
        const float xi = (float)(p[i].x - ox), yi = (float)(p[i].y - oy);
        const float xj = (float)(p[j].x - ox), yj = (float)(p[j].y - oy);
        acc += __builtin_fmaf(xi, yj, -(xj * yi));
    }
    return 0.5f * fabsf(acc);
}
// Quad vs unit pixel centred on (x, y): four Sutherland-Hodgman passes
// (overlap.c:45-70, 93-116).  A convex quad clipped by 4 half-planes has at
// most 8 vertices; the buffers hold 16 like the reference's.
__device__ float g_overlap(const DPt q[4], int x, int y) {
    DPt a[16], b[16];
    DPt *in = a, *out = b;
    int n = 4;
    for (int j = 0; j < 4; ++j) in[j] = q[j];
    for (int s = 0; s < 4; ++s) {
        const double v = (s == 0) ? (double)x - 0.5 : (s == 1) ? (double)x + 0.5 : (s == 2) ? (double)y - 0.5
                                                                                              : (double)y + 0.5;
        int m = 0;
        DPt prev = in[n - 1];
        bool pin = g_in(prev, s, v);
        for (int i = 0; i < n; ++i) {
            const DPt cur = in[i];
            const bool cin = g_in(cur, s, v);
            if (cin) {
                if (!pin && m < 16) out[m++] = g_cross(prev, cur, s, v);
                if (m < 16) out[m++] = cur;
            } else if (pin) {
                if (m < 16) out[m++] = g_cross(prev, cur, s, v);
            }
            prev = cur;
            pin = cin;
        }
        n = m;
        if (n < 3) return 0.0f;
        DPt *t = in; in = out; out = t;
    }
    return g_area(in, n);
}
// Pixel quad, full area and clamped candidate box (updc_core.c:512-528).
__device__ bool g_pixel(const double *__restrict__ v, int iw, int W, int H, int x, int y, DPt q[4], float &full,
                        int &x0, int &x1, int &y0, int &y1) {
    const size_t a = 2 * ((size_t)y * (size_t)(iw + 1) + (size_t)x), b = a + 2 * (size_t)(iw + 1);
    q[0] = {v[a], v[a + 1]};
    q[1] = {v[b], v[b + 1]};
    q[2] = {v[b + 2], v[b + 3]};
    q[3] = {v[a + 2], v[a + 3]};
    full = g_area(q, 4);
    if (!isfinite(full) || full <= 1.0e-12f) return false;
    double mnx = q[0].x, mxx = mnx, mny = q[0].y, mxy = mny;
    for (int j = 1; j < 4; ++j) {
        mnx = q[j].x < mnx ? q[j].x : mnx;
        mxx = q[j].x > mxx ? q[j].x : mxx;
        mny = q[j].y < mny ? q[j].y : mny;
        mxy = q[j].y > mxy ? q[j].y : mxy;
    }
    if (!isfinite(mnx) || !isfinite(mxx) || !isfinite(mny) || !isfinite(mxy) || mxx < -.5 || mxy < -.5 ||
        mnx > W - .5 || mny > H - .5)
        return false;
    x0 = (int)floor(mnx + .5); x1 = (int)floor(mxx + .5);
    y0 = (int)floor(mny + .5); y1 = (int)floor(mxy + .5);
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 >= W) x1 = W - 1;
    if (y1 >= H) y1 = H - 1;
    return x0 <= x1 && y0 <= y1;
}

__global__ __launch_bounds__(256) void k_lm_count(const double *__restrict__ vert, int iw, int ih, int W, int H,
                                                  uint32_t *__restrict__ cnt, uint32_t *__restrict__ first,
                                                  unsigned int *flags) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (size_t)iw * ih) return;
    const int x = (int)(p % iw), y = (int)(p / iw);
    DPt q[4];
    float full;
    int x0, x1, y0, y1;
    uint32_t n = 0, f = 0, l = 0;
    if (g_pixel(vert, iw, W, H, x, y, q, full, x0, x1, y0, y1)) {
        for (int ty = y0; ty <= y1; ++ty)
            for (int tx = x0; tx <= x1; ++tx) {
                const float area = g_overlap(q, tx, ty);
                if (area <= 1.0e-5f) continue;
                const uint32_t t = (uint32_t)((size_t)ty * W + tx);
                if (n == 0) f = t;
                l = t;
                ++n;
            }
    }
    cnt[p] = n;
    first[p] = n ? f : 0u;
    if (n && l - f > 0xffffu) atomicOr(flags, 1u);
}

__global__ __launch_bounds__(256) void k_lm_fill(const double *__restrict__ vert, int iw, int ih, int W, int H,
                                                 const uint32_t *__restrict__ off, const uint32_t *__restrict__ base,
                                                 uint32_t encoding, uint16_t *__restrict__ d16,
                                                 uint32_t *__restrict__ t32, float *__restrict__ frac,
                                                 unsigned int *flags) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (size_t)iw * ih) return;
    const int x = (int)(p % iw), y = (int)(p / iw);
    DPt q[4];
    float full;
    int x0, x1, y0, y1;
    uint32_t z = off[p];
    const uint32_t zend = off[p + 1];
    if (g_pixel(vert, iw, W, H, x, y, q, full, x0, x1, y0, y1)) {
        for (int ty = y0; ty <= y1; ++ty)
            for (int tx = x0; tx <= x1; ++tx) {
                const float area = g_overlap(q, tx, ty);
                if (area <= 1.0e-5f) continue;
                const uint32_t t = (uint32_t)((size_t)ty * W + tx);
                if (z >= zend) { atomicOr(flags, 2u); return; }
                if (encoding == UPDC_LM_DELTA16) {
                    const uint32_t d = t - base[p];
                    if (d > 0xffffu) { atomicOr(flags, 4u); return; }
                    d16[z] = (uint16_t)d;
                } else {
                    t32[z] = t;
                }
                frac[z] = area / full;
                ++z;
            }
    }
    if (z != zend) atomicOr(flags, 8u);
}

hipError_t lm_count(const double *vert, int iw, int ih, int W, int H, uint32_t *cnt, uint32_t *first,
                    unsigned int *flags, hipStream_t st) {
    const size_t n = (size_t)iw * ih;
    hipLaunchKernelGGL(k_lm_count, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, vert, iw, ih, W, H, cnt,
                       first, flags);
    return hipGetLastError();
}
hipError_t lm_fill(const double *vert, int iw, int ih, int W, int H, const uint32_t *off, const uint32_t *base,
                   uint32_t encoding, uint16_t *d16, uint32_t *t32, float *frac, unsigned int *flags,
                   hipStream_t st) {
    const size_t n = (size_t)iw * ih;
    hipLaunchKernelGGL(k_lm_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, vert, iw, ih, W, H, off,
                       base, encoding, d16, t32, frac, flags);
    return hipGetLastError();
}

// ---------------------------------------------------------------- PSF projector
// The PSF push-forward of psf_projector.c:82-146: input pixel (x, y) with a
// non-zero value and mapped area > 1e-12 links to every output pixel (tx, ty)
// of its clamped bounding box in [-hx, hx] x [-hy, hy] whose overlap area
// exceeds 1e-8, with frac = area / full_area; output index
// (ty + hy) * (2 hx + 1) + (tx + hx).  Two passes (count, then fill) into an
// ABS32 link list in (ty, tx) order per pixel; the weighted sum per output
// pixel then runs over the transposed list in ascending input pixel order,
// the reference's fmaf(value, frac, out) sequence.
__device__ bool pp_pixel(const double *__restrict__ v, const float *__restrict__ val, int iw, int hx, int hy, int x,
                         int y, DPt q[4], float &full, int &x0, int &x1, int &y0, int &y1) {
    if (val[(size_t)y * iw + x] == 0.0f) return false;
    const size_t a = 2 * ((size_t)y * (size_t)(iw + 1) + (size_t)x), b = a + 2 * (size_t)(iw + 1);
    q[0] = {v[a], v[a + 1]};
    q[1] = {v[b], v[b + 1]};
    q[2] = {v[b + 2], v[b + 3]};
    q[3] = {v[a + 2], v[a + 3]};
    full = g_area(q, 4);
    if (!isfinite(full) || !(full > 1.0e-12f)) return false;
    double mnx = q[0].x, mxx = mnx, mny = q[0].y, mxy = mny;
    for (int j = 1; j < 4; ++j) {
        mnx = q[j].x < mnx ? q[j].x : mnx;
        mxx = q[j].x > mxx ? q[j].x : mxx;
        mny = q[j].y < mny ? q[j].y : mny;
        mxy = q[j].y > mxy ? q[j].y : mxy;
    }
    x0 = (int)floor(mnx + .5); x1 = (int)floor(mxx + .5);
    y0 = (int)floor(mny + .5); y1 = (int)floor(mxy + .5);
    if (x0 < -hx) x0 = -hx;
    if (x1 > hx) x1 = hx;
    if (y0 < -hy) y0 = -hy;
    if (y1 > hy) y1 = hy;
    return true;
}
__global__ __launch_bounds__(256) void k_pp_count(const double *__restrict__ vert, const float *__restrict__ val, int iw,
                                                  int ih, int hx, int hy, uint32_t *__restrict__ cnt) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (size_t)iw * ih) return;
    DPt q[4];
    float full;
    int x0, x1, y0, y1;
    uint32_t n = 0;
    if (pp_pixel(vert, val, iw, hx, hy, (int)(p % iw), (int)(p / iw), q, full, x0, x1, y0, y1))
        for (int ty = y0; ty <= y1; ++ty)
            for (int tx = x0; tx <= x1; ++tx)
                if (g_overlap(q, tx, ty) > 1.0e-8f) ++n;
    cnt[p] = n;
}
__global__ __launch_bounds__(256) void k_pp_fill(const double *__restrict__ vert, const float *__restrict__ val, int iw,
                                                 int ih, int hx, int hy, const uint32_t *__restrict__ off,
                                                 uint32_t *__restrict__ t32, float *__restrict__ frac,
                                                 unsigned int *flags) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (size_t)iw * ih) return;
    DPt q[4];
    float full;
    int x0, x1, y0, y1;
    uint32_t z = off[p];
    const uint32_t zend = off[p + 1];
    const int ow = 2 * hx + 1;
    if (pp_pixel(vert, val, iw, hx, hy, (int)(p % iw), (int)(p / iw), q, full, x0, x1, y0, y1))
        for (int ty = y0; ty <= y1; ++ty)
            for (int tx = x0; tx <= x1; ++tx) {
                const float area = g_overlap(q, tx, ty);
                if (!(area > 1.0e-8f)) continue;
                if (z >= zend) { atomicOr(flags, 2u); return; }
                t32[z] = (uint32_t)((ty + hy) * ow + (tx + hx));
                frac[z] = area / full;
                ++z;
            }
    if (z != zend) atomicOr(flags, 8u);
}
hipError_t pp_count(const double *vert, const float *val, int iw, int ih, int hx, int hy, uint32_t *cnt,
                    hipStream_t st) {
    const size_t n = (size_t)iw * ih;
    hipLaunchKernelGGL(k_pp_count, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, vert, val, iw, ih, hx, hy, cnt);
    return hipGetLastError();
}
hipError_t pp_fill(const double *vert, const float *val, int iw, int ih, int hx, int hy, const uint32_t *off,
                   uint32_t *t32, float *frac, unsigned int *flags, hipStream_t st) {
    const size_t n = (size_t)iw * ih;
    hipLaunchKernelGGL(k_pp_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, vert, val, iw, ih, hx, hy, off,
                       t32, frac, flags);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_lm_validate(const uint32_t *__restrict__ off, const uint32_t *__restrict__ base,
                                                     const uint16_t *__restrict__ d16, const uint32_t *__restrict__ t32,
                                                     const float *__restrict__ frac, uint32_t encoding, uint32_t nin,
                                                     uint64_t nlinks, uint64_t nout, unsigned int *flags) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= nin) return;
    const uint32_t z0 = off[p], z1 = off[p + 1];
    if (z1 < z0 || (uint64_t)z1 > nlinks) {
        atomicOr(flags, 1u);
        return;
    }
    for (uint32_t z = z0; z < z1; ++z) {
        const uint64_t t = encoding == UPDC_LM_DELTA16 ? (uint64_t)base[p] + d16[z] : (uint64_t)t32[z];
        if (t >= nout) atomicOr(flags, 2u);
        if (!isfinite(frac[z])) atomicOr(flags, 4u);
    }
}
hipError_t lm_validate(const uint32_t *off, const uint32_t *base, const uint16_t *d16, const uint32_t *t32,
                       const float *frac, uint32_t encoding, uint32_t nin, uint64_t nlinks, uint64_t nout,
                       unsigned int *flags, hipStream_t st) {
    hipLaunchKernelGGL(k_lm_validate, dim3((nin + 255) / 256), dim3(256), 0, st, off, base, d16, t32, frac,
                       encoding, nin, nlinks, nout, flags);
    return hipGetLastError();
}

// ===================================================================== scan
// Three-phase exclusive scan: per-block sums (64-bit), a one-block scan of
// the block sums, then the per-block local scan plus its offset.
static constexpr int SCAN_NT = 256, SCAN_IT = 8, SCAN_TILE = SCAN_NT * SCAN_IT;

__device__ __forceinline__ unsigned long long block_excl_scan(unsigned long long v, unsigned long long *sh,
                                                              unsigned long long *total) {
    // Hillis-Steele over 256 entries in LDS (log2(256) = 8 steps).
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < SCAN_NT; o <<= 1) {
        const unsigned long long a = t >= o ? sh[t - o] : 0ull;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    const unsigned long long incl = sh[t];
    if (total) *total = sh[SCAN_NT - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_sums(const uint32_t *__restrict__ in, size_t n,
                                                       unsigned long long *__restrict__ bsum) {
    __shared__ unsigned long long sh[SCAN_NT];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE;
    unsigned long long s = 0;
    for (int i = 0; i < SCAN_IT; ++i) {
        const size_t idx = base + (size_t)i * SCAN_NT + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    unsigned long long tot;
    block_excl_scan(s, sh, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}
__global__ __launch_bounds__(SCAN_NT) void k_scan_top(unsigned long long *bsum, size_t nb,
                                                      unsigned long long *total64) {
    __shared__ unsigned long long sh[SCAN_NT];
    unsigned long long carry = 0;
    for (size_t base = 0; base < nb; base += SCAN_NT) {
        const size_t idx = base + threadIdx.x;
        const unsigned long long v = idx < nb ? bsum[idx] : 0ull;
        unsigned long long tot;
        const unsigned long long ex = block_excl_scan(v, sh, &tot);
        if (idx < nb) bsum[idx] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total64 = carry;
}
__global__ __launch_bounds__(SCAN_NT) void k_scan_apply(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                        size_t n, const unsigned long long *__restrict__ bsum,
                                                        const unsigned long long *__restrict__ total64) {
    __shared__ unsigned long long sh[SCAN_NT];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_IT;
    // each thread owns SCAN_IT consecutive entries
    uint32_t v[SCAN_IT];
    unsigned long long s = 0;
    for (int i = 0; i < SCAN_IT; ++i) {
        const size_t idx = base + i;
        v[i] = idx < n ? in[idx] : 0u;
        s += v[i];
    }
    unsigned long long run = bsum[blockIdx.x] + block_excl_scan(s, sh, nullptr);
    for (int i = 0; i < SCAN_IT; ++i) {
        const size_t idx = base + i;
        if (idx < n) out[idx] = (uint32_t)run;
        run += v[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = (uint32_t)*total64;
}

size_t scan_scratch_bytes(size_t n) { return ((n + SCAN_TILE - 1) / SCAN_TILE + 1) * sizeof(unsigned long long); }

hipError_t scan_u32(const uint32_t *in, uint32_t *out, size_t n, void *scratch, unsigned long long *total64,
                    hipStream_t st) {
    const size_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    unsigned long long *bsum = (unsigned long long *)scratch;
    if (nb == 0) {
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_NT), 0, st, bsum, (size_t)0, total64);
        hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(SCAN_NT), 0, st, in, out, n, bsum, total64);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nb), dim3(SCAN_NT), 0, st, in, n, bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_NT), 0, st, bsum, nb, total64);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SCAN_NT), 0, st, in, out, n, bsum, total64);
    return hipGetLastError();
}

// ===================================================================== CSC
// Target-major transpose of S_k for the atomic-free adjoint.  Counting and
// placement use integer atomics (order-free); the per-target sort by source
// pixel then fixes the order, so the result is deterministic and the gather
// reproduces the reference's ascending-p fmaf chain exactly.
template <class TW>
__global__ __launch_bounds__(256) void k_csc_count(const ExpDev *__restrict__ e_dev, uint32_t nin,
                                                   const TW *__restrict__ w, uint32_t *__restrict__ tcnt) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= nin) return;
    const ExpDev &e = *e_dev;
    const uint32_t z0 = e.off[p], z1 = e.off[p + 1];
    if (!(w[p] > (TW)0) || z0 == z1) return;
    for (uint32_t z = z0; z < z1; ++z) atomicAdd(&tcnt[lm_target(e, p, z)], 1u);
}
template <class TW>
__global__ __launch_bounds__(256) void k_csc_fill(const ExpDev *__restrict__ e_dev, uint32_t nin,
                                                  const TW *__restrict__ w, uint32_t *__restrict__ cursor,
                                                  uint2 *__restrict__ tent) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= nin) return;
    const ExpDev &e = *e_dev;
    const uint32_t z0 = e.off[p], z1 = e.off[p + 1];
    if (!(w[p] > (TW)0) || z0 == z1) return;
    for (uint32_t z = z0; z < z1; ++z) {
        const uint32_t pos = atomicAdd(&cursor[lm_target(e, p, z)], 1u);
        tent[pos] = make_uint2(p, __float_as_uint(e.frac[z]));
    }
}
__global__ __launch_bounds__(256) void k_csc_sort(const uint32_t *__restrict__ toff, uint2 *__restrict__ tent,
                                                  size_t nout) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nout) return;
    const uint32_t a = toff[t], b = toff[t + 1];
    for (uint32_t i = a + 1; i < b; ++i) {
        const uint2 key = tent[i];
        uint32_t j = i;
        while (j > a && tent[j - 1].x > key.x) {
            tent[j] = tent[j - 1];
            --j;
        }
        tent[j] = key;
    }
}
// The compact copy of a sorted CSC, one wave per 64-target tile: the tile's
// first entry and smallest source pixel, each target's count, each entry's
// pixel minus that smallest pixel and its fraction (same order, so a gather
// over it repeats the ascending-p chain).  flags bit 0: a target with more
// than 255 entries; bit 1: a tile whose pixels span more than 16 bits.
__global__ __launch_bounds__(256) void k_csc16_build(const uint32_t *__restrict__ toff, const uint2 *__restrict__ tent,
                                                     size_t nout, uint32_t *__restrict__ c_toff,
                                                     uint32_t *__restrict__ c_pb, uint8_t *__restrict__ c_cnt,
                                                     uint16_t *__restrict__ c_dp, float *__restrict__ c_fr,
                                                     unsigned int *flags) {
    const size_t ntiles = (nout + 63) / 64;
    const size_t tile = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile >= ntiles) return;  // uniform across the wave
    const int lane = threadIdx.x & 63;
    const size_t t = tile * 64 + lane;
    uint32_t a = 0, b = 0;
    if (t < nout) {
        a = toff[t];
        b = toff[t + 1];
    }
    // entries of a target are sorted by pixel: first and last bound its range
    uint32_t lo = b > a ? tent[a].x : 0xffffffffu, hi = b > a ? tent[b - 1].x : 0u;
    for (int o = 32; o >= 1; o >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, o));
        hi = max(hi, (uint32_t)__shfl_xor((int)hi, o));
    }
    const uint32_t base = lo == 0xffffffffu ? 0u : lo;
    if (lane == 0) {
        c_toff[tile] = toff[tile * 64];
        c_pb[tile] = base;
        if (tile == ntiles - 1) c_toff[ntiles] = toff[nout];
        if (lo != 0xffffffffu && hi - lo > 0xffffu) atomicOr(flags, 2u);
    }
    if (t < nout) c_cnt[t] = (uint8_t)min(b - a, 255u);
    if (b - a > 255u) atomicOr(flags, 1u);
    for (uint32_t i = a; i < b; ++i) {
        const uint2 en = tent[i];
        c_dp[i] = (uint16_t)(en.x - base);
        c_fr[i] = __uint_as_float(en.y);
    }
}
hipError_t csc16_build(const uint32_t *toff, const uint2 *tent, size_t nout, uint32_t *c_toff, uint32_t *c_pb,
                       uint8_t *c_cnt, uint16_t *c_dp, float *c_fr, unsigned int *flags, hipStream_t st) {
    const size_t ntiles = (nout + 63) / 64;
    hipLaunchKernelGGL(k_csc16_build, dim3((unsigned)((ntiles + 3) / 4)), dim3(256), 0, st, toff, tent, nout, c_toff,
                       c_pb, c_cnt, c_dp, c_fr, flags);
    return hipGetLastError();
}
template <class TW>
hipError_t csc_count(const ExpDev *e_dev, uint32_t nin, const TW *w, uint32_t *tcnt, hipStream_t st) {
    hipLaunchKernelGGL((k_csc_count<TW>), dim3((nin + 255) / 256), dim3(256), 0, st, e_dev, nin, w, tcnt);
    return hipGetLastError();
}
template <class TW>
hipError_t csc_fill(const ExpDev *e_dev, uint32_t nin, const TW *w, uint32_t *cursor, uint2 *tent, hipStream_t st) {
    hipLaunchKernelGGL((k_csc_fill<TW>), dim3((nin + 255) / 256), dim3(256), 0, st, e_dev, nin, w, cursor, tent);
    return hipGetLastError();
}
hipError_t csc_sort(const uint32_t *toff, uint2 *tent, size_t nout, hipStream_t st) {
    hipLaunchKernelGGL(k_csc_sort, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, st, toff, tent, nout);
    return hipGetLastError();
}

// ===================================================================== weights
template <class TW>
__global__ __launch_bounds__(256) void k_normalize(const float *__restrict__ sci_raw, const float *__restrict__ w_raw,
                                                   size_t n, float scale, double scale_d, float ped_f, double ped_d,
                                                   TW *__restrict__ sci, TW *__restrict__ w,
                                                   unsigned long long *masked, unsigned int *err) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    unsigned int m = 0;
    if (p < n) {
        const float in = sci_raw[p], wt = w_raw[p];
        if (!isfinite(in) || !isfinite(wt) || !(wt > 0)) {
            sci[p] = (TW)0;
            w[p] = (TW)0;
            m = 1;
        } else if constexpr (sizeof(TW) == 4) {
            const float wn = wt / scale;
            if (!isfinite(wn) || !(wn > 0)) atomicOr(err, UPDC_E_WEIGHT);
            const float y = in + ped_f;
            if (!isfinite(y) || y < 0.0f) atomicOr(err, UPDC_E_YEFF);
            sci[p] = in;
            w[p] = wn;
        } else {
            const double wn = (double)wt / scale_d;
            if (!isfinite(wn) || !(wn > 0)) atomicOr(err, UPDC_E_WEIGHT);
            const double y = (double)in + ped_d;
            if (!isfinite(y) || y < 0.0) atomicOr(err, UPDC_E_YEFF);
            sci[p] = (TW)in;
            w[p] = (TW)wn;
        }
    }
    // wave-level count, one atomic per wave
    const unsigned long long bal = __ballot(m);
    if ((threadIdx.x & 63) == 0 && bal) atomicAdd(masked, (unsigned long long)__popcll(bal));
}
template <class TW>
hipError_t normalize_weights(const float *sci_raw, const float *w_raw, size_t n, float scale, double scale_d,
                             float ped_f, double ped_d, TW *sci, TW *w, unsigned long long *masked,
                             unsigned int *err, hipStream_t st) {
    hipLaunchKernelGGL((k_normalize<TW>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, sci_raw, w_raw, n,
                       scale, scale_d, ped_f, ped_d, sci, w, masked, err);
    return hipGetLastError();
}

// ===================================================================== drizzle
// One thread per target pixel.  Exposures are visited k-major and pixels in
// ascending order inside each (the CSC order), which is the summation order of
// the serial branch (updc_core.c:950-964).  With nworkers > 1 each worker's
// partial (exposures k = w mod nworkers) starts at 0 and the partials are added
// in worker order from 0 (updc_core.c:885-903, 929-934).
template <class TF>
__global__ __launch_bounds__(256) void k_drizzle(const ExpDev *__restrict__ ex, int nexp, const TF *__restrict__ sci,
                                                 const TF *__restrict__ w, int nworkers, size_t nout,
                                                 TF *__restrict__ est, TF *__restrict__ cov, unsigned int *err) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nout) return;
    TF c = (TF)0, d = (TF)0, bk = (TF)0;
    for (int wk = 0; wk < nworkers; ++wk) {
        TF cw = (TF)0, dw = (TF)0, bw = (TF)0;
        for (int k = wk; k < nexp; k += nworkers) {
            const ExpDev &e = ex[k];
            const TF bg = sizeof(TF) == 4 ? (TF)e.background_f : (TF)e.background;
            const uint32_t a0 = e.toff[t], a1 = e.toff[t + 1];
            for (uint32_t a = a0; a < a1; ++a) {
                const uint2 en = e.tent[a];
                const TF q = (TF)__uint_as_float(en.y) * w[e.pix + en.x];
                cw += q;
                dw = fmat(q, sci[e.pix + en.x], dw);
                bw = fmat(q, bg, bw);
            }
        }
        if (nworkers == 1) { c = cw; d = dw; bk = bw; }
        else { c += cw; d += dw; bk += bw; }
    }
    cov[t] = c;
    const TF thr = sizeof(TF) == 4 ? (TF)1e-12f : (TF)1e-12;
    if (c > thr) {
        const TF x = d / c - bk / c;
        if (!isfinite(x)) atomicOr(err, UPDC_E_DRIZZLE);
        est[t] = x;
    } else {
        est[t] = (TF)0;
    }
}
template <class TF>
hipError_t drizzle(const ExpDev *ex, int nexp, const TF *sci, const TF *w, int nworkers, size_t nout, TF *estimate,
                   TF *coverage, unsigned int *err, hipStream_t st) {
    hipLaunchKernelGGL((k_drizzle<TF>), dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, st, ex, nexp, sci, w,
                       nworkers, nout, estimate, coverage, err);
    return hipGetLastError();
}

// ===================================================================== detector
// One thread per detector pixel, every exposure of the batch in one launch
// (blockIdx.y = exposure).  Forward gather pred = sum_z fma(frac, blur[t], .)
// in link order from 0; raw_mu = pred + background + pedestal; permanent
// freeze when !(raw_mu > 0); ratio 1 when frozen else (sci + pedestal) /
// raw_mu; wr = w * ratio (updc_core.c:1152-1170).  F32: every step in f32 as
// the reference.  MIXED: pred in f32, raw_mu, the test and the ratio in f64.
// F64: everything in f64.
// Grid-stride over chunks of 256 consecutive pixels of one exposure
// (blockIdx.y).  A chunk owns one contiguous link range [off[p0], off[p0+256]).
// Each chunk costs two dependent memory round trips: (1) the chunk's link
// words and fractions, fetched as aligned 16-byte blocks and issued together
// with the lanes' own pixel operands (off, base, w, sci, frozen); (2) the blur
// gathers, addressed from LDS.  The next chunk's link range is fetched one
// chunk ahead.  Chunks whose link range exceeds the LDS capacity read their
// links from HBM directly.  FreezeMu counters and the minimum raw_mu are
// reduced per block and committed with one atomic per block (a per-wave
// atomic on the exposure's counters serialises the whole launch).
// GATHER_CAP keeps the block's LDS at 24 KiB and the FP32-blur kernels fit 80
// VGPRs, so six blocks share a CU (6 waves per SIMD); the grid is one wave of
// six blocks on each of the 256 CUs.  With 8 exposures of 2048² on a 4096²
// grid (≈ 2300 links per chunk): 0.67 ms, against 0.85 ms at 4000 links
// (32 KiB: four blocks per CU), 12 blurs per batch and five blocks per CU.
static constexpr int GATHER_NT = 256, GATHER_CAP = 3000, GATHER_BLOCKS = 1536, GATHER_UNROLL = 10;
// Waves per SIMD the launch bounds ask for: FP64 blur values would spill at 80 VGPRs.
template <class TF> constexpr int gather_waves() { return sizeof(TF) == 4 ? 6 : 4; }
// 16-byte blocks of LDS for link words (u16 deltas or u32 targets) and for
// fractions: GATHER_CAP links plus one block of misalignment slack each.
static constexpr int GATHER_LV = GATHER_CAP / 4 + 1, GATHER_FV = GATHER_CAP / 4 + 1;
static constexpr int GATHER_SLOTS = (GATHER_LV + GATHER_NT - 1) / GATHER_NT;

template <class TF, class TA>
__global__ __launch_bounds__(GATHER_NT, (gather_waves<TF>())) void k_gather(const ExpDev *__restrict__ ex, int kbeg,
                                                      const TF *__restrict__ blur, size_t blur_bstride,
                                                      const TF *__restrict__ sci, const TF *__restrict__ w,
                                                      uint8_t *__restrict__ frozen, TF *__restrict__ wr,
                                                      ExpDiag *diag, IterDiag *idiag) {
    __shared__ uint4 s_lv[GATHER_LV];
    __shared__ uint4 s_fv[GATHER_FV];
    __shared__ unsigned int r_new[GATHER_NT / 64], r_tot[GATHER_NT / 64];
    __shared__ TA r_min[GATHER_NT / 64];
    // an earlier iteration of this call failed: change nothing (uniform exit)
    if (idiag->halt) return;
    const int k = kbeg + blockIdx.y;
    const ExpDev &e = ex[k];
    const uint32_t nin = e.nin;
    const bool d16 = e.encoding == UPDC_LM_DELTA16;
    const uintptr_t lbase = d16 ? (uintptr_t)e.d16 : (uintptr_t)e.t32;
    const uint32_t lsz = d16 ? 2u : 4u;
    // the record's arrays as global pointers (loads through generic pointers
    // would be flat loads, which also hold up every LDS read behind them)
    const auto off = gp(e.off);
    const auto frac = gp(e.frac);
    const auto ebase = gp(e.base);
    const auto ed16 = gp(e.d16);
    const auto et32 = gp(e.t32);
    const TF *bl = blur + (size_t)k * blur_bstride;
    const uint16_t *s16 = (const uint16_t *)s_lv;
    const uint32_t *s32 = (const uint32_t *)s_lv;
    const float *sfr = (const float *)s_fv;
    unsigned int nnew = 0, ntot = 0, bad = 0;
    unsigned long long badkey = ~0ull;
    TA mn = (TA)INFINITY;
    const uint32_t step = gridDim.x * GATHER_NT;
    uint32_t p0 = blockIdx.x * GATHER_NT;
    // Software pipeline over the block's chunks: while chunk i's blur gathers
    // run, chunk i+1's staging blocks and pixel operands are in flight, and
    // chunk i+2's link range is being read.
    const auto offc = cp(e.off);
    auto range = [&](uint32_t q0, uint32_t &A0, uint32_t &A1) {
        if (q0 < nin) {
            A0 = offc[q0];
            A1 = offc[min(q0 + (uint32_t)GATHER_NT, nin)];
        }
    };
    uint4 lv[GATHER_SLOTS], fv[GATHER_SLOTS];
    // staging state of the chunk whose loads are in flight
    uint32_t n_nl = 0, n_nf = 0, n_lsh = 0, n_fsh = 0, n_Z0 = 0, n_z0 = 0, n_z1 = 0, n_b0 = 0;
    bool n_staged = false;
    TF n_wp = (TF)0, n_sv = (TF)0;
    uint8_t n_fz = 0;
    auto stage = [&](uint32_t q0, uint32_t A0, uint32_t A1) __attribute__((always_inline)) {
        // aligned 16-byte block ranges covering the chunk's links and fractions
        const uintptr_t la = lbase + (uintptr_t)A0 * lsz, fa = (uintptr_t)(frac + A0);
        const uintptr_t la0 = la & ~(uintptr_t)15, fa0 = fa & ~(uintptr_t)15;
        const uint32_t nlv = (uint32_t)((((lbase + (uintptr_t)A1 * lsz) + 15) & ~(uintptr_t)15) - la0) >> 4;
        const uint32_t nfv = (uint32_t)((((uintptr_t)(frac + A1) + 15) & ~(uintptr_t)15) - fa0) >> 4;
        n_staged = nlv <= (uint32_t)GATHER_LV && nfv <= (uint32_t)GATHER_FV;
        n_lsh = (uint32_t)(la - la0) / lsz;
        n_fsh = (uint32_t)(fa - fa0) >> 2;
        n_Z0 = A0;
        // the staging blocks and the lane's pixel operands, issued together
        n_nl = n_staged ? nlv : 0u;
        n_nf = n_staged ? nfv : 0u;
#pragma unroll
        for (int s = 0; s < GATHER_SLOTS; ++s) {
            const uint32_t i = threadIdx.x + s * GATHER_NT;
            lv[s] = fv[s] = uint4{};
            if (i < n_nl) lv[s] = ((gptr_t<const uint4>)la0)[i];
            if (i < n_nf) fv[s] = ((gptr_t<const uint4>)fa0)[i];
        }
        // the lane's pixel operands, issued by every lane (a lane past the
        // exposure's end reads its last pixel and never uses the values)
        const uint32_t q = min(q0 + threadIdx.x, nin - 1);
        const size_t gq = e.pix + q;
        n_z0 = off[q];
        n_z1 = off[q + 1];
        n_b0 = d16 ? ebase[q] : 0u;
        n_wp = w[gq];
        n_sv = sci[gq];
        n_fz = frozen[gq];
    };
    uint32_t Zn0 = 0, Zn1 = 0;
    if (p0 < nin) {
        uint32_t A0 = 0, A1 = 0;
        range(p0, A0, A1);
        stage(p0, A0, A1);
    }
    range(p0 + step, Zn0, Zn1);
    for (; p0 < nin; p0 += step) {
        const uint32_t p = p0 + threadIdx.x;
        const uint32_t pend = min(p0 + (uint32_t)GATHER_NT, nin);
        // this chunk's staging state
        const bool staged = n_staged;
        const uint32_t lsh = n_lsh, fsh = n_fsh, Z0 = n_Z0, nl = n_nl, nf = n_nf;
        const uint32_t z0 = n_z0, z1 = n_z1, b0 = n_b0;
        const TF wp = n_wp, sv = n_sv;
        uint8_t fz = n_fz;
        const size_t gp = e.pix + p;
        __syncthreads();  // previous chunk's LDS reads are done
#pragma unroll
        for (int s = 0; s < GATHER_SLOTS; ++s) {
            const uint32_t i = threadIdx.x + s * GATHER_NT;
            if (i < nl) s_lv[i] = lv[s];
            if (i < nf) s_fv[i] = fv[s];
        }
        __syncthreads();
        // the next chunk's loads, then the range of the one after
        const uint32_t pn = p0 + step;
        if (pn < nin) stage(pn, Zn0, Zn1);
        range(pn + step, Zn0, Zn1);
        if (p < pend) {
            TF out = (TF)0;
            if (wp > (TF)0 && z0 != z1) {
                TF pred = (TF)0;
                if (staged) {
                    // round trip 2: the first GATHER_UNROLL blur loads issued together,
                    // then the fmaf chain in link order (same operations as a plain loop)
                    const uint32_t nz = z1 - z0, l0 = z0 - Z0 + lsh, f0 = z0 - Z0 + fsh;
                    TF vals[GATHER_UNROLL];
#pragma unroll
                    for (int j = 0; j < GATHER_UNROLL; ++j) {
                        vals[j] = (TF)0;
                        if ((uint32_t)j < nz) vals[j] = bl[d16 ? b0 + (uint32_t)s16[l0 + j] : s32[l0 + j]];
                    }
#pragma unroll
                    for (int j = 0; j < GATHER_UNROLL; ++j)
                        if ((uint32_t)j < nz) pred = fmat((TF)sfr[f0 + j], vals[j], pred);
                    for (uint32_t j = GATHER_UNROLL; j < nz; ++j) {
                        const uint32_t t = d16 ? b0 + (uint32_t)s16[l0 + j] : s32[l0 + j];
                        pred = fmat((TF)sfr[f0 + j], bl[t], pred);
                    }
                } else if (d16) {
                    for (uint32_t z = z0; z < z1; ++z) pred = fmat((TF)frac[z], bl[b0 + (uint32_t)ed16[z]], pred);
                } else {
                    for (uint32_t z = z0; z < z1; ++z) pred = fmat((TF)frac[z], bl[et32[z]], pred);
                }
                TA raw;
                if constexpr (sizeof(TA) == 4) raw = pred + e.mean_offset_f;
                else raw = (TA)pred + (TA)e.mean_offset;
                unsigned int b = 0;
                if (!isfinite(raw)) b |= UPDC_E_MU;
                mn = raw < mn ? raw : mn;
                // the reference fails before its freeze test on a non-finite mean
                if (!fz && !b && !(raw > (TA)0)) {
                    fz = 1;
                    frozen[gp] = 1;
                    ++nnew;
                }
                TA ratio;
                if (fz) {
                    ++ntot;
                    ratio = (TA)1;
                } else {
                    TA y;
                    if constexpr (sizeof(TA) == 4) y = sv + e.pedestal_f;
                    else y = (TA)sv + (TA)e.pedestal;
                    ratio = y / raw;
                    if (!isfinite(ratio)) b |= UPDC_E_RATIO;
                }
                out = (TF)((TA)wp * ratio);
                if (b) {
                    bad |= b;
                    const unsigned long long key = err_key_detector((unsigned)k, p, (b & UPDC_E_MU) ? 0u : 1u);
                    badkey = badkey < key ? badkey : key;
                }
            }
            wr[gp] = out;
        }
    }
    // block reduction, one atomic per block and quantity
    for (int o = 32; o >= 1; o >>= 1) {
        nnew += __shfl_xor(nnew, o);
        ntot += __shfl_xor(ntot, o);
        const TA other = __shfl_xor(mn, o);
        mn = other < mn ? other : mn;
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        r_new[wv] = nnew;
        r_tot[wv] = ntot;
        r_min[wv] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int a = 0, t = 0;
        TA m = (TA)INFINITY;
        for (int i = 0; i < GATHER_NT / 64; ++i) {
            a += r_new[i];
            t += r_tot[i];
            m = r_min[i] < m ? r_min[i] : m;
        }
        if (a) atomicAdd(&diag[k].n_new, (unsigned long long)a);
        if (t) atomicAdd(&diag[k].n_total, (unsigned long long)t);
        if (m < (TA)INFINITY) atomicMin(&diag[k].min_key, dkey((double)m));
    }
    if (bad) {
        atomicOr(&idiag->err, bad);
        atomicMin(&idiag->err_key, badkey);
    }
}
template <class TF, class TA>
hipError_t gather_freeze(const ExpDev *ex, int kbeg, int kend, uint32_t max_nin, const TF *blur,
                         size_t blur_bstride, const TF *sci, const TF *w, uint8_t *frozen, TF *wr, ExpDiag *diag,
                         IterDiag *idiag, hipStream_t st) {
    if (kend <= kbeg) return hipSuccess;
    const int nk = kend - kbeg;
    const unsigned chunks = (max_nin + GATHER_NT - 1) / GATHER_NT;
    const unsigned gb = std::max(1u, std::min(chunks, (unsigned)std::max(1, GATHER_BLOCKS / nk)));
    hipLaunchKernelGGL((k_gather<TF, TA>), dim3(gb, nk), dim3(GATHER_NT), 0, st, ex, kbeg, blur,
                       blur_bstride, sci, w, frozen, wr, diag, idiag);
    return hipGetLastError();
}

template <class TF>
__global__ __launch_bounds__(256) void k_predict(const ExpDev *__restrict__ ex, const TF *__restrict__ blur,
                                                 size_t blur_bstride, TF *__restrict__ pred) {
    const int k = blockIdx.y;
    const ExpDev &e = ex[k];
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= e.nin) return;
    const TF *bl = blur + (size_t)k * blur_bstride;
    TF acc = (TF)0;
    for (uint32_t z = e.off[p]; z < e.off[p + 1]; ++z) acc = fmat((TF)e.frac[z], bl[lm_target(e, p, z)], acc);
    pred[e.pix + p] = acc;
}
template <class TF>
hipError_t predict(const ExpDev *ex, int nexp, uint32_t max_nin, const TF *blur, size_t blur_bstride, TF *pred,
                   hipStream_t st) {
    hipLaunchKernelGGL((k_predict<TF>), dim3((max_nin + 255) / 256, nexp), dim3(256), 0, st, ex, blur, blur_bstride,
                       pred);
    return hipGetLastError();
}
template hipError_t predict<float>(const ExpDev *, int, uint32_t, const float *, size_t, float *, hipStream_t);
template hipError_t predict<double>(const ExpDev *, int, uint32_t, const double *, size_t, double *, hipStream_t);

// ===================================================================== update
// x <- x * max(C/D, 0) where D > 1e-8, else 0, with the reference's guards
// (updc_core.c:1439-1456).  The signed sum and the negative count are reduced
// in a fixed order: per-block partials, then the last block to finish sums
// them in block order (a ticket counter it resets for the next launch).
static constexpr int UPD_NT = 256, UPD_BLOCKS = 1024;
int update_blocks() { return UPD_BLOCKS; }

template <class TF, class TA>
__global__ __launch_bounds__(UPD_NT) void k_update(TF *__restrict__ x, const TA *__restrict__ C,
                                                   const TA *__restrict__ D, size_t nout, TA max_sens,
                                                   IterDiag *idiag, double *partials, unsigned long long *negs,
                                                   unsigned int *ticket) {
    __shared__ double sh[UPD_NT];
    __shared__ unsigned long long shn[UPD_NT];
    __shared__ bool last;
    double sum = 0.0;
    unsigned long long nneg = 0;
    unsigned int bad = 0;
    unsigned long long badkey = ~0ull;
    const TA thr = sizeof(TA) == 4 ? (TA)1e-8f : (TA)1e-8;
    const TA neg = sizeof(TA) == 4 ? (TA)(-1.0e-6f) : (TA)(-1.0e-6);
    // the detector stage of this iteration, or an earlier iteration, failed:
    // the reference never reaches this update, so x stays as it is
    const bool hold = idiag->halt || (idiag->err & (UPDC_E_MU | UPDC_E_RATIO));
    for (size_t t = (size_t)blockIdx.x * UPD_NT + threadIdx.x; t < nout && !hold; t += (size_t)UPD_BLOCKS * UPD_NT) {
        const TA d = D[t];
        TF xv = x[t];
        unsigned int b = 0;
        if (d > thr) {
            const TA c = C[t];
            TA f = c / d;
            if (!isfinite(f)) b = b ? b : UPDC_E_FACTOR;
            if (f < (TA)0) {
                if (c < neg * max_sens) b = b ? b : UPDC_E_NEGCORR;
                f = (TA)0;
            }
            xv = (TF)((TA)xv * f);
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
    }
    if (bad) {
        atomicOr(&idiag->err, bad);
        atomicMin(&idiag->err_key, badkey);
    }
    sh[threadIdx.x] = sum;
    shn[threadIdx.x] = nneg;
    __syncthreads();
    for (int o = UPD_NT / 2; o > 0; o >>= 1) {
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
        last = atomicAdd(ticket, 1u) == UPD_BLOCKS - 1;
    }
    __syncthreads();
    if (last) {
        __threadfence();
        double s = 0.0;
        unsigned long long n = 0;
        for (int i = threadIdx.x; i < UPD_BLOCKS; i += UPD_NT) {
            s += __builtin_nontemporal_load(&partials[i]);
            n += __builtin_nontemporal_load(&negs[i]);
        }
        sh[threadIdx.x] = s;
        shn[threadIdx.x] = n;
        __syncthreads();
        for (int o = UPD_NT / 2; o > 0; o >>= 1) {
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
template <class TF, class TA>
hipError_t update(TF *x, const TA *C, const TA *D, size_t nout, TA max_sens, IterDiag *idiag, double *partials,
                  unsigned long long *negs, unsigned int *ticket, hipStream_t st) {
    hipLaunchKernelGGL((k_update<TF, TA>), dim3(UPD_BLOCKS), dim3(UPD_NT), 0, st, x, C, D, nout, max_sens, idiag,
                       partials, negs, ticket);
    return hipGetLastError();
}

// ===================================================================== misc
template <class T> __global__ void k_fill(T *p, T v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = v;
}
template <class T> hipError_t fill(T *p, T v, size_t n, hipStream_t st) {
    if (!n) return hipSuccess;
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL((k_fill<T>), dim3(g), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}
template <class TD> __global__ void k_max_key(const TD *p, size_t n, unsigned long long *key) {
    unsigned long long m = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const unsigned long long k = dkey((double)p[i]);
        m = k > m ? k : m;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        const unsigned long long other = __shfl_xor(m, o);
        m = other > m ? other : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(key, m);
}
template <class TD> hipError_t max_key(const TD *p, size_t n, unsigned long long *key, hipStream_t st) {
    const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 4096));
    hipLaunchKernelGGL((k_max_key<TD>), dim3(g), dim3(256), 0, st, p, n, key);
    return hipGetLastError();
}
// Normalised PSF wrapped with its centre on (0,0) (updc_core.c:813-821).
template <class TF>
__global__ void k_psf_wrap(const float *__restrict__ psf, int pw, int ph, int transposed, int FW, int FH, TF inv_sum,
                           TF *__restrict__ image) {
    const int fw = transposed ? ph : pw, fh = transposed ? pw : ph;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)fw * fh) return;
    const int x = (int)(i % fw), y = (int)(i / fw);
    const int yy = (y - fh / 2 + FH) % FH, xx = (x - fw / 2 + FW) % FW;
    const size_t src = transposed ? (size_t)x * pw + y : (size_t)y * pw + x;
    image[(size_t)yy * FW + xx] = (TF)psf[src] * inv_sum;
}
template <class TF>
hipError_t psf_wrap(const float *psf, int pw, int ph, int transposed, int FW, int FH, TF inv_sum, TF *image,
                    hipStream_t st) {
    const size_t n = (size_t)pw * ph;
    hipLaunchKernelGGL((k_psf_wrap<TF>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, psf, pw, ph,
                       transposed, FW, FH, inv_sum, image);
    return hipGetLastError();
}
template <class TS, class TD> __global__ void k_convert(const TS *s, TD *d, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) d[i] = (TD)s[i];
}
template <class TS, class TD> hipError_t convert(const TS *s, TD *d, size_t n, hipStream_t st) {
    if (!n) return hipSuccess;
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL((k_convert<TS, TD>), dim3(g), dim3(256), 0, st, s, d, n);
    return hipGetLastError();
}
__global__ void k_init_diag(ExpDiag *d, int n, IterDiag *it, int nit) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        d[i].n_new = 0;
        d[i].n_total = 0;
        d[i].min_key = ~0ull;
    }
    if (i < nit) {
        it[i].signed_sum = 0.0;
        it[i].n_negative = 0;
        it[i].err = 0;
        it[i].halt = 0;
        it[i].err_key = ~0ull;
    }
}
hipError_t init_diag(ExpDiag *d, int n, IterDiag *it, int nit, hipStream_t st) {
    const int m = std::max(n, nit);
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_init_diag, dim3((m + 255) / 256), dim3(256), 0, st, d, n, it, nit);
    return hipGetLastError();
}

__global__ void k_commit_diag(ExpDiag *cur, int nexp, IterDiag *icur, ExpDiag *hist, IterDiag *ihist, int cap,
                              unsigned int *ctr) {
    const unsigned int it = *ctr;
    __syncthreads();
    const int i = threadIdx.x;
    if (it < (unsigned int)cap) {
        for (int k = i; k < nexp; k += blockDim.x) hist[(size_t)it * nexp + k] = cur[k];
        if (i == 0) ihist[it] = *icur;
    }
    __syncthreads();
    for (int k = i; k < nexp; k += blockDim.x) {
        cur[k].n_new = 0;
        cur[k].n_total = 0;
        cur[k].min_key = ~0ull;
    }
    if (i == 0) {
        if (icur->err) icur->halt = 1u;  // sticky for the rest of the call
        icur->signed_sum = 0.0;
        icur->n_negative = 0;
        icur->err = 0;
        icur->err_key = ~0ull;
        *ctr = it + 1;
    }
}
hipError_t commit_diag(ExpDiag *cur, int nexp, IterDiag *icur, ExpDiag *hist, IterDiag *ihist, int cap,
                       unsigned int *ctr, hipStream_t st) {
    hipLaunchKernelGGL(k_commit_diag, dim3(1), dim3(256), 0, st, cur, nexp, icur, hist, ihist, cap, ctr);
    return hipGetLastError();
}
template <class TF> __global__ void k_csc_apply(const ExpDev *__restrict__ e_dev, const TF *__restrict__ v,
                                                size_t nout, TF *__restrict__ out) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nout) return;
    const ExpDev &e = *e_dev;
    TF acc = (TF)0;
    for (uint32_t a = e.toff[t]; a < e.toff[t + 1]; ++a) {
        const uint2 en = e.tent[a];
        acc = fmat((TF)__uint_as_float(en.y), v[e.pix + en.x], acc);
    }
    out[t] = acc;
}
template <class TF> hipError_t csc_apply(const ExpDev *e_dev, const TF *v, size_t nout, TF *out, hipStream_t st) {
    hipLaunchKernelGGL((k_csc_apply<TF>), dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, st, e_dev, v, nout, out);
    return hipGetLastError();
}

// Streaming copy of known size: the scale reference for HBM byte counters.
__global__ __launch_bounds__(256) void k_copy_ref(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}
hipError_t copy_ref(const void *src, void *dst, size_t bytes, hipStream_t st) {
    const size_t n = bytes / 16;
    hipLaunchKernelGGL(k_copy_ref, dim3(8192), dim3(256), 0, st, (const uint4 *)src, (uint4 *)dst, n);
    return hipGetLastError();
}

// ---- explicit instantiations
// Batched adjoint image: out[k * nout + t] = sum over the CSC entries of
// target t of fma(frac, v[pix_k + p], .) in ascending p (the reference's
// scatter order, updc_core.c:1171-1176), for every exposure k of the batch
// (blockIdx.y).  Chunks of CSCI_G * CSCI_NT targets, XCD by XCD; a lane
// takes CSCI_G targets CSCI_NT apart (coalesced offset loads and stores) and
// issues the first CSCI_U entries of all of them, then their v loads, together.
static constexpr int CSCI_NT = 256, CSCI_U = 4;
// 3 targets per lane (2 in FP64) keep the kernel at 62 VGPRs, 8 waves per SIMD.
// With 8 exposures of 2048² on a 4096² grid: 0.94 ms, against 1.00 (1 target),
// 1.13 (4) and 1.29 (6); 0.91 ms with the XCD-aware chunk order.
template <class TF> constexpr int csci_g() { return sizeof(TF) == 4 ? 3 : 2; }
template <class TF>
__global__ __launch_bounds__(CSCI_NT) void k_csc_image(const ExpDev *__restrict__ ex, const TF *__restrict__ v,
                                                       size_t nout, TF *__restrict__ out, int skip_c16) {
    constexpr int G = csci_g<TF>();
    const ExpDev &e = ex[blockIdx.y];
    if (skip_c16 && e.c16) return;  // that exposure runs in k_csc_image16
    const TF *vk = v + e.pix;
    TF *ok = out + (size_t)blockIdx.y * nout;
    // XCD-aware: the blocks of one XCD (blockIdx.x % 8; the host makes
    // gridDim.x a multiple of 8) walk one contiguous eighth of the chunks, so
    // the v lines that neighbouring chunks share stay in that XCD's L2
    const size_t chunk = (size_t)CSCI_NT * G, nch = (nout + chunk - 1) / chunk, per = (nch + 7) / 8;
    const size_t c_end = min(nch, (size_t)(blockIdx.x % 8 + 1) * per);
    for (size_t ci = (size_t)(blockIdx.x % 8) * per + blockIdx.x / 8; ci < c_end; ci += gridDim.x / 8) {
        const size_t t0 = ci * chunk + threadIdx.x;
        uint32_t a[G], n[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const size_t t = t0 + (size_t)g * CSCI_NT;
            a[g] = t < nout ? e.toff[t] : 0u;
            n[g] = t < nout ? e.toff[t + 1] : 0u;
        }
        uint2 en[G][CSCI_U];
        TF wv[G][CSCI_U];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            n[g] -= a[g];
#pragma unroll
            for (int j = 0; j < CSCI_U; ++j) en[g][j] = (uint32_t)j < n[g] ? e.tent[a[g] + j] : make_uint2(0u, 0u);
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int j = 0; j < CSCI_U; ++j) wv[g][j] = (uint32_t)j < n[g] ? vk[en[g][j].x] : (TF)0;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const size_t t = t0 + (size_t)g * CSCI_NT;
            TF acc = (TF)0;
#pragma unroll
            for (int j = 0; j < CSCI_U; ++j)
                if ((uint32_t)j < n[g]) acc = fmat((TF)__uint_as_float(en[g][j].y), wv[g][j], acc);
            for (uint32_t j = CSCI_U; j < n[g]; ++j) {
                const uint2 x = e.tent[a[g] + j];
                acc = fmat((TF)__uint_as_float(x.y), vk[x.x], acc);
            }
            if (t < nout) ok[t] = acc;
        }
    }
}
// The same gather over the compact CSC (exposures with c16 set), entry-major.
// A wave takes one 64-target tile at a time; the tile's first entry and pixel
// base are wave-uniform, and each lane's first entry is the exclusive prefix
// sum of the counts of the lanes before it.  The wave's lanes load the tile's
// entries 64 consecutive ones at a time (so the u16 offsets and f32 fractions
// arrive in whole lines, with no lane idle on a short target), gather v at
// each entry's pixel, and leave the (fraction, value) pairs in the wave's LDS
// window.  Each lane then runs its own target's fmaf chain over its entries
// there, in entry order, i.e. ascending pixel: the chain of the 8-byte form.
// LDS accesses of one wave run in issue order, so the window needs no barrier.
// Software-pipelined: the next tile's entries (and the counts of the one
// after) are in flight while the current tile's values are gathered.  Every
// load of the pipelined path is issued by every lane, at a clamped index,
// and its value selected afterwards: loads under a branch would make the
// compiler wait for all outstanding loads at the loop head.  The first
// C16_WIN entries of a tile go through the window at once; a longer tile
// walks further windows.  C16_G tiles per wave and iteration.  With 8
// exposures of 2048² on a 4096² grid (about 144 entries per tile): 0.56 ms
// with a 192-entry window (three rounds of loads, every one issued), within
// 2 % for one to four tiles per iteration; 0.59-0.60 ms with 256 entries.
static constexpr int C16_U = 3, C16_WIN = C16_U * 64, C16_G = 2;
template <class TF> struct C16Ent { float f; TF v; };
template <class TF>
__global__ __launch_bounds__(CSCI_NT) void k_csc_image16(const ExpDev *__restrict__ ex, const TF *__restrict__ v,
                                                         size_t nout, TF *__restrict__ out) {
    constexpr int U = C16_U, G = C16_G;
    __shared__ C16Ent<TF> win_all[CSCI_NT / 64][G][C16_WIN];
    const ExpDev &e = ex[blockIdx.y];
    if (!e.c16) return;
    // the record's fields, read once
    const auto cnt = gp(e.c_cnt);
    const auto toff = gp(e.c_toff);
    const auto pbase = gp(e.c_pb);
    const auto dps = gp(e.c_dp);
    const auto frs = gp(e.c_fr);
    const auto vk = gp(v + e.pix);
    const auto ok = gp(out + (size_t)blockIdx.y * nout);
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // chunks of G * CSCI_NT targets; wave w takes tiles g * CSCI_NT / 64 + w of a chunk
    constexpr size_t CH = (size_t)G * CSCI_NT;
    const size_t ntiles = (nout + 63) / 64, nch = (nout + CH - 1) / CH, per = (nch + 7) / 8;
    // the last entry of the map and the last pixel of the exposure: bounds of
    // every clamped read (an empty map reads the allocation's spare slot 0)
    const uint32_t nent = toff[ntiles], ent_last = nent ? nent - 1 : 0u, pix_last = e.nin ? e.nin - 1 : 0u;
    const size_t c_end = min(nch, (size_t)(blockIdx.x % 8 + 1) * per), step = gridDim.x / 8;
    // Raw loads of a chunk ahead: the lane's count and the tile's first entry
    // and base.  Values are selected where they are used, one iteration after
    // their loads: a select next to its load would make the loop wait for it.
    uint32_t m_cnt[G], m_fst[G], m_pb[G];
    auto meta = [&](size_t cj) {  // (a chunk past the end reads the last one)
        cj = min(cj, nch - 1);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const size_t tb = cj * CH + (size_t)g * CSCI_NT + (size_t)wave * 64, t = tb + lane;
            const size_t tile = min(tb >> 6, ntiles - 1);
            m_cnt[g] = cnt[min(t, nout - 1)];
            m_fst[g] = toff[tile];  // a tile past the end (tb >= nout) has no targets: its counts are zero
            m_pb[g] = pbase[tile];
        }
    };
    // counts of chunk cj (loaded by meta) -> lane offsets (wave prefix sum) ->
    // raw loads of each tile's first C16_WIN entries (entries past the tile's
    // last read its last one, or for an empty tile the allocation's spare tail)
    struct St {
        uint32_t n[G], s[G], fst[G], pb[G], tot[G], dp[G][U];
        float fr[G][U];
    };
    auto take = [&](size_t cj, St &a) __attribute__((always_inline)) {
        cj = min(cj, nch - 1);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const size_t t = cj * CH + (size_t)g * CSCI_NT + (size_t)wave * 64 + lane;
            a.n[g] = t < nout ? m_cnt[g] : 0u;
            const uint32_t x = wave_incl_scan(a.n[g]);
            a.s[g] = x - a.n[g];
            a.tot[g] = (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
            a.fst[g] = m_fst[g];
            a.pb[g] = m_pb[g];
            const uint32_t last = a.tot[g] ? a.tot[g] - 1 : 0u;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                // (clamped to the map's entries as well: a tile's range never
                // leaves them, and a read never leaves the allocation)
                const uint32_t jc = min(a.fst[g] + min((uint32_t)(u * 64 + lane), last), ent_last);
                a.dp[g][u] = dps[jc];
                a.fr[g][u] = frs[jc];
            }
        }
    };
    size_t ci = (size_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (ci >= c_end) return;
    St c;
    meta(ci);
    take(ci, c);
    meta(ci + step);
    for (; ci < c_end; ci += step) {
        // fractions of the current tiles into the windows (every slot is
        // written; slots past a tile's entries are never read)
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int u = 0; u < U; ++u) win_all[wave][g][u * 64 + lane].f = c.fr[g][u];
        // values (an entry past the tile reads v at the tile's base)
        TF wv[G][U];
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int u = 0; u < U; ++u)
                wv[g][u] = vk[min(c.pb[g] + ((uint32_t)(u * 64 + lane) < c.tot[g] ? c.dp[g][u] : 0u), pix_last)];
        // the next tiles' loads, unconditionally (past the end they are unused):
        // the values' wait below leaves them in flight
        St nx;
        take(ci + step, nx);
        meta(ci + 2 * step);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            C16Ent<TF> *wl = win_all[wave][g];
#pragma unroll
            for (int u = 0; u < U; ++u) wl[u * 64 + lane].v = wv[g][u];
            const uint32_t j0 = c.s[g], j1 = c.s[g] + c.n[g], tot = c.tot[g], fst = c.fst[g], pb = c.pb[g];
            TF acc = (TF)0;
            for (uint32_t j = j0; j < min(j1, (uint32_t)C16_WIN); ++j) {
                const C16Ent<TF> q = wl[j];
                acc = fmat((TF)q.f, q.v, acc);
            }
            // tiles beyond one window (wave-uniform)
            for (uint32_t wb = C16_WIN; wb < tot; wb += C16_WIN) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t j = wb + (uint32_t)(u * 64 + lane);
                    if (j < tot) wl[u * 64 + lane] = C16Ent<TF>{frs[fst + j], vk[pb + (uint32_t)dps[fst + j]]};
                }
                for (uint32_t j = max(j0, wb); j < min(j1, wb + (uint32_t)C16_WIN); ++j) {
                    const C16Ent<TF> q = wl[j - wb];
                    acc = fmat((TF)q.f, q.v, acc);
                }
            }
            const size_t t = ci * CH + (size_t)g * CSCI_NT + threadIdx.x;
            if (t < nout) ok[t] = acc;
        }
        c = nx;
    }
}
template <class TF>
hipError_t csc_image(const ExpDev *ex, int nexp, const TF *v, size_t nout, TF *out, bool any_c16, bool any_wide,
                     hipStream_t st) {
    if (nexp <= 0) return hipSuccess;
    // grid-stride over chunks of CSCI_NT * G targets, at most 16384 blocks in all,
    // a multiple of 8 per exposure (the kernels' XCD ranges)
    auto grid = [&](int g) {
        const size_t chunk = (size_t)CSCI_NT * g;
        return (unsigned)((std::min<size_t>((nout + chunk - 1) / chunk, (size_t)std::max(1, 16384 / nexp)) + 7) / 8 * 8);
    };
    const unsigned gx = grid(csci_g<TF>());
    if (any_c16)
        hipLaunchKernelGGL((k_csc_image16<TF>), dim3(grid(C16_G), nexp), dim3(CSCI_NT), 0, st, ex, v, nout, out);
    if (any_wide)
        hipLaunchKernelGGL((k_csc_image<TF>), dim3(gx, nexp), dim3(CSCI_NT), 0, st, ex, v, nout, out, (int)any_c16);
    return hipGetLastError();
}
template hipError_t csc_image<float>(const ExpDev *, int, const float *, size_t, float *, bool, bool, hipStream_t);
template hipError_t csc_image<double>(const ExpDev *, int, const double *, size_t, double *, bool, bool,
                                      hipStream_t);

template hipError_t csc_apply<float>(const ExpDev *, const float *, size_t, float *, hipStream_t);
template hipError_t csc_apply<double>(const ExpDev *, const double *, size_t, double *, hipStream_t);
template hipError_t csc_count<float>(const ExpDev *, uint32_t, const float *, uint32_t *, hipStream_t);
template hipError_t csc_count<double>(const ExpDev *, uint32_t, const double *, uint32_t *, hipStream_t);
template hipError_t csc_fill<float>(const ExpDev *, uint32_t, const float *, uint32_t *, uint2 *, hipStream_t);
template hipError_t csc_fill<double>(const ExpDev *, uint32_t, const double *, uint32_t *, uint2 *, hipStream_t);
template hipError_t normalize_weights<float>(const float *, const float *, size_t, float, double, float, double,
                                             float *, float *, unsigned long long *, unsigned int *, hipStream_t);
template hipError_t normalize_weights<double>(const float *, const float *, size_t, float, double, float, double,
                                              double *, double *, unsigned long long *, unsigned int *, hipStream_t);
template hipError_t drizzle<float>(const ExpDev *, int, const float *, const float *, int, size_t, float *, float *,
                                   unsigned int *, hipStream_t);
template hipError_t drizzle<double>(const ExpDev *, int, const double *, const double *, int, size_t, double *,
                                    double *, unsigned int *, hipStream_t);
template hipError_t gather_freeze<float, float>(const ExpDev *, int, int, uint32_t, const float *, size_t,
                                                const float *, const float *, uint8_t *, float *, ExpDiag *,
                                                IterDiag *, hipStream_t);
template hipError_t gather_freeze<float, double>(const ExpDev *, int, int, uint32_t, const float *, size_t,
                                                 const float *, const float *, uint8_t *, float *, ExpDiag *,
                                                 IterDiag *, hipStream_t);
template hipError_t gather_freeze<double, double>(const ExpDev *, int, int, uint32_t, const double *, size_t,
                                                  const double *, const double *, uint8_t *, double *, ExpDiag *,
                                                  IterDiag *, hipStream_t);
template hipError_t update<float, float>(float *, const float *, const float *, size_t, float, IterDiag *, double *,
                                         unsigned long long *, unsigned int *, hipStream_t);
template hipError_t update<float, double>(float *, const double *, const double *, size_t, double, IterDiag *,
                                          double *, unsigned long long *, unsigned int *, hipStream_t);
template hipError_t update<double, double>(double *, const double *, const double *, size_t, double, IterDiag *,
                                           double *, unsigned long long *, unsigned int *, hipStream_t);
template hipError_t fill<float>(float *, float, size_t, hipStream_t);
template hipError_t fill<double>(double *, double, size_t, hipStream_t);
template hipError_t fill<uint32_t>(uint32_t *, uint32_t, size_t, hipStream_t);
template hipError_t fill<uint8_t>(uint8_t *, uint8_t, size_t, hipStream_t);
template hipError_t fill<unsigned long long>(unsigned long long *, unsigned long long, size_t, hipStream_t);
template hipError_t max_key<float>(const float *, size_t, unsigned long long *, hipStream_t);
template hipError_t max_key<double>(const double *, size_t, unsigned long long *, hipStream_t);
template hipError_t psf_wrap<float>(const float *, int, int, int, int, int, float, float *, hipStream_t);
template hipError_t psf_wrap<double>(const float *, int, int, int, int, int, double, double *, hipStream_t);
template hipError_t convert<float, double>(const float *, double *, size_t, hipStream_t);
template hipError_t convert<double, float>(const double *, float *, size_t, hipStream_t);
template hipError_t convert<double, double>(const double *, double *, size_t, hipStream_t);
template hipError_t convert<float, float>(const float *, float *, size_t, hipStream_t);

}  // namespace updc
