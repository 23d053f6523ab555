/* CPU oracle for the UPDC v17 reconstruction iteration -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C11 compiled with -ffp-contract=off, the float32
 * arithmetic of the reference core (updc/c_core/updc_core.c and
 * updc/c_core/overlap.c).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.  The product library
 * (updc_amd/csrc) never links or calls it.
 *
 * Every function names the reference lines it follows.  Operation order and
 * the explicit fmaf() calls are kept identical, because the reference builds
 * with -std=c11 (GCC then defaults to -ffp-contract=off) and the parity tests
 * compare the GPU kernels with these functions bit for bit.
 *
 * Pinning (how each function is checked without running the reference core):
 *   orc_polygon_area / orc_overlap_area  -> the reference's own overlap.c
 *       compiled into oracle/_ref, the self-test constants
 *       (updc_core.c:1545-1557), closed-form areas and an exact rational
 *       brute force (tests/test_oracle_geometry.py).
 *   orc_linkmap_*   -> ABI constants (updc/cache_abi.py:9-12), encoding rules,
 *       and a per-pixel recomputation through the _ref overlap area.
 *   orc_fused_detector / orc_sensitivity_scatter -> hand-checkable scalar
 *       cases, the adjoint identity <S u, v> = <u, S^T v>, FreezeMu invariants.
 *   orc_update / orc_drizzle / orc_weight_scale / orc_psf_* -> closed forms
 *       on hand-made inputs (tests/test_oracle_stages.py).
 *   orc_spectrum_multiply / orc_accumulate_adjoint -> exact complex products
 *       of the same float32 inputs within the rounding bound, with a
 *       conjugation error shown to fall far outside it; the adjoint identity
 *       of the whole blur / adjoint pair built on them
 *       (tests/test_oracle_adjoint.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* LinkMap v2 ABI constants (updc_core.c:44-54, cache_abi.py:9-12).    */
/* ------------------------------------------------------------------ */
#define ORC_LM_MAGIC 0x324d4c55u
#define ORC_LM_VERSION 2u
#define ORC_LM_ABS32 1u
#define ORC_LM_DELTA16 2u
#define ORC_LM_HEADER 64u

typedef struct { double x, y; } OrcPt;

/* ------------------------------------------------------------------ */
/* Geometry: overlap.c:14-116                                          */
/* ------------------------------------------------------------------ */
/* side: 0 left (x >= v), 1 right (x <= v), 2 bottom (y >= v), 3 top (y <= v)
 * (overlap.c:14-22). */
static int orc_in(OrcPt p, int side, double v) {
    if (side == 0) return p.x >= v;
    if (side == 1) return p.x <= v;
    if (side == 2) return p.y >= v;
    return p.y <= v;
}

/* Segment/boundary crossing, overlap.c:24-43: the parameter t is computed
 * with one division and the other coordinate as a + t*(b-a) (two roundings,
 * no contraction).  A near-degenerate segment (|delta| <= 1e-15) returns a. */
static OrcPt orc_cross(OrcPt a, OrcPt b, int side, double v) {
    OrcPt r = a;
    if (side < 2) {
        double d = b.x - a.x;
        if (fabs(d) > 1.0e-15) {
            double t = (v - a.x) / d;
            r.x = v;
            r.y = a.y + t * (b.y - a.y);
        }
    } else {
        double d = b.y - a.y;
        if (fabs(d) > 1.0e-15) {
            double t = (v - a.y) / d;
            r.x = a.x + t * (b.x - a.x);
            r.y = v;
        }
    }
    return r;
}

/* One Sutherland-Hodgman pass, overlap.c:45-70.  Walks edges (prev -> cur),
 * starting with prev = last vertex. */
static int orc_clip(const OrcPt *in, int n, OrcPt *out, int side, double v) {
    int m = 0;
    OrcPt prev = in[n - 1];
    int pin = orc_in(prev, side, v);
    for (int i = 0; i < n; ++i) {
        OrcPt cur = in[i];
        int cin = orc_in(cur, side, v);
        if (cin) {
            if (!pin) out[m++] = orc_cross(prev, cur, side, v);
            out[m++] = cur;
        } else if (pin) {
            out[m++] = orc_cross(prev, cur, side, v);
        }
        prev = cur;
        pin = cin;
    }
    return m;
}

/* Shoelace area, overlap.c:72-91: local origin at vertex 0 (f64 subtraction,
 * then cast to f32), f32 accumulation of fmaf(xi, yj, -(xj*yi)),
 * result 0.5f*|sum|. */
static float orc_poly_area(const OrcPt *p, int n) {
    if (n < 3) return 0.0f;
    const double ox = p[0].x, oy = p[0].y;
    float acc = 0.0f;
    for (int i = 0; i < n; ++i) {
        int j = (i + 1) % n;
        float xi = (float)(p[i].x - ox), yi = (float)(p[i].y - oy);
        float xj = (float)(p[j].x - ox), yj = (float)(p[j].y - oy);
        acc += fmaf(xi, yj, -(xj * yi));
    }
    return 0.5f * fabsf(acc);
}

/* Overlap of a quad with the unit target pixel centred on (x, y),
 * overlap.c:93-116: clip left, right, bottom, top at x-.5, x+.5, y-.5, y+.5;
 * any pass leaving fewer than three vertices gives 0. */
static float orc_quad_overlap(const OrcPt q[4], int x, int y) {
    OrcPt a[16], b[16];
    OrcPt *in = a, *out = b;
    const double lim[4] = {(double)x - 0.5, (double)x + 0.5, (double)y - 0.5, (double)y + 0.5};
    int n = 4;
    for (int j = 0; j < 4; ++j) in[j] = q[j];
    for (int s = 0; s < 4; ++s) {
        n = orc_clip(in, n, out, s, lim[s]);
        if (n < 3) return 0.0f;
        OrcPt *t = in; in = out; out = t;
    }
    return orc_poly_area(in, n);
}

/* C-ABI wrappers over packed (x0,y0,x1,y1,...) double arrays. */
float orc_polygon_area(const double *xy, int n) { return orc_poly_area((const OrcPt *)xy, n); }
float orc_overlap_area(const double *quad_xy, int x, int y) { return orc_quad_overlap((const OrcPt *)quad_xy, x, y); }

/* ------------------------------------------------------------------ */
/* FFT shape helpers: updc_core.c:117-126, 170-179, 264-276            */
/* ------------------------------------------------------------------ */
int orc_next_fast_size(int n) {
    for (;; ++n) {
        int q = n;
        while (q % 2 == 0) q /= 2;
        while (q % 3 == 0) q /= 3;
        while (q % 5 == 0) q /= 5;
        while (q % 7 == 0) q /= 7;
        if (q == 1) return n;
    }
}

/* First `cap` 7-smooth sizes >= minimum (updc_core.c:170-179). */
int orc_fast_candidates(int minimum, int *out, int cap) {
    int n = 0, v = minimum;
    while (n < cap) {
        v = orc_next_fast_size(v);
        out[n++] = v;
        if (v == 2147483647) break;
        v++;
    }
    return n;
}

/* Candidate shape list of the autotuner (updc_core.c:264-276): the diagonal
 * (w_i, h_i), then (w2,h0), (w0,h2), (w4,h0), (w0,h4), without duplicates.
 * Writes (w,h) pairs into out[2*i], returns the count (<= 16). */
int orc_fft_shape_list(int min_w, int min_h, int *out) {
    int wc[5], hc[5], ns = 0;
    int nw = orc_fast_candidates(min_w, wc, 5), nh = orc_fast_candidates(min_h, hc, 5);
    int cand[9][2], nc = 0;
    int nd = nw < nh ? nw : nh;
    for (int i = 0; i < nd; ++i) { cand[nc][0] = wc[i]; cand[nc][1] = hc[i]; nc++; }
    if (nw > 2) { cand[nc][0] = wc[2]; cand[nc][1] = hc[0]; nc++; }
    if (nh > 2) { cand[nc][0] = wc[0]; cand[nc][1] = hc[2]; nc++; }
    if (nw > 4) { cand[nc][0] = wc[4]; cand[nc][1] = hc[0]; nc++; }
    if (nh > 4) { cand[nc][0] = wc[0]; cand[nc][1] = hc[4]; nc++; }
    for (int i = 0; i < nc; ++i) {
        int dup = 0;
        for (int j = 0; j < ns; ++j) if (out[2 * j] == cand[i][0] && out[2 * j + 1] == cand[i][1]) dup = 1;
        if (!dup && ns < 16) { out[2 * ns] = cand[i][0]; out[2 * ns + 1] = cand[i][1]; ns++; }
    }
    return ns;
}

/* ------------------------------------------------------------------ */
/* LinkMap builder: updc_core.c:463-609                                */
/* ------------------------------------------------------------------ */
/* Detector pixel (x,y) -> quad (v[y][x], v[y+1][x], v[y+1][x+1], v[y][x+1]),
 * full area, clamped candidate bbox (updc_core.c:512-528). */
static int orc_pixel_quad(const double *v, int iw, int W, int H, int x, int y, OrcPt q[4],
                          float *full, int *x0, int *x1, int *y0, int *y1) {
    size_t a = 2 * ((size_t)y * (size_t)(iw + 1) + (size_t)x), b = a + 2 * (size_t)(iw + 1);
    q[0].x = v[a]; q[0].y = v[a + 1];
    q[1].x = v[b]; q[1].y = v[b + 1];
    q[2].x = v[b + 2]; q[2].y = v[b + 3];
    q[3].x = v[a + 2]; q[3].y = v[a + 3];
    *full = orc_poly_area(q, 4);
    if (!isfinite(*full) || *full <= 1.0e-12f) return 0;
    double mnx = q[0].x, mxx = mnx, mny = q[0].y, mxy = mny;
    for (int j = 1; j < 4; ++j) {
        if (q[j].x < mnx) mnx = q[j].x;
        if (q[j].x > mxx) mxx = q[j].x;
        if (q[j].y < mny) mny = q[j].y;
        if (q[j].y > mxy) mxy = q[j].y;
    }
    if (!isfinite(mnx) || !isfinite(mxx) || !isfinite(mny) || !isfinite(mxy) ||
        mxx < -.5 || mxy < -.5 || mnx > W - .5 || mny > H - .5) return 0;
    *x0 = (int)floor(mnx + .5); *x1 = (int)floor(mxx + .5);
    *y0 = (int)floor(mny + .5); *y1 = (int)floor(mxy + .5);
    if (*x0 < 0) *x0 = 0;
    if (*y0 < 0) *y0 = 0;
    if (*x1 >= W) *x1 = W - 1;
    if (*y1 >= H) *y1 = H - 1;
    return *x0 <= *x1 && *y0 <= *y1;
}

static size_t orc_align4(size_t v) { return (v + 3u) & ~(size_t)3u; }

/* Exact file size for a header (updc_core.c:469-483). */
size_t orc_linkmap_bytes(uint64_t nin, uint64_t nlinks, uint32_t encoding) {
    size_t pos = ORC_LM_HEADER + (size_t)(nin + 1) * 4u;
    if (encoding == ORC_LM_DELTA16) {
        pos += (size_t)nin * 4u;
        pos += (size_t)nlinks * 2u;
        pos = orc_align4(pos);
        pos += (size_t)nlinks * 4u;
    } else if (encoding == ORC_LM_ABS32) {
        pos += (size_t)nlinks * 8u;
    } else {
        return 0;
    }
    return pos;
}

/* Pass 1 (updc_core.c:549-570): counts, per-pixel first target, encoding.
 * offsets: u32[nin+1], base: u32[nin] (0 for inactive pixels).
 * Returns 0 on success, -1 if the link count overflows u32. */
int orc_linkmap_pass1(const double *v, int iw, int ih, int W, int H, uint32_t *offsets, uint32_t *base,
                      uint64_t *nlinks, uint32_t *encoding) {
    uint64_t count = 0;
    int all16 = 1;
    OrcPt q[4];
    for (int y = 0; y < ih; ++y)
        for (int x = 0; x < iw; ++x) {
            size_t p = (size_t)y * iw + x;
            if (count > 0xffffffffu) return -1;
            offsets[p] = (uint32_t)count;
            base[p] = 0;
            float full;
            int x0, x1, y0, y1;
            if (!orc_pixel_quad(v, iw, W, H, x, y, q, &full, &x0, &x1, &y0, &y1)) continue;
            uint64_t before = count;
            uint32_t first = 0, last = 0;
            for (int ty = y0; ty <= y1; ++ty)
                for (int tx = x0; tx <= x1; ++tx) {
                    float area = orc_quad_overlap(q, tx, ty);
                    if (area <= 1.0e-5f) continue;
                    if (count >= 0xffffffffu) return -1;
                    uint32_t t = (uint32_t)((size_t)ty * W + tx);
                    if (count == before) first = t;
                    last = t;
                    ++count;
                }
            if (count != before) {
                base[p] = first;
                if (last - first > 0xffffu) all16 = 0;
            }
        }
    offsets[(size_t)iw * ih] = (uint32_t)count;
    *nlinks = count;
    *encoding = all16 ? ORC_LM_DELTA16 : ORC_LM_ABS32;
    return 0;
}

/* Complete ULM2 image (header + arrays) into `out` (orc_linkmap_bytes()
 * bytes, zero-initialised by the caller so the DELTA16 pad is zero, as the
 * reference's ftruncate'd file is).  Pass 2 is updc_core.c:586-597. */
int orc_linkmap_write(const double *v, int iw, int ih, int W, int H, const uint32_t *offsets,
                      const uint32_t *base, uint64_t nlinks, uint32_t encoding, uint8_t *out) {
    const size_t nin = (size_t)iw * ih;
    uint32_t hdr[16];
    memset(hdr, 0, sizeof(hdr));
    hdr[0] = ORC_LM_MAGIC; hdr[1] = ORC_LM_VERSION; hdr[2] = encoding; hdr[3] = ORC_LM_HEADER;
    uint64_t n64 = nin;
    memcpy(&hdr[4], &n64, 8);
    memcpy(&hdr[6], &nlinks, 8);
    hdr[8] = (uint32_t)W; hdr[9] = (uint32_t)H;
    memcpy(out, hdr, ORC_LM_HEADER);
    size_t pos = ORC_LM_HEADER;
    memcpy(out + pos, offsets, (nin + 1) * 4u);
    pos += (nin + 1) * 4u;
    uint16_t *d16 = NULL;
    uint32_t *t32 = NULL;
    if (encoding == ORC_LM_DELTA16) {
        memcpy(out + pos, base, nin * 4u);
        pos += nin * 4u;
        d16 = (uint16_t *)(out + pos);
        pos += (size_t)nlinks * 2u;
        pos = orc_align4(pos);
    } else {
        t32 = (uint32_t *)(out + pos);
        pos += (size_t)nlinks * 4u;
    }
    float *fr = (float *)(out + pos);
    OrcPt q[4];
    for (int y = 0; y < ih; ++y)
        for (int x = 0; x < iw; ++x) {
            size_t p = (size_t)y * iw + x;
            uint32_t z = offsets[p];
            float full;
            int x0, x1, y0, y1;
            if (!orc_pixel_quad(v, iw, W, H, x, y, q, &full, &x0, &x1, &y0, &y1)) continue;
            for (int ty = y0; ty <= y1; ++ty)
                for (int tx = x0; tx <= x1; ++tx) {
                    float area = orc_quad_overlap(q, tx, ty);
                    if (area <= 1.0e-5f) continue;
                    uint32_t t = (uint32_t)((size_t)ty * W + tx);
                    if (z >= offsets[p + 1]) return -1;
                    if (d16) {
                        uint32_t d = t - base[p];
                        if (d > 0xffffu) return -2;
                        d16[z] = (uint16_t)d;
                    } else {
                        t32[z] = t;
                    }
                    fr[z] = area / full;
                    ++z;
                }
            if (z != offsets[p + 1]) return -3;
        }
    return 0;
}

/* ------------------------------------------------------------------ */
/* LinkMap views (updc_core.c:76-79, 645-662)                          */
/* ------------------------------------------------------------------ */
typedef struct {
    uint32_t encoding;
    const uint32_t *off, *base, *t32;
    const uint16_t *d16;
    const float *frac;
} OrcLM;

static OrcLM orc_view(const uint8_t *img, size_t nin) {
    OrcLM m;
    uint64_t nl;
    memcpy(&m.encoding, img + 8, 4);
    memcpy(&nl, img + 24, 8);
    size_t pos = ORC_LM_HEADER;
    m.off = (const uint32_t *)(img + pos);
    pos += (nin + 1) * 4u;
    m.base = NULL; m.t32 = NULL; m.d16 = NULL;
    if (m.encoding == ORC_LM_DELTA16) {
        m.base = (const uint32_t *)(img + pos);
        pos += nin * 4u;
        m.d16 = (const uint16_t *)(img + pos);
        pos += (size_t)nl * 2u;
        pos = orc_align4(pos);
    } else {
        m.t32 = (const uint32_t *)(img + pos);
        pos += (size_t)nl * 4u;
    }
    m.frac = (const float *)(img + pos);
    return m;
}
static inline uint32_t orc_tgt(const OrcLM *m, size_t p, uint32_t z) {
    return m->encoding == ORC_LM_DELTA16 ? m->base[p] + (uint32_t)m->d16[z] : m->t32[z];
}

/* ------------------------------------------------------------------ */
/* Detector stage: fused_exposure, updc_core.c:1150-1179               */
/* ------------------------------------------------------------------ */
/* Returns 0, or -1 for a non-finite forward mean (:1160), -2 for a
 * non-finite ratio (:1168).  scatter (nout floats) is zeroed first. */
int orc_fused_detector(const uint8_t *lm_img, size_t nin, size_t nout, const float *sci, const float *w,
                       uint8_t *frozen, const float *blur, float mean_offset, float pedestal, float *scatter,
                       uint64_t *n_new, uint64_t *n_total, float *min_raw_mu, float *wr_out) {
    OrcLM m = orc_view(lm_img, nin);
    memset(scatter, 0, nout * sizeof(float));
    uint64_t nnew = 0, ntot = 0;
    float mn = INFINITY;
    for (size_t p = 0; p < nin; ++p) {
        if (wr_out) wr_out[p] = 0.0f;
        if (!(w[p] > 0) || m.off[p] == m.off[p + 1]) continue;
        const uint32_t z0 = m.off[p], z1 = m.off[p + 1];
        float pred = 0.0f;
        for (uint32_t z = z0; z < z1; ++z) pred = fmaf(m.frac[z], blur[orc_tgt(&m, p, z)], pred);
        const float raw = pred + mean_offset;
        if (!isfinite(raw)) return -1;
        if (raw < mn) mn = raw;
        if (!frozen[p] && !(raw > 0.0f)) { frozen[p] = 1; ++nnew; }
        float ratio;
        if (frozen[p]) { ++ntot; ratio = 1.0f; }
        else {
            const float y = sci[p] + pedestal;
            ratio = y / raw;
            if (!isfinite(ratio)) return -2;
        }
        const float wr = w[p] * ratio;
        if (wr_out) wr_out[p] = wr;
        for (uint32_t z = z0; z < z1; ++z) {
            const uint32_t t = orc_tgt(&m, p, z);
            scatter[t] = fmaf(m.frac[z], wr, scatter[t]);
        }
    }
    *n_new = nnew; *n_total = ntot; *min_raw_mu = mn;
    return 0;
}

/* S_k^T w scatter of the sensitivity (updc_core.c:1059-1066). */
void orc_sensitivity_scatter(const uint8_t *lm_img, size_t nin, size_t nout, const float *w, float *scatter) {
    OrcLM m = orc_view(lm_img, nin);
    memset(scatter, 0, nout * sizeof(float));
    for (size_t p = 0; p < nin; ++p) {
        if (!(w[p] > 0) || m.off[p] == m.off[p + 1]) continue;
        const float wp = w[p];
        for (uint32_t z = m.off[p]; z < m.off[p + 1]; ++z) {
            const uint32_t t = orc_tgt(&m, p, z);
            scatter[t] = fmaf(m.frac[z], wp, scatter[t]);
        }
    }
}

/* Forward gather only: pred[p] = sum_z frac*blur[t] in z order (the first
 * loop of updc_core.c:1154-1158), for every pixel with links (weights are not
 * consulted).  Used by the synthetic generator's noise-free model. */
void orc_forward_gather(const uint8_t *lm_img, size_t nin, const float *blur, float *pred_out) {
    OrcLM m = orc_view(lm_img, nin);
    for (size_t p = 0; p < nin; ++p) {
        float pred = 0.0f;
        for (uint32_t z = m.off[p]; z < m.off[p + 1]; ++z) pred = fmaf(m.frac[z], blur[orc_tgt(&m, p, z)], pred);
        pred_out[p] = pred;
    }
}

/* ------------------------------------------------------------------ */
/* Initial signed drizzle: updc_core.c:885-973                         */
/* ------------------------------------------------------------------ */
/* nworkers == 1 follows the serial branch (:948-971); nworkers > 1 follows
 * the worker branch: worker w owns exposures k = w (mod nworkers) in private
 * partial sums (:885-903) that are combined in worker order from 0.0f
 * (:929-940).  Returns 0 or -1 for a non-finite estimate. */
int orc_drizzle(const uint8_t *const *lm_imgs, int nexp, size_t nin, size_t nout, const float *sci_all,
                const float *w_all, const float *background, int nworkers, float *estimate, float *coverage) {
    if (nworkers > nexp) nworkers = nexp;
    if (nworkers < 1) nworkers = 1;
    float *cov = calloc((size_t)nworkers * nout, sizeof(float));
    float *dat = calloc((size_t)nworkers * nout, sizeof(float));
    float *bkg = calloc((size_t)nworkers * nout, sizeof(float));
    if (!cov || !dat || !bkg) { free(cov); free(dat); free(bkg); return -3; }
    for (int k = 0; k < nexp; ++k) {
        const int wk = nworkers > 1 ? k % nworkers : 0;
        float *c = cov + (size_t)wk * nout, *d = dat + (size_t)wk * nout, *b = bkg + (size_t)wk * nout;
        OrcLM m = orc_view(lm_imgs[k], nin);
        const float *in = sci_all + (size_t)k * nin, *w = w_all + (size_t)k * nin;
        for (size_t p = 0; p < nin; ++p) {
            if (!(w[p] > 0) || m.off[p] == m.off[p + 1]) continue;
            const float wp = w[p];
            for (uint32_t z = m.off[p]; z < m.off[p + 1]; ++z) {
                const uint32_t t = orc_tgt(&m, p, z);
                const float q = m.frac[z] * wp;
                c[t] += q;
                d[t] = fmaf(q, in[p], d[t]);
                b[t] = fmaf(q, background[k], b[t]);
            }
        }
    }
    int rc = 0;
    for (size_t t = 0; t < nout; ++t) {
        float c, d, b;
        if (nworkers == 1) {
            c = cov[t]; d = dat[t]; b = bkg[t];
        } else {
            c = 0.0f; d = 0.0f; b = 0.0f;
            for (int wk = 0; wk < nworkers; ++wk) {
                c += cov[(size_t)wk * nout + t];
                d += dat[(size_t)wk * nout + t];
                b += bkg[(size_t)wk * nout + t];
            }
        }
        coverage[t] = c;
        if (c > 1e-12f) {
            estimate[t] = d / c - b / c;
            if (!isfinite(estimate[t])) rc = -1;
        } else {
            estimate[t] = 0.0f;
        }
    }
    free(cov); free(dat); free(bkg);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Weights: updc_core.c:127-129, 378-424                               */
/* ------------------------------------------------------------------ */
static int orc_cmp(const void *a, const void *b) {
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

/* Median of w[p] for p % 64 == 0 that are finite and > 0, over all
 * exposures in order; sample[n/2] after sorting.  Returns NAN if empty. */
float orc_weight_scale(const float *w_all, int nexp, size_t nin) {
    float *s = malloc(((size_t)nexp * (nin / 64 + 1)) * sizeof(float));
    size_t n = 0;
    for (int k = 0; k < nexp; ++k)
        for (size_t p = 0; p < nin; p += 64) {
            float v = w_all[(size_t)k * nin + p];
            if (isfinite(v) && v > 0) s[n++] = v;
        }
    if (!n) { free(s); return NAN; }
    qsort(s, n, sizeof(float), orc_cmp);
    float r = s[n / 2];
    free(s);
    return r;
}

/* Masking and normalisation in place.  Returns 0, -1 for an invalid
 * normalised weight, -2 for a negative / non-finite y_eff.  *masked counts
 * the pixels set to sci = w = 0. */
int orc_normalize_weights(float *sci_all, float *w_all, int nexp, size_t nin, float scale,
                          const float *pedestal, uint64_t *masked) {
    uint64_t nm = 0;
    for (int k = 0; k < nexp; ++k) {
        float *in = sci_all + (size_t)k * nin, *w = w_all + (size_t)k * nin;
        for (size_t p = 0; p < nin; ++p) {
            if (!isfinite(in[p]) || !isfinite(w[p]) || !(w[p] > 0)) {
                in[p] = 0.0f; w[p] = 0.0f; ++nm; continue;
            }
            w[p] /= scale;
            if (!isfinite(w[p]) || !(w[p] > 0)) return -1;
            const float y = in[p] + pedestal[k];
            if (!isfinite(y) || y < 0.0f) return -2;
        }
    }
    *masked = nm;
    return 0;
}

/* ------------------------------------------------------------------ */
/* PSF preparation: updc_core.c:777-821                                */
/* ------------------------------------------------------------------ */
/* Neumaier compensated f32 sum.  Returns NAN for an invalid pixel. */
float orc_psf_sum(const float *psf, size_t n) {
    float s = 0.0f, c = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        const float v = psf[i];
        if (!isfinite(v) || v < 0.0f) return NAN;
        const float t = s + v;
        if (fabsf(s) >= fabsf(v)) c += (s - t) + v;
        else c += (v - t) + s;
        s = t;
    }
    return s + c;
}

/* Normalised PSF wrapped so its centre lands on (0,0) of an FH x FW grid
 * (image is zeroed here).  `transposed` reads the stored raster as [x][y]. */
void orc_psf_wrap(const float *psf, int pw, int ph, int transposed, int FW, int FH, float inv_sum, float *image) {
    memset(image, 0, (size_t)FW * FH * sizeof(float));
    const int fw = transposed ? ph : pw, fh = transposed ? pw : ph;
    for (int y = 0; y < fh; ++y)
        for (int x = 0; x < fw; ++x) {
            const int yy = (y - fh / 2 + FH) % FH, xx = (x - fw / 2 + FW) % FW;
            const size_t src = transposed ? (size_t)x * pw + y : (size_t)y * pw + x;
            image[(size_t)yy * FW + xx] = psf[src] * inv_sum;
        }
}

/* ------------------------------------------------------------------ */
/* Spectrum products: updc_core.c:753-760, 988-995                     */
/* ------------------------------------------------------------------ */
/* Interleaved complex64 arrays of nc elements; products without FMA. */
void orc_spectrum_multiply(const float *src, const float *ker, float *out, size_t nc) {
    for (size_t p = 0; p < nc; ++p) {
        const float a = src[2 * p], b = src[2 * p + 1], c = ker[2 * p], d = ker[2 * p + 1];
        out[2 * p] = a * c - b * d;
        out[2 * p + 1] = a * d + b * c;
    }
}
void orc_accumulate_adjoint(float *dest, const float *src, const float *ker, size_t nc) {
    for (size_t p = 0; p < nc; ++p) {
        const float a = src[2 * p], b = src[2 * p + 1], c = ker[2 * p], d = ker[2 * p + 1];
        dest[2 * p] += a * c + b * d;
        dest[2 * p + 1] += b * c - a * d;
    }
}
/* Crop [0,H)x[0,W) of an FH x FW image and multiply by 1.0f/(float)(FW*FH)
 * (updc_core.c:741-743). */
void orc_crop_scale(const float *image, int FW, int FH, int W, int H, float *out) {
    const float s = 1.0f / (float)((size_t)FW * FH);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) out[(size_t)y * W + x] = image[(size_t)y * FW + x] * s;
}

/* ------------------------------------------------------------------ */
/* Multiplicative update: updc_core.c:1438-1456                        */
/* ------------------------------------------------------------------ */
/* Returns 0, -1 non-finite factor, -2 substantially negative correction,
 * -3 non-finite estimate; *bad receives the failing target index. */
int orc_update(float *x, const float *C, const float *D, size_t nout, float max_sens, float *sum_out,
               uint64_t *nneg_out, uint64_t *bad) {
    float sum = 0.0f;
    uint64_t nneg = 0;
    for (size_t t = 0; t < nout; ++t) {
        if (D[t] > 1e-8f) {
            float f = C[t] / D[t];
            if (!isfinite(f)) { *bad = t; return -1; }
            if (f < 0) {
                if (C[t] < -1.0e-6f * max_sens) { *bad = t; return -2; }
                f = 0.0f;
            }
            x[t] *= f;
        } else {
            x[t] = 0;
        }
        if (!isfinite(x[t])) { *bad = t; return -3; }
        sum += x[t];
        if (x[t] < 0) nneg++;
    }
    *sum_out = sum;
    *nneg_out = nneg;
    return 0;
}

/* ------------------------------------------------------------------ */
/* PSF polygon push-forward (psf_projector.c:71-146).                 */
/* Pinned against the reference's own psf_projector.c built into      */
/* oracle/_ref (tests/test_oracle_psf_projector.py), plus the identity */
/* mapping and flux conservation.                                      */
/* ------------------------------------------------------------------ */
static int orc_half_extent(double max_abs) {
    /* psf_projector.c:71-80 */
    double needed = max_abs - 0.5 - 1.0e-7;
    if (needed <= 0.0) return 0;
    return (int)ceil(needed);
}

/* Output half extents from the vertex lattice ((ih+1) x (iw+1) x 2). */
int orc_psf_extent(const double *v, int iw, int ih, int *hx, int *hy) {
    double mx = 0.0, my = 0.0;
    const size_t n = (size_t)(iw + 1) * (size_t)(ih + 1);
    for (size_t i = 0; i < n; ++i) {
        const double x = v[2 * i], y = v[2 * i + 1];
        if (!isfinite(x) || !isfinite(y)) return -1;
        if (fabs(x) > mx) mx = fabs(x);
        if (fabs(y) > my) my = fabs(y);
    }
    *hx = orc_half_extent(mx);
    *hy = orc_half_extent(my);
    return 0;
}

/* out ((2hy+1) x (2hx+1), zeroed here) += psf pushed through the quads,
 * psf_projector.c:107-146; returns the overlap count. */
uint64_t orc_psf_project(const float *psf, int iw, int ih, const double *v, int hx, int hy, float *out) {
    const int ow = 2 * hx + 1, oh = 2 * hy + 1;
    memset(out, 0, (size_t)ow * (size_t)oh * sizeof(float));
    uint64_t overlaps = 0;
    for (int y = 0; y < ih; ++y)
        for (int x = 0; x < iw; ++x) {
            const float value = psf[(size_t)y * iw + x];
            if (value == 0.0f) continue;
            const size_t a = 2 * ((size_t)y * (size_t)(iw + 1) + (size_t)x);
            const size_t b = a + 2 * (size_t)(iw + 1);
            OrcPt q[4] = {{v[a], v[a + 1]}, {v[b], v[b + 1]}, {v[b + 2], v[b + 3]}, {v[a + 2], v[a + 3]}};
            const float full = orc_poly_area(q, 4);
            if (!isfinite(full) || !(full > 1.0e-12f)) continue;
            double mnx = q[0].x, mxx = q[0].x, mny = q[0].y, mxy = q[0].y;
            for (int j = 1; j < 4; ++j) {
                if (q[j].x < mnx) mnx = q[j].x;
                if (q[j].x > mxx) mxx = q[j].x;
                if (q[j].y < mny) mny = q[j].y;
                if (q[j].y > mxy) mxy = q[j].y;
            }
            int x0 = (int)floor(mnx + 0.5), x1 = (int)floor(mxx + 0.5);
            int y0 = (int)floor(mny + 0.5), y1 = (int)floor(mxy + 0.5);
            if (x0 < -hx) x0 = -hx;
            if (x1 > hx) x1 = hx;
            if (y0 < -hy) y0 = -hy;
            if (y1 > hy) y1 = hy;
            for (int ty = y0; ty <= y1; ++ty)
                for (int tx = x0; tx <= x1; ++tx) {
                    const float area = orc_quad_overlap(q, tx, ty);
                    if (!(area > 1.0e-8f)) continue;
                    const float frac = area / full;
                    const size_t o = (size_t)(ty + hy) * ow + (size_t)(tx + hx);
                    out[o] = fmaf(value, frac, out[o]);
                    ++overlaps;
                }
        }
    (void)oh;
    return overlaps;
}
