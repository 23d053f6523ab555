// Context, setup and iteration driver of the MI355X UPDC library (C ABI).
//
// One updc_ctx holds every exposure's device data for the whole solve, so the
// iteration never touches the host: LinkMaps and their CSC transposes, PSF
// spectra, normalised SCI/WHT, FreezeMu masks, the estimate and the
// denominator stay resident in HBM (288 GB per MI355X holds every workload in
// the survey with room to spare).  One iteration is eight batched launches
// (every exposure inside each launch) plus one bookkeeping kernel, optionally
// captured once as a hipGraph and replayed.
#include "updc_hip.h"
#include "fft.hpp"
#include "stages.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <unistd.h>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace updc {

// ------------------------------------------------------------------ errors
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) throw Error(UPDC_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

thread_local std::string g_last_error;

template <class T> struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    T *alloc(size_t count) {
        reset();
        if (count == 0) count = 1;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            throw Error(UPDC_ERR_NOMEM, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
        return p;
    }
};

// ------------------------------------------------------------------ host geometry
// Host copy of the polygon policy (overlap.c:72-116) for the C ABI symbols and
// the self-test; the GPU builder has its own device copy in stages.hip.
namespace geo {
struct P { double x, y; };
static bool in(P p, int s, double v) { return s == 0 ? p.x >= v : s == 1 ? p.x <= v : s == 2 ? p.y >= v : p.y <= v; }
static P cross(P a, P b, int s, double v) {
    P r = a;
    if (s < 2) {
        const double d = b.x - a.x;
        if (std::fabs(d) > 1.0e-15) { const double t = (v - a.x) / d; r.x = v; r.y = a.y + t * (b.y - a.y); }
    } else {
        const double d = b.y - a.y;
        if (std::fabs(d) > 1.0e-15) { const double t = (v - a.y) / d; r.x = a.x + t * (b.x - a.x); r.y = v; }
    }
    return r;
}
static float area(const P *p, int n) {
    if (n < 3) return 0.0f;
    const double ox = p[0].x, oy = p[0].y;
    float acc = 0.0f;
    for (int i = 0; i < n; ++i) {
        const int j = (i + 1) % n;
        const float xi = (float)(p[i].x - ox), yi = (float)(p[i].y - oy);
        const float xj = (float)(p[j].x - ox), yj = (float)(p[j].y - oy);
        acc += std::fmaf(xi, yj, -(xj * yi));
    }
    return 0.5f * std::fabs(acc);
}
static float overlap(const P q[4], int x, int y) {
    P a[16], b[16];
    P *inp = a, *out = b;
    int n = 4;
    for (int j = 0; j < 4; ++j) inp[j] = q[j];
    const double lim[4] = {(double)x - 0.5, (double)x + 0.5, (double)y - 0.5, (double)y + 0.5};
    for (int s = 0; s < 4; ++s) {
        int m = 0;
        P prev = inp[n - 1];
        bool pin = in(prev, s, lim[s]);
        for (int i = 0; i < n; ++i) {
            const P cur = inp[i];
            const bool cin = in(cur, s, lim[s]);
            if (cin) {
                if (!pin && m < 16) out[m++] = cross(prev, cur, s, lim[s]);
                if (m < 16) out[m++] = cur;
            } else if (pin && m < 16) {
                out[m++] = cross(prev, cur, s, lim[s]);
            }
            prev = cur;
            pin = cin;
        }
        n = m;
        if (n < 3) return 0.0f;
        std::swap(inp, out);
    }
    return area(inp, n);
}
}  // namespace geo

// ------------------------------------------------------------------ helpers
static double host_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
static bool seven_smooth(long n) {
    if (n < 1) return false;
    for (int p : {2, 3, 5, 7}) while (n % p == 0) n /= p;
    return n == 1;
}
static int next_even_smooth(int n) {
    if (n < 2) n = 2;
    for (;; ++n)
        if (!(n & 1) && seven_smooth(n)) return n;
}
static int next_smooth(int n) {
    if (n < 1) n = 1;
    for (;; ++n)
        if (seven_smooth(n)) return n;
}
// Neumaier-compensated sum of the raw PSF (updc_core.c:777-789): f32 as the
// reference, f64 for the FP64 tier.  A negative or non-finite pixel fails as
// the reference does
template <class T> static T psf_sum(const float *psf, size_t n) {
    T s = 0, c = 0;
    for (size_t i = 0; i < n; ++i) {
        const T v = (T)psf[i];
        if (!std::isfinite(psf[i]) || psf[i] < 0.0f) throw Error(UPDC_ERR_NUMERIC, "invalid PSF pixel");
        const T t = s + v;
        if (std::fabs(s) >= std::fabs(v)) c += (s - t) + v;
        else c += (v - t) + s;
        s = t;
    }
    return s + c;
}
static size_t align4(size_t v) { return (v + 3u) & ~(size_t)3u; }
static size_t linkmap_bytes(uint64_t nin, uint64_t nlinks, uint32_t enc) {
    size_t pos = 64 + (size_t)(nin + 1) * 4u;
    if (enc == UPDC_LM_DELTA16) return align4(pos + nin * 4u + nlinks * 2u) + nlinks * 4u;
    if (enc == UPDC_LM_ABS32) return pos + nlinks * 8u;
    return 0;
}

// The gather stages links with aligned 16-byte loads, whose last block may
// extend up to 15 bytes past the image: allocate whole 16-byte blocks.
static size_t lm_alloc_bytes(size_t bytes) { return (bytes + 15) & ~(size_t)15; }

struct HostExp {
    int iw = 0, ih = 0;
    size_t nin = 0, pix = 0;
    std::vector<float> sci, wht, psf;
    int pw = 0, ph = 0, psf_t = 0;
    double background = 0, pedestal = 0;
    float background_f = 0, pedestal_f = 0, mean_offset_f = 0;
    DevBuf<uint8_t> lm;
    size_t lm_bytes = 0;
    uint32_t encoding = 0;
    uint64_t nlinks = 0;
    DevBuf<uint32_t> toff;
    DevBuf<uint2> tent;
    uint64_t ncsc = 0;
    // the compact copy of the CSC the iteration's adjoint gather reads (c16)
    bool c16 = false;
    DevBuf<uint32_t> c_toff, c_pb;
    DevBuf<uint8_t> c_cnt;
    DevBuf<uint16_t> c_dp;
    DevBuf<float> c_fr;
    size_t csc_bytes(size_t nout) const {
        return c16 ? ((nout + 63) / 64 + 1) * 4 + (nout + 63) / 64 * 4 + nout + ncsc * 6 : (nout + 1) * 4 + ncsc * 8;
    }
    ExpDev view() const {
        ExpDev e{};
        e.nin = (uint32_t)nin;
        e.encoding = encoding;
        const uint8_t *b = lm.p;
        size_t pos = 64;
        e.off = (const uint32_t *)(b + pos);
        pos += (nin + 1) * 4u;
        if (encoding == UPDC_LM_DELTA16) {
            e.base = (const uint32_t *)(b + pos);
            pos += nin * 4u;
            e.d16 = (const uint16_t *)(b + pos);
            pos = align4(pos + nlinks * 2u);
        } else {
            e.t32 = (const uint32_t *)(b + pos);
            pos += nlinks * 4u;
        }
        e.frac = (const float *)(b + pos);
        e.toff = toff.p;
        e.tent = tent.p;
        e.c16 = c16 ? 1u : 0u;
        e.c_toff = c_toff.p;
        e.c_pb = c_pb.p;
        e.c_cnt = c_cnt.p;
        e.c_dp = c_dp.p;
        e.c_fr = c_fr.p;
        e.pix = pix;
        e.background = background;
        e.pedestal = pedestal;
        e.mean_offset = background + pedestal;
        e.background_f = background_f;
        e.pedestal_f = pedestal_f;
        e.mean_offset_f = mean_offset_f;
        return e;
    }
};

struct Ctx;
struct PipeBase {
    virtual ~PipeBase() = default;
    virtual void prepare(Ctx &c) = 0;
    virtual void iterate(Ctx &c, int n, updc_freeze_row *rows, updc_iter_info *info) = 0;
    virtual void get_estimate(Ctx &c, double *d, float *f) = 0;
    virtual void set_estimate(Ctx &c, const double *d) = 0;
    virtual void get_wht(Ctx &c, float *out) = 0;
    virtual void get_denominator(Ctx &c, double *out) = 0;
    virtual void get_initial(Ctx &c, double *est, double *cov) = 0;
    virtual void get_psf_spectrum(Ctx &c, int k, double *out) = 0;
    virtual void stage_forward(Ctx &c, const double *x, double *blur) = 0;
    virtual void stage_model(Ctx &c, const double *x, double *pred) = 0;
    virtual void stage_detector(Ctx &c, int k, const double *blur, double *wr, double *scatter, double *stats) = 0;
    virtual void stage_adjoint(Ctx &c, const double *v, double *out) = 0;
    virtual void stage_scatter(Ctx &c, int k, const double *v, double *out) = 0;
    virtual void stage_adjoint_images(Ctx &c, const double *v, double *out) = 0;
    virtual void fft_r2c(Ctx &c, const double *img, double *spec) = 0;
    virtual void fft_c2r(Ctx &c, const double *spec, double *img) = 0;
    virtual void get_frozen(Ctx &c, int k, uint8_t *out) = 0;
    virtual void reset_frozen(Ctx &c) = 0;
    virtual void invalidate_graph() = 0;  // after a flag change that alters the launches
    virtual int spectrum_layout(const Ctx &c) const = 0;  // SpecLayout of the batched iteration's spectra
    // The plain form's extra spectra, allocated here (at prepare or when the
    // flag is switched on), never inside launch_iteration_plain, which may
    // run under stream capture where hipMalloc is not allowed.
    virtual void ensure_plain_buffers(Ctx &c) = 0;
    // device-pointer operators (element type TF), on the caller's stream
    virtual void op_r2c(Ctx &c, const void *img, int rows, int cols, int pitch, void *spec, hipStream_t st) = 0;
    // op_c2r with a caller-provided column-inverse workspace (device memory,
    // op_c2r_workspace_bytes() of it); allocates nothing
    virtual void op_c2r_ws(Ctx &c, const void *spec, void *img, int rows, int cols, int pitch, double scale,
                           void *ws, hipStream_t st) = 0;
    virtual size_t op_c2r_workspace_bytes() const = 0;
    // op_c2r_ws of `rows` rows would allocate nothing (four-step scratch reserved)
    virtual bool op_c2r_ws_ready(int rows) const = 0;
    virtual void op_c2r(Ctx &c, const void *spec, void *img, int rows, int cols, int pitch, double scale,
                        hipStream_t st) = 0;
    virtual void op_gather(Ctx &c, int k, const void *img, void *out, hipStream_t st) = 0;
    virtual void op_scatter(Ctx &c, int k, const void *v, void *img, hipStream_t st) = 0;
    double max_sens = 0;
    std::vector<std::string> stage_names;
    std::vector<double> stage_ms;
    std::vector<int> stage_n;
};

// ------------------------------------------------------------------ set-up caches
// The reference's persistent set-up caches (updc_core.c:145-158, 790-843,
// 905-976, 1024-1083): raw little-endian files in the cache root, used only
// when their size is exact, written through a temporary file and a rename.
static bool load_exact(const std::string &path, void *dst, size_t bytes) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    bool ok = std::fseek(f, 0, SEEK_END) == 0 && (size_t)std::ftell(f) == bytes && std::fseek(f, 0, SEEK_SET) == 0 &&
              std::fread(dst, 1, bytes, f) == bytes;
    std::fclose(f);
    return ok;
}
static void save_atomic(const std::string &path, const void *src, size_t bytes) {
    const std::string tmp = path + ".tmp." + std::to_string((long)getpid());
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) throw Error(UPDC_ERR_IO, "cannot create cache file " + tmp);
    const bool ok = std::fwrite(src, 1, bytes, f) == bytes;
    if (std::fclose(f) != 0 || !ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        throw Error(UPDC_ERR_IO, "cache write failed: " + path);
    }
}

struct TuneRow {
    int w, h, edge;  // shape and column kernel (1 edge-stage, 0 LDS-only)
    double ms;       // the iteration's transforms; < 0 when the shape is unsupported
};

struct Ctx {
    int W = 0, H = 0, FW = 0, FH = 0, prec = 0, device = 0;
    unsigned flags = 0;
    int workers = 1;
    int col_edge = -1;           // column kernel: -1 automatic, 0 LDS-only, 1 edge-stage
    std::string fft_cache_dir;   // autotune cache directory ("" = no cache)
    bool fft_retune = false;     // ignore a cached autotune result
    std::vector<TuneRow> tune;   // the last autotune's measurements
    bool tune_cached = false;
    std::string setup_cache_dir; // PSF spectrum / initial estimate / denominator caches ("" = none)
    int setup_cache_hits = 0;    // caches read by the last prepare (bit 0 PSF spectra, 1 initial, 2 D)
    std::vector<std::unique_ptr<HostExp>> ex;
    hipStream_t st = nullptr;
    bool prepared = false;
    int iter_done = 0;  // the last updc_iterate: iterations completed (estimate updated)
    int iter_rows = 0;  // ... and iterations whose freeze rows are valid
    updc_setup_times setup{};                 // set-up stage times (seconds)
    std::vector<updc_iter_times> iter_times;  // per iteration of the last updc_iterate
    float weight_scale = 0;
    unsigned long long masked = 0;
    std::unique_ptr<PipeBase> pipe;
    size_t nout() const { return (size_t)W * H; }
    size_t npix() const {
        size_t s = 0;
        for (auto &e : ex) s += e->nin;
        return s;
    }
    bool quiet() const { return flags & UPDC_FLAG_QUIET; }
};

// ------------------------------------------------------------------ pipeline
template <class TF, class TA> struct Pipe : PipeBase {
    // Mixed tier (TA wider than TF): the summed adjoint spectrum stays in TA
    // and its inverse transform, columns then rows, runs in TA (fa), so C is
    // never rounded to TF.  The per-exposure transforms stay in TF.
    static constexpr bool kWide = sizeof(TA) > sizeof(TF);
    Fft2<TF> fft;
    Fft2<TA> fa;  // kWide: the TA transform of the summed adjoint spectrum (and of D)
    DevBuf<ExpDev> d_ex;
    std::vector<ExpDev> h_ex;
    DevBuf<TF> sci, w, wr, blur, x, cov;
    DevBuf<TF> x_init;  // the initial estimate, kept for updc_get_initial_f64
    DevBuf<TA> C, D;
    DevBuf<uint8_t> frozen;
    DevBuf<Cx<TF>> psf_spec, spec_x, spec_k, spec_tmp;
    DevBuf<Cx<TA>> spec_acc;  // summed adjoint spectrum (per-exposure form; kWide: every form), pitch rp(FH)
    DevBuf<Cx<TA>> spec_xa;   // kWide: the sum after its TA column inverse, rows [0, H)
    bool acc_fused = false;   // kWide: that inverse runs inside the accumulating column pass
    bool fused_update = false;  // the final row C2R applies the update (persistent row pass available)
    int lay = SL_COL;           // layout of the spectra the batched iteration's column passes write
    bool corr_pair = false;     // kWide: the summed correction also leaves its column pass in lay
    bool any_c16 = false, any_wide = false;  // exposures with / without the compact CSC
    DevBuf<ExpDiag> diag_cur, diag_hist;
    DevBuf<IterDiag> idiag_cur, idiag_hist;
    DevBuf<double> partials;
    DevBuf<unsigned long long> negs, key;
    DevBuf<unsigned int> ticket, ctr, err;
    int hist_cap = 0;
    uint32_t max_nin = 0;
    hipGraphExec_t graph = nullptr;
    int graph_cap = -1;

    ~Pipe() override {
        if (graph) (void)hipGraphExecDestroy(graph);
        fft.release();
        fa.release();
    }

    int N(Ctx &c) const { return (int)c.ex.size(); }
    TF scale(Ctx &c) const { return (TF)1 / (TF)((size_t)c.FW * c.FH); }
    TA scale_a(Ctx &c) const { return (TA)1 / (TA)((size_t)c.FW * c.FH); }

    // C = crop C2R(sum_k conj(P^_k) spec_k) / nf from the row spectra spec_k
    // of the n exposures (updc_core.c:985-995, 1280-1287).  Records one event
    // after each launch but the last.  With update_x the final row C2R also
    // applies the update to x (updc_core.c:1437-1457) where the persistent
    // row pass allows it, and returns true; C is then not stored.
    // l: the layout the column pass writes its output in (spec_k stays SL_COL).
    template <class Mark> bool adjoint_sum(Ctx &c, int n, Mark mark, bool update_x, int l) {
        const FftShape &s = fft.s;
        const size_t kstride = fft.spec_elems(c.H), fstride = fft.spec_elems(s.FH);
        if constexpr (kWide) {
            // the sum leaves the accumulating pass in layout la
            const int la = acc_fused && corr_pair ? l : SL_COL;
            if (acc_fused) {
                // the column inverse of the sum runs in TA inside the accumulating pass
                HIPCHK(fft.template colacc<TA>(spec_k.p, kstride, c.H, spec_xa.p, c.H, c.H, psf_spec.p, fstride, c.H,
                                               n, &fft.col_wide, c.st, la));
            } else {
                HIPCHK(fft.template colacc<TA>(spec_k.p, kstride, c.H, spec_acc.p, s.FH, s.FH, psf_spec.p, fstride,
                                               c.H, n, nullptr, c.st, SL_COL));
                mark();
                HIPCHK(fa.template colpass<TA>(COL_INV, spec_acc.p, 0, s.FH, spec_xa.p, 0, c.H, nullptr, 0, s.FH, c.H,
                                               1, c.st));
            }
            mark();
            if (update_x && fused_update) {
                HIPCHK(fa.template row_c2r_update<TF>(spec_xa.p, c.H, x.p, D.p, c.W, c.H, scale_a(c), (TA)max_sens,
                                                      idiag_cur.p, partials.p, negs.p, ticket.p, c.st, la));
                return true;
            }
            HIPCHK(fa.row_c2r(spec_xa.p, 0, c.H, C.p, 0, c.W, c.W, scale_a(c), c.H, 1, c.st, la));
        } else {
            HIPCHK(fft.template colpass<TA>(COL_ADJ, spec_k.p, kstride, c.H, spec_x.p, 0, c.H, psf_spec.p, fstride,
                                            c.H, c.H, n, c.st, l));
            mark();
            if (update_x && fused_update) {
                HIPCHK(fft.template row_c2r_update<TF>(spec_x.p, c.H, x.p, (const TF *)D.p, c.W, c.H, scale(c),
                                                       (TF)max_sens, idiag_cur.p, partials.p, negs.p, ticket.p,
                                                       c.st, l));
                return true;
            }
            HIPCHK(fft.row_c2r(spec_x.p, 0, c.H, C.p, 0, c.W, c.W, scale(c), c.H, 1, c.st, l));
        }
        return false;
    }
    // The layout of the spectra the batched iteration's column passes write,
    // under the context's flags: SL_PAIR where prepare found the passes able
    // to use it, unless UPDC_FLAG_COLUMN_SPECTRA asks for SL_COL.
    int iter_layout(const Ctx &c) const { return (c.flags & UPDC_FLAG_COLUMN_SPECTRA) ? SL_COL : lay; }

    // ---- PSF spectra P^_k for every exposure into `out` (column-major,
    // N x f.spec_elems(FH), pitch rp(FH)).
    template <class TD> static void psf_spectra(Ctx &c, Fft2<TD> &f, Cx<TD> *out) {
        const FftShape &s = f.s;
        DevBuf<TD> img;
        img.alloc((size_t)s.FW * s.FH);
        DevBuf<float> raw;
        for (size_t k = 0; k < c.ex.size(); ++k) {
            HostExp &e = *c.ex[k];
            const size_t cnt = (size_t)e.pw * e.ph;
            TD inv;
            if (sizeof(TD) == 4) {
                const float sum = psf_sum<float>(e.psf.data(), cnt);
                if (!(sum > 0.0f) || !std::isfinite(sum)) throw Error(UPDC_ERR_NUMERIC, "PSF sum is zero/non-finite");
                inv = (TD)(1.0f / sum);
            } else {
                const double sum = psf_sum<double>(e.psf.data(), cnt);
                if (!(sum > 0.0) || !std::isfinite(sum)) throw Error(UPDC_ERR_NUMERIC, "PSF sum is zero/non-finite");
                inv = (TD)(1.0 / sum);
            }
            raw.alloc(cnt);
            HIPCHK(hipMemcpyAsync(raw.p, e.psf.data(), cnt * sizeof(float), hipMemcpyHostToDevice, c.st));
            HIPCHK(fill<TD>(img.p, (TD)0, img.n, c.st));
            HIPCHK(psf_wrap<TD>(raw.p, e.pw, e.ph, e.psf_t, s.FW, s.FH, inv, img.p, c.st));
            Cx<TD> *o = out + k * f.spec_elems(s.FH);
            HIPCHK(f.row_r2c(ROW_PLAIN, img.p, 0, s.FW, s.FW, nullptr, nullptr, 0, o, 0, s.FH, s.FH, 1, c.st));
            HIPCHK(f.template colpass<TD>(COL_FWD, o, 0, s.FH, o, 0, s.FH, nullptr, 0, s.FH, s.FH, 1, c.st));
            HIPCHK(hipStreamSynchronize(c.st));
        }
    }

    // ---- D = crop(C2R(sum_k conj(P^_k) R2C(pad(S_k^T w_k)))) / nf (updc_core.c:1026-1086).
    template <class TD>
    static void denominator(Ctx &c, Fft2<TD> &f, const Cx<TD> *pspec, const ExpDev *dex, const TD *wts, TD *Dout) {
        const FftShape &s = f.s;
        const int n = (int)c.ex.size();
        DevBuf<Cx<TD>> sk, sx;
        const size_t ks = f.spec_elems(c.H);
        sk.alloc((size_t)n * ks);
        sx.alloc(ks);
        HIPCHK(f.row_r2c(ROW_GATHER, nullptr, 0, 0, 0, dex, wts, c.W, sk.p, ks, c.H, c.H, n, c.st));
        HIPCHK(f.template colpass<TD>(COL_ADJ, sk.p, ks, c.H, sx.p, 0, c.H, pspec, f.spec_elems(s.FH), c.H, c.H, n,
                                      c.st));
        const TD sc = (TD)1 / (TD)((size_t)s.FW * s.FH);
        HIPCHK(f.row_c2r(sx.p, 0, c.H, Dout, 0, c.W, c.W, sc, c.H, 1, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
    }

    void build_csc(Ctx &c) {
        const size_t nout = c.nout();
        DevBuf<uint32_t> tcnt, cursor;
        tcnt.alloc(nout);
        cursor.alloc(nout + 1);
        DevBuf<uint8_t> scratch;
        scratch.alloc(scan_scratch_bytes(nout));
        DevBuf<unsigned long long> tot;
        tot.alloc(1);
        DevBuf<ExpDev> one;
        one.alloc(1);
        for (size_t k = 0; k < c.ex.size(); ++k) {
            HostExp &e = *c.ex[k];
            e.toff.alloc(nout + 1);
            ExpDev v = e.view();
            HIPCHK(hipMemcpyAsync(one.p, &v, sizeof(ExpDev), hipMemcpyHostToDevice, c.st));
            HIPCHK(fill<uint32_t>(tcnt.p, 0u, nout, c.st));
            HIPCHK(csc_count<TF>(one.p, (uint32_t)e.nin, w.p + e.pix, tcnt.p, c.st));
            HIPCHK(scan_u32(tcnt.p, e.toff.p, nout, scratch.p, tot.p, c.st));
            unsigned long long total = 0;
            HIPCHK(hipMemcpyAsync(&total, tot.p, sizeof(total), hipMemcpyDeviceToHost, c.st));
            HIPCHK(hipStreamSynchronize(c.st));
            if (total > 0xffffffffull) throw Error(UPDC_ERR_ARG, "CSC transpose exceeds uint32 capacity");
            e.ncsc = total;
            e.tent.alloc(total);
            HIPCHK(hipMemcpyAsync(cursor.p, e.toff.p, nout * sizeof(uint32_t), hipMemcpyDeviceToDevice, c.st));
            v = e.view();
            HIPCHK(hipMemcpyAsync(one.p, &v, sizeof(ExpDev), hipMemcpyHostToDevice, c.st));
            HIPCHK(csc_fill<TF>(one.p, (uint32_t)e.nin, w.p + e.pix, cursor.p, e.tent.p, c.st));
            HIPCHK(csc_sort(e.toff.p, e.tent.p, nout, c.st));
            // the compact copy for the iteration's adjoint gather, where it fits
            const size_t ntiles = (nout + 63) / 64;
            e.c_toff.alloc(ntiles + 1);
            e.c_pb.alloc(ntiles);
            e.c_cnt.alloc(nout);
            // 16 bytes to spare: the adjoint gather reads the entry runs in
            // aligned 16-byte pieces
            e.c_dp.alloc(total + 8);
            e.c_fr.alloc(total + 4);
            HIPCHK(fill<unsigned int>(err.p, 0u, 1, c.st));
            HIPCHK(csc16_build(e.toff.p, e.tent.p, nout, e.c_toff.p, e.c_pb.p, e.c_cnt.p, e.c_dp.p, e.c_fr.p, err.p,
                               c.st));
            unsigned int fl = 0;
            HIPCHK(hipMemcpyAsync(&fl, err.p, 4, hipMemcpyDeviceToHost, c.st));
            HIPCHK(hipStreamSynchronize(c.st));
            e.c16 = fl == 0;
            if (!e.c16) {
                e.c_toff.reset();
                e.c_pb.reset();
                e.c_cnt.reset();
                e.c_dp.reset();
                e.c_fr.reset();
            }
        }
        HIPCHK(fill<unsigned int>(err.p, 0u, 1, c.st));
        any_c16 = any_wide = false;
        for (auto &x : c.ex) (x->c16 ? any_c16 : any_wide) = true;
    }

    void upload_ex(Ctx &c) {
        h_ex.clear();
        for (auto &e : c.ex) h_ex.push_back(e->view());
        d_ex.alloc(h_ex.size());
        HIPCHK(hipMemcpyAsync(d_ex.p, h_ex.data(), h_ex.size() * sizeof(ExpDev), hipMemcpyHostToDevice, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
    }

    void prepare(Ctx &c) override {
        const int n = N(c);
        const size_t nout = c.nout(), npix = c.npix();
        double tm = host_seconds();
        // seconds since the last mark, the stream drained first
        auto lap = [&]() {
            HIPCHK(hipStreamSynchronize(c.st));
            const double t1 = host_seconds(), d = t1 - tm;
            tm = t1;
            return d;
        };
        fft.allow_edge = c.col_edge != 0;
        std::string e = fft.init(c.FW, c.FH);
        if (!e.empty()) throw Error(UPDC_ERR_ARG, e);
        if (kWide) {
            // edge-stage columns where possible: the FP64 LDS-only kernel stops at FH 6144
            fa.release();
            e = fa.init(c.FW, c.FH);
            if (!e.empty()) throw Error(UPDC_ERR_ARG, e);
            HIPCHK(fa.long_reserve(n));
        }
        HIPCHK(fft.long_reserve(n));
        c.setup.fft_plan = lap();
        const FftShape &s = fft.s;
        max_nin = 0;
        for (auto &x : c.ex) max_nin = std::max<uint32_t>(max_nin, (uint32_t)x->nin);

        // weights (updc_core.c:378-424): median of the p % 64 == 0 sample.
        std::vector<float> sample;
        for (auto &x : c.ex)
            for (size_t p = 0; p < x->nin; p += 64)
                if (std::isfinite(x->wht[p]) && x->wht[p] > 0) sample.push_back(x->wht[p]);
        if (sample.empty()) throw Error(UPDC_ERR_NUMERIC, "all input WHT pixels are zero");
        std::sort(sample.begin(), sample.end());
        c.weight_scale = sample[sample.size() / 2];
        if (!(c.weight_scale > 0) || !std::isfinite(c.weight_scale))
            throw Error(UPDC_ERR_NUMERIC, "invalid WHT normalization");
        sci.alloc(npix);
        w.alloc(npix);
        wr.alloc(npix);
        frozen.alloc(npix);
        err.alloc(1);
        key.alloc(1);
        DevBuf<unsigned long long> masked;
        masked.alloc(1);
        HIPCHK(fill<unsigned int>(err.p, 0u, 1, c.st));
        HIPCHK(fill<unsigned long long>(masked.p, 0ull, 1, c.st));
        {
            DevBuf<float> rs, rw;
            for (auto &x : c.ex) {
                rs.alloc(x->nin);
                rw.alloc(x->nin);
                HIPCHK(hipMemcpyAsync(rs.p, x->sci.data(), x->nin * 4, hipMemcpyHostToDevice, c.st));
                HIPCHK(hipMemcpyAsync(rw.p, x->wht.data(), x->nin * 4, hipMemcpyHostToDevice, c.st));
                if (sizeof(TF) == 4)
                    HIPCHK(normalize_weights<TF>(rs.p, rw.p, x->nin, c.weight_scale, (double)c.weight_scale,
                                                 x->pedestal_f, x->pedestal, sci.p + x->pix, w.p + x->pix, masked.p,
                                                 err.p, c.st));
                else
                    HIPCHK(normalize_weights<TF>(rs.p, rw.p, x->nin, c.weight_scale, (double)c.weight_scale,
                                                 x->pedestal_f, x->pedestal, sci.p + x->pix, w.p + x->pix, masked.p,
                                                 err.p, c.st));
                HIPCHK(hipStreamSynchronize(c.st));
            }
        }
        unsigned int ebits = 0;
        HIPCHK(hipMemcpy(&ebits, err.p, 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&c.masked, masked.p, 8, hipMemcpyDeviceToHost));
        if (ebits & UPDC_E_WEIGHT) throw Error(UPDC_ERR_NUMERIC, "invalid normalized WHT");
        if (ebits & UPDC_E_YEFF) throw Error(UPDC_ERR_NUMERIC, "negative/non-finite y_eff reached solver");
        c.setup.weight_scale = lap();

        build_csc(c);
        upload_ex(c);
        c.setup.csc_transpose = lap();

        // PSF spectra (cached for the F32 tier in the reference's FFTW layout:
        // FH x (FW/2+1) complex64, row-major)
        const bool cache = !c.setup_cache_dir.empty() && c.prec == UPDC_PREC_F32;
        c.setup_cache_hits = 0;
        psf_spec.alloc((size_t)n * fft.spec_elems(s.FH));
        DevBuf<Cx<TA>> psf_wide;  // kWide: the TA spectra, for D
        {
            const size_t P = rp(s.FH), one = fft.spec_elems(s.FH), nc = (size_t)s.FH * s.NC;
            std::vector<std::string> paths(n);
            std::vector<Cx<float>> file, cm;
            bool all = cache;
            for (int k = 0; k < n; ++k) {
                char nm[96];
                std::snprintf(nm, sizeof nm, "/psf_spectrum_%02d_%dx%d_t%d.c64", k + 1, s.FW, s.FH, c.ex[k]->psf_t);
                paths[k] = c.setup_cache_dir + nm;
            }
            if (all) {
                file.resize(nc);
                cm.assign(one, Cx<float>{0.f, 0.f});
                for (int k = 0; k < n && all; ++k) {
                    if (!load_exact(paths[k], file.data(), nc * sizeof(Cx<float>))) { all = false; break; }
                    for (int r = 0; r < s.FH; ++r)
                        for (int v = 0; v < s.NC; ++v) cm[(size_t)v * P + r] = file[(size_t)r * s.NC + v];
                    HIPCHK(hipMemcpy(psf_spec.p + (size_t)k * one, cm.data(), one * sizeof(Cx<float>),
                                     hipMemcpyHostToDevice));
                }
            }
            if (all) {
                c.setup_cache_hits |= 1;
            } else if (kWide) {
                // the TA spectra rounded once to TF: the per-exposure products
                // use the nearest TF values of the spectra D is built from
                psf_wide.alloc((size_t)n * fa.spec_elems(s.FH));
                psf_spectra<TA>(c, fa, psf_wide.p);
                HIPCHK((convert<TA, TF>((const TA *)psf_wide.p, (TF *)psf_spec.p, 2 * psf_spec.n, c.st)));
            } else {
                psf_spectra<TF>(c, fft, psf_spec.p);
                if (cache) {
                    HIPCHK(hipStreamSynchronize(c.st));
                    file.resize(nc);
                    cm.resize(one);
                    for (int k = 0; k < n; ++k) {
                        HIPCHK(hipMemcpy(cm.data(), psf_spec.p + (size_t)k * one, one * sizeof(Cx<float>),
                                         hipMemcpyDeviceToHost));
                        for (int r = 0; r < s.FH; ++r)
                            for (int v = 0; v < s.NC; ++v) file[(size_t)r * s.NC + v] = cm[(size_t)v * P + r];
                        save_atomic(paths[k], file.data(), nc * sizeof(Cx<float>));
                    }
                }
            }
        }

        c.setup.psf_spectra = lap();
        // initial drizzle (updc_core.c:905-983)
        x.alloc(nout);
        cov.alloc(nout);
        {
            const std::string pe = c.setup_cache_dir + "/initial_estimate.f32";
            const std::string pc = c.setup_cache_dir + "/initial_coverage.f32";
            std::vector<float> he(cache ? nout : 0), hc(cache ? nout : 0);
            if (cache && load_exact(pe, he.data(), nout * 4) && load_exact(pc, hc.data(), nout * 4)) {
                HIPCHK(hipMemcpy(x.p, he.data(), nout * 4, hipMemcpyHostToDevice));
                HIPCHK(hipMemcpy(cov.p, hc.data(), nout * 4, hipMemcpyHostToDevice));
                c.setup_cache_hits |= 2;
            } else {
                HIPCHK(fill<unsigned int>(err.p, 0u, 1, c.st));
                HIPCHK(drizzle<TF>(d_ex.p, n, sci.p, w.p, std::max(1, std::min(c.workers, n)), nout, x.p, cov.p,
                                   err.p, c.st));
                HIPCHK(hipMemcpyAsync(&ebits, err.p, 4, hipMemcpyDeviceToHost, c.st));
                HIPCHK(hipStreamSynchronize(c.st));
                if (ebits & UPDC_E_DRIZZLE) throw Error(UPDC_ERR_NUMERIC, "non-finite signed initial estimate");
                if (cache) {
                    HIPCHK(hipMemcpy(he.data(), x.p, nout * 4, hipMemcpyDeviceToHost));
                    HIPCHK(hipMemcpy(hc.data(), cov.p, nout * 4, hipMemcpyDeviceToHost));
                    save_atomic(pe, he.data(), nout * 4);
                    save_atomic(pc, hc.data(), nout * 4);
                }
            }
        }
        x_init.alloc(nout);
        HIPCHK(hipMemcpyAsync(x_init.p, x.p, nout * sizeof(TF), hipMemcpyDeviceToDevice, c.st));
        c.setup.initial_drizzle = lap();

        // denominator
        D.alloc(nout);
        char dname[64];
        std::snprintf(dname, sizeof dname, "/denominator_%dx%d.f32", s.FW, s.FH);
        const std::string pd = c.setup_cache_dir + dname;
        std::vector<float> hd(cache ? nout : 0);
        if (cache && load_exact(pd, hd.data(), nout * 4)) {
            HIPCHK(hipMemcpy(D.p, hd.data(), nout * 4, hipMemcpyHostToDevice));
            c.setup_cache_hits |= 4;
        } else if (sizeof(TA) == sizeof(TF)) {
            denominator<TF>(c, fft, psf_spec.p, d_ex.p, w.p, (TF *)D.p);
            if (cache) {
                HIPCHK(hipStreamSynchronize(c.st));
                HIPCHK(hipMemcpy(hd.data(), D.p, nout * 4, hipMemcpyDeviceToHost));
                save_atomic(pd, hd.data(), nout * 4);
            }
        } else {
            // mixed tier: the denominator is computed entirely in FP64
            DevBuf<TA> wd;
            wd.alloc(npix);
            HIPCHK((convert<TF, TA>(w.p, wd.p, npix, c.st)));
            denominator<TA>(c, fa, psf_wide.p, d_ex.p, wd.p, D.p);
            psf_wide.reset();
        }
        HIPCHK(fill<unsigned long long>(key.p, 0ull, 1, c.st));
        HIPCHK(max_key<TA>(D.p, nout, key.p, c.st));
        unsigned long long kk = 0;
        HIPCHK(hipMemcpyAsync(&kk, key.p, 8, hipMemcpyDeviceToHost, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        max_sens = std::max(0.0, dkey_inv(kk));
        if (!(max_sens > 0) || !std::isfinite(max_sens)) throw Error(UPDC_ERR_NUMERIC, "invalid H^T W sensitivity");
        c.setup.sensitivity = lap();

        // iteration buffers
        blur.alloc((size_t)n * nout);
        C.alloc(nout);
        spec_x.alloc(fft.spec_elems(std::max(c.H, s.FH)));
        spec_k.alloc((size_t)n * fft.spec_elems(c.H));
        fused_update = kWide ? fa.row_persistent : fft.row_persistent;
        if (kWide) {
            acc_fused = fft.acc_inverse_ok(fft.col_wide);
            if (!acc_fused) spec_acc.alloc(fft.spec_elems(s.FH));
            spec_xa.alloc(fa.spec_elems(c.H));
        }
        // the spectra a column pass writes for a row C2R in the pair layout
        // where both allow it (kWide: the summed correction only where the
        // accumulating pass also inverts it and the TA row C2R takes pairs;
        // a separate TA inverse writes column-major spectra)
        lay = fft.pair_ok() ? SL_PAIR : SL_COL;
        corr_pair = kWide && acc_fused && fa.row_pair_ok();
        HIPCHK(fill<uint8_t>(frozen.p, (uint8_t)0, npix, c.st));
        diag_cur.alloc(n);
        idiag_cur.alloc(1);
        // per-iteration history, sized so a captured graph survives most calls
        hist_cap = 1024;
        diag_hist.alloc((size_t)hist_cap * n);
        idiag_hist.alloc(hist_cap);
        const int nred = std::max({update_blocks(), fft.row_grid, fa.row_grid});  // reduction slots per launch
        partials.alloc(nred);
        negs.alloc(nred);
        ticket.alloc(1);
        ctr.alloc(1);
        HIPCHK(fill<unsigned int>(ticket.p, 0u, 1, c.st));
        if (c.flags & UPDC_FLAG_PER_EXPOSURE) ensure_plain_buffers(c);
        HIPCHK(hipStreamSynchronize(c.st));
    }

    // Plain form (UPDC_FLAG_PER_EXPOSURE): the reference's dataflow, one launch
    // per exposure and stage (updc_core.c:1140-1185, 1255-1287): X^ once;
    // per exposure blur = crop C2R(X^ P^_k), detector, R2C(scatter),
    // Â += conj(P^_k) F(scatter); one inverse of Â; update.
    void launch_iteration_plain(Ctx &c, hipEvent_t *ev) {
        const int n = N(c);
        const FftShape &s = fft.s;
        const size_t nout = c.nout();
        const size_t kstride = fft.spec_elems(c.H), fstride = fft.spec_elems(s.FH);
        const TF sc = scale(c);
        auto mark = [&](int i) { if (ev) HIPCHK(hipEventRecord(ev[i], c.st)); };
        if (!spec_tmp.p || !spec_acc.p) throw Error(UPDC_ERR_STATE, "per-exposure form buffers not allocated");
        Cx<TF> *xh = spec_x.p;  // full FH-row spectrum (pitch rp(FH)) after the column pass
        mark(0);
        HIPCHK(fft.row_r2c(ROW_PLAIN, x.p, 0, c.W, c.W, nullptr, nullptr, 0, xh, 0, s.FH, c.H, 1, c.st));
        HIPCHK(fft.template colpass<TA>(COL_FWD, xh, 0, s.FH, xh, 0, s.FH, nullptr, 0, c.H, s.FH, 1, c.st));
        HIPCHK(fill<TA>((TA *)spec_acc.p, (TA)0, 2 * fstride, c.st));
        mark(1);
        for (int k = 0; k < n; ++k) {
            const Cx<TF> *pk = psf_spec.p + (size_t)k * fstride;
            Cx<TF> *sk = spec_k.p + (size_t)k * kstride;
            HIPCHK(fft.spec_mul(xh, pk, spec_tmp.p, c.st));
            HIPCHK(fft.template colpass<TA>(COL_INV, spec_tmp.p, 0, s.FH, sk, 0, c.H, nullptr, 0, s.FH, c.H, 1,
                                            c.st));
            HIPCHK(fft.row_c2r(sk, 0, c.H, blur.p + (size_t)k * nout, 0, c.W, c.W, sc, c.H, 1, c.st));
            HIPCHK((gather_freeze<TF, TA>(d_ex.p, k, k + 1, (uint32_t)c.ex[k]->nin, blur.p, nout, sci.p, w.p,
                                          frozen.p, wr.p, diag_cur.p, idiag_cur.p, c.st)));
            HIPCHK(fft.row_r2c(ROW_GATHER, nullptr, 0, 0, 0, d_ex.p + k, wr.p, c.W, sk, 0, c.H, c.H, 1, c.st));
            HIPCHK(fft.template colpass<TA>(COL_FWD, sk, 0, c.H, spec_tmp.p, 0, s.FH, nullptr, 0, c.H, s.FH, 1,
                                            c.st));
            HIPCHK(fft.template spec_acc<TA>(spec_tmp.p, pk, spec_acc.p, c.st));
        }
        mark(2);
        if constexpr (kWide) {
            // the column inverse of the sum with the batched form's plan (the
            // accumulate pass's own), so both forms agree bit for bit
            if (acc_fused)
                HIPCHK(fft.template acc_inverse<TA>(spec_acc.p, spec_xa.p, c.H, c.H, fft.col_wide, c.st));
            else
                HIPCHK(fa.template colpass<TA>(COL_INV, spec_acc.p, 0, s.FH, spec_xa.p, 0, c.H, nullptr, 0, s.FH,
                                               c.H, 1, c.st));
            HIPCHK(fa.row_c2r(spec_xa.p, 0, c.H, C.p, 0, c.W, c.W, scale_a(c), c.H, 1, c.st));
        } else {
            HIPCHK(fft.template spec_cast<TA>(spec_acc.p, spec_tmp.p, c.st));
            HIPCHK(fft.template colpass<TA>(COL_INV, spec_tmp.p, 0, s.FH, spec_x.p, 0, c.H, nullptr, 0, s.FH, c.H, 1,
                                            c.st));
            HIPCHK(fft.row_c2r(spec_x.p, 0, c.H, C.p, 0, c.W, c.W, sc, c.H, 1, c.st));
        }
        mark(3);
        HIPCHK((update<TF, TA>(x.p, C.p, D.p, nout, (TA)max_sens, idiag_cur.p, partials.p, negs.p, ticket.p, c.st)));
        mark(4);
        HIPCHK(commit_diag(diag_cur.p, n, idiag_cur.p, diag_hist.p, idiag_hist.p, hist_cap, ctr.p, c.st));
    }

    // One iteration's launches (identical eager or under graph capture).
    void launch_iteration(Ctx &c, hipEvent_t *ev) {
        if (c.flags & UPDC_FLAG_PER_EXPOSURE) return launch_iteration_plain(c, ev);
        const int n = N(c);
        const FftShape &s = fft.s;
        const size_t nout = c.nout();
        const size_t kstride = fft.spec_elems(c.H);
        const TF sc = scale(c);
        const int l = iter_layout(c);
        auto mark = [&](int i) { if (ev) HIPCHK(hipEventRecord(ev[i], c.st)); };
        mark(0);
        HIPCHK(fft.row_r2c(ROW_PLAIN, x.p, 0, c.W, c.W, nullptr, nullptr, 0, spec_x.p, 0, c.H, c.H, 1, c.st));
        mark(1);
        HIPCHK(fft.template colpass<TA>(COL_BCAST, spec_x.p, 0, c.H, spec_k.p, kstride, c.H, psf_spec.p,
                                        fft.spec_elems(s.FH), c.H, c.H, n, c.st, l));
        mark(2);
        HIPCHK(fft.row_c2r(spec_k.p, kstride, c.H, blur.p, nout, c.W, c.W, sc, c.H, n, c.st, l));
        mark(3);
        HIPCHK((gather_freeze<TF, TA>(d_ex.p, 0, n, max_nin, blur.p, nout, sci.p, w.p, frozen.p, wr.p, diag_cur.p,
                                      idiag_cur.p, c.st)));
        mark(4);
        int m = 5;
        if (!(c.flags & UPDC_FLAG_FUSED_ADJOINT)) {
            // the default: S_k^T wr_k for every exposure into the (now dead) blur
            // images, then the plain row R2C of those images.  The standalone
            // gather runs at 8 waves per SIMD; fused into the row R2C it shares
            // the CU with the transform tile at 4, and the fused launch is slower
            // than the two (1.58 ms against 0.91 + 0.60 ms with 8 exposures of
            // 2048² on a 4096² grid)
            HIPCHK(csc_image<TF>(d_ex.p, n, wr.p, nout, blur.p, any_c16, any_wide, c.st));
            mark(m++);
            HIPCHK(fft.row_r2c(ROW_PLAIN, blur.p, nout, c.W, c.W, nullptr, nullptr, 0, spec_k.p, kstride, c.H, c.H, n,
                               c.st));
        } else {
            HIPCHK(fft.row_r2c(ROW_GATHER, nullptr, 0, 0, 0, d_ex.p, wr.p, c.W, spec_k.p, kstride, c.H, c.H, n, c.st));
        }
        mark(m++);
        if (!adjoint_sum(c, n, [&]() { mark(m++); }, true, l)) {
            mark(m++);
            HIPCHK((update<TF, TA>(x.p, C.p, D.p, nout, (TA)max_sens, idiag_cur.p, partials.p, negs.p, ticket.p,
                                   c.st)));
        }
        mark(m++);
        HIPCHK(commit_diag(diag_cur.p, n, idiag_cur.p, diag_hist.p, idiag_hist.p, hist_cap, ctr.p, c.st));
    }
    void invalidate_graph() override {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }
    int spectrum_layout(const Ctx &c) const override { return iter_layout(c); }
    void ensure_plain_buffers(Ctx &c) override {
        const size_t fstride = fft.spec_elems(fft.s.FH);
        if (!spec_tmp.p) spec_tmp.alloc(fstride);
        if (!spec_acc.p) spec_acc.alloc(fstride);
        (void)c;
    }

    void iterate(Ctx &c, int niter, updc_freeze_row *rows, updc_iter_info *info) override {
        const int n = N(c);
        if (niter > hist_cap) {
            diag_hist.alloc((size_t)niter * n);
            idiag_hist.alloc(niter);
            hist_cap = niter;
            if (graph) { (void)hipGraphExecDestroy(graph); graph = nullptr; }
        }
        HIPCHK(init_diag(diag_cur.p, n, idiag_cur.p, 1, c.st));
        HIPCHK(init_diag(diag_hist.p, niter * n, idiag_hist.p, niter, c.st));
        HIPCHK(fill<unsigned int>(ctr.p, 0u, 1, c.st));
        const bool use_graph = (c.flags & UPDC_FLAG_GRAPH) != 0;
        std::vector<hipEvent_t> ev;
        const int nev = 12;
        hipEvent_t it0, it1;
        HIPCHK(hipEventCreate(&it0));
        HIPCHK(hipEventCreate(&it1));
        std::vector<double> acc(8, 0.0);
        if (!use_graph) {
            ev.resize((size_t)nev * niter);
            for (auto &e : ev) HIPCHK(hipEventCreate(&e));
        }
        try {
            if (use_graph && !graph) {
                hipGraph_t g;
                HIPCHK(hipStreamBeginCapture(c.st, hipStreamCaptureModeThreadLocal));
                try {
                    launch_iteration(c, nullptr);
                } catch (...) {
                    hipGraph_t dummy;
                    (void)hipStreamEndCapture(c.st, &dummy);
                    throw;
                }
                HIPCHK(hipStreamEndCapture(c.st, &g));
                HIPCHK(hipGraphInstantiate(&graph, g, nullptr, nullptr, 0));
                HIPCHK(hipGraphDestroy(g));
            }
            HIPCHK(hipEventRecord(it0, c.st));
            for (int it = 0; it < niter; ++it) {
                if (use_graph) HIPCHK(hipGraphLaunch(graph, c.st));
                else launch_iteration(c, ev.data() + (size_t)it * nev);
            }
            HIPCHK(hipEventRecord(it1, c.st));
            HIPCHK(hipStreamSynchronize(c.st));
        } catch (...) {
            for (auto &e : ev) (void)hipEventDestroy(e);
            (void)hipEventDestroy(it0);
            (void)hipEventDestroy(it1);
            throw;
        }
        float total_ms = 0;
        HIPCHK(hipEventElapsedTime(&total_ms, it0, it1));
        const bool plain = (c.flags & UPDC_FLAG_PER_EXPOSURE) != 0;
        const bool split = (c.flags & UPDC_FLAG_FUSED_ADJOINT) == 0;
        if (plain)
            stage_names = {"plain_xhat", "plain_exposures", "plain_adjoint_inverse", "update"};
        else {
            stage_names = {"row_r2c_x", "col_fwd_bcast", "row_c2r_blur", "gather_freeze"};
            if (split) stage_names.insert(stage_names.end(), {"adjoint_image", "row_r2c_adj"});
            else stage_names.push_back("row_r2c_gather");
            stage_names.push_back("col_adj_accum");
            if (kWide && !acc_fused) stage_names.push_back("col_inv_corr");
            if (fused_update) stage_names.push_back("row_c2r_update");
            else stage_names.insert(stage_names.end(), {"row_c2r_corr", "update"});
        }
        const int nst = (int)stage_names.size();
        stage_ms.assign(nst + 1, 0.0);
        stage_n.assign(nst + 1, 0);
        // per iteration, the launches grouped as the reference's timed stages
        // (updc_core.c:1411-1478): seconds, NaN where graph replay hides them
        c.iter_times.assign(niter, updc_iter_times{});
        const double nan = std::nan("");
        for (int it = 0; it < niter; ++it) {
            updc_iter_times &r = c.iter_times[it];
            r.iteration = it + 1;
            r.eager = use_graph ? 0 : 1;
            r.estimate_fft = r.fused_exposure_wall = r.forward_ifft_critical = r.fused_detector_critical =
                r.scatter_fft_critical = r.adjoint_reduce_ifft = r.update = nan;
            r.total = total_ms * 1e-3 / std::max(1, niter);
        }
        if (!use_graph) {
            for (int it = 0; it < niter; ++it) {
                double sec[16] = {0};
                for (int s = 0; s < nst; ++s) {
                    float ms = 0;
                    HIPCHK(hipEventElapsedTime(&ms, ev[(size_t)it * nev + s], ev[(size_t)it * nev + s + 1]));
                    stage_ms[s] += ms;
                    stage_n[s] += 1;
                    sec[s] = ms * 1e-3;
                }
                updc_iter_times &r = c.iter_times[it];
                auto sum = [&](std::initializer_list<const char *> names) {
                    double t = 0;
                    for (int s = 0; s < nst; ++s)
                        for (const char *nm : names)
                            if (stage_names[s] == nm) t += sec[s];
                    return t;
                };
                r.total = 0;
                for (int s = 0; s < nst; ++s) r.total += sec[s];
                if (plain) {
                    // one launch sequence per exposure: blur, detector and
                    // adjoint transform interleave, only their sum is timed
                    r.estimate_fft = sum({"plain_xhat"});
                    r.fused_exposure_wall = sum({"plain_exposures"});
                    r.adjoint_reduce_ifft = sum({"plain_adjoint_inverse"});
                    r.update = sum({"update"});
                } else {
                    // the X^ column pass runs inside the broadcast column pass
                    // (forward); the gathers are the detector loop; the adjoint
                    // row transforms and the summed column pass the scatter
                    // transforms; the correction's inverse may carry the update
                    r.estimate_fft = sum({"row_r2c_x"});
                    r.forward_ifft_critical = sum({"col_fwd_bcast", "row_c2r_blur"});
                    r.fused_detector_critical = sum({"gather_freeze", "adjoint_image"});
                    r.scatter_fft_critical = sum({"row_r2c_adj", "row_r2c_gather", "col_adj_accum"});
                    r.adjoint_reduce_ifft = sum({"col_inv_corr", "row_c2r_corr", "row_c2r_update"});
                    r.update = sum({"update"});
                    r.fused_exposure_wall = r.forward_ifft_critical + r.fused_detector_critical + r.scatter_fft_critical;
                }
            }
            for (int s = 0; s < nst; ++s) stage_ms[s] /= std::max(1, stage_n[s]);
        }
        stage_ms[nst] = total_ms / std::max(1, niter);
        stage_n[nst] = niter;
        for (auto &e : ev) (void)hipEventDestroy(e);
        (void)hipEventDestroy(it0);
        (void)hipEventDestroy(it1);

        std::vector<ExpDiag> hd((size_t)niter * n);
        std::vector<IterDiag> hi(niter);
        HIPCHK(hipMemcpy(hd.data(), diag_hist.p, hd.size() * sizeof(ExpDiag), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hi.data(), idiag_hist.p, hi.size() * sizeof(IterDiag), hipMemcpyDeviceToHost));
        // The first failing iteration f (0-based), if any.  The reference stops
        // there: a detector-stage failure before writing f's freeze rows
        // (updc_core.c:1160-1169, rows at :1425-1431), an update failure after
        // them (:1441-1453).  Later iterations did not run on the device either
        // (the halt flag).
        int fail_it = niter;
        for (int it = 0; it < niter; ++it)
            if (hi[it].err) { fail_it = it; break; }
        const bool det_fail = fail_it < niter && (hi[fail_it].err & (UPDC_E_MU | UPDC_E_RATIO));
        c.iter_done = fail_it;
        c.iter_rows = fail_it < niter && !det_fail ? fail_it + 1 : fail_it;
        if (rows) std::memset((void *)rows, 0, (size_t)niter * n * sizeof(updc_freeze_row));
        if (info) std::memset((void *)info, 0, (size_t)niter * sizeof(updc_iter_info));
        for (int it = 0; it < c.iter_rows; ++it) {
            uint64_t nn = 0, nt = 0;
            double mn = INFINITY;
            for (int k = 0; k < n; ++k) {
                const ExpDiag &d = hd[(size_t)it * n + k];
                const double m = d.min_key == ~0ull ? INFINITY : dkey_inv(d.min_key);
                if (rows) {
                    updc_freeze_row &r = rows[(size_t)it * n + k];
                    r.iteration = it + 1;
                    r.exposure = k + 1;
                    r.n_new_frozen = d.n_new;
                    r.n_total_frozen = d.n_total;
                    r.min_raw_mu = m;
                }
                nn += d.n_new;
                nt += d.n_total;
                mn = std::min(mn, m);
            }
            if (info && it < c.iter_done) {
                info[it].iteration = it + 1;
                info[it].signed_sum = hi[it].signed_sum;
                info[it].n_negative = hi[it].n_negative;
                info[it].n_new_frozen = nn;
                info[it].n_total_frozen = nt;
                info[it].min_raw_mu = mn;
                info[it].ms_total = total_ms / std::max(1, niter);
            }
        }
        if (fail_it < niter) {
            // the kind and place the reference's serial loops meet first
            const unsigned long long key = hi[fail_it].err_key;
            const unsigned rank = (unsigned)(key & 3u);
            char msg[256];
            if (det_fail) {
                std::snprintf(msg, sizeof msg, "%s (iteration %d, exposure %u, pixel %llu)",
                              rank == 0 ? "non-finite forward mean" : "non-finite y/raw_mu ratio", fail_it + 1,
                              (unsigned)(key >> 34) + 1u, (unsigned long long)((key >> 2) & 0xffffffffull));
            } else {
                std::snprintf(msg, sizeof msg, "%s (iteration %d, target %llu)",
                              rank == 0   ? "non-finite multiplicative correction"
                              : rank == 1 ? "substantially negative multiplicative correction"
                                          : "non-finite signed UPDC estimate",
                              fail_it + 1, (unsigned long long)(key >> 2));
            }
            throw Error(UPDC_ERR_NUMERIC, msg);
        }
    }

    // ---- transfers
    template <class T> void to_host_f64(const T *d, size_t cnt, double *out, Ctx &c) {
        if (sizeof(T) == 8) {
            HIPCHK(hipMemcpy(out, d, cnt * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<T> tmp(cnt);
            HIPCHK(hipMemcpy(tmp.data(), d, cnt * sizeof(T), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < cnt; ++i) out[i] = (double)tmp[i];
        }
    }
    template <class T> void from_host_f64(const double *in, size_t cnt, T *d, Ctx &c) {
        if (sizeof(T) == 8) {
            HIPCHK(hipMemcpy(d, in, cnt * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<T> tmp(cnt);
            for (size_t i = 0; i < cnt; ++i) tmp[i] = (T)in[i];
            HIPCHK(hipMemcpy(d, tmp.data(), cnt * sizeof(T), hipMemcpyHostToDevice));
        }
    }
    void get_estimate(Ctx &c, double *d, float *f) override {
        if (d) to_host_f64(x.p, c.nout(), d, c);
        if (f) {
            if (sizeof(TF) == 4) HIPCHK(hipMemcpy(f, x.p, c.nout() * 4, hipMemcpyDeviceToHost));
            else {
                std::vector<double> tmp(c.nout());
                to_host_f64(x.p, c.nout(), tmp.data(), c);
                for (size_t i = 0; i < tmp.size(); ++i) f[i] = (float)tmp[i];
            }
        }
    }
    void set_estimate(Ctx &c, const double *d) override { from_host_f64(d, c.nout(), x.p, c); }
    void get_wht(Ctx &c, float *out) override {
        std::vector<double> tmp(c.nout());
        to_host_f64(cov.p, c.nout(), tmp.data(), c);
        // coverage * weight_scale in f32 (updc_core.c:1482)
        for (size_t i = 0; i < tmp.size(); ++i) out[i] = (float)tmp[i] * c.weight_scale;
    }
    void get_denominator(Ctx &c, double *out) override { to_host_f64(D.p, c.nout(), out, c); }
    void get_initial(Ctx &c, double *est, double *covo) override {
        if (est) to_host_f64(x_init.p, c.nout(), est, c);
        if (covo) to_host_f64(cov.p, c.nout(), covo, c);
    }
    // column-major (pitch rp(rows)) device spectrum -> row-major rows x NC host array
    void spectrum_to_host(Ctx &c, const Cx<TF> *d, int rows, double *out) {
        const FftShape &s = fft.s;
        const int P = rp(rows);
        std::vector<double> cm((size_t)s.NC * P * 2);
        to_host_f64((const TF *)d, cm.size(), cm.data(), c);
        for (int v = 0; v < s.NC; ++v)
            for (int r = 0; r < rows; ++r) {
                out[((size_t)r * s.NC + v) * 2] = cm[((size_t)v * P + r) * 2];
                out[((size_t)r * s.NC + v) * 2 + 1] = cm[((size_t)v * P + r) * 2 + 1];
            }
    }
    void get_psf_spectrum(Ctx &c, int k, double *out) override {
        spectrum_to_host(c, psf_spec.p + (size_t)k * fft.spec_elems(fft.s.FH), fft.s.FH, out);
    }
    void stage_forward(Ctx &c, const double *xin, double *bl) override {
        const int n = N(c);
        const FftShape &s = fft.s;
        const size_t nout = c.nout();
        DevBuf<TF> xs;
        xs.alloc(nout);
        from_host_f64(xin, nout, xs.p, c);
        const size_t kstride = fft.spec_elems(c.H);
        const int l = iter_layout(c);
        HIPCHK(fft.row_r2c(ROW_PLAIN, xs.p, 0, c.W, c.W, nullptr, nullptr, 0, spec_x.p, 0, c.H, c.H, 1, c.st));
        HIPCHK(fft.template colpass<TA>(COL_BCAST, spec_x.p, 0, c.H, spec_k.p, kstride, c.H, psf_spec.p,
                                        fft.spec_elems(s.FH), c.H, c.H, n, c.st, l));
        HIPCHK(fft.row_c2r(spec_k.p, kstride, c.H, blur.p, nout, c.W, c.W, scale(c), c.H, n, c.st, l));
        HIPCHK(hipStreamSynchronize(c.st));
        to_host_f64(blur.p, (size_t)n * nout, bl, c);
    }
    void stage_model(Ctx &c, const double *xin, double *pred) override {
        const int n = N(c);
        const FftShape &s = fft.s;
        const size_t nout = c.nout();
        DevBuf<TF> xs;
        xs.alloc(nout);
        from_host_f64(xin, nout, xs.p, c);
        const size_t kstride = fft.spec_elems(c.H);
        const int l = iter_layout(c);
        HIPCHK(fft.row_r2c(ROW_PLAIN, xs.p, 0, c.W, c.W, nullptr, nullptr, 0, spec_x.p, 0, c.H, c.H, 1, c.st));
        HIPCHK(fft.template colpass<TA>(COL_BCAST, spec_x.p, 0, c.H, spec_k.p, kstride, c.H, psf_spec.p,
                                        fft.spec_elems(s.FH), c.H, c.H, n, c.st, l));
        HIPCHK(fft.row_c2r(spec_k.p, kstride, c.H, blur.p, nout, c.W, c.W, scale(c), c.H, n, c.st, l));
        HIPCHK(predict<TF>(d_ex.p, n, max_nin, blur.p, nout, wr.p, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        to_host_f64(wr.p, c.npix(), pred, c);
    }
    void stage_detector(Ctx &c, int k, const double *bl, double *wro, double *scat, double *stats) override {
        const size_t nout = c.nout();
        HostExp &e = *c.ex[k];
        from_host_f64(bl, nout, blur.p + (size_t)k * nout, c);
        HIPCHK(init_diag(diag_cur.p, N(c), idiag_cur.p, 1, c.st));
        HIPCHK((gather_freeze<TF, TA>(d_ex.p, k, k + 1, (uint32_t)e.nin, blur.p, nout, sci.p, w.p, frozen.p, wr.p,
                                      diag_cur.p, idiag_cur.p, c.st)));
        DevBuf<TF> out;
        out.alloc(nout);
        HIPCHK(csc_apply<TF>(d_ex.p + k, wr.p, nout, out.p, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        if (wro) to_host_f64(wr.p + e.pix, e.nin, wro, c);
        if (scat) to_host_f64(out.p, nout, scat, c);
        ExpDiag d;
        IterDiag id;
        HIPCHK(hipMemcpy(&d, diag_cur.p + k, sizeof d, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&id, idiag_cur.p, sizeof id, hipMemcpyDeviceToHost));
        if (stats) {
            stats[0] = (double)d.n_new;
            stats[1] = (double)d.n_total;
            stats[2] = d.min_key == ~0ull ? INFINITY : dkey_inv(d.min_key);
        }
        if (id.err) throw Error(UPDC_ERR_NUMERIC, (id.err & UPDC_E_MU) ? "non-finite forward mean" : "non-finite y/raw_mu ratio");
    }
    void stage_adjoint(Ctx &c, const double *v, double *out) override {
        const int n = N(c);
        const size_t kstride = fft.spec_elems(c.H);
        from_host_f64(v, c.npix(), wr.p, c);
        HIPCHK(fft.row_r2c(ROW_GATHER, nullptr, 0, 0, 0, d_ex.p, wr.p, c.W, spec_k.p, kstride, c.H, c.H, n, c.st));
        adjoint_sum(c, n, []() {}, false, iter_layout(c));
        HIPCHK(hipStreamSynchronize(c.st));
        to_host_f64(C.p, c.nout(), out, c);
    }
    void stage_scatter(Ctx &c, int k, const double *v, double *out) override {
        HostExp &e = *c.ex[k];
        from_host_f64(v, e.nin, wr.p + e.pix, c);
        DevBuf<TF> o;
        o.alloc(c.nout());
        HIPCHK(csc_apply<TF>(d_ex.p + k, wr.p, c.nout(), o.p, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        to_host_f64(o.p, c.nout(), out, c);
    }
    // The iteration's own adjoint-image launch (csc_image over every exposure,
    // as in launch_iteration), written into the blur images it uses there.
    void stage_adjoint_images(Ctx &c, const double *v, double *out) override {
        const size_t nout = c.nout();
        from_host_f64(v, c.npix(), wr.p, c);
        HIPCHK(csc_image<TF>(d_ex.p, N(c), wr.p, nout, blur.p, any_c16, any_wide, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        to_host_f64(blur.p, (size_t)N(c) * nout, out, c);
    }
    void fft_r2c(Ctx &c, const double *img, double *spec) override {
        const FftShape &s = fft.s;
        DevBuf<TF> im;
        im.alloc((size_t)s.FW * s.FH);
        from_host_f64(img, (size_t)s.FW * s.FH, im.p, c);
        DevBuf<Cx<TF>> sp;
        sp.alloc(fft.spec_elems(s.FH));
        HIPCHK(fft.row_r2c(ROW_PLAIN, im.p, 0, s.FW, s.FW, nullptr, nullptr, 0, sp.p, 0, s.FH, s.FH, 1, c.st));
        HIPCHK(fft.template colpass<TA>(COL_FWD, sp.p, 0, s.FH, sp.p, 0, s.FH, nullptr, 0, s.FH, s.FH, 1, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        spectrum_to_host(c, sp.p, s.FH, spec);
    }
    void fft_c2r(Ctx &c, const double *spec, double *img) override {
        const FftShape &s = fft.s;
        const int P = rp(s.FH);
        std::vector<double> cm((size_t)s.NC * P * 2, 0.0);
        for (int r = 0; r < s.FH; ++r)
            for (int v = 0; v < s.NC; ++v) {
                cm[((size_t)v * P + r) * 2] = spec[((size_t)r * s.NC + v) * 2];
                cm[((size_t)v * P + r) * 2 + 1] = spec[((size_t)r * s.NC + v) * 2 + 1];
            }
        DevBuf<Cx<TF>> sp;
        sp.alloc(fft.spec_elems(s.FH));
        from_host_f64(cm.data(), cm.size(), (TF *)sp.p, c);
        HIPCHK(fft.template colpass<TA>(COL_INV, sp.p, 0, s.FH, sp.p, 0, s.FH, nullptr, 0, s.FH, s.FH, 1, c.st));
        DevBuf<TF> im;
        im.alloc((size_t)s.FW * s.FH);
        HIPCHK(fft.row_c2r(sp.p, 0, s.FH, im.p, 0, s.FW, s.FW, (TF)1, s.FH, 1, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        to_host_f64(im.p, (size_t)s.FW * s.FH, img, c);
    }
    void get_frozen(Ctx &c, int k, uint8_t *out) override {
        HIPCHK(hipMemcpy(out, frozen.p + c.ex[k]->pix, c.ex[k]->nin, hipMemcpyDeviceToHost));
    }
    void reset_frozen(Ctx &c) override {
        HIPCHK(fill<uint8_t>(frozen.p, (uint8_t)0, c.npix(), c.st));
        HIPCHK(hipStreamSynchronize(c.st));
    }

    // ---- device-pointer operators.  Spectra are column-major, NC columns of
    // pitch rp(FH) (element (r, v) at v * rp(FH) + r), as everywhere here.
    DevBuf<Cx<TF>> op_ws;  // column-inverse workspace of op_c2r
    void op_r2c(Ctx &c, const void *img, int rows, int cols, int pitch, void *spec, hipStream_t st) override {
        (void)c;
        const FftShape &s = fft.s;
        Cx<TF> *sp = (Cx<TF> *)spec;
        HIPCHK(fft.row_r2c(ROW_PLAIN, (const TF *)img, 0, pitch, cols, nullptr, nullptr, 0, sp, 0, s.FH, rows, 1, st));
        HIPCHK(fft.template colpass<TA>(COL_FWD, sp, 0, s.FH, sp, 0, s.FH, nullptr, 0, rows, s.FH, 1, st));
    }
    size_t op_c2r_workspace_bytes() const override { return fft.spec_elems(fft.s.FH) * sizeof(Cx<TF>); }
    bool op_c2r_ws_ready(int rows) const override { return fft.long_covers(rows); }
    void op_c2r_ws(Ctx &c, const void *spec, void *img, int rows, int cols, int pitch, double sc, void *ws,
                   hipStream_t st) override {
        (void)c;
        const FftShape &s = fft.s;
        Cx<TF> *w = (Cx<TF> *)ws;
        // the column inverse writes rows [0, rows) of every column, all the row pass reads
        HIPCHK(fft.template colpass<TA>(COL_INV, (const Cx<TF> *)spec, 0, s.FH, w, 0, s.FH, nullptr, 0, s.FH, rows,
                                        1, st));
        HIPCHK(fft.row_c2r(w, 0, s.FH, (TF *)img, 0, pitch, cols, (TF)sc, rows, 1, st));
    }
    void op_c2r(Ctx &c, const void *spec, void *img, int rows, int cols, int pitch, double sc,
                hipStream_t st) override {
        if (!op_ws.p) op_ws.alloc(fft.spec_elems(fft.s.FH));
        op_c2r_ws(c, spec, img, rows, cols, pitch, sc, op_ws.p, st);
    }
    void op_gather(Ctx &c, int k, const void *img, void *out, hipStream_t st) override {
        // the kernels index per-pixel arrays at pix_k + p: shift the caller's
        // nin_k buffer by -pix_k (only [pix_k, pix_k + nin_k) is touched)
        const HostExp &e = *c.ex[k];
        HIPCHK(predict<TF>(d_ex.p + k, 1, (uint32_t)e.nin, (const TF *)img, 0, (TF *)out - e.pix, st));
    }
    void op_scatter(Ctx &c, int k, const void *v, void *img, hipStream_t st) override {
        const HostExp &e = *c.ex[k];
        HIPCHK(csc_apply<TF>(d_ex.p + k, (const TF *)v - e.pix, c.nout(), (TF *)img, st));
    }
};

// ------------------------------------------------------------------ exposures
static void check_dims(Ctx &c, int iw, int ih, int pw, int ph) {
    if (iw < 1 || ih < 1) throw Error(UPDC_ERR_ARG, "invalid input dimensions");
    if (pw < 1 || ph < 1) throw Error(UPDC_ERR_ARG, "invalid PSF dimensions in manifest");
    if ((size_t)iw * ih > 0xffffffffull) throw Error(UPDC_ERR_ARG, "detector exceeds 32-bit pixel index capacity");
    if (c.prepared) throw Error(UPDC_ERR_STATE, "exposures must be added before updc_prepare");
}

static HostExp &new_exposure(Ctx &c, const float *sci, const float *wht, int iw, int ih, const float *psf, int pw,
                             int ph, int pt, double bg, double ped) {
    if (!std::isfinite(bg) || !std::isfinite(ped) || ped < 0.0) throw Error(UPDC_ERR_ARG, "invalid background/pedestal in manifest");
    auto e = std::make_unique<HostExp>();
    e->iw = iw;
    e->ih = ih;
    e->nin = (size_t)iw * ih;
    e->pix = c.npix();
    e->sci.assign(sci, sci + e->nin);
    e->wht.assign(wht, wht + e->nin);
    e->psf.assign(psf, psf + (size_t)pw * ph);
    e->pw = pw;
    e->ph = ph;
    e->psf_t = pt ? 1 : 0;
    e->background = bg;
    e->pedestal = ped;
    e->background_f = (float)bg;
    e->pedestal_f = (float)ped;
    e->mean_offset_f = e->background_f + e->pedestal_f;
    if (!std::isfinite(e->mean_offset_f) || e->pedestal_f < 0.0f)
        throw Error(UPDC_ERR_ARG, "invalid background/pedestal in manifest");
    c.ex.push_back(std::move(e));
    return *c.ex.back();
}

static void validate_linkmap(Ctx &c, const uint8_t *img, size_t bytes, size_t nin) {
    // updc_core.c:485-496 and :652
    if (bytes < 64) throw Error(UPDC_ERR_FORMAT, "invalid or obsolete LinkMap cache");
    uint32_t h[16];
    std::memcpy(h, img, 64);
    uint64_t hn, hl;
    std::memcpy(&hn, img + 16, 8);
    std::memcpy(&hl, img + 24, 8);
    if (h[0] != UPDC_LM_MAGIC || h[1] != UPDC_LM_VERSION || h[8] != (uint32_t)c.W || h[9] != (uint32_t)c.H ||
        h[3] != 64 || hn != nin || hl > 0xffffffffull)
        throw Error(UPDC_ERR_FORMAT, "invalid or obsolete LinkMap cache");
    const size_t expect = linkmap_bytes(hn, hl, h[2]);
    if (!expect || expect != bytes) throw Error(UPDC_ERR_FORMAT, "invalid or obsolete LinkMap cache");
    uint32_t last;
    std::memcpy(&last, img + 64 + nin * 4, 4);
    if (last != (uint32_t)hl) throw Error(UPDC_ERR_FORMAT, "LinkMap offset/link-count mismatch");
}

static void add_linkmap_image(Ctx &c, HostExp &e, const void *img, size_t bytes) {
    validate_linkmap(c, (const uint8_t *)img, bytes, e.nin);
    std::memcpy(&e.encoding, (const uint8_t *)img + 8, 4);
    std::memcpy(&e.nlinks, (const uint8_t *)img + 24, 8);
    e.lm_bytes = bytes;
    e.lm.alloc(lm_alloc_bytes(bytes));
    HIPCHK(hipMemcpy(e.lm.p, img, bytes, hipMemcpyHostToDevice));
    if (e.encoding != UPDC_LM_DELTA16 && e.encoding != UPDC_LM_ABS32) throw Error(UPDC_ERR_FORMAT, "invalid LinkMap encoding");
    // the offsets are checked on the device before any kernel follows them
    DevBuf<unsigned int> flags;
    flags.alloc(1);
    HIPCHK(fill<unsigned int>(flags.p, 0u, 1, c.st));
    const ExpDev v = e.view();
    HIPCHK(lm_validate(v.off, v.base, v.d16, v.t32, v.frac, e.encoding, (uint32_t)e.nin, e.nlinks, c.nout(), flags.p,
                       c.st));
    unsigned int fl = 0;
    HIPCHK(hipMemcpyAsync(&fl, flags.p, 4, hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    if (fl & 1u) throw Error(UPDC_ERR_FORMAT, "LinkMap offsets are not monotone / exceed the link count");
    if (fl & 2u) throw Error(UPDC_ERR_FORMAT, "LinkMap target outside the target grid");
    if (fl & 4u) throw Error(UPDC_ERR_FORMAT, "LinkMap fraction is not finite");
}

// GPU LinkMap build from a vertex lattice (updc_core.c:530-609), two passes.
static void build_linkmap(Ctx &c, HostExp &e, const double *vert) {
    const size_t nv = 2 * (size_t)(e.iw + 1) * (e.ih + 1);
    DevBuf<double> dv;
    dv.alloc(nv);
    HIPCHK(hipMemcpyAsync(dv.p, vert, nv * 8, hipMemcpyHostToDevice, c.st));
    DevBuf<uint32_t> cnt, first;
    cnt.alloc(e.nin);
    first.alloc(e.nin);
    DevBuf<unsigned int> flags;
    flags.alloc(1);
    HIPCHK(fill<unsigned int>(flags.p, 0u, 1, c.st));
    HIPCHK(lm_count(dv.p, e.iw, e.ih, c.W, c.H, cnt.p, first.p, flags.p, c.st));
    DevBuf<uint32_t> off;
    off.alloc(e.nin + 1);
    DevBuf<uint8_t> scratch;
    scratch.alloc(scan_scratch_bytes(e.nin));
    DevBuf<unsigned long long> tot;
    tot.alloc(1);
    HIPCHK(scan_u32(cnt.p, off.p, e.nin, scratch.p, tot.p, c.st));
    unsigned int fl = 0;
    unsigned long long total = 0;
    HIPCHK(hipMemcpyAsync(&fl, flags.p, 4, hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipMemcpyAsync(&total, tot.p, 8, hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    if (total >= 0xffffffffull) throw Error(UPDC_ERR_ARG, "LinkMap exceeds uint32 offset capacity");
    e.encoding = (fl & 1u) ? UPDC_LM_ABS32 : UPDC_LM_DELTA16;
    e.nlinks = total;
    e.lm_bytes = linkmap_bytes(e.nin, total, e.encoding);
    e.lm.alloc(lm_alloc_bytes(e.lm_bytes));
    HIPCHK(hipMemsetAsync(e.lm.p, 0, e.lm_bytes, c.st));
    uint32_t h[16] = {0};
    h[0] = UPDC_LM_MAGIC; h[1] = UPDC_LM_VERSION; h[2] = e.encoding; h[3] = 64;
    const uint64_t nin64 = e.nin;
    std::memcpy(&h[4], &nin64, 8);
    std::memcpy(&h[6], &total, 8);
    h[8] = (uint32_t)c.W; h[9] = (uint32_t)c.H;
    HIPCHK(hipMemcpyAsync(e.lm.p, h, 64, hipMemcpyHostToDevice, c.st));
    HIPCHK(hipMemcpyAsync(e.lm.p + 64, off.p, (e.nin + 1) * 4, hipMemcpyDeviceToDevice, c.st));
    ExpDev v = e.view();
    if (e.encoding == UPDC_LM_DELTA16)
        HIPCHK(hipMemcpyAsync((void *)v.base, first.p, e.nin * 4, hipMemcpyDeviceToDevice, c.st));
    HIPCHK(fill<unsigned int>(flags.p, 0u, 1, c.st));
    HIPCHK(lm_fill(dv.p, e.iw, e.ih, c.W, c.H, v.off, v.base, e.encoding, (uint16_t *)v.d16, (uint32_t *)v.t32,
                   (float *)v.frac, flags.p, c.st));
    HIPCHK(hipMemcpyAsync(&fl, flags.p, 4, hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    if (fl & 2u) throw Error(UPDC_ERR_ARG, "LinkMap second-pass count overflow");
    if (fl & 4u) throw Error(UPDC_ERR_ARG, "DELTA16 LinkMap invariant failed");
    if (fl & 8u) throw Error(UPDC_ERR_ARG, "LinkMap second-pass count mismatch");
}

}  // namespace updc

using namespace updc;

struct updc_ctx : Ctx {};

template <class F> static int guarded(F &&f) {
    try {
        f();
        g_last_error.clear();
        return UPDC_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "out of memory";
        return UPDC_ERR_NOMEM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return UPDC_ERR_ARG;
    }
}
static void need(bool ok, int code, const char *msg) {
    if (!ok) throw Error(code, msg);
}

// ------------------------------------------------------------------ FFT shape autotuner
// Candidate shapes, after the reference's selection (updc_core.c:170-179,
// 242-291): the first five admissible sizes per axis from the minimum (here
// widths even and 7-smooth, heights 7-smooth, the row pass needing an even
// width), the diagonal pairs, then (w2,h0), (w0,h2), (w4,h0), (w0,h4).
static int shape_candidates(int min_w, int min_h, int (*out)[2], int cap) {
    int wc[5], hc[5];
    for (int i = 0, v = min_w; i < 5; ++i) v = (wc[i] = next_even_smooth(v)) + 1;
    for (int i = 0, v = min_h; i < 5; ++i) v = (hc[i] = next_smooth(v)) + 1;
    int n = 0;
    auto add = [&](int w, int h) {
        for (int i = 0; i < n; ++i)
            if (out[i][0] == w && out[i][1] == h) return;
        if (n < cap) { out[n][0] = w; out[n][1] = h; ++n; }
    };
    for (int i = 0; i < 5; ++i) add(wc[i], hc[i]);
    add(wc[2], hc[0]);
    add(wc[0], hc[2]);
    add(wc[4], hc[0]);
    add(wc[0], hc[4]);
    return n;
}

// The transforms of one batched iteration at shape FW x FH, timed (best of
// three after a warm-up, in ms; < 0 when the shape does not fit the kernels):
// X^ (row R2C of one image, then the broadcast column pass with the PSF
// product and inverse for all n exposures), the n blurred rows back, the n
// adjoint images forward (plain row R2C standing in for the gather-fed one),
// the summed adjoint column pass and the final row C2R.  Values are zero; the
// timing depends only on the shape.  This is the work the shape choice
// changes, where the reference times one R2C + C2R round trip
// (updc_core.c:225-240); the two rank shapes differently once the column
// passes carry the exposure loop and the PSF products.
template <class T> static double time_iteration_transforms(Ctx &c, int FW, int FH, bool edge, bool *edge_used) {
    Fft2<T> f;
    f.allow_edge = edge;
    if (!f.init(FW, FH).empty()) {
        f.release();
        return -1.0;
    }
    *edge_used = f.col_re != 0;
    const int n = (int)c.ex.size();
    const size_t nimg = (size_t)c.W * c.H;
    double best = -1.0;
    hipEvent_t a = nullptr, b = nullptr;
    try {
        DevBuf<T> img;
        DevBuf<Cx<T>> sk, sx, ps;
        const size_t kh = f.spec_elems(c.H), kf = f.spec_elems(FH);
        img.alloc((size_t)n * nimg);
        sk.alloc((size_t)n * kh);
        sx.alloc(f.spec_elems(std::max(c.H, FH)));
        ps.alloc((size_t)n * kf);
        HIPCHK(hipMemsetAsync(img.p, 0, (size_t)n * nimg * sizeof(T), c.st));
        HIPCHK(hipMemsetAsync(ps.p, 0, (size_t)n * kf * sizeof(Cx<T>), c.st));
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        const int l = f.pair_ok() ? SL_PAIR : SL_COL;  // the layout the iteration takes at this shape
        auto trip = [&]() {
            HIPCHK(f.row_r2c(ROW_PLAIN, img.p, 0, c.W, c.W, nullptr, nullptr, 0, sx.p, 0, c.H, c.H, 1, c.st));
            HIPCHK(f.template colpass<T>(COL_BCAST, sx.p, 0, c.H, sk.p, kh, c.H, ps.p, kf, c.H, c.H, n, c.st, l));
            HIPCHK(f.row_c2r(sk.p, kh, c.H, img.p, nimg, c.W, c.W, (T)1, c.H, n, c.st, l));
            HIPCHK(f.row_r2c(ROW_PLAIN, img.p, nimg, c.W, c.W, nullptr, nullptr, 0, sk.p, kh, c.H, c.H, n, c.st));
            HIPCHK(f.template colpass<T>(COL_ADJ, sk.p, kh, c.H, sx.p, 0, c.H, ps.p, kf, c.H, c.H, n, c.st, l));
            HIPCHK(f.row_c2r(sx.p, 0, c.H, img.p, 0, c.W, c.W, (T)1, c.H, 1, c.st, l));
        };
        trip();
        for (int r = 0; r < 3; ++r) {
            HIPCHK(hipEventRecord(a, c.st));
            trip();
            HIPCHK(hipEventRecord(b, c.st));
            HIPCHK(hipEventSynchronize(b));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, a, b));
            if (best < 0 || ms < best) best = ms;
        }
    } catch (...) {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        f.release();
        throw;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    f.release();
    return best;
}

static std::string device_signature(int dev) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return "unknown-device";
    return std::string(p.name) + "|" + p.gcnArchName + "|cu" + std::to_string(p.multiProcessorCount) + "|updc_hip 1";
}

static std::string tune_cache_path(const Ctx &c, int min_w, int min_h) {
    const char *pn = c.prec == UPDC_PREC_F64 ? "f64" : c.prec == UPDC_PREC_MIXED ? "mixed" : "f32";
    return c.fft_cache_dir + "/fft_autotune_gpu_v2_" + pn + "_n" + std::to_string(c.ex.size()) + "_" +
           std::to_string(c.W) + "x" + std::to_string(c.H) + "_" + std::to_string(min_w) + "x" +
           std::to_string(min_h) + ".txt";
}

static bool load_tune_cache(Ctx &c, int min_w, int min_h) {
    if (c.fft_cache_dir.empty() || c.fft_retune) return false;
    FILE *f = std::fopen(tune_cache_path(c, min_w, min_h).c_str(), "r");
    if (!f) return false;
    char sig[512] = {0};
    int w = 0, h = 0, edge = 0;
    double ms = 0;
    const bool ok = std::fscanf(f, "signature=%511[^\n]\nwidth=%d\nheight=%d\ncolumn_edge=%d\ntransforms_ms=%lf",
                                sig, &w, &h, &edge, &ms) == 5;
    std::fclose(f);
    if (!ok || device_signature(c.device) != sig || w < min_w || h < min_h || (w & 1) || !seven_smooth(w) ||
        !seven_smooth(h))
        return false;
    c.FW = w;
    c.FH = h;
    c.col_edge = edge ? 1 : 0;
    c.tune.assign(1, TuneRow{w, h, c.col_edge, ms});
    c.tune_cached = true;
    return true;
}

static void save_tune_cache(const Ctx &c, int min_w, int min_h, double ms) {
    if (c.fft_cache_dir.empty()) return;
    const std::string path = tune_cache_path(c, min_w, min_h);
    const std::string tmp = path + ".tmp." + std::to_string((long)getpid());
    FILE *f = std::fopen(tmp.c_str(), "w");
    if (!f) return;
    std::fprintf(f, "signature=%s\nwidth=%d\nheight=%d\ncolumn_edge=%d\ntransforms_ms=%.6f\n",
                 device_signature(c.device).c_str(), c.FW, c.FH, c.col_edge, ms);
    if (std::fclose(f) != 0 || std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
}

static void autotune_shape(Ctx &c, int min_w, int min_h) {
    if (load_tune_cache(c, min_w, min_h)) {
        if (!c.quiet())
            std::printf("FFT autotune persistent-cache hit: %dx%d, %s column kernel (iteration transforms %.3f ms)\n", c.FW,
                        c.FH, c.col_edge ? "edge-stage" : "LDS-only", c.tune[0].ms);
        return;
    }
    int shapes[16][2];
    const int ns = shape_candidates(min_w, min_h, shapes, 16);
    c.tune.clear();
    c.tune_cached = false;
    double best = -1;
    int bw = 0, bh = 0, be = 0;
    if (!c.quiet()) std::printf("FFT autotune: %d candidate shapes, %zu exposures\n", ns, c.ex.size());
    for (int i = 0; i < ns; ++i) {
        for (int edge = 1; edge >= 0; --edge) {
            bool used = false;
            const bool f64 = c.prec == UPDC_PREC_F64;
            const double ms = f64 ? time_iteration_transforms<double>(c, shapes[i][0], shapes[i][1], edge != 0, &used)
                                  : time_iteration_transforms<float>(c, shapes[i][0], shapes[i][1], edge != 0, &used);
            if (edge && !used && ms >= 0) {
                // no edge radix for this height: the LDS-only kernel ran; record it once
                c.tune.push_back(TuneRow{shapes[i][0], shapes[i][1], 0, ms});
                if (best < 0 || ms < best) { best = ms; bw = shapes[i][0]; bh = shapes[i][1]; be = 0; }
                if (!c.quiet()) std::printf("  FFT candidate %dx%d LDS-only transforms %.3f ms\n", shapes[i][0], shapes[i][1], ms);
                break;
            }
            c.tune.push_back(TuneRow{shapes[i][0], shapes[i][1], edge, ms});
            if (!c.quiet()) {
                if (ms < 0) std::printf("  FFT candidate %dx%d unsupported by the LDS-resident passes\n", shapes[i][0], shapes[i][1]);
                else std::printf("  FFT candidate %dx%d %s transforms %.3f ms\n", shapes[i][0], shapes[i][1],
                                 edge ? "edge-stage" : "LDS-only", ms);
            }
            if (ms < 0) break;
            if (best < 0 || ms < best) { best = ms; bw = shapes[i][0]; bh = shapes[i][1]; be = edge; }
        }
    }
    if (best < 0) throw Error(UPDC_ERR_ARG, "no FFT shape candidate fits the LDS-resident passes");
    c.FW = bw;
    c.FH = bh;
    c.col_edge = be;
    if (!c.quiet())
        std::printf("FFT autotune selected %dx%d, %s column kernel (iteration transforms %.3f ms)\n", bw, bh,
                    be ? "edge-stage" : "LDS-only", best);
    save_tune_cache(c, min_w, min_h, best);
}

extern "C" {

const char *updc_last_error(void) { return g_last_error.c_str(); }

float updc_polygon_area(const double *xy, int count) { return geo::area((const geo::P *)xy, count); }
float updc_overlap_area(const double *quad_xy, int x, int y) { return geo::overlap((const geo::P *)quad_xy, x, y); }

int updc_self_test(int verbose) {
    const double unit[8] = {1000.0, 2000.0, 1000.0, 2001.0, 1001.0, 2001.0, 1001.0, 2000.0};
    const float a = updc_polygon_area(unit, 4);
    if (std::fabs(a - 1.0f) > 2.0e-6f) {
        std::fprintf(stderr, "self-test polygon area failed: %.9g\n", (double)a);
        return 1;
    }
    const double half[8] = {-0.5, -0.5, -0.5, 0.5, 0.0, 0.5, 0.0, -0.5};
    const float o = updc_overlap_area(half, 0, 0);
    if (std::fabs(o - 0.5f) > 2.0e-6f) {
        std::fprintf(stderr, "self-test overlap failed: %.9g\n", (double)o);
        return 1;
    }
    struct Hdr { uint32_t a, b, c, d; uint64_t e, f; uint32_t g, h; uint32_t r[6]; };
    struct Tmp { uint32_t t; float f; };
    if (sizeof(Hdr) != 64 || sizeof(Tmp) != 8) {
        std::fprintf(stderr, "self-test LinkMap ABI failed\n");
        return 1;
    }
    if (verbose) std::printf("UPDC C core self-test: OK\n");
    return 0;
}

static void *scratch_on(int dev);

int updc_ctx_create(updc_ctx **out, int width, int height, int fft_width, int fft_height, int precision,
                    unsigned flags, int device) {
    return guarded([&] {
        need(out != nullptr, UPDC_ERR_ARG, "null context pointer");
        *out = nullptr;
        need(width >= 1 && height >= 1, UPDC_ERR_ARG, "invalid dimensions/exposures/iterations/workers/save_every");
        need((size_t)width * height <= 0xffffffffull, UPDC_ERR_ARG, "target grid exceeds 32-bit link index capacity");
        need(precision >= UPDC_PREC_F32 && precision <= UPDC_PREC_MIXED, UPDC_ERR_ARG, "unknown precision");
        need(fft_width >= 0 && fft_height >= 0, UPDC_ERR_ARG, "invalid FFT shape");
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        need(device >= 0 && device < ndev, UPDC_ERR_HIP, "no such HIP device");
        HIPCHK(hipSetDevice(device));
        (void)scratch_on(device);  // the allocation-free entry point's workspace address on this device
        auto c = std::make_unique<updc_ctx>();
        c->W = width;
        c->H = height;
        c->FW = fft_width;
        c->FH = fft_height;
        c->prec = precision;
        c->flags = flags;
        c->device = device;
        HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
        *out = c.release();
    });
}

int updc_ctx_destroy(updc_ctx *ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        ctx->pipe.reset();
        ctx->ex.clear();
        if (ctx->st) (void)hipStreamDestroy(ctx->st);
        delete ctx;
    });
}

int updc_ctx_set_flags(updc_ctx *ctx, unsigned flags) {
    return guarded([&] {
        need(ctx != nullptr, UPDC_ERR_ARG, "null context");
        const unsigned shape = UPDC_FLAG_PER_EXPOSURE | UPDC_FLAG_FUSED_ADJOINT | UPDC_FLAG_COLUMN_SPECTRA;
        if (((ctx->flags ^ flags) & shape) && ctx->pipe) ctx->pipe->invalidate_graph();
        if ((flags & UPDC_FLAG_PER_EXPOSURE) && ctx->prepared) {
            HIPCHK(hipSetDevice(ctx->device));
            ctx->pipe->ensure_plain_buffers(*ctx);
        }
        ctx->flags = flags;
    });
}

int updc_ctx_set_preprocess_workers(updc_ctx *ctx, int workers) {
    return guarded([&] {
        need(ctx && workers >= 1, UPDC_ERR_ARG, "invalid worker count");
        ctx->workers = workers;
    });
}

int updc_fft_shape_candidates(int min_width, int min_height, int *widths, int *heights, int capacity, int *count) {
    return guarded([&] {
        need(min_width >= 2 && min_height >= 1 && count && capacity >= 0, UPDC_ERR_ARG, "invalid arguments");
        int shapes[16][2];
        const int n = shape_candidates(min_width, min_height, shapes, 16);
        for (int i = 0; i < n && i < capacity; ++i) {
            if (widths) widths[i] = shapes[i][0];
            if (heights) heights[i] = shapes[i][1];
        }
        *count = n;
    });
}

int updc_ctx_set_fft_cache(updc_ctx *ctx, const char *dir, int retune) {
    return guarded([&] {
        need(ctx != nullptr, UPDC_ERR_ARG, "null context");
        need(!ctx->prepared, UPDC_ERR_STATE, "context already prepared");
        ctx->fft_cache_dir = dir ? dir : "";
        ctx->fft_retune = retune != 0;
    });
}

// ---- PSF projector (psf_projector.c:82-172)
// Half extent of the output grid: the smallest h with max |coordinate| <= h +
// 0.5 (+1e-7 tolerance), psf_projector.c:71-80.
static int pp_half_extent(double max_abs) {
    if (!(max_abs >= 0.0) || !std::isfinite(max_abs)) throw Error(UPDC_ERR_NUMERIC, "invalid PSF mapped extent");
    const double needed = max_abs - 0.5 - 1.0e-7;
    if (needed <= 0.0) return 0;
    if (needed > 1000000.0) throw Error(UPDC_ERR_ARG, "unreasonable PSF mapped extent");
    return (int)std::ceil(needed);
}
int updc_psf_project(int device, const float *psf, int width, int height, const double *vertices, float *out,
                     size_t out_capacity, int *out_width, int *out_height, double *input_sum, double *output_sum,
                     uint64_t *overlaps) {
    return guarded([&] {
        need(psf && vertices && out_width && out_height && width >= 1 && height >= 1, UPDC_ERR_ARG,
             "invalid PSF projection arguments");
        const size_t nin = (size_t)width * height, nvert = (size_t)(width + 1) * (height + 1);
        double isum = 0.0;
        for (size_t i = 0; i < nin; ++i) {
            if (!std::isfinite(psf[i]) || psf[i] < 0.0f)
                throw Error(UPDC_ERR_NUMERIC, "PSF contains negative/non-finite pixels");
            isum += (double)psf[i];
        }
        if (!(isum > 0.0)) throw Error(UPDC_ERR_NUMERIC, "PSF input sum is zero");
        double mx = 0.0, my = 0.0;
        for (size_t i = 0; i < nvert; ++i) {
            const double x = vertices[2 * i], y = vertices[2 * i + 1];
            if (!std::isfinite(x) || !std::isfinite(y)) throw Error(UPDC_ERR_NUMERIC, "non-finite PSF mapped vertex");
            mx = std::max(mx, std::fabs(x));
            my = std::max(my, std::fabs(y));
        }
        const int hx = pp_half_extent(mx), hy = pp_half_extent(my);
        const int ow = 2 * hx + 1, oh = 2 * hy + 1;
        const size_t nout = (size_t)ow * oh;
        *out_width = ow;
        *out_height = oh;
        if (!out) return;  // size query
        need(out_capacity >= nout, UPDC_ERR_ARG, "PSF projection output buffer too small");
        HIPCHK(hipSetDevice(device));
        hipStream_t st = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        try {
            DevBuf<double> dv;
            DevBuf<float> dval, dfrac, dout;
            DevBuf<uint32_t> cnt, off, t32, tcnt, toff, cursor;
            DevBuf<uint2> tent;
            DevBuf<uint8_t> scratch;
            DevBuf<unsigned long long> tot;
            DevBuf<unsigned int> flags;
            DevBuf<ExpDev> one;
            dv.alloc(2 * nvert);
            dval.alloc(nin);
            HIPCHK(hipMemcpyAsync(dv.p, vertices, 2 * nvert * sizeof(double), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(dval.p, psf, nin * sizeof(float), hipMemcpyHostToDevice, st));
            cnt.alloc(nin);
            off.alloc(nin + 1);
            scratch.alloc(scan_scratch_bytes(std::max(nin, nout)));
            tot.alloc(1);
            flags.alloc(1);
            HIPCHK(fill<unsigned int>(flags.p, 0u, 1, st));
            HIPCHK(pp_count(dv.p, dval.p, width, height, hx, hy, cnt.p, st));
            HIPCHK(scan_u32(cnt.p, off.p, nin, scratch.p, tot.p, st));
            unsigned long long nl = 0;
            HIPCHK(hipMemcpyAsync(&nl, tot.p, sizeof nl, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            if (nl > 0xffffffffull) throw Error(UPDC_ERR_ARG, "PSF projection exceeds uint32 link capacity");
            t32.alloc(nl);
            dfrac.alloc(nl);
            HIPCHK(pp_fill(dv.p, dval.p, width, height, hx, hy, off.p, t32.p, dfrac.p, flags.p, st));
            // transposed (per output pixel, ascending input pixel) list, then the ordered sums
            ExpDev e{};
            e.nin = (uint32_t)nin;
            e.encoding = UPDC_LM_ABS32;
            e.off = off.p;
            e.t32 = t32.p;
            e.frac = dfrac.p;
            e.pix = 0;
            one.alloc(1);
            HIPCHK(hipMemcpyAsync(one.p, &e, sizeof e, hipMemcpyHostToDevice, st));
            tcnt.alloc(nout);
            toff.alloc(nout + 1);
            HIPCHK(fill<uint32_t>(tcnt.p, 0u, nout, st));
            HIPCHK(csc_count<float>(one.p, (uint32_t)nin, dval.p, tcnt.p, st));
            HIPCHK(scan_u32(tcnt.p, toff.p, nout, scratch.p, tot.p, st));
            unsigned long long ne = 0;
            HIPCHK(hipMemcpyAsync(&ne, tot.p, sizeof ne, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            tent.alloc(ne);
            cursor.alloc(nout + 1);
            HIPCHK(hipMemcpyAsync(cursor.p, toff.p, nout * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
            HIPCHK(csc_fill<float>(one.p, (uint32_t)nin, dval.p, cursor.p, tent.p, st));
            HIPCHK(csc_sort(toff.p, tent.p, nout, st));
            e.toff = toff.p;
            e.tent = tent.p;
            HIPCHK(hipMemcpyAsync(one.p, &e, sizeof e, hipMemcpyHostToDevice, st));
            dout.alloc(nout);
            HIPCHK(csc_apply<float>(one.p, dval.p, nout, dout.p, st));
            unsigned int fl = 0;
            HIPCHK(hipMemcpyAsync(&fl, flags.p, 4, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(out, dout.p, nout * sizeof(float), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            if (fl) throw Error(UPDC_ERR_FORMAT, "PSF projection link fill mismatch");
            double osum = 0.0;
            for (size_t i = 0; i < nout; ++i) {
                if (!std::isfinite(out[i]) || out[i] < 0.0f)
                    throw Error(UPDC_ERR_NUMERIC, "PSF contains negative/non-finite pixels");
                osum += (double)out[i];
            }
            if (input_sum) *input_sum = isum;
            if (output_sum) *output_sum = osum;
            if (overlaps) *overlaps = nl;
            if (std::fabs((osum - isum) / isum) > 5.0e-5)
                throw Error(UPDC_ERR_NUMERIC, "PSF polygon projection is not flux-conserving");
        } catch (...) {
            (void)hipStreamDestroy(st);
            throw;
        }
        HIPCHK(hipStreamDestroy(st));
    });
}

int updc_ctx_set_setup_cache(updc_ctx *ctx, const char *dir) {
    return guarded([&] {
        need(ctx != nullptr, UPDC_ERR_ARG, "null context");
        need(!ctx->prepared, UPDC_ERR_STATE, "context already prepared");
        ctx->setup_cache_dir = dir ? dir : "";
    });
}
int updc_get_setup_cache_hits(updc_ctx *ctx, int *hits) {
    return guarded([&] {
        need(ctx != nullptr && hits, UPDC_ERR_ARG, "null argument");
        *hits = ctx->setup_cache_hits;
    });
}

int updc_get_spectrum_layout(updc_ctx *ctx, int *layout) {
    return guarded([&] {
        need(ctx != nullptr && layout, UPDC_ERR_ARG, "null argument");
        need(ctx->prepared, UPDC_ERR_STATE, "context not prepared");
        *layout = ctx->pipe->spectrum_layout(*ctx);
    });
}

int updc_ctx_set_column_kernel(updc_ctx *ctx, int kind) {
    return guarded([&] {
        need(ctx != nullptr && kind >= -1 && kind <= 1, UPDC_ERR_ARG, "invalid column kernel kind");
        need(!ctx->prepared, UPDC_ERR_STATE, "context already prepared");
        ctx->col_edge = kind;
    });
}

int updc_get_autotune(updc_ctx *ctx, int *widths, int *heights, int *column_edge, double *ms, int capacity,
                      int *count, int *cached) {
    return guarded([&] {
        need(ctx != nullptr && count && capacity >= 0, UPDC_ERR_ARG, "invalid arguments");
        const int n = (int)ctx->tune.size();
        for (int i = 0; i < n && i < capacity; ++i) {
            if (widths) widths[i] = ctx->tune[i].w;
            if (heights) heights[i] = ctx->tune[i].h;
            if (column_edge) column_edge[i] = ctx->tune[i].edge;
            if (ms) ms[i] = ctx->tune[i].ms;
        }
        *count = n;
        if (cached) *cached = ctx->tune_cached ? 1 : 0;
    });
}

int updc_add_exposure(updc_ctx *ctx, const float *sci, const float *wht, int iw, int ih, const void *linkmap,
                      size_t linkmap_bytes, const float *psf, int pw, int ph, int psf_t, double bg, double ped) {
    return guarded([&] {
        need(ctx && sci && wht && psf && linkmap, UPDC_ERR_ARG, "null argument");
        HIPCHK(hipSetDevice(ctx->device));
        check_dims(*ctx, iw, ih, pw, ph);
        HostExp &e = new_exposure(*ctx, sci, wht, iw, ih, psf, pw, ph, psf_t, bg, ped);
        const double t0 = host_seconds();
        try {
            add_linkmap_image(*ctx, e, linkmap, linkmap_bytes);
            ctx->setup.linkmaps += host_seconds() - t0;
        } catch (...) {
            ctx->ex.pop_back();
            throw;
        }
    });
}

int updc_add_exposure_vertices(updc_ctx *ctx, const float *sci, const float *wht, int iw, int ih,
                               const double *vert, const float *psf, int pw, int ph, int psf_t, double bg,
                               double ped) {
    return guarded([&] {
        need(ctx && sci && wht && psf && vert, UPDC_ERR_ARG, "null argument");
        HIPCHK(hipSetDevice(ctx->device));
        check_dims(*ctx, iw, ih, pw, ph);
        HostExp &e = new_exposure(*ctx, sci, wht, iw, ih, psf, pw, ph, psf_t, bg, ped);
        const double t0 = host_seconds();
        try {
            build_linkmap(*ctx, e, vert);
            ctx->setup.linkmaps += host_seconds() - t0;
        } catch (...) {
            ctx->ex.pop_back();
            throw;
        }
    });
}

int updc_linkmap_size(updc_ctx *ctx, int k, size_t *bytes) {
    return guarded([&] {
        need(ctx && bytes && k >= 0 && k < (int)ctx->ex.size(), UPDC_ERR_ARG, "invalid exposure index");
        *bytes = ctx->ex[k]->lm_bytes;
    });
}
int updc_get_linkmap(updc_ctx *ctx, int k, void *out, size_t cap) {
    return guarded([&] {
        need(ctx && out && k >= 0 && k < (int)ctx->ex.size(), UPDC_ERR_ARG, "invalid exposure index");
        need(cap >= ctx->ex[k]->lm_bytes, UPDC_ERR_ARG, "buffer too small for LinkMap");
        HIPCHK(hipSetDevice(ctx->device));
        HIPCHK(hipMemcpy(out, ctx->ex[k]->lm.p, ctx->ex[k]->lm_bytes, hipMemcpyDeviceToHost));
    });
}

int updc_prepare(updc_ctx *ctx) {
    return guarded([&] {
        need(ctx != nullptr, UPDC_ERR_ARG, "null context");
        need(!ctx->prepared, UPDC_ERR_STATE, "context already prepared");
        need(!ctx->ex.empty(), UPDC_ERR_ARG, "no exposures");
        HIPCHK(hipSetDevice(ctx->device));
        int pad_w = 0, pad_h = 0;
        for (auto &e : ctx->ex) {
            pad_w = std::max(pad_w, e->psf_t ? e->ph : e->pw);
            pad_h = std::max(pad_h, e->psf_t ? e->pw : e->ph);
        }
        const int min_w = ctx->W + pad_w - 1, min_h = ctx->H + pad_h - 1;
        const double ta = host_seconds();
        if (ctx->FW == 0 && ctx->FH == 0 && (ctx->flags & UPDC_FLAG_AUTOTUNE)) autotune_shape(*ctx, min_w, min_h);
        ctx->setup.fft_autotune = host_seconds() - ta;
        if (ctx->FW == 0) ctx->FW = next_even_smooth(min_w);
        if (ctx->FH == 0) ctx->FH = next_smooth(min_h);
        need(ctx->FW >= min_w && ctx->FH >= min_h, UPDC_ERR_ARG, "FFT shape smaller than target + PSF - 1");
        need(!(ctx->FW & 1) && seven_smooth(ctx->FW) && seven_smooth(ctx->FH), UPDC_ERR_ARG,
             "FFT shape must be 7-smooth with an even width");
        if (ctx->prec == UPDC_PREC_F32) ctx->pipe = std::make_unique<Pipe<float, float>>();
        else if (ctx->prec == UPDC_PREC_F64) ctx->pipe = std::make_unique<Pipe<double, double>>();
        else ctx->pipe = std::make_unique<Pipe<float, double>>();
        try {
            ctx->pipe->prepare(*ctx);
            // the context's stream does not order against other streams: finish the
            // set-up before operators run on the caller's stream
            HIPCHK(hipStreamSynchronize(ctx->st));
        } catch (...) {
            ctx->pipe.reset();
            throw;
        }
        // host copies are no longer needed
        for (auto &e : ctx->ex) {
            std::vector<float>().swap(e->sci);
            std::vector<float>().swap(e->wht);
        }
        ctx->prepared = true;
    });
}

#define NEED_PREPARED(ctx) need((ctx) && (ctx)->prepared, UPDC_ERR_STATE, "context not prepared")

int updc_iterate(updc_ctx *ctx, int n, updc_freeze_row *rows, updc_iter_info *info) {
    return guarded([&] {
        NEED_PREPARED(ctx);
        need(n >= 1, UPDC_ERR_ARG, "invalid iteration count");
        ctx->iter_done = ctx->iter_rows = 0;
        HIPCHK(hipSetDevice(ctx->device));
        ctx->pipe->iterate(*ctx, n, rows, info);
    });
}
// Leaves updc_last_error() as it is: callers read it after a failed iterate.
int updc_get_setup_times(updc_ctx *ctx, updc_setup_times *out) {
    return guarded([&] {
        need(ctx && out, UPDC_ERR_ARG, "null argument");
        *out = ctx->setup;
    });
}
int updc_get_iteration_times(updc_ctx *ctx, updc_iter_times *out, int capacity, int *count) {
    return guarded([&] {
        need(ctx && count && (out || capacity == 0), UPDC_ERR_ARG, "null argument");
        const int n = (int)ctx->iter_times.size();
        for (int i = 0; i < std::min(n, capacity); ++i) out[i] = ctx->iter_times[i];
        *count = n;
    });
}
int updc_get_iterate_progress(updc_ctx *ctx, int *completed, int *freeze_rows) {
    if (!ctx) return UPDC_ERR_ARG;
    if (completed) *completed = ctx->iter_done;
    if (freeze_rows) *freeze_rows = ctx->iter_rows;
    return UPDC_OK;
}
int updc_get_estimate(updc_ctx *ctx, float *out) {
    return guarded([&] { NEED_PREPARED(ctx); need(out, UPDC_ERR_ARG, "null"); ctx->pipe->get_estimate(*ctx, nullptr, out); });
}
int updc_get_estimate_f64(updc_ctx *ctx, double *out) {
    return guarded([&] { NEED_PREPARED(ctx); need(out, UPDC_ERR_ARG, "null"); ctx->pipe->get_estimate(*ctx, out, nullptr); });
}
int updc_set_estimate_f64(updc_ctx *ctx, const double *in) {
    return guarded([&] { NEED_PREPARED(ctx); need(in, UPDC_ERR_ARG, "null"); ctx->pipe->set_estimate(*ctx, in); });
}
int updc_get_wht(updc_ctx *ctx, float *out) {
    return guarded([&] { NEED_PREPARED(ctx); need(out, UPDC_ERR_ARG, "null"); ctx->pipe->get_wht(*ctx, out); });
}
int updc_get_denominator_f64(updc_ctx *ctx, double *out) {
    return guarded([&] { NEED_PREPARED(ctx); need(out, UPDC_ERR_ARG, "null"); ctx->pipe->get_denominator(*ctx, out); });
}
int updc_get_initial_f64(updc_ctx *ctx, double *est, double *cov) {
    return guarded([&] { NEED_PREPARED(ctx); ctx->pipe->get_initial(*ctx, est, cov); });
}
int updc_get_psf_spectrum_f64(updc_ctx *ctx, int k, double *out) {
    return guarded([&] {
        NEED_PREPARED(ctx);
        need(out && k >= 0 && k < (int)ctx->ex.size(), UPDC_ERR_ARG, "invalid exposure index");
        ctx->pipe->get_psf_spectrum(*ctx, k, out);
    });
}
int updc_get_scalars(updc_ctx *ctx, double *ws, double *ms, int *fw, int *fh, uint64_t *masked) {
    return guarded([&] {
        NEED_PREPARED(ctx);
        if (ws) *ws = ctx->weight_scale;
        if (ms) *ms = ctx->pipe->max_sens;
        if (fw) *fw = ctx->FW;
        if (fh) *fh = ctx->FH;
        if (masked) *masked = ctx->masked;
    });
}
int updc_get_frozen(updc_ctx *ctx, int k, uint8_t *out) {
    return guarded([&] {
        NEED_PREPARED(ctx);
        need(out && k >= 0 && k < (int)ctx->ex.size(), UPDC_ERR_ARG, "invalid exposure index");
        ctx->pipe->get_frozen(*ctx, k, out);
    });
}
int updc_get_exposure_info(updc_ctx *ctx, int k, uint64_t *nin, uint64_t *nlinks, uint32_t *encoding,
                           uint64_t *ncsc) {
    return guarded([&] {
        need(ctx && k >= 0 && k < (int)ctx->ex.size(), UPDC_ERR_ARG, "invalid exposure index");
        const HostExp &e = *ctx->ex[k];
        if (nin) *nin = e.nin;
        if (nlinks) *nlinks = e.nlinks;
        if (encoding) *encoding = e.encoding;
        if (ncsc) *ncsc = e.ncsc;
    });
}
int updc_get_csc_layout(updc_ctx *ctx, int k, int *compact, uint64_t *bytes) {
    return guarded([&] {
        NEED_PREPARED(ctx);
        need(k >= 0 && k < (int)ctx->ex.size(), UPDC_ERR_ARG, "invalid exposure index");
        const HostExp &e = *ctx->ex[k];
        if (compact) *compact = e.c16 ? 1 : 0;
        if (bytes) *bytes = e.csc_bytes(ctx->nout());
    });
}
int updc_reset_frozen(updc_ctx *ctx) {
    return guarded([&] { NEED_PREPARED(ctx); ctx->pipe->reset_frozen(*ctx); });
}
int updc_stage_forward(updc_ctx *ctx, const double *x, double *blur) {
    return guarded([&] { NEED_PREPARED(ctx); need(x && blur, UPDC_ERR_ARG, "null"); ctx->pipe->stage_forward(*ctx, x, blur); });
}
int updc_stage_model(updc_ctx *ctx, const double *x, double *pred) {
    return guarded([&] { NEED_PREPARED(ctx); need(x && pred, UPDC_ERR_ARG, "null"); ctx->pipe->stage_model(*ctx, x, pred); });
}
int updc_stage_detector(updc_ctx *ctx, int k, const double *blur, double *wr, double *scatter, double *stats) {
    return guarded([&] {
        NEED_PREPARED(ctx);
        need(blur && k >= 0 && k < (int)ctx->ex.size(), UPDC_ERR_ARG, "invalid exposure index");
        ctx->pipe->stage_detector(*ctx, k, blur, wr, scatter, stats);
    });
}
int updc_stage_adjoint(updc_ctx *ctx, const double *v, double *out) {
    return guarded([&] { NEED_PREPARED(ctx); need(v && out, UPDC_ERR_ARG, "null"); ctx->pipe->stage_adjoint(*ctx, v, out); });
}
int updc_stage_scatter(updc_ctx *ctx, int k, const double *v, double *out) {
    return guarded([&] {
        NEED_PREPARED(ctx);
        need(v && out && k >= 0 && k < (int)ctx->ex.size(), UPDC_ERR_ARG, "invalid exposure index");
        ctx->pipe->stage_scatter(*ctx, k, v, out);
    });
}
int updc_stage_adjoint_images(updc_ctx *ctx, const double *v, double *out) {
    return guarded([&] {
        NEED_PREPARED(ctx);
        need(v && out, UPDC_ERR_ARG, "null");
        HIPCHK(hipSetDevice(ctx->device));
        ctx->pipe->stage_adjoint_images(*ctx, v, out);
    });
}

// ---- device-pointer operators
int updc_op_layout(updc_ctx *ctx, int *fft_width, int *fft_height, int *spectrum_columns, int *spectrum_pitch,
                   int *element_bytes) {
    return guarded([&] {
        NEED_PREPARED(ctx);
        if (fft_width) *fft_width = ctx->FW;
        if (fft_height) *fft_height = ctx->FH;
        if (spectrum_columns) *spectrum_columns = ctx->FW / 2 + 1;
        if (spectrum_pitch) *spectrum_pitch = rp(ctx->FH);
        if (element_bytes) *element_bytes = ctx->prec == UPDC_PREC_F64 ? 8 : 4;
    });
}
int updc_op_r2c(updc_ctx *ctx, const void *image, int rows, int cols, int pitch, void *spectrum, void *stream) {
    return guarded([&] {
        NEED_PREPARED(ctx);
        need(image && spectrum && rows >= 1 && rows <= ctx->FH && cols >= 1 && cols <= ctx->FW && pitch >= cols,
             UPDC_ERR_ARG, "invalid operator arguments");
        HIPCHK(hipSetDevice(ctx->device));
        ctx->pipe->op_r2c(*ctx, image, rows, cols, pitch, spectrum, (hipStream_t)stream);
    });
}
int updc_op_c2r(updc_ctx *ctx, const void *spectrum, void *image, int rows, int cols, int pitch, double scale,
                void *stream) {
    return guarded([&] {
        NEED_PREPARED(ctx);
        need(image && spectrum && rows >= 1 && rows <= ctx->FH && cols >= 1 && cols <= ctx->FW && pitch >= cols,
             UPDC_ERR_ARG, "invalid operator arguments");
        HIPCHK(hipSetDevice(ctx->device));
        ctx->pipe->op_c2r(*ctx, spectrum, image, rows, cols, pitch, scale, (hipStream_t)stream);
    });
}
// ---- allocation-free C2R (updc_op_c2r_noalloc)
// Its column-inverse workspace: a file-scope device array, one copy on each
// GPU that loads the library, held until the process ends.  Host code never
// passes the array's name to a kernel or a copy (on the host it is only a
// stand-in); it uses the per-device address hipGetSymbolAddress returns.
static __device__ __align__(256) unsigned char scratch[UPDC_NOALLOC_WORKSPACE_BYTES];
static constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

// Address of `scratch` on the current device `dev` (looked up once per device).
static void *scratch_on(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw Error(UPDC_ERR_ARG, "device ordinal beyond the workspace table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(UPDC_ERR_HIP, "no device address for the allocation-free workspace");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

int updc_op_c2r_noalloc(updc_ctx *ctx, const void *spectrum, void *image, int rows, int cols, int pitch,
                        double scale, void *stream) {
    return guarded([&] {
        NEED_PREPARED(ctx);
        need(image && spectrum && rows >= 1 && rows <= ctx->FH && cols >= 1 && cols <= ctx->FW && pitch >= cols,
             UPDC_ERR_ARG, "invalid operator arguments");
        const size_t need_bytes = ctx->pipe->op_c2r_workspace_bytes();
        if (need_bytes > (size_t)UPDC_NOALLOC_WORKSPACE_BYTES)
            throw Error(UPDC_ERR_ARG, "spectrum of " + std::to_string(need_bytes) +
                                          " bytes exceeds the allocation-free workspace of " +
                                          std::to_string((size_t)UPDC_NOALLOC_WORKSPACE_BYTES) + " bytes");
        // a four-step FFT axis works in the scratch updc_prepare reserved
        need(ctx->pipe->op_c2r_ws_ready(rows), UPDC_ERR_STATE, "four-step FFT scratch not reserved");
        HIPCHK(hipSetDevice(ctx->device));
        void *ws = scratch_on(ctx->device);
        ctx->pipe->op_c2r_ws(*ctx, spectrum, image, rows, cols, pitch, scale, ws, (hipStream_t)stream);
    });
}
int updc_op_gather(updc_ctx *ctx, int k, const void *image, void *out, void *stream) {
    return guarded([&] {
        NEED_PREPARED(ctx);
        need(image && out && k >= 0 && k < (int)ctx->ex.size(), UPDC_ERR_ARG, "invalid operator arguments");
        HIPCHK(hipSetDevice(ctx->device));
        ctx->pipe->op_gather(*ctx, k, image, out, (hipStream_t)stream);
    });
}
int updc_op_scatter(updc_ctx *ctx, int k, const void *v, void *image, void *stream) {
    return guarded([&] {
        NEED_PREPARED(ctx);
        need(v && image && k >= 0 && k < (int)ctx->ex.size(), UPDC_ERR_ARG, "invalid operator arguments");
        HIPCHK(hipSetDevice(ctx->device));
        ctx->pipe->op_scatter(*ctx, k, v, image, (hipStream_t)stream);
    });
}

int updc_fft_r2c_f64(updc_ctx *ctx, const double *image, double *spectrum) {
    return guarded([&] { NEED_PREPARED(ctx); need(image && spectrum, UPDC_ERR_ARG, "null"); ctx->pipe->fft_r2c(*ctx, image, spectrum); });
}
int updc_fft_c2r_f64(updc_ctx *ctx, const double *spectrum, double *image) {
    return guarded([&] { NEED_PREPARED(ctx); need(image && spectrum, UPDC_ERR_ARG, "null"); ctx->pipe->fft_c2r(*ctx, spectrum, image); });
}
int updc_copy_calibration(size_t bytes, int reps, int device, double *ms) {
    return guarded([&] {
        need(bytes >= 16 && reps >= 1, UPDC_ERR_ARG, "invalid copy size");
        HIPCHK(hipSetDevice(device));
        DevBuf<uint8_t> a, b;
        a.alloc(bytes);
        b.alloc(bytes);
        hipStream_t st;
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(fill<uint8_t>(a.p, (uint8_t)1, bytes, st));
        HIPCHK(copy_ref(a.p, b.p, bytes, st));
        HIPCHK(hipEventRecord(e0, st));
        for (int i = 0; i < reps; ++i) HIPCHK(copy_ref(a.p, b.p, bytes, st));
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        if (ms) *ms = t / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(st);
    });
}

int updc_get_stage_times(updc_ctx *ctx, char names[][32], double *ms, int *launches, int *count) {
    return guarded([&] {
        NEED_PREPARED(ctx);
        need(count != nullptr, UPDC_ERR_ARG, "null");
        PipeBase &p = *ctx->pipe;
        const int have = (int)p.stage_ms.size();
        const int m = std::min(*count, have);
        for (int i = 0; i < m; ++i) {
            const std::string nm = i < (int)p.stage_names.size() ? p.stage_names[i] : std::string("iteration");
            if (names) {
                std::strncpy(names[i], nm.c_str(), 31);
                names[i][31] = 0;
            }
            if (ms) ms[i] = p.stage_ms[i];
            if (launches) launches[i] = p.stage_n[i];
        }
        *count = m;
    });
}

}  // extern "C"
