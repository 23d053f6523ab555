// updc_all: the drop-in replacement of the reference's UPDC_ALL
// (updc/c_core/updc_core.c:1289-1543) over the context API.
//
// File contract kept as-is (updc_core.c:298-315, 426-461, 463-663, 1404-1508):
//   manifest.txt rows  science \t weight \t psf \t psf_w \t psf_h \t vertices \t background \t pedestal
//   raw planes         little-endian float32, exact size (no trailing bytes)
//   <vertices>.links   ULM2 LinkMap, reused when UPDC_PERSIST_CACHE is set and the
//                      header/size validate, rebuilt (on the GPU) otherwise
//   <prefix>_%02d.bin  saved iterations (save_every rule, last always saved)
//   <wht_path>         coverage * weight_scale, float32
//   <prefix>_mufreeze.csv, <prefix>_timing.csv
// Errors return a status and a message instead of exiting.
#include "updc_hip.h"

#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

namespace {

struct DErr : std::runtime_error {
    int code;
    DErr(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Row {
    std::string sci, wht, psf, vert, links;
    int pw = 0, ph = 0;
    float bg_f = 0, ped_f = 0;
    double bg = 0, ped = 0;
};

// updc_core.c:426-461
std::vector<Row> parse_manifest(const char *path, int n) {
    FILE *f = std::fopen(path, "r");
    if (!f) throw DErr(UPDC_ERR_IO, "cannot open manifest");
    std::vector<Row> rows;
    char line[8192];
    for (int k = 0; k < n; ++k) {
        if (!std::fgets(line, sizeof line, f)) {
            std::fclose(f);
            throw DErr(UPDC_ERR_ARG, "manifest has fewer rows than expected");
        }
        char *save = nullptr, *field[8];
        for (int j = 0; j < 8; ++j) field[j] = strtok_r(j ? nullptr : line, "\t\r\n", &save);
        for (int j = 0; j < 8; ++j)
            if (!field[j]) {
                std::fclose(f);
                throw DErr(UPDC_ERR_ARG,
                           "manifest requires science, weight, psf, psf_width, psf_height, vertices, background, pedestal");
            }
        Row r;
        r.sci = field[0];
        r.wht = field[1];
        r.psf = field[2];
        r.pw = (int)std::strtol(field[3], nullptr, 10);
        r.ph = (int)std::strtol(field[4], nullptr, 10);
        r.vert = field[5];
        // the reference parses these with strtof (f32); the FP64 tiers keep strtod
        r.bg_f = std::strtof(field[6], nullptr);
        r.ped_f = std::strtof(field[7], nullptr);
        r.bg = std::strtod(field[6], nullptr);
        r.ped = std::strtod(field[7], nullptr);
        r.links = r.vert + ".links";
        if (r.pw < 1 || r.ph < 1) {
            std::fclose(f);
            throw DErr(UPDC_ERR_ARG, "invalid PSF dimensions in manifest");
        }
        const float mo = r.bg_f + r.ped_f;
        if (!std::isfinite(r.bg_f) || !std::isfinite(r.ped_f) || r.ped_f < 0.0f || !std::isfinite(mo)) {
            std::fclose(f);
            throw DErr(UPDC_ERR_ARG, "invalid background/pedestal in manifest");
        }
        rows.push_back(r);
    }
    if (std::fgets(line, sizeof line, f)) {
        std::fclose(f);
        throw DErr(UPDC_ERR_ARG, "manifest has more rows than expected");
    }
    std::fclose(f);
    return rows;
}

// Raw float32 planes (updc_core.c:298-315: the label and path on stderr,
// then the failure, exact size) and the f64 vertex map (:536-539: its own
// messages, no trailing-byte check).
template <class T> std::vector<T> read_exact(const std::string &path, size_t count, const char *label) {
    constexpr bool vertex = sizeof(T) == sizeof(double);
    auto fail = [&](const char *f32msg, const char *vmsg) {
        if (!vertex) std::fprintf(stderr, "%s: %s\n", label, path.c_str());
        return DErr(UPDC_ERR_IO, vertex ? vmsg : f32msg);
    };
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw fail("cannot open float32 input", "cannot open vertex map");
    std::vector<T> v(count);
    if (std::fread(v.data(), sizeof(T), count, f) != count) {
        std::fclose(f);
        throw fail("short float32 input", "short vertex map");
    }
    if (!vertex && std::fgetc(f) != EOF) {
        std::fclose(f);
        throw fail("float32 input has unexpected trailing bytes", "");
    }
    std::fclose(f);
    return v;
}

bool read_file(const std::string &path, std::vector<uint8_t> &out) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    struct stat st;
    if (fstat(fileno(f), &st) != 0) {
        std::fclose(f);
        return false;
    }
    out.resize((size_t)st.st_size);
    const bool ok = std::fread(out.data(), 1, out.size(), f) == out.size();
    std::fclose(f);
    return ok;
}

// tmp -> fsync -> rename (updc_core.c:151-157, 574-601)
void write_atomic(const std::string &path, const void *data, size_t bytes, const char *what) {
    const std::string tmp = path + ".build." + std::to_string((long)getpid());
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) throw DErr(UPDC_ERR_IO, std::string("cannot create ") + what);
    if (bytes && std::fwrite(data, 1, bytes, f) != bytes) {
        std::fclose(f);
        throw DErr(UPDC_ERR_IO, std::string(what) + " write failed");
    }
    if (std::fflush(f) || fsync(fileno(f)) || std::fclose(f)) throw DErr(UPDC_ERR_IO, std::string(what) + " close failed");
    if (std::rename(tmp.c_str(), path.c_str())) throw DErr(UPDC_ERR_IO, std::string(what) + " rename failed");
}

void check(int rc) {
    if (rc != UPDC_OK) throw DErr(rc, updc_last_error());
}

struct CtxGuard {
    updc_ctx *c = nullptr;
    ~CtxGuard() {
        if (c) updc_ctx_destroy(c);
    }
};

// The cache root is the manifest's directory (updc_core.c:130-138).
std::string cache_root(const char *manifest) {
    const std::string m = manifest;
    const size_t slash = m.rfind('/');
    return slash == std::string::npos ? std::string(".") : m.substr(0, slash);
}

int env_int(const char *name, int dflt) {
    const char *v = std::getenv(name);
    return (v && *v) ? std::atoi(v) : dflt;
}

void run_all(const char *manifest, int W, int H, int iw, int ih, int n, int psf_t, int iterations,
             const char *prefix, const char *wht_out, int fftw, int expw, int prew, int plan_measure, int autotune,
             int retune, const char *wisdom, int save_every) {
    (void)fftw; (void)plan_measure; (void)wisdom;
    if (W < 1 || H < 1 || iw < 1 || ih < 1 || n < 1 || iterations < 1 || fftw < 1 || expw < 1 || prew < 1 ||
        save_every < 0)
        throw DErr(UPDC_ERR_ARG, "invalid dimensions/exposures/iterations/workers/save_every");
    const double t_start = now();
    const size_t nout = (size_t)W * H, nin = (size_t)iw * ih;
    if (nout > 0xffffffffull) throw DErr(UPDC_ERR_ARG, "target grid exceeds 32-bit link index capacity");
    const char *pc = std::getenv("UPDC_PERSIST_CACHE");
    const bool persist = pc && std::strcmp(pc, "0") != 0;
    const bool quiet = env_int("UPDC_HIP_QUIET", 0) != 0;
    std::printf("Persistent C preprocessing cache: %s\n", persist ? "enabled" : "disabled");
    if (expw > 1)
        std::printf("exposure_workers=%d: the GPU sums the exposures' adjoint spectra in exposure order, not as "
                    "per-worker partial sums (updc_core.c:1214-1216, 1280-1285)\n", expw);
    std::vector<Row> rows = parse_manifest(manifest, n);

    // precision tier of the drop-in: reference float32 by default
    int prec = UPDC_PREC_F32;
    if (const char *p = std::getenv("UPDC_HIP_PRECISION")) {
        if (!std::strcmp(p, "f64")) prec = UPDC_PREC_F64;
        else if (!std::strcmp(p, "mixed")) prec = UPDC_PREC_MIXED;
    }
    // eager launches by default, so every per-iteration stage of the timing
    // CSV is measured; UPDC_HIP_GRAPH=1 replays a hipGraph (per-iteration
    // totals only, the stage rows then read nan)
    unsigned flags = env_int("UPDC_HIP_GRAPH", 0) ? UPDC_FLAG_GRAPH : 0u;
    if (quiet) flags |= UPDC_FLAG_QUIET;
    // FFT shape: autotuned on the GPU when asked, with its result cached in
    // the manifest's directory when the persistent cache is on and no retune
    // is requested (updc_core.c:130-138, 187-223, 242-291)
    if (autotune) flags |= UPDC_FLAG_AUTOTUNE;
    CtxGuard g;
    check(updc_ctx_create(&g.c, W, H, 0, 0, prec, flags, env_int("UPDC_HIP_DEVICE", 0)));
    check(updc_ctx_set_preprocess_workers(g.c, prew));
    if (autotune && persist) check(updc_ctx_set_fft_cache(g.c, cache_root(manifest).c_str(), retune));
    if (persist) check(updc_ctx_set_setup_cache(g.c, cache_root(manifest).c_str()));

    double t_links = 0, t_data = 0;
    for (int k = 0; k < n; ++k) {
        const Row &r = rows[k];
        // F32 keeps the reference's strtof values exactly; the FP64 tiers use strtod
        const double bg = prec == UPDC_PREC_F32 ? (double)r.bg_f : r.bg;
        const double ped = prec == UPDC_PREC_F32 ? (double)r.ped_f : r.ped;
        double td = now();
        std::vector<float> sci = read_exact<float>(r.sci, nin, "SCI raw");
        std::vector<float> wht = read_exact<float>(r.wht, nin, "WHT raw");
        std::vector<float> psf = read_exact<float>(r.psf, (size_t)r.pw * r.ph, "PSF raw");
        t_data += now() - td;
        double tl = now();
        std::vector<uint8_t> img;
        bool hit = false;
        if (persist && read_file(r.links, img)) {
            const int rc = updc_add_exposure(g.c, sci.data(), wht.data(), iw, ih, img.data(), img.size(),
                                             psf.data(), r.pw, r.ph, psf_t, bg, ped);
            if (rc == UPDC_OK) hit = true;
            else if (rc != UPDC_ERR_FORMAT) check(rc);
        }
        if (!hit) {
            std::vector<double> vert =
                read_exact<double>(r.vert, 2 * (size_t)(iw + 1) * (ih + 1), "vertex map");
            check(updc_add_exposure_vertices(g.c, sci.data(), wht.data(), iw, ih, vert.data(), psf.data(), r.pw,
                                             r.ph, psf_t, bg, ped));
            size_t bytes = 0;
            check(updc_linkmap_size(g.c, k, &bytes));
            img.resize(bytes);
            check(updc_get_linkmap(g.c, k, img.data(), bytes));
            write_atomic(r.links, img.data(), bytes, "LinkMap");
        }
        uint32_t enc = 0;
        std::memcpy(&enc, img.data() + 8, 4);
        uint64_t nl = 0;
        std::memcpy(&nl, img.data() + 24, 8);
        std::printf("AreaLink exposure %d: %s%llu overlaps; encoding=%s\n", k + 1, hit ? "persistent cache hit, " : "",
                    (unsigned long long)nl, enc == UPDC_LINKMAP_DELTA16 ? "DELTA16" : "ABS32");
        t_links += now() - tl;
    }
    check(updc_prepare(g.c));
    double ws = 0, ms = 0;
    int fw = 0, fh = 0;
    uint64_t masked = 0;
    check(updc_get_scalars(g.c, &ws, &ms, &fw, &fh, &masked));
    std::printf("Background-aware UPDC (MI355X): %d exposures, input %dx%d, target %dx%d, FFT %dx%d, iterations %d\n",
                n, iw, ih, W, H, fw, fh, iterations);
    std::printf("WHT median normalization %.9g; %llu masked pixels; max H^T W sensitivity %.9g\n", ws,
                (unsigned long long)masked, ms);
    std::fflush(stdout);

    const std::string freeze_path = std::string(prefix) + "_mufreeze.csv";
    FILE *fcsv = std::fopen(freeze_path.c_str(), "w");
    if (!fcsv) throw DErr(UPDC_ERR_IO, "cannot write freeze CSV");
    std::fprintf(fcsv, "iteration,exposure,n_new_frozen,n_total_frozen,min_raw_mu,mu_marker,frozen_ratio\n");
    std::vector<float> est(nout);
    std::vector<double> iter_s(iterations, 0.0), out_s(iterations, 0.0);
    std::vector<updc_iter_times> stage(iterations);
    int done = 0;
    try {
        while (done < iterations) {
            // run up to the next saved iteration in one call
            int next = iterations;
            if (save_every > 0) next = std::min(iterations, (done / save_every + 1) * save_every);
            const int chunk = next - done;
            std::vector<updc_freeze_row> fr((size_t)chunk * n);
            std::vector<updc_iter_info> info(chunk);
            const double tc = now();
            const int rc = updc_iterate(g.c, chunk, fr.data(), info.data());
            const double el = now() - tc;
            const std::string failure = rc != UPDC_OK ? updc_last_error() : "";
            // after a numeric failure only the iterations the reference would
            // have finished print and log (updc_core.c:1425-1431, 1473-1478)
            int completed = chunk, with_rows = chunk;
            if (rc != UPDC_OK) check(updc_get_iterate_progress(g.c, &completed, &with_rows));
            for (int i = 0; i < with_rows; ++i)
                for (int k = 0; k < n; ++k) {
                    const updc_freeze_row &r = fr[(size_t)i * n + k];
                    std::fprintf(fcsv, "%d,%d,%llu,%llu,%.9g,%.9g,1\n", done + i + 1, k + 1,
                                 (unsigned long long)r.n_new_frozen, (unsigned long long)r.n_total_frozen,
                                 (double)(float)r.min_raw_mu, (double)1.0e-12f);
                }
            int cnt = 0;
            const int trc = updc_get_iteration_times(g.c, stage.data() + done, chunk, &cnt);
            for (int i = 0; i < completed; ++i) iter_s[done + i] = el / chunk;
            auto line = [&](int i, bool saved) {
                if (quiet) return;
                const updc_iter_times &t = stage[done + i];
                std::printf("UPDC iteration %d/%d: signed sum %.9g, negative %llu/%zu, frozen +%llu total %llu, "
                            "min raw mu %.9g, wall %.3f s [estimateFFT %.3f fusedWall %.3f forwardIFFTcrit %.3f "
                            "detectorCrit %.3f scatterFFTcrit %.3f adjointReduceIFFT %.3f update %.3f output %.3f]%s\n",
                            done + i + 1, iterations, info[i].signed_sum, (unsigned long long)info[i].n_negative,
                            nout, (unsigned long long)info[i].n_new_frozen,
                            (unsigned long long)info[i].n_total_frozen, (double)(float)info[i].min_raw_mu,
                            iter_s[done + i] + out_s[done + i], t.estimate_fft, t.fused_exposure_wall,
                            t.forward_ifft_critical, t.fused_detector_critical, t.scatter_fft_critical,
                            t.adjoint_reduce_ifft, t.update, out_s[done + i], saved ? " SAVE" : "");
            };
            if (rc != UPDC_OK) {
                for (int i = 0; i < completed; ++i) line(i, false);
                throw DErr(rc, failure);
            }
            check(trc);
            for (int i = 0; i + 1 < chunk; ++i) line(i, false);
            const double to = now();
            check(updc_get_estimate(g.c, est.data()));
            char path[4096];
            if (std::snprintf(path, sizeof path, "%s_%02d.bin", prefix, next) >= (int)sizeof path)
                throw DErr(UPDC_ERR_ARG, "iteration output path is too long");
            FILE *f = std::fopen(path, "wb");
            if (!f) throw DErr(UPDC_ERR_IO, "cannot create iteration output");
            if (std::fwrite(est.data(), 4, nout, f) != nout) {
                std::fclose(f);
                throw DErr(UPDC_ERR_IO, "iteration output write failed");
            }
            if (std::fclose(f)) throw DErr(UPDC_ERR_IO, "iteration output close failed");
            std::printf("UPDC iteration %d raw output: %s\n", next, path);
            out_s[next - 1] = now() - to;
            line(chunk - 1, true);
            std::fflush(stdout);
            done = next;
        }
    } catch (...) {
        std::fclose(fcsv);
        throw;
    }
    if (std::fclose(fcsv)) throw DErr(UPDC_ERR_IO, "freeze CSV close failed");

    std::vector<float> wht(nout);
    check(updc_get_wht(g.c, wht.data()));
    {
        FILE *f = std::fopen(wht_out, "wb");
        if (!f) throw DErr(UPDC_ERR_IO, "cannot create WHT output");
        if (std::fwrite(wht.data(), 4, nout, f) != nout) {
            std::fclose(f);
            throw DErr(UPDC_ERR_IO, "WHT output write failed");
        }
        if (std::fclose(f)) throw DErr(UPDC_ERR_IO, "WHT output close failed");
    }
    // timing CSV with the reference's rows (updc_core.c:1487-1508): set-up
    // stages, then nine rows per iteration; every value in seconds.  No
    // memory-mapped files here (LinkMaps live in HBM): persistent_mmap is the
    // time spent reading and writing the LinkMap cache files, mmap_prewarm 0.
    // csc_transpose (the transposed LinkMaps) has no reference counterpart.
    updc_setup_times su{};
    check(updc_get_setup_times(g.c, &su));
    const std::string tpath = std::string(prefix) + "_timing.csv";
    FILE *tf = std::fopen(tpath.c_str(), "w");
    if (!tf) throw DErr(UPDC_ERR_IO, "cannot write timing CSV");
    std::fprintf(tf,
                 "stage,wall_seconds\nfft_autotune,%.6f\nfft_plan,%.6f\npsf_cache,%.6f\npolygon_links,%.6f\n"
                 "persistent_mmap,%.6f\nmmap_prewarm,%.6f\ndata_cache,%.6f\nweight_scale,%.6f\n"
                 "initial_drizzle,%.6f\nsensitivity,%.6f\ncsc_transpose,%.6f\n",
                 su.fft_autotune, su.fft_plan, su.psf_spectra, su.linkmaps, std::max(0.0, t_links - su.linkmaps), 0.0,
                 t_data, su.weight_scale, su.initial_drizzle, su.sensitivity, su.csc_transpose);
    for (int i = 0; i < iterations; ++i) {
        const updc_iter_times &t = stage[i];
        std::fprintf(tf, "iteration_%02d_estimate_fft,%.6f\n", i + 1, t.estimate_fft);
        std::fprintf(tf, "iteration_%02d_fused_exposure_wall,%.6f\n", i + 1, t.fused_exposure_wall);
        std::fprintf(tf, "iteration_%02d_forward_ifft_critical,%.6f\n", i + 1, t.forward_ifft_critical);
        std::fprintf(tf, "iteration_%02d_fused_detector_critical,%.6f\n", i + 1, t.fused_detector_critical);
        std::fprintf(tf, "iteration_%02d_scatter_fft_critical,%.6f\n", i + 1, t.scatter_fft_critical);
        std::fprintf(tf, "iteration_%02d_adjoint_reduce_ifft,%.6f\n", i + 1, t.adjoint_reduce_ifft);
        std::fprintf(tf, "iteration_%02d_update,%.6f\n", i + 1, t.update);
        std::fprintf(tf, "iteration_%02d_output,%.6f\n", i + 1, out_s[i]);
        std::fprintf(tf, "iteration_%02d_total,%.6f\n", i + 1, iter_s[i] + out_s[i]);
    }
    std::fprintf(tf, "total_solver,%.6f\n", now() - t_start);
    if (std::fclose(tf)) throw DErr(UPDC_ERR_IO, "timing CSV close failed");
    std::printf("FFT transforms per iteration: %d (2N+2)\n", 2 * n + 2);
    std::fflush(stdout);
}

}  // namespace

namespace updc { extern thread_local std::string g_last_error; }

extern "C" int updc_all(const char *manifest, int width, int height, int input_width, int input_height,
                        int exposures, int transpose_psf, int iterations, const char *iter_prefix,
                        const char *weight_output, int fft_worker_count, int exposure_worker_count,
                        int preprocess_worker_count, int plan_measure, int autotune, int retune_fft,
                        const char *wisdom_path, int save_every) {
    int code = UPDC_OK;
    try {
        if (!manifest || !iter_prefix || !weight_output) throw DErr(UPDC_ERR_ARG, "null path");
        run_all(manifest, width, height, input_width, input_height, exposures, transpose_psf, iterations,
                iter_prefix, weight_output, fft_worker_count, exposure_worker_count, preprocess_worker_count,
                plan_measure, autotune, retune_fft, wisdom_path ? wisdom_path : "", save_every);
        updc::g_last_error.clear();
    } catch (const DErr &e) {
        updc::g_last_error = e.what();
        code = e.code;
    } catch (const std::bad_alloc &) {
        updc::g_last_error = "out of memory";
        code = UPDC_ERR_NOMEM;
    } catch (const std::exception &e) {
        updc::g_last_error = e.what();
        code = UPDC_ERR_ARG;
    }
    std::fflush(stdout);
    return code;
}
