// updc_hip_psf_projector: drop-in for the reference's updc_psf_projector
// (updc/c_core/psf_projector.c:174-268), the engine's `psf_projector_binary`
// (updc/_engine.py:437-470).  Same command line (`manifest workers results.csv`,
// `--self-test`), manifest (input \t width \t height \t vertices \t output \t
// label), raw float32 outputs and results CSV; a failure prints
// "UPDC_PSF_PROJECT: <message>" on stderr and exits 1.  The projection itself
// runs on the GPU (updc_psf_project); `workers` is accepted and unused.
#include "updc_hip.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

[[noreturn]] void die(const std::string &msg) {
    std::fprintf(stderr, "UPDC_PSF_PROJECT: %s\n", msg.c_str());
    std::exit(EXIT_FAILURE);
}

template <class T> std::vector<T> read_exact(const std::string &path, size_t count, const char *what) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) die(std::string("cannot open ") + what);
    std::vector<T> v(count);
    if (std::fread(v.data(), sizeof(T), count, f) != count) die(std::string("short ") + what);
    if (std::fgetc(f) != EOF) die(std::string(what) + " has trailing bytes");
    std::fclose(f);
    return v;
}

struct Job {
    std::string input, vertices, output;
    int width = 0, height = 0, ow = 0, oh = 0;
    double in_sum = 0, out_sum = 0, rel = 0;
    unsigned long long overlaps = 0;
};

}  // namespace

int main(int argc, char **argv) {
    if (argc == 2 && std::strcmp(argv[1], "--self-test") == 0) {
        const double q[8] = {-0.5, -0.5, -0.5, 0.5, 0.5, 0.5, 0.5, -0.5};
        if (std::fabs(updc_polygon_area(q, 4) - 1.0f) > 2.0e-6f) return 1;
        if (std::fabs(updc_overlap_area(q, 0, 0) - 1.0f) > 2.0e-6f) return 1;
        std::printf("UPDC PSF projector self-test: OK\n");
        return 0;
    }
    if (argc != 4) die("usage: updc_psf_projector manifest workers results.csv");
    FILE *mf = std::fopen(argv[1], "r");
    if (!mf) die("cannot open PSF projection manifest");
    std::vector<Job> jobs;
    char line[8192];
    while (std::fgets(line, sizeof line, mf)) {
        if (line[0] == '\0' || line[0] == '\n' || line[0] == '#') continue;
        char *save = nullptr, *f[6];
        for (int j = 0; j < 6; ++j) f[j] = strtok_r(j ? nullptr : line, "\t\r\n", &save);
        for (int j = 0; j < 6; ++j)
            if (!f[j]) die("PSF manifest needs input,width,height,vertices,output,label");
        Job jb;
        jb.input = f[0];
        jb.width = (int)std::strtol(f[1], nullptr, 10);
        jb.height = (int)std::strtol(f[2], nullptr, 10);
        jb.vertices = f[3];
        jb.output = f[4];
        if (jb.width < 1 || jb.height < 1) die("invalid PSF input dimensions");
        jobs.push_back(jb);
    }
    std::fclose(mf);
    if (jobs.empty()) die("empty PSF projection manifest");
    const int device = std::getenv("UPDC_HIP_DEVICE") ? std::atoi(std::getenv("UPDC_HIP_DEVICE")) : 0;
    for (size_t i = 0; i < jobs.size(); ++i) {
        Job &jb = jobs[i];
        const size_t nin = (size_t)jb.width * jb.height;
        const size_t nv = 2 * (size_t)(jb.width + 1) * (jb.height + 1);
        std::vector<float> psf = read_exact<float>(jb.input, nin, "PSF raw input");
        std::vector<double> vert = read_exact<double>(jb.vertices, nv, "PSF vertex map");
        if (updc_psf_project(device, psf.data(), jb.width, jb.height, vert.data(), nullptr, 0, &jb.ow, &jb.oh,
                             nullptr, nullptr, nullptr) != UPDC_OK)
            die(updc_last_error());
        std::vector<float> out((size_t)jb.ow * jb.oh);
        uint64_t ov = 0;
        const int rc = updc_psf_project(device, psf.data(), jb.width, jb.height, vert.data(), out.data(), out.size(),
                                        &jb.ow, &jb.oh, &jb.in_sum, &jb.out_sum, &ov);
        jb.overlaps = ov;
        jb.rel = jb.in_sum > 0 ? (jb.out_sum - jb.in_sum) / jb.in_sum : 0.0;
        if (rc != UPDC_OK) {
            if (std::fabs(jb.rel) > 5.0e-5 && jb.in_sum > 0)
                std::fprintf(stderr, "UPDC_PSF_PROJECT: exposure %zu flux mismatch input=%.17g output=%.17g rel=%.6g\n",
                             i + 1, jb.in_sum, jb.out_sum, jb.rel);
            die(updc_last_error());
        }
        FILE *o = std::fopen(jb.output.c_str(), "wb");
        if (!o) die("cannot create projected PSF raw output");
        if (std::fwrite(out.data(), sizeof(float), out.size(), o) != out.size()) die("short PSF raw output write");
        if (std::fclose(o)) die("cannot close projected PSF raw output");
        std::printf("PSF polygon %zu/%zu: %dx%d -> %dx%d, overlaps=%llu, flux rel=%.3e\n", i + 1, jobs.size(),
                    jb.width, jb.height, jb.ow, jb.oh, jb.overlaps, jb.rel);
        std::fflush(stdout);
    }
    FILE *r = std::fopen(argv[3], "w");
    if (!r) die("cannot create PSF projection results");
    std::fprintf(r, "exposure,width,height,input_sum,output_sum,flux_rel_error,overlaps\n");
    for (size_t i = 0; i < jobs.size(); ++i) {
        const Job &j = jobs[i];
        std::fprintf(r, "%zu,%d,%d,%.17g,%.17g,%.17g,%llu\n", i + 1, j.ow, j.oh, j.in_sum, j.out_sum, j.rel,
                     j.overlaps);
    }
    if (std::fclose(r)) die("cannot close PSF projection results");
    return 0;
}
