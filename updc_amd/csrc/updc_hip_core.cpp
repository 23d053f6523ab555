// updc_hip_core: drop-in executable for the reference's updc_jwst_c
// (updc/c_core/updc_core.c:1566-1573): same 18-argument argv, same
// `--self-test`, same outputs; a failure prints "UPDC_ALL: <message>" on
// stderr and exits with status 1 (EXIT_FAILURE, updc_core.c:109).
#include "updc_hip.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

int main(int argc, char **argv) {
    if (argc == 2 && std::strcmp(argv[1], "--self-test") == 0) return updc_self_test(1) == 0 ? 0 : 1;
    if (argc != 19) {
        std::fprintf(stderr,
                     "UPDC_ALL: usage: updc_hip_core manifest width height input_width input_height exposures "
                     "psf_transposed iterations iter_prefix wht.bin fft_workers exposure_workers preprocess_workers "
                     "plan_measure autotune retune_fft wisdom_path save_every\n");
        return EXIT_FAILURE;
    }
    const int rc = updc_all(argv[1], std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]),
                            std::atoi(argv[6]), std::atoi(argv[7]), std::atoi(argv[8]), argv[9], argv[10],
                            std::atoi(argv[11]), std::atoi(argv[12]), std::atoi(argv[13]), std::atoi(argv[14]),
                            std::atoi(argv[15]), std::atoi(argv[16]), argv[17], std::atoi(argv[18]));
    if (rc != UPDC_OK) {
        std::fprintf(stderr, "UPDC_ALL: %s\n", updc_last_error());
        return EXIT_FAILURE;
    }
    return 0;
}
