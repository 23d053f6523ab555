// Device data structures shared by the kernels and the host orchestration.
#pragma once
#include "common.hpp"

// LinkMap v2 encodings (updc/c_core/updc_core.c:44-47).
#define UPDC_LM_MAGIC 0x324d4c55u
#define UPDC_LM_VERSION 2u
#define UPDC_LM_ABS32 1u
#define UPDC_LM_DELTA16 2u

// Per-exposure view handed to every batched kernel.  Detector-side arrays
// are the ULM2 CSR arrays exactly as stored (no re-encoding), plus the
// target-major (CSC) transpose built once in updc_prepare.
struct ExpDev {
    uint32_t nin;       // detector pixels
    uint32_t encoding;  // UPDC_LM_DELTA16 / UPDC_LM_ABS32
    const uint32_t *off;   // [nin+1]
    const uint32_t *base;  // [nin]      (DELTA16)
    const uint16_t *d16;   // [nlinks]   (DELTA16)
    const uint32_t *t32;   // [nlinks]   (ABS32)
    const float *frac;     // [nlinks]
    const uint32_t *toff;  // [nout+1]   CSC offsets
    const uint2 *tent;     // [nlinks']  CSC entries {pixel, frac bits}, ascending pixel per target
    // Compact copy of the same CSC, read by the iteration's adjoint-image
    // gather when every 64-target tile fits it (c16 != 0): per tile the
    // offset of its first entry and the smallest source pixel, per target
    // its entry count, per entry the pixel minus the tile's smallest and the
    // fraction.  6 bytes per entry and 1 per target instead of 8 and 4.
    uint32_t c16;
    const uint32_t *c_toff;  // [ntiles+1] first entry of each 64-target tile
    const uint32_t *c_pb;    // [ntiles]   smallest source pixel of the tile
    const uint8_t *c_cnt;    // [nout]     entries of each target (<= 255)
    const uint16_t *c_dp;    // [nlinks']  source pixel - c_pb[tile]
    const float *c_fr;       // [nlinks']  fraction
    size_t pix;            // start of this exposure in the flattened per-pixel arrays
    double mean_offset, pedestal, background;     // manifest values (f64 parse)
    float mean_offset_f, pedestal_f, background_f; // the reference's f32 values (updc_core.c:447-449)
};

__device__ __forceinline__ uint32_t lm_target(const ExpDev &e, uint32_t p, uint32_t z) {
    return e.encoding == UPDC_LM_DELTA16 ? e.base[p] + (uint32_t)e.d16[z] : e.t32[z];
}

// Per-(iteration, exposure) FreezeMu diagnostics written by the gather kernel
// (the mufreeze CSV row, updc_core.c:1425-1431).
struct ExpDiag {
    unsigned long long n_new;
    unsigned long long n_total;
    unsigned long long min_key;  // dkey(min raw_mu)
};
// Per-iteration diagnostics of the update kernel (updc_core.c:1438-1456) and
// error word.  err bits: 1 non-finite forward mean, 2 non-finite ratio,
// 4 non-finite correction factor, 8 substantially negative correction,
// 16 non-finite estimate.  err_key is the smallest failure key (atomicMin):
// the failure the reference's serial loops meet first, with the kind the
// reference checks first at that element.
//   detector stage: (exposure << 34) | (pixel << 2) | rank, rank 0 non-finite
//                   forward mean, 1 non-finite ratio (updc_core.c:1160-1169);
//   update:         (target << 2) | rank, rank 0 non-finite factor, 1
//                   substantially negative, 2 non-finite estimate
//                   (updc_core.c:1441-1453).
// halt is sticky within one updc_iterate call: the bookkeeping kernel sets it
// once an iteration has failed, and the detector and update kernels of later
// iterations then change nothing (no freeze, no estimate write), so the state
// stops at the failing iteration, as the reference stops there.
struct IterDiag {
    double signed_sum;
    unsigned long long n_negative;
    unsigned int err;
    unsigned int halt;
    unsigned long long err_key;
};
__host__ __device__ __forceinline__ unsigned long long err_key_detector(unsigned k, unsigned p, unsigned rank) {
    return ((unsigned long long)k << 34) | ((unsigned long long)p << 2) | rank;
}
__host__ __device__ __forceinline__ unsigned long long err_key_update(unsigned long long t, unsigned rank) {
    return (t << 2) | rank;
}
#define UPDC_E_MU 1u
#define UPDC_E_RATIO 2u
#define UPDC_E_FACTOR 4u
#define UPDC_E_NEGCORR 8u
#define UPDC_E_ESTIMATE 16u
#define UPDC_E_DRIZZLE 32u
#define UPDC_E_WEIGHT 64u
#define UPDC_E_YEFF 128u
