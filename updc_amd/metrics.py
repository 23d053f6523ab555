"""Image comparison metrics and the algorithmic byte model of one iteration.

stats() follows the reference's tools/compare_fits.py:10-29 (max_abs, rms,
max_rel_masked with |a| > 1e-8 max|a|, sums, negative and non-finite counts)
and adds the normalised L-infinity max|d| / max|a|.
"""
from __future__ import annotations

import numpy as np


def stats(ref: np.ndarray, new: np.ndarray) -> dict:
    a = np.asarray(ref, np.float64)
    b = np.asarray(new, np.float64)
    d = b - a
    scale = max(float(np.nanmax(np.abs(a))), 1e-30)
    mask = np.abs(a) > 1e-8 * scale
    rel = np.abs(d[mask] / a[mask]) if np.any(mask) else np.array([0.0])
    return dict(max_abs=float(np.nanmax(np.abs(d))), rms=float(np.sqrt(np.nanmean(d * d))),
                max_rel_masked=float(np.nanmax(rel)), ref_sum=float(np.nansum(a)), new_sum=float(np.nansum(b)),
                ref_negative=int(np.sum(a < 0)), new_negative=int(np.sum(b < 0)),
                nonfinite_ref=int(np.sum(~np.isfinite(a))), nonfinite_new=int(np.sum(~np.isfinite(b))),
                norm_linf=float(np.nanmax(np.abs(d)) / scale),
                rel_rms=float(np.sqrt(np.nanmean(d * d)) / max(np.sqrt(np.nanmean(a * a)), 1e-30)))


def stage_bytes(W: int, H: int, FW: int, FH: int, nin: list[int], nlinks: list[int], ncsc: list[int],
                lm_bytes: list[int], e: int, c: int, e_d: int, e_a: int | None = None,
                fused_inverse: bool = False, adj_csc_bytes: list[int] | None = None) -> dict:
    """Algorithmic HBM bytes per launch of each iteration kernel: every
    operand read once and every result written once (FFT-internal LDS passes
    and re-reads excluded).  e/c = real/complex element bytes of the FFT
    type, e_d = bytes of the denominator element.  e_a: real element bytes of
    a summed adjoint spectrum wider than the FFT type (the mixed tier), which
    leaves the column pass whole (FH rows) and is inverted in that type by
    col_inv_corr and row_c2r_corr; with fused_inverse the column inverse runs
    inside col_adj_accum, which then writes the H inverted rows.
    adj_csc_bytes: the transposed maps the adjoint-image gather reads (the
    compact copies where the exposures have them); default the 8-byte form."""
    N = len(nin)
    NC = FW // 2 + 1
    nout = W * H
    spec_x = H * NC * c
    psf = N * FH * NC * c
    spec_k = N * H * NC * c
    pix = sum(nin)
    lm = sum(lm_bytes)
    csc = sum(4 * (nout + 1) + 8 * n for n in ncsc)
    out = {
        "row_r2c_x": nout * e + spec_x,
        "col_fwd_bcast": spec_x + psf + spec_k,
        "row_c2r_blur": spec_k + N * nout * e,
        # LinkMap + blur + sci + w (e) + frozen r/w (1+1) + wr (e)
        "gather_freeze": lm + N * nout * e + pix * (3 * e + 2),
        # CSC + wr + spectrum rows
        "row_r2c_gather": csc + pix * e + spec_k,
        # split form: CSC + wr in, N adjoint images out; then those images in, spectra out
        "adjoint_image": (csc if adj_csc_bytes is None else sum(adj_csc_bytes)) + pix * e + N * nout * e,
        "row_r2c_adj": N * nout * e + spec_k,
        "col_adj_accum": spec_k + psf + spec_x,
        "row_c2r_corr": spec_x + nout * e,
        "update": nout * (3 * e + e_d),
        # the final row C2R with the update applied: x in and out, D, no C
        "row_c2r_update": spec_x + nout * (2 * e + e_d),
    }
    if e_a is not None and e_a > e:
        acc_full = FH * NC * 2 * e_a
        acc_h = H * NC * 2 * e_a
        out.update({
            "col_adj_accum": spec_k + psf + (acc_h if fused_inverse else acc_full),
            "col_inv_corr": acc_full + acc_h,
            "row_c2r_corr": acc_h + nout * e_a,
            # x read and written, C in the wide type, D
            "update": nout * (2 * e + e_a + e_d),
            "row_c2r_update": acc_h + nout * (2 * e + e_d),
        })
    return out


def iteration_bytes(W: int, H: int, FW: int, FH: int, nin: list[int], lm_bytes: list[int], e: int = 4,
                    c: int = 8) -> int:
    """Algorithmic HBM bytes of one whole iteration, stage by stage, each
    stage reading its operands once and writing its results once (FFT-internal
    passes excluded), with the real LinkMap sizes:

      S1 X^ = R2C(x)                       nout e + nc c
      S2 blur_k = crop C2R(X^ P^_k)        N (2 nc c + nout e)
      S3 detector (LinkMap, sci, w,        sum_k (LM_k + 10 nin_k) + N 2 nout e
         frozen r/w; blur in, scatter out)
      S4 sum_k conj(P^_k) R2C(scatter_k)   N (nout e + nc c) + nc c
      S5 C = crop C2R(sum)                 nc c + nout e
      S6 update (x, C, D in; x out)        4 nout e

    nc = FH (FW/2 + 1) complex elements; e/c the real/complex element bytes
    (4/8 for float32 storage)."""
    N = len(nin)
    nout = W * H
    nc = FH * (FW // 2 + 1)
    s1 = nout * e + nc * c
    s2 = N * (2 * nc * c + nout * e)
    s3 = sum(lm + 10 * n for lm, n in zip(lm_bytes, nin)) + N * 2 * nout * e
    s4 = N * (nout * e + nc * c) + nc * c
    s5 = nc * c + nout * e
    s6 = 4 * nout * e
    return s1 + s2 + s3 + s4 + s5 + s6
