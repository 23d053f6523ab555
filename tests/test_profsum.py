"""The rocprofv3 summaries (updc_amd/profsum.py) on hand-made CSV traces: stage
classification by launch order, the iteration window, and the copy-kernel
calibration of the HBM counters (CPU only)."""
import csv

import pytest

from updc_amd import profsum as P

# one iteration of the mixed tier's fused form, in launch order
FUSED = [
    "void updc::k_row_r2c_p<float, 1024, 16, 4, 4>(FftLen, ...)",
    "void updc::k_col_e<float, float, 512, 16, 5, 2, true>(FftLen, ...)",
    "void updc::k_row_c2r_p<float, 1024, 16, 4, 12>(FftLen, ...)",
    "void updc::k_gather<float, double>(ExpDev const*, ...)",
    "void updc::k_row_r2c<float, 1024, 16, 4, 1>(FftLen, ...)",
    "void updc::k_col_e<float, double, 512, 16, 5, 4, true>(FftLen, ...)",
    "void updc::k_col_e<double, double, 512, 16, 5, 1, false>(FftLen, ...)",
    "void updc::k_row_c2r_p<double, 512, 16, 2, 16>(FftLen, ...)",
    "void updc::k_update<float, double>(float*, ...)",
    "void updc::k_row_c2r_upd_p<double, float, 1024, 10, 2, 6>(FftLen, ...)",
    "updc::k_commit_diag(ExpDiag const*, ...)",
]
FUSED_STAGES = ["row_r2c_x", "col_fwd_bcast", "row_c2r_blur", "gather_freeze", "row_r2c_gather", "col_adj_accum",
                "col_inv_corr", "row_c2r_corr", "update", "row_c2r_update"]
# the split form replaces the fused gather R2C by the adjoint image and a plain R2C
SPLIT = FUSED[:4] + ["void updc::k_csc_image<float>(ExpDev const*, ...)",
                     "void updc::k_row_r2c_p<float, 1024, 16, 4, 4>(FftLen, ...)"] + FUSED[5:]
SPLIT_STAGES = FUSED_STAGES[:4] + ["adjoint_image", "row_r2c_adj"] + FUSED_STAGES[5:]
SETUP = ["updc::k_lm_count(double const*, ...)", "void updc::k_row_r2c<float, 1024, 16, 4, 0>(FftLen, ...)",
         "void updc::k_col_e<float, float, 512, 16, 5, 0, true>(FftLen, ...)"]


def _write_trace(path, names, t0=1000, dur=10):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        for i, n in enumerate(names):
            w.writerow([i + 1, n, t0 + 100 * i, t0 + 100 * i + dur * (1 + i % 3)])


def test_stage_order_fused_and_split():
    rows = [{"Dispatch_Id": str(i + 1), "Kernel_Name": n} for i, n in enumerate(SETUP + FUSED + SPLIT)]
    got = [st for st, _ in P._iteration_stages(rows, "Kernel_Name")]
    assert got == FUSED_STAGES + SPLIT_STAGES


def test_compact_adjoint_image_kernel_is_the_adjoint_stage():
    # the compact-map gather (k_csc_image16) is the same stage as the 8-byte one
    split16 = [n.replace("k_csc_image<", "k_csc_image16<") for n in SPLIT]
    rows = [{"Dispatch_Id": str(i + 1), "Kernel_Name": n} for i, n in enumerate(SETUP + split16)]
    assert [st for st, _ in P._iteration_stages(rows, "Kernel_Name")] == SPLIT_STAGES


def test_setup_kernels_before_the_first_gather_are_skipped():
    # the set-up's own plain R2C and column FFT precede the window
    rows = [{"Dispatch_Id": str(i + 1), "Kernel_Name": n} for i, n in enumerate(SETUP + FUSED)]
    first = P._iteration_stages(rows, "Kernel_Name")[0][1]
    assert first["Dispatch_Id"] == str(len(SETUP) + 1)


def test_trace_summary_counts_and_durations(tmp_path):
    d = tmp_path / "trace"
    d.mkdir()
    _write_trace(d / "run_kernel_trace.csv", SETUP + FUSED * 3)
    text = P.trace_summary(d, ["rocprofv3 --kernel-trace -- python3 bench.py --profile"])
    lines = {l.split()[0]: l.split() for l in text.splitlines() if l and l.split()[0] in P.STAGES}
    assert set(lines) == set(FUSED_STAGES)
    assert all(int(v[1]) == 3 for v in lines.values())
    assert "iterations: 3" in text


def test_traffic_calibration(tmp_path):
    copy_bytes = 1 << 30
    for counter, sub, copy_kib, per_kib in (("FETCH_SIZE", "f", copy_bytes / 2048, 1000.0),
                                            ("WRITE_SIZE", "w", copy_bytes / 1024, 500.0)):
        d = tmp_path / sub
        d.mkdir()
        with open(d / "run_counter_collection.csv", "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
            w.writerow([1, "updc::k_copy_ref(uint4 const*, uint4*, unsigned long)", counter, copy_kib])
            for i, n in enumerate(FUSED):
                w.writerow([i + 2, n, counter, per_kib])
    text = P.traffic_summary(tmp_path / "f", tmp_path / "w", copy_bytes, ["a", "b"])
    path = tmp_path / "t.txt"
    path.write_text(text)
    tr = P.read_traffic(path)
    # FETCH_SIZE reads half the copy's bytes (scale 2), WRITE_SIZE all of them (scale 1)
    assert "scale 2.0000" in text and "scale 1.0000" in text
    assert tr["update"] == dict(fetch=2048000, write=512000, total=2560000)
    assert set(tr) == set(FUSED_STAGES)


def test_traffic_needs_the_copy_kernel(tmp_path):
    for sub in ("f", "w"):
        d = tmp_path / sub
        d.mkdir()
        with open(d / "run_counter_collection.csv", "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
            w.writerow([1, FUSED[3], "FETCH_SIZE" if sub == "f" else "WRITE_SIZE", 1.0])
    with pytest.raises(SystemExit):
        P.traffic_summary(tmp_path / "f", tmp_path / "w", 1 << 30, [])
