"""The drop-in boundary: the engine's rules (CPU) and the drop-in executable
end to end (GPU).

CPU: manifest line format (updc/_engine.py:1300-1310), the 18-argument order
(_engine.py:1381-1389), the psf_transposed rule (:1384), the output-iteration
rule (:1409-1415), the FreezeMu CSV row check (:1392-1407) and
UPDC_PERSIST_CACHE (:1369-1379).

GPU: updc_hip_core --self-test; a 2-iteration solve through files, read back
with the engine's rules and compared bit for bit with the in-memory context;
FreezeMu forcing by a background of -100 with the reference's own invariants
(tests/test_smoke.py:117-136); the error path (exit 1, "UPDC_ALL:" on stderr);
the persistent LinkMap and autotune caches.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from updc_amd import engine as E


def test_manifest_line_format(tmp_path):
    r = E.ManifestRow(Path("/d/sci_1.bin"), Path("/d/wht_1.bin"), Path("/d/psf_1.bin"), 9, 7,
                      Path("/d/vertices_1.bin"), 0.1, 50.0)
    assert r.line() == "/d/sci_1.bin\t/d/wht_1.bin\t/d/psf_1.bin\t9\t7\t/d/vertices_1.bin\t" \
                       "0.10000000000000001\t50\n"
    # %.17g round-trips every double
    for v in (1e-300, -100.0, 2.0 / 3.0, 51.327478872070316, np.nextafter(1.0, 2.0)):
        r.background = v
        assert float(r.line().split("\t")[6]) == v
    p = E.write_manifest(tmp_path / "manifest.txt", [r, r])
    assert p.read_text().count("\n") == 2


def test_write_exposure_files(tmp_path):
    sci = np.arange(12, dtype=np.float32).reshape(3, 4)
    v = np.zeros((4, 5, 2))
    row = E.write_exposure(tmp_path, 1, sci, sci + 1, np.ones((7, 9), np.float32), v, -1.5, 2.25)
    assert row.psf_width == 9 and row.psf_height == 7
    assert np.array_equal(np.fromfile(row.science_raw, "<f4").reshape(3, 4), sci)
    assert row.vertices.stat().st_size == 4 * 5 * 2 * 8
    with pytest.raises(ValueError):
        E.write_exposure(tmp_path, 2, sci, sci, np.ones((3, 3)), np.zeros((3, 4, 2)), 0, 0)


def test_core_argv_order():
    a = E.core_argv("core", "m.txt", 96, 80, 48, 40, 2, True, 20, "it", "w.bin", 4, 2, 3, "measure", False, True,
                    "wis", 5)
    assert a == ["core", "m.txt", "96", "80", "48", "40", "2", "1", "20", "it", "w.bin", "4", "2", "3", "1", "0",
                 "1", "wis", "5"]
    assert len(a) == 19


def test_psf_transposed_rule():
    assert E.psf_transposed("target", "stored-transposed")
    assert not E.psf_transposed("target", "native")
    assert not E.psf_transposed("detector", "stored-transposed")


@pytest.mark.parametrize("iters,every,expect", [(2, 1, [1, 2]), (40, 0, [40]), (10, 3, [3, 6, 9, 10]),
                                                (9, 3, [3, 6, 9]), (2, 5, [2])])
def test_output_iterations(iters, every, expect):
    assert E.output_iterations(iters, every) == expect


def test_freeze_csv_row_check(tmp_path):
    hdr = "iteration,exposure,n_new_frozen,n_total_frozen,min_raw_mu,mu_marker,frozen_ratio\n"
    p = tmp_path / "f.csv"
    p.write_text(hdr + "1,1,0,0,1.5,1e-12,1\n1,2,3,3,-0.25,1e-12,1\n2,1,0,0,1,1e-12,1\n")
    with pytest.raises(ValueError):
        E.read_freeze_csv(p, 2, 2)
    p.write_text(hdr + "1,1,0,0,1.5,1e-12,1\n1,2,3,3,-0.25,1e-12,1\n2,1,0,0,1,1e-12,1\n2,2,1,4,-1,1e-12,1\n")
    rows = E.read_freeze_csv(p, 2, 2)
    assert rows[1]["n_new_frozen"] == 3 and rows[3]["n_total_frozen"] == 4 and rows[1]["min_raw_mu"] == -0.25


def test_solver_env():
    assert E.solver_env(True, {})["UPDC_PERSIST_CACHE"] == "1"
    assert E.solver_env(False, {"X": "y"}) == {"X": "y", "UPDC_PERSIST_CACHE": "0"}


# ----------------------------------------------------------------------- GPU
def _files(ds, work):
    rows = [E.write_exposure(work, k + 1, ds.sci[k], ds.wht[k], ds.psfs[k], ds.vertices(k), ds.background[k],
                             ds.pedestal[k]) for k in range(len(ds.geoms))]
    return E.write_manifest(work / "manifest.txt", rows), rows


def _argv(ds, work, manifest, iterations, save_every, prefix="iter", autotune=False):
    from updc_amd._lib import EXE_PATH
    g = ds.geoms[0]
    return E.core_argv(EXE_PATH, manifest, ds.W, ds.H, g.iw, g.ih, len(ds.geoms), False, iterations,
                       work / prefix, work / f"{prefix}_wht.bin", fft_autotune=autotune, save_every=save_every)


@pytest.mark.gpu
def test_dropin_executable_end_to_end(small_ds, tmp_path):
    from updc_amd._lib import EXE_PATH
    from updc_amd.solver import Context, Exposure
    ds = small_ds
    print("case: drop-in self-test", flush=True)
    st = subprocess.run([str(EXE_PATH), "--self-test"], capture_output=True, text=True, timeout=60)
    assert st.returncode == 0 and "UPDC C core self-test: OK" in st.stdout

    print("case: drop-in 2 iterations, save_every 1", flush=True)
    manifest, rows = _files(ds, tmp_path)
    env = E.solver_env(False)
    env["UPDC_HIP_QUIET"] = "1"
    res = E.run_core(_argv(ds, tmp_path, manifest, 2, 1), env=env, timeout=120)
    assert sorted(res.iterations) == [1, 2]
    assert len(res.freeze_rows) == 2 * len(ds.geoms)
    assert np.isfinite(res.iterations[2]).all() and np.isfinite(res.wht).all()
    # the reference's timing rows (updc_core.c:1487-1508), every value seconds
    setup = ["fft_autotune", "fft_plan", "psf_cache", "polygon_links", "persistent_mmap", "mmap_prewarm",
             "data_cache", "weight_scale", "initial_drizzle", "sensitivity"]
    per_it = ["estimate_fft", "fused_exposure_wall", "forward_ifft_critical", "fused_detector_critical",
              "scatter_fft_critical", "adjoint_reduce_ifft", "update", "output", "total"]
    names = list(res.timing)
    assert names[0] == "fft_autotune" and names[-1] == "total_solver"
    assert all(n in res.timing for n in setup)
    for i in (1, 2):
        rows_i = [f"iteration_{i:02d}_{n}" for n in per_it]
        assert [n for n in names if n.startswith(f"iteration_{i:02d}_")] == rows_i
        # eager launches (the drop-in's default): every stage measured
        t = {n: res.timing[f"iteration_{i:02d}_{n}"] for n in per_it}
        assert all(np.isfinite(v) and v >= 0 for v in t.values()), t
        assert t["fused_exposure_wall"] == pytest.approx(
            t["forward_ifft_critical"] + t["fused_detector_critical"] + t["scatter_fft_critical"], abs=2e-6)
        assert t["estimate_fft"] > 0 and t["fused_detector_critical"] > 0 and t["total"] >= t["output"]
    assert res.timing["weight_scale"] > 0 and res.timing["sensitivity"] > 0 and res.timing["polygon_links"] > 0

    print("case: drop-in, hipGraph replay, stdout lines", flush=True)
    env_g = E.solver_env(False)
    env_g["UPDC_HIP_GRAPH"] = "1"
    argv = _argv(ds, tmp_path, manifest, 2, 0, prefix="graph")
    out = subprocess.run([str(a) for a in argv], capture_output=True, text=True, timeout=120, env=env_g)
    assert out.returncode == 0, out.stderr
    lines = [l for l in out.stdout.splitlines() if l.startswith("UPDC iteration") and "/2:" in l]
    assert len(lines) == 2 and lines[1].endswith(" SAVE") and not lines[0].endswith(" SAVE")
    assert "[estimateFFT nan fusedWall nan" in lines[0] and "output " in lines[0] and ", wall " in lines[0]
    gres = E.run_core(argv, env=env_g, timeout=120)
    assert np.isnan(gres.timing["iteration_01_estimate_fft"]) and gres.timing["iteration_01_total"] > 0
    assert np.array_equal(gres.iterations[2], res.iterations[2])

    # the same solve in memory (F32 tier, LinkMaps built on the GPU from the vertices)
    c = Context(ds.W, ds.H, "f32", graph=True)
    for k in range(len(ds.geoms)):
        c.add_exposure(Exposure(ds.sci[k], ds.wht[k], ds.psfs[k], ds.background[k], ds.pedestal[k],
                                vertices=ds.vertices(k)))
    c.prepare()
    fr, _ = c.iterate(2)
    assert np.array_equal(c.estimate(), res.iterations[2])
    assert np.array_equal(c.wht(), res.wht)
    assert [(r[2], r[3]) for r in fr] == [(r["n_new_frozen"], r["n_total_frozen"]) for r in res.freeze_rows]
    c.close()
    # LinkMaps were written next to the vertices, in the ULM2 format
    assert all(Path(str(r.vertices) + ".links").exists() for r in rows)

    print("case: FITS products of the core's outputs", flush=True)
    from updc_amd import fitsout as F
    n = len(ds.geoms)
    metrics = [dict(source=f"exp{k}.fits", background=float(ds.background[k]), background_sigma=0.0,
                    poisson_pedestal=float(ds.pedestal[k]), auto_positivity_pedestal=0.0,
                    nonpositive_valid_before_pedestal=0, sky_var_poisson=0.0) for k in range(n)]
    psf = [dict(anchor_x=0.0, anchor_y=0.0, oversample=1, center_target_x=0.0, center_target_y=0.0, extent_x=0.0,
                extent_y=0.0, input_width=p.shape[1], input_height=p.shape[0], mapped_width=p.shape[1],
                mapped_height=p.shape[0], output_width=p.shape[1], output_height=p.shape[0],
                input_sum=float(p.sum()), output_sum=float(p.sum()), crop_retained_flux=1.0,
                crop_discarded_flux=0.0, flux_rel_error=0.0, overlaps=0, method="NONE", wcs_used="")
           for p in ds.psfs]
    files = E.write_iteration_products(res, tmp_path / "products", "run", [("BUNIT", "MJy/sr", "")],
                                       [("NCOMBINE", n, "")], metrics, "ivm", psf,
                                       [f"psf{k}.raw" for k in range(n)])
    assert [f.name for f in files] == ["run_iter01_i2d.fits", "run_iter02_i2d.fits"]
    raw = files[1].read_bytes()
    hd = F.read_hdus(raw)
    assert np.array_equal(F.image_data(raw, hd[1]), res.iterations[2])
    assert hd[1].header["MUFRTOT"] == sum(r["n_total_frozen"] for r in res.freeze_rows if r["iteration"] == 2)


@pytest.mark.gpu
def test_dropin_freeze_forcing(small_ds, tmp_path):
    """Background -100 forces non-positive forward means (test_smoke.py:117-136)."""
    ds = small_ds
    print("case: drop-in FreezeMu forcing", flush=True)
    manifest, _ = _files(ds, tmp_path)
    lines = []
    for line in manifest.read_text().splitlines():
        f = line.split("\t")
        f[6] = "-100"
        lines.append("\t".join(f))
    manifest.write_text("\n".join(lines) + "\n")
    res = E.run_core(_argv(ds, tmp_path, manifest, 2, 1, prefix="forced"), env=E.solver_env(False), timeout=120)
    rows = res.freeze_rows
    assert any(r["n_new_frozen"] > 0 for r in rows)
    for k in range(1, len(ds.geoms) + 1):
        tot = [r["n_total_frozen"] for r in rows if r["exposure"] == k]
        new = [r["n_new_frozen"] for r in rows if r["exposure"] == k]
        assert tot[0] == new[0]
        assert tot[1] == tot[0] + new[1]


@pytest.mark.gpu
def test_dropin_error_path(small_ds, tmp_path):
    ds = small_ds
    print("case: drop-in error path", flush=True)
    manifest, rows = _files(ds, tmp_path)
    rows[0].science_raw = tmp_path / "missing.bin"
    E.write_manifest(manifest, rows)
    with pytest.raises(subprocess.CalledProcessError) as ei:
        E.run_core(_argv(ds, tmp_path, manifest, 1, 0), env=E.solver_env(False), timeout=120)
    assert ei.value.returncode == 1
    assert "UPDC_ALL:" in ei.value.stderr


@pytest.mark.gpu
def test_dropin_persistent_caches(small_ds, tmp_path):
    ds = small_ds
    print("case: drop-in persistent caches", flush=True)
    manifest, _ = _files(ds, tmp_path)
    env = E.solver_env(True)
    a = E.run_core(_argv(ds, tmp_path, manifest, 2, 0, prefix="a", autotune=True), env=env, timeout=120)
    assert list(tmp_path.glob("fft_autotune_gpu_v2_*.txt"))
    out = subprocess.run(_argv(ds, tmp_path, manifest, 2, 0, prefix="b", autotune=True), env=env,
                         capture_output=True, text=True, timeout=120, check=True).stdout
    assert "persistent cache hit" in out and "FFT autotune persistent-cache hit" in out
    b = E.run_core(_argv(ds, tmp_path, manifest, 2, 0, prefix="b", autotune=True), env=env, timeout=120)
    assert np.array_equal(a.iterations[2], b.iterations[2])


@pytest.mark.gpu
@pytest.mark.parametrize("defect,message", [("short", "short float32 input"),
                                            ("trailing", "float32 input has unexpected trailing bytes")])
def test_dropin_raw_plane_size_errors(small_ds, tmp_path, defect, message):
    """Raw float32 planes must hold exactly iw*ih values (updc_core.c:298-315):
    the label and path go to stderr, then UPDC_ALL: <the reference's message>,
    exit 1."""
    ds = small_ds
    print(f"case: drop-in raw plane {defect}", flush=True)
    manifest, rows = _files(ds, tmp_path)
    raw = rows[1].weight_raw.read_bytes()
    rows[1].weight_raw.write_bytes(raw[:-4] if defect == "short" else raw + b"\0")
    with pytest.raises(subprocess.CalledProcessError) as ei:
        E.run_core(_argv(ds, tmp_path, manifest, 1, 0), env=E.solver_env(False), timeout=120)
    assert ei.value.returncode == 1
    assert f"UPDC_ALL: {message}" in ei.value.stderr
    assert f"WHT raw: {rows[1].weight_raw}" in ei.value.stderr


@pytest.mark.gpu
def test_dropin_vertex_map_trailing_bytes_accepted(small_ds, tmp_path):
    """The reference reads the vertex map with one fread of (ih+1)(iw+1)*2
    doubles and no trailing-byte check (updc_core.c:536-539): extra bytes
    after the lattice are ignored, and the solve equals the clean one."""
    ds = small_ds
    print("case: drop-in vertex map with trailing bytes", flush=True)
    manifest, rows = _files(ds, tmp_path)
    clean = E.run_core(_argv(ds, tmp_path, manifest, 2, 0, prefix="clean"), env=E.solver_env(False), timeout=120)
    for r in rows:
        r.vertices.write_bytes(r.vertices.read_bytes() + b"\0" * 24)
        links = Path(str(r.vertices) + ".links")
        if links.exists():
            links.unlink()
    res = E.run_core(_argv(ds, tmp_path, manifest, 2, 0, prefix="extra"), env=E.solver_env(False), timeout=120)
    assert np.array_equal(res.iterations[2], clean.iterations[2])


@pytest.mark.gpu
@pytest.mark.parametrize("defect,message", [
    ("fewer", "manifest has fewer rows than expected"),
    ("more", "manifest has more rows than expected"),
    ("fields", "manifest requires science, weight, psf, psf_width, psf_height, vertices, background, pedestal"),
    ("psf_width", "invalid PSF dimensions in manifest"),
    ("pedestal", "invalid background/pedestal in manifest")])
def test_dropin_manifest_errors(small_ds, tmp_path, defect, message):
    """The manifest rules of updc_core.c:426-461: exactly N rows (any further
    line, even a blank one, is an extra row), eight tab-separated fields, PSF
    sizes >= 1, finite background, finite non-negative pedestal."""
    ds = small_ds
    print(f"case: drop-in manifest {defect}", flush=True)
    manifest, _ = _files(ds, tmp_path)
    lines = manifest.read_text().splitlines()
    if defect == "fewer":
        lines = lines[:-1]
    elif defect == "more":
        lines = lines + [""]
    elif defect == "fields":
        lines[0] = "\t".join(lines[0].split("\t")[:7])
    elif defect == "psf_width":
        f = lines[0].split("\t")
        f[3] = "0"
        lines[0] = "\t".join(f)
    else:
        f = lines[1].split("\t")
        f[7] = "-1"
        lines[1] = "\t".join(f)
    manifest.write_text("\n".join(lines) + "\n")
    with pytest.raises(subprocess.CalledProcessError) as ei:
        E.run_core(_argv(ds, tmp_path, manifest, 1, 0), env=E.solver_env(False), timeout=120)
    assert ei.value.returncode == 1
    assert f"UPDC_ALL: {message}" in ei.value.stderr
