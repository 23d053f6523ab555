"""Pins the oracle's PSF projector restatement (psf_projector.c:71-146).

Against the reference's own psf_projector.c, compiled into oracle/_ref from
the sources under the reference tree (skipped where that build is absent):
bit-identical projected PSFs, shapes, sums and overlap counts through the
reference's manifest / results CSV.  Closed forms: the identity mapping
returns the input; the flux of any mapping is conserved to the reference's
5e-5 check.
"""
import csv
import subprocess

import numpy as np
import pytest

from oracle import oracle as O


def psf_geometry(iw, ih, scale=1.0, theta=0.0, shift=(0.0, 0.0), seed=0, jitter=0.0):
    """Corners of a PSF's pixels in target-pixel offsets from its centre."""
    g = np.random.Generator(np.random.PCG64(seed))
    u = np.arange(iw + 1) - 0.5 - (iw - 1) / 2.0
    w = np.arange(ih + 1) - 0.5 - (ih - 1) / 2.0
    uu, ww = np.meshgrid(u, w)
    c, s = np.cos(theta), np.sin(theta)
    x = scale * (c * uu - s * ww) + shift[0]
    y = scale * (s * uu + c * ww) + shift[1]
    if jitter:
        x = x + g.normal(0, jitter, x.shape)
        y = y + g.normal(0, jitter, y.shape)
    return np.ascontiguousarray(np.stack([x, y], -1))


def seq_sum(a):
    """Sequential double sum in raster order, as psf_projector.c:63-70."""
    t = 0.0
    for x in np.asarray(a, np.float32).ravel():
        t += float(x)
    return t


def psf_image(iw, ih, seed=0):
    g = np.random.Generator(np.random.PCG64(seed))
    yy, xx = np.mgrid[0:ih, 0:iw]
    p = np.exp(-((xx - (iw - 1) / 2) ** 2 + (yy - (ih - 1) / 2) ** 2) / (2 * 1.7 ** 2))
    p = p + 0.01 * g.uniform(0, 1, p.shape)
    p[g.uniform(0, 1, p.shape) < 0.05] = 0.0
    return (p / p.sum()).astype(np.float32)


CASES = [(9, 7, 1.0, 0.0, (0, 0), 0.0), (7, 9, 2.0, 0.3, (0.2, -0.1), 0.0), (15, 15, 1.5, 1.1, (0.4, 0.3), 0.01),
         (31, 25, 0.7, -0.4, (0.0, 0.5), 0.02), (11, 11, 3.0, 0.05, (-0.5, 0.5), 0.0)]


@pytest.mark.skipif(not O.REF_PROJECTOR.exists(), reason="oracle/_ref projector (reference source) not built here")
def test_oracle_projector_bitexact_vs_reference(tmp_path):
    lines = []
    exp = []
    for i, (iw, ih, sc, th, sh, jit) in enumerate(CASES):
        p = psf_image(iw, ih, seed=i)
        v = psf_geometry(iw, ih, sc, th, sh, seed=i, jitter=jit)
        p.tofile(tmp_path / f"in{i}.f32")
        v.tofile(tmp_path / f"v{i}.f64")
        lines.append(f"{tmp_path / f'in{i}.f32'}\t{iw}\t{ih}\t{tmp_path / f'v{i}.f64'}\t{tmp_path / f'out{i}.f32'}\tpsf{i}\n")
        exp.append(O.psf_project(p, v))
    (tmp_path / "m.txt").write_text("".join(lines))
    subprocess.run([str(O.REF_PROJECTOR), str(tmp_path / "m.txt"), "2", str(tmp_path / "r.csv")], check=True,
                   capture_output=True, timeout=60)
    rows = list(csv.DictReader(open(tmp_path / "r.csv")))
    for i, (out, n) in enumerate(exp):
        r = rows[i]
        oh, ow = out.shape
        assert int(r["width"]) == ow and int(r["height"]) == oh
        ref = np.fromfile(tmp_path / f"out{i}.f32", np.float32).reshape(oh, ow)
        assert np.array_equal(ref, out), i
        assert int(r["overlaps"]) == n
        assert float(r["output_sum"]) == seq_sum(out)


def test_identity_mapping_returns_the_input():
    for iw, ih in [(9, 7), (7, 9), (5, 5)]:
        p = psf_image(iw, ih)
        out, n = O.psf_project(p, psf_geometry(iw, ih))
        assert out.shape == (ih, iw)
        assert np.array_equal(out, p)
        assert n == int((p != 0).sum())


def test_flux_conserved():
    for i, (iw, ih, sc, th, sh, jit) in enumerate(CASES):
        p = psf_image(iw, ih, seed=i)
        out, _ = O.psf_project(p, psf_geometry(iw, ih, sc, th, sh, seed=i, jitter=jit))
        s_in, s_out = float(p.astype(np.float64).sum()), float(out.astype(np.float64).sum())
        assert abs(s_out - s_in) / s_in <= 5e-5


# ----------------------------------------------------------------------- GPU
def _gpu_project(p, v):
    import ctypes as C
    from updc_amd import _lib
    lib = _lib.load()
    ih, iw = p.shape
    p = np.ascontiguousarray(p, np.float32)
    v = np.ascontiguousarray(v, np.float64)
    ow, oh = C.c_int(), C.c_int()
    rc = lib.updc_psf_project(0, p.ctypes.data_as(C.POINTER(C.c_float)), iw, ih,
                              v.ctypes.data_as(C.POINTER(C.c_double)), None, 0, C.byref(ow), C.byref(oh),
                              None, None, None)
    if rc != 0:
        return rc, None, 0.0, 0.0, 0
    out = np.empty((oh.value, ow.value), np.float32)
    si, so, n = C.c_double(), C.c_double(), C.c_uint64()
    rc = lib.updc_psf_project(0, p.ctypes.data_as(C.POINTER(C.c_float)), iw, ih,
                              v.ctypes.data_as(C.POINTER(C.c_double)), out.ctypes.data_as(C.POINTER(C.c_float)),
                              out.size, C.byref(ow), C.byref(oh), C.byref(si), C.byref(so), C.byref(n))
    return rc, out, si.value, so.value, n.value


@pytest.mark.gpu
def test_gpu_projector_bitexact_vs_oracle():
    print("case: GPU PSF projector vs oracle", flush=True)
    cases = CASES + [(255, 255, 2.0, 0.2, (0.3, -0.2), 0.005), (101, 63, 1.3, -0.7, (0.1, 0.1), 0.0)]
    for i, (iw, ih, sc, th, sh, jit) in enumerate(cases):
        p = psf_image(iw, ih, seed=i)
        v = psf_geometry(iw, ih, sc, th, sh, seed=i, jitter=jit)
        ref, n = O.psf_project(p, v)
        rc, out, si, so, ng = _gpu_project(p, v)
        assert rc == 0
        assert out.shape == ref.shape and np.array_equal(out, ref), i
        assert ng == n
        assert si == seq_sum(p) and so == seq_sum(ref)


@pytest.mark.gpu
def test_gpu_projector_errors():
    from updc_amd import _lib
    print("case: GPU PSF projector error paths", flush=True)
    p = psf_image(9, 9)
    v = psf_geometry(9, 9)
    bad = p.copy()
    bad[3, 3] = -1.0
    numeric = next(k for k, v_ in _lib.ERR_NAMES.items() if v_ == "NUMERIC")
    rc, *_ = _gpu_project(bad, v)
    assert rc == numeric
    vnan = v.copy()
    vnan[2, 2, 0] = np.nan
    rc, *_ = _gpu_project(p, vnan)
    assert rc != 0
    rc, *_ = _gpu_project(np.zeros((5, 5), np.float32), psf_geometry(5, 5))
    assert rc != 0


@pytest.mark.gpu
def test_dropin_projector_end_to_end(tmp_path):
    """updc_hip_psf_projector through the reference's manifest and results CSV;
    outputs equal the oracle's and, where oracle/_ref holds the reference's own
    projector, that program's files and CSV."""
    from updc_amd import _lib
    print("case: drop-in PSF projector", flush=True)
    lines, exp = [], []
    for i, (iw, ih, sc, th, sh, jit) in enumerate(CASES):
        p = psf_image(iw, ih, seed=i)
        v = psf_geometry(iw, ih, sc, th, sh, seed=i, jitter=jit)
        p.tofile(tmp_path / f"in{i}.f32")
        v.tofile(tmp_path / f"v{i}.f64")
        lines.append(f"{tmp_path / f'in{i}.f32'}\t{iw}\t{ih}\t{tmp_path / f'v{i}.f64'}\t{tmp_path / f'out{i}.f32'}\t{i}\n")
        exp.append(O.psf_project(p, v))
    (tmp_path / "m.txt").write_text("".join(lines))
    r = subprocess.run([str(_lib.PROJECTOR_PATH), str(tmp_path / "m.txt"), "1", str(tmp_path / "r.csv")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = list(csv.DictReader(open(tmp_path / "r.csv")))
    for i, (out, n) in enumerate(exp):
        got = np.fromfile(tmp_path / f"out{i}.f32", np.float32).reshape(out.shape)
        assert np.array_equal(got, out)
        assert int(rows[i]["overlaps"]) == n
    if O.REF_PROJECTOR.exists():
        ref_lines = "".join(l.replace("/out", "/ref") for l in lines)
        (tmp_path / "mr.txt").write_text(ref_lines)
        subprocess.run([str(O.REF_PROJECTOR), str(tmp_path / "mr.txt"), "1", str(tmp_path / "rr.csv")], check=True,
                       capture_output=True, timeout=60)
        assert (tmp_path / "rr.csv").read_text() == (tmp_path / "r.csv").read_text()
        for i in range(len(CASES)):
            assert (tmp_path / f"ref{i}.f32").read_bytes() == (tmp_path / f"out{i}.f32").read_bytes()
