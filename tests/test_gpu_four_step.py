"""Four-step transforms: FFT rows of more than M = 4096 points and columns of
more than 8192, which one workgroup cannot hold, run as two passes over a
scratch buffer (L = L1 x L2 per line).  The reference pads any target by the
PSF (updc_core.c:1330-1339), so these shapes are reachable by any wide or tall
target.

- the transforms alone against numpy's FFT at long rows (8400 x 112), long
  columns (112 x 8400) and both (8400 x 8400), FP32 and FP64;
- whole iterations at the long shapes against the FP64 oracle, each tier
  within its usual bound, the plain (per-exposure) form equal to the batched
  form bit for bit, a graph replay equal to eager launches bit for bit;
- targets wider / taller than 8192 - PSF + 1, whose automatic FFT shape
  takes the four-step rows / columns, against the FP64 oracle.
"""
import numpy as np
import pytest

from conftest import oracle_dataset, oracle_exposures

pytestmark = pytest.mark.gpu


def _ctx(ds, precision, fft=None, graph=False):
    from updc_amd.solver import Context, Exposure
    c = Context(ds.W, ds.H, precision, fft_shape=fft, graph=graph)
    for k in range(len(ds.geoms)):
        c.add_exposure(Exposure(ds.sci[k], ds.wht[k], ds.psfs[k], ds.background[k], ds.pedestal[k],
                                linkmap=ds.linkmaps[k]))
    return c


@pytest.mark.parametrize("shape", [(8400, 112), (112, 8400), (8400, 8400)])
def test_four_step_fft_against_numpy(small_ds, shape):
    print(f"case: four-step FFT {shape} vs numpy", flush=True)
    for prec, tol in (("f32", 5e-6), ("f64", 1e-13)):
        c = _ctx(small_ds, prec, fft=shape)
        c.prepare()
        FW, FH = shape
        rng = np.random.default_rng(17)
        img = rng.normal(size=(FH, FW))
        ref = np.fft.rfft2(img)
        spec = c.fft_r2c(img)
        err = np.max(np.abs(spec - ref)) / np.max(np.abs(ref))
        print(f"  {prec}: r2c {err:.2e}", flush=True)
        assert err < tol, (prec, err)
        del spec
        back = c.fft_c2r(ref)
        err2 = np.max(np.abs(back - img * FW * FH)) / (FW * FH * np.max(np.abs(img)))
        print(f"  {prec}: c2r {err2:.2e}", flush=True)
        assert err2 < tol, (prec, err2)
        c.close()


@pytest.mark.parametrize("shape", [(8400, 112), (112, 8400)])
def test_four_step_iterations_against_oracle(small_ds, shape):
    from oracle import oracle as O
    from updc_amd._lib import FLAG_PER_EXPOSURE, check
    print(f"case: iterations at FFT {shape} (four-step)", flush=True)
    ref = O.RefSolver(small_ds.W, small_ds.H, oracle_exposures(small_ds), "f64", fft_shape=shape)
    ref.iterate(3)
    xs = {}
    for prec, tol in (("f64", 1e-11), ("mixed", 1e-4), ("f32", 1e-4)):
        c = _ctx(small_ds, prec, fft=shape)
        c.prepare()
        c.iterate(3)
        xs[prec] = c.estimate(f64=True)
        st = O.compare_stats(ref.x, xs[prec])
        print(f"  {prec}: norm_linf {st['norm_linf']:.3e}", flush=True)
        assert st["norm_linf"] < tol, (prec, st)
        c.close()
    p = _ctx(small_ds, "mixed", fft=shape)
    check(p.lib.updc_ctx_set_flags(p.h, p.flags | FLAG_PER_EXPOSURE))
    p.prepare()
    p.iterate(3)
    assert np.array_equal(p.estimate(f64=True), xs["mixed"])
    p.close()
    g = _ctx(small_ds, "f32", fft=shape, graph=True)
    g.prepare()
    g.iterate(2)
    g.iterate(1)
    assert np.array_equal(g.estimate(f64=True), xs["f32"])
    g.close()


def _strip(W, H, det_w, det_h, seed):
    """A strip target observed by one detector at twice the pixel scale."""
    from updc_amd import synth
    g = np.random.Generator(np.random.PCG64(seed))
    truth = synth.truth_image(W, H, 40, 0.5, g, sigma_range=(0.6, 2.0))
    geoms = [synth.Geometry(det_w, det_h, 2.0, 0.2, (0.3, 0.2), None, ((W - 1) / 2.0, (H - 1) / 2.0)),
             synth.Geometry(det_w, det_h, 2.0, -0.3, (0.7, 0.6), None, ((W - 1) / 2.0, (H - 1) / 2.0))]
    psfs = [synth.gaussian_psf(9, 2.0), synth.gaussian_psf(9, 2.2, 0.1, -0.05)]
    return oracle_dataset(synth.Dataset(W, H, geoms, psfs, truth, [0.0, 0.0], 2.0, 0.01, None, 5, "strip"))


@pytest.mark.parametrize("case", ["wide", "tall"])
def test_target_beyond_one_workgroup_lines(case):
    from oracle import oracle as O
    print(f"case: {case} target, automatic four-step FFT shape", flush=True)
    ds = _strip(8200, 24, 4100, 12, 31) if case == "wide" else _strip(24, 8300, 12, 4150, 32)
    ref = O.RefSolver(ds.W, ds.H, oracle_exposures(ds), "f64")
    ref.iterate(3)
    for prec, tol in (("f64", 1e-11), ("f32", 1e-4)):
        c = _ctx(ds, prec)
        c.prepare()
        s = c.scalars()
        FW, FH = s["fft_width"], s["fft_height"]
        assert (FW, FH) == (ref.FW, ref.FH)
        assert (FW // 2 > 4096) if case == "wide" else (FH > 8192), (FW, FH)
        c.iterate(3)
        st = O.compare_stats(ref.x, c.estimate(f64=True))
        print(f"  {prec} FFT {FW}x{FH}: norm_linf {st['norm_linf']:.3e}", flush=True)
        assert st["norm_linf"] < tol, (prec, st)
        c.close()
