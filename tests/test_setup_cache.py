"""Persistent set-up caches (the reference's psf_spectrum_*.c64,
initial_estimate.f32 / initial_coverage.f32 and denominator_*.f32,
updc/c_core/updc_core.c:145-158, 790-843, 905-976, 1024-1083)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ctx(ds, prec, cache):
    from updc_amd.solver import Context, Exposure
    c = Context(ds.W, ds.H, prec, setup_cache_dir=str(cache))
    for k in range(len(ds.geoms)):
        c.add_exposure(Exposure(ds.sci[k], ds.wht[k], ds.psfs[k], ds.background[k], ds.pedestal[k],
                                linkmap=ds.linkmaps[k]))
    return c


def test_setup_caches_write_then_read(small_ds, tmp_path):
    ds = small_ds
    print("case: set-up caches, write then read", flush=True)
    a = _ctx(ds, "f32", tmp_path)
    a.prepare()
    assert a.setup_cache_hits() == dict(psf_spectra=False, initial=False, denominator=False)
    sc = a.scalars()
    FW, FH = sc["fft_width"], sc["fft_height"]
    nout = ds.W * ds.H
    files = {f"psf_spectrum_{k + 1:02d}_{FW}x{FH}_t0.c64": FH * (FW // 2 + 1) * 8 for k in range(len(ds.geoms))}
    files.update({"initial_estimate.f32": nout * 4, "initial_coverage.f32": nout * 4,
                  f"denominator_{FW}x{FH}.f32": nout * 4})
    for name, size in files.items():
        assert (tmp_path / name).stat().st_size == size, name
    # the PSF spectrum file is FFTW's R2C layout: numpy's rfft2 of the wrapped, normalised PSF
    for k in range(len(ds.geoms)):
        spec = np.fromfile(tmp_path / f"psf_spectrum_{k + 1:02d}_{FW}x{FH}_t0.c64", np.complex64).reshape(FH, -1)
        p = ds.psfs[k].astype(np.float64)
        img = np.zeros((FH, FW))
        h, w = p.shape
        img[np.ix_((np.arange(h) - h // 2) % FH, (np.arange(w) - w // 2) % FW)] = p / p.sum()
        ref = np.fft.rfft2(img)
        assert np.max(np.abs(spec - ref)) <= 2e-6 * np.max(np.abs(ref))
    a.iterate(3)
    xa, D0 = a.estimate(f64=True), a.denominator()
    x0a, cova = a.initial()
    a.close()

    b = _ctx(ds, "f32", tmp_path)
    b.prepare()
    assert b.setup_cache_hits() == dict(psf_spectra=True, initial=True, denominator=True)
    x0b, covb = b.initial()
    assert np.array_equal(x0a, x0b) and np.array_equal(cova, covb)
    assert np.array_equal(b.denominator(), D0)
    b.iterate(3)
    assert np.array_equal(b.estimate(f64=True), xa)
    b.close()


def test_setup_cache_wrong_size_is_recomputed(small_ds, tmp_path):
    ds = small_ds
    print("case: set-up caches, truncated file", flush=True)
    a = _ctx(ds, "f32", tmp_path)
    a.prepare()
    a.close()
    sc_name = next(tmp_path.glob("denominator_*.f32"))
    size = sc_name.stat().st_size
    sc_name.write_bytes(sc_name.read_bytes()[:-4])
    b = _ctx(ds, "f32", tmp_path)
    b.prepare()
    assert b.setup_cache_hits() == dict(psf_spectra=True, initial=True, denominator=False)
    assert sc_name.stat().st_size == size
    b.close()


@pytest.mark.parametrize("prec", ["mixed", "f64"])
def test_setup_caches_only_for_the_f32_tier(small_ds, tmp_path, prec):
    print(f"case: set-up caches ignored by {prec}", flush=True)
    c = _ctx(small_ds, prec, tmp_path)
    c.prepare()
    assert not any(tmp_path.iterdir())
    c.close()
