"""Uneven batching: exposures of different detector and PSF sizes in one
launch per stage, against the FP64 oracle and against per-exposure launches.

The reference core takes one input size for all exposures (argv iw ih,
updc_core.c:1566-1573); the context API batches differently sized exposures
(subarrays) natively, each with its own LinkMap and PSF.  The FP64 GPU form
must match the FP64 oracle to rounding, and the batched launches must equal
the per-exposure launches bit for bit in every tier.
"""
import numpy as np
import pytest

from conftest import oracle_dataset, oracle_exposures

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def uneven_ds():
    from updc_amd import synth
    g = np.random.Generator(np.random.PCG64([synth.BASE_SEED, 31]))
    target = 96
    sizes = [(48, 48), (24, 48), (20, 20)]
    centres = [(47.5, 47.5), (30.0, 50.0), (62.0, 35.0)]
    geoms = [synth.Geometry(iw, ih, 2.0, float(g.uniform(-2, 2)), (float(g.uniform(0, 3)), float(g.uniform(0, 3))),
                            None, c) for (iw, ih), c in zip(sizes, centres)]
    psfs = [synth.gaussian_psf(9, 2.0), synth.gaussian_psf(5, 1.8, 0.2, 0.0), synth.gaussian_psf(11, 2.6, 0, -0.1)]
    truth = synth.truth_image(target, target, 15, 0.5, g, sersic_fraction=0.3, sigma_range=(0.6, 2.0))
    ds = synth.Dataset(target, target, geoms, psfs, truth, [0.0, 0.3, -0.2], 2.0, 0.02, None, 5, "uneven")
    return oracle_dataset(ds)


def _ctx(ds, prec, plain=False, graph=False):
    from updc_amd.solver import Context, Exposure
    c = Context(ds.W, ds.H, prec, graph=graph)
    if plain:
        c.set_per_exposure(True)
    for k in range(len(ds.geoms)):
        c.add_exposure(Exposure(ds.sci[k], ds.wht[k], ds.psfs[k], ds.background[k], ds.pedestal[k],
                                linkmap=ds.linkmaps[k]))
    c.prepare()
    return c


def test_uneven_f64_matches_oracle(uneven_ds):
    from oracle import oracle as O
    from updc_amd import metrics
    ds = uneven_ds
    print("case: uneven sizes, f64 vs oracle f64, 5 iterations", flush=True)
    ref = O.RefSolver(ds.W, ds.H, oracle_exposures(ds), "f64")
    c = _ctx(ds, "f64")
    x0, _ = c.initial()
    assert np.max(np.abs(x0 - ref.x)) <= 1e-12 * max(1.0, np.max(np.abs(ref.x)))
    rows_ref, _ = ref.iterate(5)
    rows, _ = c.iterate(5)
    st = metrics.stats(ref.x, c.estimate(f64=True))
    print(f"  norm_linf {st['norm_linf']:.3e}", flush=True)
    assert st["norm_linf"] < 1e-11
    assert [(r[2], r[3]) for r in rows] == [(r[2], r[3]) for r in rows_ref]
    c.close()


@pytest.mark.parametrize("prec", ["f32", "mixed", "f64"])
def test_uneven_batched_equals_per_exposure(uneven_ds, prec):
    print(f"case: uneven sizes, batched vs per-exposure {prec}", flush=True)
    out = []
    for plain, graph in ((False, True), (True, False)):
        c = _ctx(uneven_ds, prec, plain=plain, graph=graph)
        rows, _ = c.iterate(4)
        out.append((c.estimate(f64=True), [(r[2], r[3]) for r in rows]))
        c.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


def test_uneven_f32_close_to_f64(uneven_ds):
    from updc_amd import metrics
    print("case: uneven sizes, mixed vs f64", flush=True)
    xs = []
    for prec in ("mixed", "f64"):
        c = _ctx(uneven_ds, prec)
        c.iterate(5)
        xs.append(c.estimate(f64=True))
        c.close()
    st = metrics.stats(xs[1], xs[0])
    print(f"  norm_linf {st['norm_linf']:.3e}", flush=True)
    assert st["norm_linf"] < 1e-4
