"""A detector finer than the target grid (scale 0.6: a target pixel overlaps up
to 9 detector pixels, a detector pixel 1-4 target pixels).

The other datasets use a detector twice as coarse as the target, where no
target has more than 4 CSC entries. Here most have more, so these cases run the
gathers' long-chain paths: entries 5 and beyond of the row gather
(`k_row_r2c`, ROW_GATHER) and of the standalone CSC gathers, and a forward
gather with few links per pixel. The chains are the reference's
(updc_core.c:1173-1176, :1064): ascending p, one fmaf per (p, t) pair, from 0.
"""
import numpy as np
import pytest

from conftest import oracle_dataset, oracle_exposures


@pytest.fixture(scope="module")
def fine_ds():
    from updc_amd import synth
    return oracle_dataset(synth.small_config(det=160, target=96, seed=21))


def _ctx(ds, precision, graph=False):
    from updc_amd.solver import Context, Exposure
    c = Context(ds.W, ds.H, precision, graph=graph)
    for k in range(len(ds.geoms)):
        c.add_exposure(Exposure(ds.sci[k], ds.wht[k], ds.psfs[k], ds.background[k], ds.pedestal[k],
                                vertices=ds.vertices(k)))
    return c


def test_fine_detector_has_long_csc_columns(fine_ds):
    """The dataset does what it is for: targets with 5-9 active source pixels."""
    from oracle import oracle as O
    for k in range(len(fine_ds.geoms)):
        p = O.parse_linkmap(fine_ds.linkmaps[k])
        keep = fine_ds.wht[k].ravel()[p["pixel"]] > 0
        cnt = np.bincount(p["target"][keep], minlength=fine_ds.W * fine_ds.H)
        assert cnt.max() == 9 and (cnt > 4).sum() > 1000
        assert np.diff(p["offset"].astype(np.int64)).max() <= 4


@pytest.mark.gpu
def test_fine_detector_linkmap_and_detector_stage(fine_ds):
    """LinkMap bytes from the vertices, then the detector stage given the same
    blur (wr, scatter, FreezeMu counters) bit-identical to the oracle."""
    from oracle import oracle as O
    from updc_amd.solver import Context
    print("case: fine detector, LinkMap + detector stage f32", flush=True)
    ref = O.RefSolver(fine_ds.W, fine_ds.H, oracle_exposures(fine_ds), "f32")
    c = _ctx(fine_ds, "f32")
    for k in range(len(fine_ds.geoms)):
        assert c.linkmap(k) == fine_ds.linkmaps[k]
    c.prepare()
    blur = ref.blur()
    for k in range(len(fine_ds.geoms)):
        e = ref.ex[k]
        bl = blur[k].copy()
        bl[: bl.shape[0] // 4] *= np.float32(-2.0)  # some pixels through FreezeMu
        frozen = np.zeros(e["s"].size, np.uint8)
        wr_o, sc_o, st_o = O.fused_detector(e["linkmap"], ref.nout, e["s"], e["w"], frozen, bl,
                                            float(e["mo_f"]), float(e["ped_f"]))
        wr_g, sc_g, st_g = c.detector(k, bl)
        assert np.array_equal(wr_g.astype(np.float32), wr_o)
        assert np.array_equal(sc_g.astype(np.float32).ravel(), sc_o)
        assert st_g[0] == st_o[0] > 0 and st_g[1] == st_o[1]
        assert np.float32(st_g[2]) == np.float32(st_o[2])
        assert np.array_equal(c.frozen(k), frozen)
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f32", "mixed", "f64"])
def test_fine_detector_fused_gather_matches_split(fine_ds, prec):
    """The row gather's long chains give the standalone CSC image gather's
    values, bit for bit: identical estimates and FreezeMu rows after 3
    iterations, fused and split."""
    print(f"case: fine detector, fused vs split adjoint {prec}", flush=True)
    out = []
    for split in (False, True):
        c = _ctx(fine_ds, prec, graph=True)
        c.set_split_adjoint(split)
        c.prepare()
        rows, _ = c.iterate(3)
        out.append((c.estimate(f64=True), [tuple(r[:4]) for r in rows]))
        c.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


@pytest.mark.gpu
def test_fine_detector_iterations_vs_oracle(fine_ds):
    """5 iterations: f64 within 1e-11 of the FP64 oracle, the mixed tier within
    2.5x the reference FP32 algorithm's own distance from FP64 (the bound the
    coarse-detector test uses); FreezeMu rows exact."""
    from oracle import oracle as O
    r64 = O.RefSolver(fine_ds.W, fine_ds.H, oracle_exposures(fine_ds), "f64")
    r32 = O.RefSolver(fine_ds.W, fine_ds.H, oracle_exposures(fine_ds), "f32")
    rows64, _ = r64.iterate(5)
    r32.iterate(5)
    base = O.compare_stats(r64.x, r32.x)["norm_linf"]
    for prec, tol in (("f64", 1e-11), ("mixed", 2.5 * base + 1e-6)):
        print(f"case: fine detector, 5 iterations {prec} (oracle f32 distance {base:.3e})", flush=True)
        c = _ctx(fine_ds, prec)
        c.prepare()
        rows, _ = c.iterate(5)
        st = O.compare_stats(r64.x, c.estimate(f64=True))
        print(f"  {prec}: norm_linf {st['norm_linf']:.3e} (tol {tol:.3e})", flush=True)
        assert st["norm_linf"] < tol, st
        for rg, ro in zip(rows, rows64):
            assert rg[:4] == ro[:4]
        c.close()
