"""The iteration's adjoint-image launch, bit for bit against the oracle.

The batched iteration turns each exposure's wr_k into S_k^T wr_k with one
CSC gather over every exposure (the adjoint-image kernel), then a row R2C of
the images.  The reference computes the same images as the scatter half of
its detector loop: scatter[t] = fmaf(frac, wr[p], scatter[t]) for p
ascending, from 0 (updc_core.c:1171-1176).  Given the oracle's wr (from the
same detector loop), the GPU images must equal the oracle's scatter exactly:

- a detector twice as coarse as the target (at most 4 entries per target:
  the kernel's batched entries);
- a detector finer than the target (5-9 entries per target: the chain
  past the batched entries);
- ABS32 LinkMaps (the CSC built from 32-bit target words);
- FreezeMu firing (wr = w on frozen pixels) and masked pixels (w = 0).

The gather reads a compact copy of the transposed map (16-bit source pixel
offsets per 64-target tile, 8-bit counts) wherever it fits, and the 8-byte
form elsewhere: a detector so fine that targets collect more than 255 links,
and a wide detector turned 90 degrees, whose target rows run down detector
columns so a tile spans more than 65535 source pixels, take the second path.
"""
import numpy as np
import pytest

from conftest import as_abs32, oracle_dataset, oracle_exposures

pytestmark = pytest.mark.gpu


def _oracle_wr_and_scatter(ds):
    """wr and scatter of every exposure from the oracle's fused detector loop,
    with a blur that freezes some pixels."""
    from oracle import oracle as O
    ref = O.RefSolver(ds.W, ds.H, oracle_exposures(ds), "f32")
    blur = ref.blur()
    wrs, scs = [], []
    for k, e in enumerate(ref.ex):
        bl = blur[k].copy()
        bl[: bl.shape[0] // 5] *= np.float32(-3.0)
        frozen = np.zeros(e["s"].size, np.uint8)
        wr, sc, (nn, _, _) = O.fused_detector(e["linkmap"], ds.W * ds.H, e["s"], e["w"], frozen, bl,
                                              float(e["mo_f"]), float(e["ped_f"]))
        assert nn > 0
        wrs.append(wr)
        scs.append(sc.reshape(ds.H, ds.W))
    return np.concatenate(wrs), np.stack(scs)


def _gpu_images(ds, prec, linkmaps, wr, compact=True):
    from updc_amd.solver import Context, Exposure
    with Context(ds.W, ds.H, prec) as c:
        for k in range(len(ds.geoms)):
            c.add_exposure(Exposure(ds.sci[k], ds.wht[k], ds.psfs[k], ds.background[k], ds.pedestal[k],
                                    linkmap=linkmaps[k]))
        c.prepare()
        assert all(c.exposure_info(k)["csc_compact"] == compact for k in range(len(ds.geoms)))
        return c.adjoint_images(wr)


@pytest.fixture(scope="module")
def fine_ds():
    from updc_amd import synth
    return oracle_dataset(synth.small_config(det=160, target=96, seed=21))


@pytest.mark.parametrize("case", ["coarse", "fine", "coarse_abs32", "fine_abs32"])
@pytest.mark.parametrize("prec", ["f32", "mixed"])
def test_adjoint_images_bitexact(small_ds, fine_ds, case, prec):
    from oracle import oracle as O
    ds = small_ds if case.startswith("coarse") else fine_ds
    lms = [as_abs32(m) for m in ds.linkmaps] if case.endswith("abs32") else ds.linkmaps
    print(f"case: adjoint images {case} {prec}", flush=True)
    wr, want = _oracle_wr_and_scatter(ds)
    if case.startswith("fine"):
        p = O.parse_linkmap(ds.linkmaps[0])
        keep = ds.wht[0].ravel()[p["pixel"]] > 0
        assert np.bincount(p["target"][keep]).max() > 4
    got = _gpu_images(ds, prec, lms, wr.astype(np.float64))
    assert got.dtype == np.float64
    assert np.array_equal(got.astype(np.float32), want), \
        f"{int((got.astype(np.float32) != want).sum())} pixels differ"
    # and the images were really exercised: non-zero on every exposure
    assert all(np.count_nonzero(want[k]) > 100 for k in range(len(ds.geoms)))


def _custom(W, H, geoms, seed):
    from updc_amd import synth
    g = np.random.Generator(np.random.PCG64(seed))
    truth = synth.truth_image(W, H, 10, 0.5, g, sigma_range=(0.6, 2.0))
    psfs = [synth.gaussian_psf(5, 2.0) for _ in geoms]
    return oracle_dataset(synth.Dataset(W, H, geoms, psfs, truth, [0.0] * len(geoms), 2.0, 0.02, None, 5, "custom"))


@pytest.mark.parametrize("case", ["very_fine", "turned_wide"])
def test_adjoint_images_bitexact_on_the_wide_form(case):
    from oracle import oracle as O
    from updc_amd import synth
    print(f"case: adjoint images, 8-byte form ({case})", flush=True)
    if case == "very_fine":
        # 20 detector pixels per target pixel along each axis: ~400 links per target
        ds = _custom(16, 16, [synth.Geometry(320, 320, 0.05, 1.0, (0.3, 0.2), None, (7.5, 7.5))], 3)
        p = O.parse_linkmap(ds.linkmaps[0])
        assert np.bincount(p["target"]).max() > 255
    else:
        # 1800 x 40 detector turned 90 degrees onto a 64-wide target: a target row
        # crosses 40 detector rows 1800 pixels apart
        ds = _custom(64, 1808, [synth.Geometry(1800, 40, 1.0, 90.0, (0.25, 0.1), None, (31.5, 903.5))], 4)
    wr, want = _oracle_wr_and_scatter(ds)
    for prec in ("f32", "mixed"):
        got = _gpu_images(ds, prec, ds.linkmaps, wr.astype(np.float64), compact=False)
        assert np.array_equal(got.astype(np.float32), want)
    assert np.count_nonzero(want) > 100
