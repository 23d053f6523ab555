"""Empty inputs on the GPU path, against the FP64 oracle.

- An exposure whose footprint lies entirely outside the target grid: its
  LinkMap has no links, so none of its pixels is active (updc_core.c:1153, a
  pixel needs w > 0 and links), its CSC is empty, and it adds nothing to the
  adjoint or the denominator.
- An exposure whose weights are all zero (fully masked): every pixel is
  inactive, and the weight scale comes from the other exposure
  (updc_core.c:378-424).

In both cases the remaining exposure must drive the solve exactly as the
oracle's does: FP64 within 1e-11, FreezeMu rows equal, the empty exposure's
rows all zero.
"""
import numpy as np
import pytest

from conftest import as_abs32, oracle_dataset, oracle_exposures

pytestmark = pytest.mark.gpu


def _solve(ds, precision, iters=3):
    from updc_amd.solver import Context, Exposure
    c = Context(ds.W, ds.H, precision)
    for k in range(len(ds.geoms)):
        c.add_exposure(Exposure(ds.sci[k], ds.wht[k], ds.psfs[k], ds.background[k], ds.pedestal[k],
                                vertices=ds.vertices(k)))
    links = [c.linkmap(k) for k in range(len(ds.geoms))]
    c.prepare()
    rows, _ = c.iterate(iters)
    x = c.estimate(f64=True)
    c.close()
    return x, rows, links


def _check(ds, empty_k):
    from oracle import oracle as O
    ref = O.RefSolver(ds.W, ds.H, oracle_exposures(ds), "f64")
    rows_o, _ = ref.iterate(3)
    x, rows, links = _solve(ds, "f64")
    assert links == ds.linkmaps
    st = O.compare_stats(ref.x, x)
    print(f"  f64: norm_linf {st['norm_linf']:.3e}", flush=True)
    assert st["norm_linf"] < 1e-11, st
    assert [r[:4] for r in rows] == [r[:4] for r in rows_o]
    assert all(r[2] == 0 and r[3] == 0 for r in rows if r[1] == empty_k + 1)
    xm, _, _ = _solve(ds, "mixed")
    stm = O.compare_stats(ref.x, xm)
    print(f"  mixed: norm_linf {stm['norm_linf']:.3e}", flush=True)
    assert stm["norm_linf"] < 1e-4, stm


def test_exposure_outside_the_target_grid():
    from oracle import oracle as O
    from updc_amd import synth
    print("case: exposure 2 entirely outside the target grid", flush=True)
    ds = synth.small_config(n_exp=2, seed=31)
    c = ds.geoms[1].center_target
    ds.geoms[1].center_target = (c[0] + 500.0, c[1] + 500.0)
    ds = oracle_dataset(ds)
    assert O.parse_linkmap(ds.linkmaps[1])["nlinks"] == 0
    assert O.parse_linkmap(ds.linkmaps[0])["nlinks"] > 0
    _check(ds, 1)


def test_fully_masked_exposure():
    from updc_amd import synth
    print("case: exposure 2 fully masked (all weights 0)", flush=True)
    ds = oracle_dataset(synth.small_config(n_exp=2, seed=32))
    ds.wht[1] = np.zeros_like(ds.wht[1])
    _check(ds, 1)


def test_psf_stored_transposed():
    """psf_transposed = 1 (updc_core.c:814-819): the stored raster is read as
    [x][y].  With non-square PSFs (9 high, 7 wide as stored) the FFT shape
    follows the transposed extents (:1335-1336).  FP64 against the oracle
    within 1e-11, FreezeMu rows equal; and the closed form: the transposed
    flag on P is the plain path on the transpose of P."""
    from oracle import oracle as O
    from updc_amd import synth
    from updc_amd.solver import Context, Exposure
    print("case: PSFs stored transposed (non-square)", flush=True)
    ds = oracle_dataset(synth.small_config(n_exp=2, seed=41))
    psfs = [np.ascontiguousarray(p[:, 1:8]) for p in ds.psfs]  # 9 rows x 7 columns as stored
    ex = oracle_exposures(ds)
    for k, e in enumerate(ex):
        e["psf"] = psfs[k]
    ref = O.RefSolver(ds.W, ds.H, ex, "f64", psf_transposed=1)
    rows_o, _ = ref.iterate(3)

    def run(precision, transposed, arrays):
        c = Context(ds.W, ds.H, precision)
        for k in range(len(ds.geoms)):
            c.add_exposure(Exposure(ds.sci[k], ds.wht[k], arrays[k], ds.background[k], ds.pedestal[k],
                                    linkmap=ds.linkmaps[k], psf_transposed=transposed))
        c.prepare()
        rows, _ = c.iterate(3)
        x = c.estimate(f64=True)
        c.close()
        return x, rows

    x_t, rows_t = run("f64", True, psfs)
    st = O.compare_stats(ref.x, x_t)
    print(f"  f64 transposed vs oracle: norm_linf {st['norm_linf']:.3e}", flush=True)
    assert st["norm_linf"] < 1e-11, st
    assert [r[:4] for r in rows_t] == [r[:4] for r in rows_o]
    x_p, rows_p = run("f64", False, [np.ascontiguousarray(p.T) for p in psfs])
    sp = O.compare_stats(x_t, x_p)
    print(f"  f64 transposed flag vs plain on P^T: norm_linf {sp['norm_linf']:.3e}", flush=True)
    assert sp["norm_linf"] < 1e-12, sp
    assert [r[:4] for r in rows_p] == [r[:4] for r in rows_t]
    # the flag matters: the plain path on P itself is a different operator
    x_n, _ = run("f64", False, psfs)
    assert O.compare_stats(x_t, x_n)["norm_linf"] > 1e-6


@pytest.mark.parametrize("prec", ["f32", "mixed"])
def test_abs32_linkmaps_iterate_like_delta16(small_ds, prec):
    """Geometry picks DELTA16 below a 65536-wide target, so ABS32 LinkMaps
    reach the iteration only as images.  Re-encoded, the same links must
    give the same solve bit for bit: the forward gather's target words and
    the CSC built from them differ only in how a link's target is stored."""
    from oracle import oracle as O
    from updc_amd.solver import Context, Exposure
    print(f"case: ABS32 LinkMap images, {prec}", flush=True)
    out = []
    for enc in ("delta16", "abs32"):
        c = Context(small_ds.W, small_ds.H, prec)
        for k in range(len(small_ds.geoms)):
            lm = small_ds.linkmaps[k] if enc == "delta16" else as_abs32(small_ds.linkmaps[k])
            if enc == "abs32":
                assert O.parse_linkmap(lm)["encoding"] == 1
            c.add_exposure(Exposure(small_ds.sci[k], small_ds.wht[k], small_ds.psfs[k], small_ds.background[k],
                                    small_ds.pedestal[k], linkmap=lm))
        c.prepare()
        rows, _ = c.iterate(3)
        out.append((c.estimate(f64=True), [tuple(r[:4]) for r in rows], c.wht()))
        c.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])


def _corrupt(img: bytes, what: str) -> bytes:
    """A DELTA16 ULM2 image with one defect (layout updc_core.c:40-54, 463-483)."""
    from oracle import oracle as O
    p = O.parse_linkmap(img)
    nin, nl = p["nin"], p["nlinks"]
    b = bytearray(img)
    pos_off = 64
    pos_base = pos_off + 4 * (nin + 1)
    pos_d16 = pos_base + 4 * nin
    pos_frac = (pos_d16 + 2 * nl + 3) & ~3
    off = np.frombuffer(b, np.uint32, nin + 1, pos_off)
    if what == "magic":
        b[0:4] = b"\0\0\0\0"
    elif what == "width":
        np.frombuffer(b, np.uint32, 16, 0)[8] += 1
    elif what == "truncated":
        return bytes(b[:-4])
    elif what == "trailing":
        return bytes(b) + b"\0"
    elif what == "last_offset":
        off[nin] -= 1
    elif what == "non_monotone":
        p0 = int(np.nonzero(np.diff(off.astype(np.int64)) > 0)[0][0])  # a pixel with links
        off[p0 + 1] = off[p0] - 1 if off[p0] > 0 else nl
        if off[p0 + 1] == nl and p0 + 2 <= nin:
            off[p0] = nl  # z0 > z1 either way
    elif what == "target":
        p0 = int(np.nonzero(np.diff(off.astype(np.int64)) > 0)[0][0])
        np.frombuffer(b, np.uint32, nin, pos_base)[p0] = p["W"] * p["H"]
    elif what == "frac":
        np.frombuffer(b, np.float32, nl, pos_frac)[0] = np.nan
    return bytes(b)


@pytest.mark.parametrize("what,message", [
    ("magic", "invalid or obsolete LinkMap"), ("width", "invalid or obsolete LinkMap"),
    ("truncated", "invalid or obsolete LinkMap"), ("trailing", "invalid or obsolete LinkMap"),
    ("last_offset", "offset/link-count mismatch"), ("non_monotone", "not monotone"),
    ("target", "outside the target grid"), ("frac", "fraction is not finite")])
def test_corrupt_linkmap_is_refused(small_ds, what, message):
    """The reference's LinkMap checks (updc_core.c:485-496, the exact-size rule
    :463-483, :652) plus the device-side walk of offsets, targets and
    fractions before any kernel follows them: each defect is refused with a
    UPDC_FORMAT error naming it."""
    from updc_amd._lib import UpdcError
    from updc_amd.solver import Context, Exposure
    print(f"case: corrupt LinkMap ({what})", flush=True)
    assert small_ds.linkmaps[0][8] == 2  # DELTA16
    bad = _corrupt(small_ds.linkmaps[0], what)
    c = Context(small_ds.W, small_ds.H, "f32")
    with pytest.raises(UpdcError, match=message):
        c.add_exposure(Exposure(small_ds.sci[0], small_ds.wht[0], small_ds.psfs[0], small_ds.background[0],
                                small_ds.pedestal[0], linkmap=bad))
    c.close()


@pytest.mark.parametrize("defect,message", [("negative", "invalid PSF pixel"), ("nan", "invalid PSF pixel"),
                                            ("zero", "PSF sum is zero/non-finite")])
def test_invalid_psf_is_refused(small_ds, defect, message):
    """The reference's PSF checks (updc_core.c:777-789, 809): every raw pixel
    finite and non-negative, the Neumaier sum positive; refused at set-up."""
    from updc_amd._lib import UpdcError
    from updc_amd.solver import Context, Exposure
    print(f"case: invalid PSF ({defect})", flush=True)
    psf = small_ds.psfs[1].copy()
    if defect == "negative":
        psf[0, 0] = -1e-3
    elif defect == "nan":
        psf[2, 3] = np.nan
    else:
        psf[:] = 0.0
    c = Context(small_ds.W, small_ds.H, "mixed")
    for k in range(len(small_ds.geoms)):
        c.add_exposure(Exposure(small_ds.sci[k], small_ds.wht[k], psf if k == 1 else small_ds.psfs[k],
                                small_ds.background[k], small_ds.pedestal[k], linkmap=small_ds.linkmaps[k]))
    with pytest.raises(UpdcError, match=message):
        c.prepare()
    c.close()
