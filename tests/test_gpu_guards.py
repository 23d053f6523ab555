"""The reference's numeric guards, forced from the host, on the GPU.

Each case is an ordinary input that the guard turns into an error; every
kernel still runs to its end (no fault, no hang):

- y_eff = sci + pedestal < 0 on a weighted pixel: set-up refuses it
  (updc_core.c:370-371);
- a NaN in the estimate: every blurred value is NaN, so the first active
  pixel of exposure 1 fails "non-finite forward mean" (:1160);
- x = 0 with background 1e-30 and one pixel of sci 1e10: that pixel's ratio
  1e10 / 1e-30 overflows float32, "non-finite y/raw_mu ratio" (:1168) in the
  reference-precision tier; the FP64 ratio of the mixed tier does not
  overflow, but the float32 wr does, and the update meets a non-finite
  correction at the first target with D > 1e-8 (:1442);
- x = 0 with background 1e-6 and one pixel of sci 1e6 (ratio 1e12): the
  correction's transform rounding leaves values below -1e-6 max(D) where the
  exact correction is zero, "substantially negative multiplicative
  correction" (:1444-1448).

After a failure the call stops as the reference does: no later iteration
runs, the estimate is left as the failing iteration found it (a detector
failure changes nothing), and the freeze rows and iteration lines cover only
what the reference writes before it exits.  The drop-in executable exits 1
with the reference's "UPDC_ALL: <message>" line and writes no rows it would
not have written.
"""
import subprocess

import numpy as np
import pytest

from conftest import oracle_dataset

pytestmark = pytest.mark.gpu

TIERS = ["f32", "mixed", "f64"]


@pytest.fixture(scope="module")
def ds():
    from updc_amd import synth
    return oracle_dataset(synth.small_config(n_exp=2, det=32, target=64, psf=7, seed=41, mask_fraction=0.0))


def _ctx(ds, prec, sci=None, wht=None, bg=None, ped=None, graph=False):
    from updc_amd.solver import Context, Exposure
    n = len(ds.geoms)
    sci = sci or ds.sci
    wht = wht or ds.wht
    bg = bg if bg is not None else ds.background
    ped = ped if ped is not None else ds.pedestal
    c = Context(ds.W, ds.H, prec, graph=graph)
    for k in range(n):
        c.add_exposure(Exposure(sci[k], wht[k], ds.psfs[k], bg[k], ped[k], linkmap=ds.linkmaps[k]))
    return c


def _first_active_pixel(ds, k):
    from oracle import oracle as O
    p = O.parse_linkmap(ds.linkmaps[k])
    act = (np.diff(p["offset"].astype(np.int64)) > 0) & (ds.wht[k].ravel() > 0)
    return int(np.flatnonzero(act)[0])


@pytest.mark.parametrize("prec", TIERS)
def test_negative_yeff_is_refused(ds, prec):
    from updc_amd._lib import UpdcError
    print(f"case: negative y_eff, {prec}", flush=True)
    sci = [s.copy() for s in ds.sci]
    wht = [w.copy() for w in ds.wht]
    sci[1][10, 11] = -100.0
    wht[1][10, 11] = 1.0
    c = _ctx(ds, prec, sci=sci, wht=wht, ped=[0.0, 0.0])
    with pytest.raises(UpdcError, match="negative/non-finite y_eff reached solver"):
        c.prepare()
    c.close()


@pytest.mark.parametrize("prec,graph", [("f32", False), ("mixed", True), ("f64", False), ("f64", True)])
def test_nonfinite_forward_mean_stops_the_call(ds, prec, graph):
    from updc_amd.solver import IterationError
    print(f"case: NaN estimate -> non-finite forward mean, {prec} graph={graph}", flush=True)
    c = _ctx(ds, prec, graph=graph)
    c.prepare()
    x0 = c.estimate(f64=True)
    x = x0.copy()
    x[20, 30] = np.nan
    c.set_estimate(x)
    with pytest.raises(IterationError) as ei:
        c.iterate(4)
    e = ei.value
    p0 = _first_active_pixel(ds, 0)
    assert f"non-finite forward mean (iteration 1, exposure 1, pixel {p0})" in str(e), str(e)
    assert e.freeze == [] and e.info == [] and e.completed == 0
    # nothing after the failing detector stage ran: no update, no freeze
    assert np.array_equal(c.estimate(f64=True), x, equal_nan=True)
    assert all(not c.frozen(k).any() for k in range(2))
    # the next call starts clean and matches a fresh context
    c.set_estimate(x0)
    rows, info = c.iterate(2)
    assert [r[0] for r in rows] == [1, 1, 2, 2] and [i["iteration"] for i in info] == [1, 2]
    fresh = _ctx(ds, prec, graph=graph)
    fresh.prepare()
    rows_f, _ = fresh.iterate(2)
    assert np.array_equal(c.estimate(f64=True), fresh.estimate(f64=True))
    assert rows == rows_f
    fresh.close()
    c.close()


def _spike(ds, value, bg):
    """x = 0 everywhere; sci = 0 except one central pixel of exposure 2."""
    sci = [np.zeros_like(s) for s in ds.sci]
    wht = [np.ones_like(w) for w in ds.wht]
    sci[1][16, 17] = value
    return sci, wht, [bg, bg]


def test_ratio_overflow_f32(ds):
    from updc_amd.solver import IterationError
    print("case: ratio overflow, f32", flush=True)
    sci, wht, bg = _spike(ds, 1e10, 1e-30)
    c = _ctx(ds, "f32", sci=sci, wht=wht, bg=bg, ped=[0.0, 0.0])
    c.prepare()
    c.set_estimate(np.zeros((ds.H, ds.W)))
    with pytest.raises(IterationError) as ei:
        c.iterate(3)
    assert f"non-finite y/raw_mu ratio (iteration 1, exposure 2, pixel {16 * 32 + 17})" in str(ei.value)
    assert ei.value.freeze == [] and ei.value.completed == 0
    assert not c.frozen(0).any() and not c.frozen(1).any()
    c.close()


def test_ratio_overflow_mixed_meets_a_nonfinite_correction(ds):
    from updc_amd.solver import IterationError
    print("case: ratio overflow, mixed -> non-finite correction", flush=True)
    sci, wht, bg = _spike(ds, 1e10, 1e-30)
    c = _ctx(ds, "mixed", sci=sci, wht=wht, bg=bg, ped=[0.0, 0.0])
    c.prepare()
    t0 = int(np.flatnonzero(c.denominator().ravel() > 1e-8)[0])
    c.set_estimate(np.zeros((ds.H, ds.W)))
    with pytest.raises(IterationError) as ei:
        c.iterate(3)
    e = ei.value
    assert f"non-finite multiplicative correction (iteration 1, target {t0})" in str(e), str(e)
    # an update failure: the failing iteration's freeze rows were written, its line was not
    assert [r[:2] for r in e.freeze] == [(1, 1), (1, 2)] and e.info == [] and e.completed == 0
    c.close()


@pytest.mark.parametrize("prec", TIERS)
def test_substantially_negative_correction(ds, prec):
    from updc_amd.solver import IterationError
    print(f"case: negative correction, {prec}", flush=True)
    sci, wht, bg = _spike(ds, 1e6, 1e-6)
    c = _ctx(ds, prec, sci=sci, wht=wht, bg=bg, ped=[0.0, 0.0])
    c.prepare()
    c.set_estimate(np.zeros((ds.H, ds.W)))
    with pytest.raises(IterationError) as ei:
        c.iterate(3)
    e = ei.value
    assert "substantially negative multiplicative correction (iteration 1, target" in str(e), str(e)
    assert [r[:2] for r in e.freeze] == [(1, 1), (1, 2)] and e.info == [] and e.completed == 0
    c.close()


def test_dropin_guard_exit_and_rows(ds, tmp_path):
    """Background 1e37: the initial estimate is about -1e37 per pixel, its
    float32 transform overflows, and iteration 1 meets a non-finite forward
    mean.  The drop-in exits 1 with the reference's line; the freeze CSV has
    its header and no row, no iteration line is printed, no iteration output,
    WHT or timing file is written (updc_core.c:109, 1410-1486)."""
    from updc_amd import engine as E
    from updc_amd._lib import EXE_PATH
    print("case: drop-in, non-finite forward mean at iteration 1", flush=True)
    rows = [E.write_exposure(tmp_path, k + 1, ds.sci[k], ds.wht[k], ds.psfs[k], ds.vertices(k), 1e37,
                             ds.pedestal[k]) for k in range(2)]
    manifest = E.write_manifest(tmp_path / "manifest.txt", rows)
    g = ds.geoms[0]
    argv = E.core_argv(EXE_PATH, manifest, ds.W, ds.H, g.iw, g.ih, 2, False, 3, tmp_path / "it",
                       tmp_path / "it_wht.bin", save_every=1)
    res = subprocess.run([str(a) for a in argv], capture_output=True, text=True, timeout=120,
                         env=E.solver_env(False))
    assert res.returncode == 1, res.stderr
    assert "UPDC_ALL: non-finite forward mean" in res.stderr
    assert "UPDC iteration" not in res.stdout
    csv = (tmp_path / "it_mufreeze.csv").read_text().splitlines()
    assert csv == ["iteration,exposure,n_new_frozen,n_total_frozen,min_raw_mu,mu_marker,frozen_ratio"]
    assert not (tmp_path / "it_01.bin").exists() and not (tmp_path / "it_wht.bin").exists()
    assert not (tmp_path / "it_timing.csv").exists()


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_plain_form_under_graph_capture(ds, prec):
    """The per-exposure form replayed as a hipGraph equals its eager launches
    bit for bit (its extra spectra are allocated before any capture)."""
    print(f"case: plain form, graph vs eager, {prec}", flush=True)
    out = []
    for graph in (False, True):
        c = _ctx(ds, prec, graph=graph)
        c.set_per_exposure(True)
        c.prepare()
        rows, _ = c.iterate(3)
        rows2, _ = c.iterate(2)
        out.append((c.estimate(f64=True), rows + rows2))
        c.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]
    # switched on after prepare as well
    c = _ctx(ds, prec, graph=True)
    c.prepare()
    c.set_per_exposure(True)
    c.iterate(3)
    c.iterate(2)
    assert np.array_equal(c.estimate(f64=True), out[0][0])
    c.close()
