"""GPU parity of the HIP path against the CPU oracle (small, seeded inputs).

Bit-exact: LinkMap bytes, initial drizzle x0 and coverage, the detector stage
(wr, FreezeMu masks and counters, the transposed-gather scatter) given the
same blur.  Within tolerance: FFT stages (the oracle uses numpy's pocketfft in
place of FFTW) and whole iterations.
"""
import numpy as np
import pytest

from conftest import oracle_dataset, oracle_exposures

pytestmark = pytest.mark.gpu


def _ctx(ds, precision, fft=None, graph=False, use_vertices=True, workers=1):
    from updc_amd.solver import Context, Exposure
    c = Context(ds.W, ds.H, precision, fft_shape=fft, graph=graph, preprocess_workers=workers)
    for k in range(len(ds.geoms)):
        e = Exposure(ds.sci[k], ds.wht[k], ds.psfs[k], ds.background[k], ds.pedestal[k])
        if use_vertices:
            e.vertices = ds.vertices(k)
        else:
            e.linkmap = ds.linkmaps[k]
        c.add_exposure(e)
    return c


def _oracle(ds, precision, workers=1):
    from oracle import oracle as O
    return O.RefSolver(ds.W, ds.H, oracle_exposures(ds), precision, workers=workers)


def test_linkmap_bytes_match_oracle(small_ds):
    print("case: GPU LinkMap build vs oracle, small_ds", flush=True)
    c = _ctx(small_ds, "f32")
    for k in range(len(small_ds.geoms)):
        assert c.linkmap(k) == small_ds.linkmaps[k]
    c.close()


@pytest.mark.parametrize("seed,rot,dist", [(11, 45.0, 0.0), (12, 0.0, 0.0), (13, 7.0, 0.05)])
def test_linkmap_geometries(seed, rot, dist):
    from oracle import oracle as O
    from updc_amd import synth
    from updc_amd.solver import Context, Exposure
    print(f"case: linkmap seed={seed} rot={rot} dist={dist}", flush=True)
    ds = synth.small_config(n_exp=1, det=40, target=80, seed=seed, rotation=rot, distort=dist, dither=9.0)
    v = ds.vertices(0)
    ref = O.build_linkmap(v, ds.W, ds.H)
    z = np.zeros((40, 40), np.float32)
    with Context(ds.W, ds.H, "f32") as c:
        c.add_exposure(Exposure(z, z + 1, ds.psfs[0], vertices=v))
        assert c.linkmap(0) == ref


def test_linkmap_abs32_wide_target():
    """Target 70000 wide: every stencil spans two target rows > 65535 -> ABS32."""
    from oracle import oracle as O
    from updc_amd import synth
    from updc_amd.solver import Context, Exposure
    print("case: ABS32 wide target", flush=True)
    g = synth.Geometry(8, 3, 2.0, 10.0, (0.3, 0.2), None, (100.0, 3.5))
    v = g.vertices()
    ref = O.build_linkmap(v, 70000, 8)
    assert O.parse_linkmap(ref)["encoding"] == 1
    z = np.zeros((3, 8), np.float32)
    with Context(70000, 8, "f32") as c:
        c.add_exposure(Exposure(z, z + 1, np.ones((3, 3), np.float32), vertices=v))
        assert c.linkmap(0) == ref


def test_fft_against_numpy(small_ds):
    print("case: FFT r2c/c2r vs numpy", flush=True)
    for prec, tol in (("f32", 2e-6), ("f64", 1e-13)):
        c = _ctx(small_ds, prec)
        c.prepare()
        s = c.scalars()
        FW, FH = s["fft_width"], s["fft_height"]
        rng = np.random.default_rng(5)
        img = rng.normal(size=(FH, FW))
        spec = c.fft_r2c(img)
        ref = np.fft.rfft2(img)
        err = np.max(np.abs(spec - ref)) / np.max(np.abs(ref))
        assert err < tol, (prec, err)
        back = c.fft_c2r(ref)
        err2 = np.max(np.abs(back - img * FW * FH)) / (FW * FH * np.max(np.abs(img)))
        assert err2 < tol, (prec, err2)
        c.close()


def test_setup_bitexact_f32(small_ds):
    """Weight scale, x0, coverage (WHT) match the f32 oracle bit for bit."""
    print("case: setup f32 bit-exact", flush=True)
    ref = _oracle(small_ds, "f32")
    c = _ctx(small_ds, "f32")
    c.prepare()
    s = c.scalars()
    assert np.float32(s["weight_scale"]) == np.float32(ref.weight_scale)
    x0, cov = c.initial()
    assert np.array_equal(x0.astype(np.float32), ref.x)
    assert np.array_equal(cov.astype(np.float32), ref.coverage)
    assert np.array_equal(c.wht(), ref.coverage * np.float32(ref.weight_scale))
    c.close()


def test_drizzle_worker_topology_bitexact(small_ds):
    print("case: drizzle with 2 workers", flush=True)
    ref = _oracle(small_ds, "f32", workers=2)
    c = _ctx(small_ds, "f32", workers=2)
    c.prepare()
    x0, cov = c.initial()
    assert np.array_equal(x0.astype(np.float32), ref.x)
    assert np.array_equal(cov.astype(np.float32), ref.coverage)
    c.close()


def test_detector_stage_bitexact_f32(small_ds):
    """Same blur in: wr, scatter, FreezeMu counters bit-identical."""
    from oracle import oracle as O
    print("case: detector stage f32 bit-exact", flush=True)
    ref = _oracle(small_ds, "f32")
    c = _ctx(small_ds, "f32")
    c.prepare()
    blur = ref.blur()
    for k in range(len(small_ds.geoms)):
        e = ref.ex[k]
        # drive some pixels through FreezeMu by scaling the blur negative
        bl = blur[k].copy()
        bl[: bl.shape[0] // 3] *= np.float32(-3.0)
        frozen = np.zeros(e["s"].size, np.uint8)
        wr_o, sc_o, st_o = O.fused_detector(e["linkmap"], ref.nout, e["s"], e["w"], frozen, bl,
                                            float(e["mo_f"]), float(e["ped_f"]))
        wr_g, sc_g, st_g = c.detector(k, bl)
        assert np.array_equal(wr_g.astype(np.float32), wr_o)
        assert np.array_equal(sc_g.astype(np.float32).ravel(), sc_o)
        assert st_g[0] == st_o[0] and st_g[1] == st_o[1]
        assert np.float32(st_g[2]) == np.float32(st_o[2])
        assert np.array_equal(c.frozen(k), frozen)
        assert st_o[0] > 0
    c.close()


def test_psf_spectrum_and_denominator(small_ds):
    print("case: PSF spectra and denominator vs oracle", flush=True)
    # mixed: FP64 transforms on weights normalised in FP32 (the reference's
    # normalisation), so its denominator carries the f32 weight rounding
    for prec, tol in (("f32", 5e-6), ("f64", 1e-12), ("mixed", 2e-7)):
        ref = _oracle(small_ds, "f64" if prec != "f32" else "f32")
        c = _ctx(small_ds, prec)
        c.prepare()
        for k in range(len(small_ds.geoms)):
            P = c.psf_spectrum(k)
            Pr = ref.ex[k]["P"]
            assert np.max(np.abs(P - Pr)) < (1e-6 if prec != "f64" else 1e-13) * np.max(np.abs(Pr))
            if prec == "mixed":
                # FP32 spectra rounded once from the FP64 transform: each
                # component within half an FP32 ulp of the FP64 value
                slack = 1e-14 * np.max(np.abs(Pr))
                for a, b in ((P.real, Pr.real), (P.imag, Pr.imag)):
                    assert np.all(np.abs(a - b) <= 2.0 ** -24 * np.abs(b) + slack)
        D = c.denominator()
        err = np.max(np.abs(D - ref.D)) / np.max(np.abs(ref.D))
        assert err < tol, (prec, err)
        c.close()


@pytest.fixture(scope="module")
def oracle_runs(small_ds):
    """FP64 oracle and the reference's FP32 algorithm (oracle f32), 5 iterations."""
    from oracle import oracle as O
    r64 = _oracle(small_ds, "f64")
    r32 = _oracle(small_ds, "f32")
    rows64, _ = r64.iterate(5)
    rows32, _ = r32.iterate(5)
    base = O.compare_stats(r64.x, r32.x)["norm_linf"]
    return r64.x.copy(), rows64, base


# The normalised drift after 5 iterations is dominated by low-coverage edge
# pixels, where D and C are small and any FP32 FFT error floor (relative to the
# image maximum) is amplified: adding white noise of 1e-7 x max to the f32
# oracle's inverse transforms multiplies its own drift by 4 (1.6e-5 -> 6.6e-5).
# An FP32 FFT within 2x of pocketfft's accuracy therefore stays within 5x.
@pytest.mark.parametrize("prec,factor", [("f64", None), ("f32", 5.0), ("mixed", 2.5)])
def test_iterations_vs_oracle(small_ds, oracle_runs, prec, factor):
    """5 iterations against the FP64 oracle.  f64: within 1e-11.  f32/mixed:
    within `factor` times the distance of the reference's own FP32 algorithm
    (the f32 oracle) from FP64 -- i.e. as accurate as the reference's float32
    arithmetic allows.  FreezeMu rows must agree exactly."""
    from oracle import oracle as O
    x64, rows64, base = oracle_runs
    print(f"case: 5 iterations {prec} vs oracle f64 (oracle f32 distance {base:.3e})", flush=True)
    c = _ctx(small_ds, prec)
    c.prepare()
    rows_g, _ = c.iterate(5)
    st = O.compare_stats(x64, c.estimate(f64=True))
    tol = 1e-11 if factor is None else factor * base + 1e-6
    print(f"  {prec}: norm_linf {st['norm_linf']:.3e} (tol {tol:.3e})", flush=True)
    assert st["norm_linf"] < tol, st
    for rg, ro in zip(rows_g, rows64):
        assert rg[:4] == ro[:4]
    c.close()


def test_stage_accuracy_vs_f64_oracle(small_ds):
    """Forward blur, adjoint and denominator of every tier against the FP64
    oracle (errors printed for the record, loose bounds asserted)."""
    print("case: stage accuracy vs f64 oracle", flush=True)
    ref = _oracle(small_ds, "f64")
    rng = np.random.default_rng(3)
    v = rng.normal(size=sum(g.iw * g.ih for g in small_ds.geoms))
    x = ref.x
    blur_o = ref.blur(x)
    adj_o = np.zeros((small_ds.H, small_ds.W))
    pos = 0
    acc = 0
    for k, e in enumerate(ref.ex):
        n = e["s"].size
        vk = np.where(e["w"] > 0, v[pos:pos + n], 0.0)
        acc = acc + ref._rfft(ref._scatter64(e, vk)) * np.conj(e["P"])
        pos += n
    adj_o = ref._irfft_crop(acc)
    for prec, tol in (("f32", 3e-6), ("mixed", 3e-6), ("f64", 1e-12)):
        c = _ctx(small_ds, prec)
        c.prepare()
        bl = c.forward(x)
        eb = max(np.max(np.abs(bl[k] - blur_o[k])) for k in range(len(blur_o))) / np.max(np.abs(blur_o))
        vm = np.concatenate([np.where(ref.ex[k]["w"] > 0, v[sum(g.iw * g.ih for g in small_ds.geoms[:k]):][:ref.ex[k]["s"].size], 0.0)
                             for k in range(len(ref.ex))])
        ad = c.adjoint(vm)
        ea = np.max(np.abs(ad - adj_o)) / np.max(np.abs(adj_o))
        ed = np.max(np.abs(c.denominator() - ref.D)) / np.max(np.abs(ref.D))
        print(f"  {prec}: blur {eb:.3e} adjoint {ea:.3e} denominator {ed:.3e}", flush=True)
        assert eb < tol and ea < tol, (prec, eb, ea)
        c.close()


def test_graph_replay_matches_eager(small_ds):
    print("case: hipGraph replay vs eager", flush=True)
    a = _ctx(small_ds, "f32", graph=False)
    b = _ctx(small_ds, "f32", graph=True)
    a.prepare()
    b.prepare()
    ra, _ = a.iterate(3)
    rb, _ = b.iterate(3)
    assert np.array_equal(a.estimate(), b.estimate())
    assert ra == rb
    # replay again after eager host work in between
    _ = a.estimate()
    a.iterate(2)
    b.iterate(2)
    assert np.array_equal(a.estimate(), b.estimate())
    a.close()
    b.close()


@pytest.mark.parametrize("prec", ["f32", "f64", "mixed"])
def test_plain_form_matches_batched(small_ds, prec):
    """The per-exposure form (one launch per exposure and stage, the
    reference's dataflow) and the batched form run the same transforms and
    products in the same order: identical estimates and FreezeMu rows."""
    from updc_amd._lib import FLAG_PER_EXPOSURE, check
    print(f"case: plain vs batched {prec}", flush=True)
    a = _ctx(small_ds, prec)
    b = _ctx(small_ds, prec)
    check(b.lib.updc_ctx_set_flags(b.h, b.flags | FLAG_PER_EXPOSURE))
    a.prepare()
    b.prepare()
    ra, _ = a.iterate(3)
    rb, _ = b.iterate(3)
    assert ra == rb
    assert np.array_equal(a.estimate(f64=True), b.estimate(f64=True))
    a.close()
    b.close()


def test_freeze_forcing_invariants():
    """Negative background forces FreezeMu (reference tests/test_smoke.py:107-136):
    n_new > 0 somewhere, total_1 == new_1, total_{i+1} == total_i + new_{i+1}."""
    from updc_amd import synth
    print("case: FreezeMu forcing", flush=True)
    ds = oracle_dataset(synth.small_config(seed=21))
    ds.background = [-100.0 for _ in ds.geoms]
    c = _ctx(ds, "f32")
    c.prepare()
    rows, _ = c.iterate(3)
    assert any(r[2] > 0 for r in rows)
    n = len(ds.geoms)
    for k in range(n):
        mine = [r for r in rows if r[1] == k + 1]
        assert mine[0][3] == mine[0][2]
        for a, b in zip(mine, mine[1:]):
            assert b[3] == a[3] + b[2]
    c.close()


@pytest.mark.parametrize("prec", ["f32", "mixed", "f64"])
def test_split_adjoint_matches_fused(small_ds, prec):
    """The adjoint gather as its own launch feeds the same values to the same
    row transform: identical estimates and FreezeMu rows."""
    print(f"case: split adjoint vs fused {prec}", flush=True)
    out = []
    for split in (False, True):
        c = _ctx(small_ds, prec, graph=True, use_vertices=False)
        c.set_split_adjoint(split)
        c.prepare()
        rows, _ = c.iterate(3)
        out.append((c.estimate(f64=True), [tuple(r[:4]) for r in rows]))
        c.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


def test_initial_estimate_survives_iterations_and_runs_repeat(small_ds):
    """updc_get_initial_f64 returns the set-up estimate after iterating, and a
    context reset to it (estimate and FreezeMu masks) repeats a run exactly."""
    print("case: initial estimate kept; repeated runs on one context", flush=True)
    c = _ctx(small_ds, "mixed", graph=True, use_vertices=False)
    c.prepare()
    x0, _ = c.initial()
    out = []
    for _ in range(2):
        c.set_estimate(x0)
        c.reset_frozen()
        rows, _ = c.iterate(3)
        out.append((c.estimate(f64=True), [tuple(r[:4]) for r in rows]))
        assert np.array_equal(c.initial()[0], x0)
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]
    c.close()


def test_odd_target_size_matches_oracle():
    """An odd target width (rows of odd pitch: the row passes take the scalar
    pair path) and odd height, FP64 against the FP64 oracle."""
    from oracle import oracle as O
    from updc_amd import synth
    print("case: odd target 95 x 97, f64 vs oracle", flush=True)
    ds = synth.small_config(n_exp=2, det=44, target=95, psf=7, seed=21)
    ds.H = 97
    ds.truth = np.pad(ds.truth, ((0, 2), (0, 0)), mode="edge")
    ds = oracle_dataset(ds)
    ref = O.RefSolver(ds.W, ds.H, oracle_exposures(ds), "f64")
    c = _ctx(ds, "f64", use_vertices=False)
    c.prepare()
    rows_ref, _ = ref.iterate(3)
    rows, _ = c.iterate(3)
    st = O.compare_stats(ref.x, c.estimate(f64=True))
    assert st["norm_linf"] < 1e-11, st
    assert [tuple(r[:4]) for r in rows] == [tuple(r[:4]) for r in rows_ref]
    c.close()


def test_mixed_precision_budget():
    """The mixed tier after 40 iterations of a quarter-scale config-2 workload
    (4 exposures of 512^2 -> 1024^2, PSF 357^2) against the f64 tier: within
    2e-6 in normalised L-inf, and within a quarter of the f32 tier's distance
    (the FP64-derived PSF spectra and the FP64 inverse of the summed adjoint
    spectrum carry it there; the per-exposure transforms stay FP32)."""
    from oracle import oracle as O
    from updc_amd import synth
    print("case: mixed precision budget, quarter-scale config 2, 40 iterations", flush=True)
    ds = oracle_dataset(synth.config2(n_exp=4, det=512, target=1024, psf_size=357, nsrc=312))
    xs = {}
    for prec in ("f64", "f32", "mixed"):
        c = _ctx(ds, prec, graph=True, use_vertices=False)
        c.prepare()
        c.iterate(40)
        xs[prec] = c.estimate(f64=True)
        c.close()
    e32 = O.compare_stats(xs["f64"], xs["f32"])["norm_linf"]
    emx = O.compare_stats(xs["f64"], xs["mixed"])["norm_linf"]
    print(f"  norm_linf vs f64: f32 {e32:.3e}, mixed {emx:.3e}", flush=True)
    assert emx < 2e-6
    assert emx < 0.25 * e32


def test_wide_rows_take_the_unfused_paths(small_ds):
    """An FFT 6400 wide (rows of M = 3200 points) is beyond the persistent
    row passes: the gather R2C, the row C2Rs and the update run as separate
    one-tile-per-workgroup launches.  Against the FP64 oracle at the same
    shape, every tier stays within its usual bound, and mixed agrees with its
    own plain (per-exposure) form bit for bit."""
    from oracle import oracle as O
    from updc_amd._lib import FLAG_PER_EXPOSURE, check
    shape = (6400, 112)
    print(f"case: FFT {shape}, unfused row paths", flush=True)
    ref = O.RefSolver(small_ds.W, small_ds.H, oracle_exposures(small_ds), "f64", fft_shape=shape)
    ref.iterate(3)
    xs = {}
    for prec, tol in (("f64", 1e-11), ("mixed", 1e-4)):
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


def test_widest_one_workgroup_row_fft(small_ds):
    """Rows of M = 4096 points (FFT 8192 wide, radix 8 only) are the widest a
    4-row tile's registers hold (two radix-8 butterflies per thread per
    stage): every tier matches the FP64 oracle at that shape.  Wider rows
    take the four-step path (test_gpu_four_step.py)."""
    from oracle import oracle as O
    shape = (8192, 112)
    print(f"case: FFT {shape}, widest rows", flush=True)
    ref = O.RefSolver(small_ds.W, small_ds.H, oracle_exposures(small_ds), "f64", fft_shape=shape)
    ref.iterate(2)
    for prec, tol in (("f64", 1e-11), ("mixed", 1e-4), ("f32", 1e-4)):
        c = _ctx(small_ds, prec, fft=shape)
        c.prepare()
        c.iterate(2)
        st = O.compare_stats(ref.x, c.estimate(f64=True))
        print(f"  {prec}: norm_linf {st['norm_linf']:.3e}", flush=True)
        assert st["norm_linf"] < tol, (prec, st)
        c.close()


def test_radix16_row_plan_matches_oracle(small_ds):
    """FFT 224 wide: rows of M = 112 = 7 x 16 points take a radix-16 stage
    (two stages instead of 8 2 7's three) in the persistent FP32 row passes.
    Every tier against the FP64 oracle at the same shape, within its usual
    bound; the plain form equals the batched form bit for bit."""
    from oracle import oracle as O
    from updc_amd._lib import FLAG_PER_EXPOSURE, check
    shape = (224, 112)
    print(f"case: FFT {shape}, radix-16 row stage", flush=True)
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
    p = _ctx(small_ds, "f32", fft=shape)
    check(p.lib.updc_ctx_set_flags(p.h, p.flags | FLAG_PER_EXPOSURE))
    p.prepare()
    p.iterate(3)
    assert np.array_equal(p.estimate(f64=True), xs["f32"])
    p.close()


def test_column_pair_layout_matches_column_major(small_ds):
    """The spectra the batched column passes write for the row C2R (blurred
    spectra, correction) in the column-pair layout (the default where the
    passes allow it) and column-major (UPDC_FLAG_COLUMN_SPECTRA) feed the
    same values to the same arithmetic: identical estimates and FreezeMu rows
    in every tier, under hipGraph replay, for an odd spectrum column count
    (FFT 112 wide: 57 columns, the last pair with a phantom partner) and an
    even one (126 wide: 64 columns); FFT height 200 runs the edge column
    kernel."""
    print("case: column-pair layout vs column-major, FFT 112 x 200 and 126 x 200", flush=True)
    for shape in ((112, 200), (126, 200)):
        for prec in ("f64", "mixed", "f32"):
            out = []
            for column in (False, True):
                c = _ctx(small_ds, prec, fft=shape, graph=True)
                c.set_column_spectra(column)
                c.prepare()
                assert c.spectrum_layout() == ("columns" if column else "pairs"), (shape, prec, column)
                rows, _ = c.iterate(3)
                out.append((c.estimate(f64=True), [tuple(r[:4]) for r in rows]))
                c.close()
            assert np.array_equal(out[0][0], out[1][0]), (shape, prec)
            assert out[0][1] == out[1][1], (shape, prec)


def test_long_row_c2r_matches_oracle(small_ds):
    """FFT 7000 x 200: rows of 3500 points are too long for the 4-row
    persistent C2R, so the FP32 tiers take its 2-row form (and the R2C stays
    persistent); every tier against the FP64 oracle within its usual bound,
    and the column-pair layout against column-major bit for bit (columns of
    200 take the edge column kernel, so the FP32 tiers write pairs)."""
    from oracle import oracle as O
    shape = (7000, 200)
    print(f"case: FFT {shape}, 2-row persistent C2R", flush=True)
    ref = O.RefSolver(small_ds.W, small_ds.H, oracle_exposures(small_ds), "f64", fft_shape=shape)
    ref.iterate(3)
    for prec, tol in (("f64", 1e-11), ("mixed", 1e-4), ("f32", 1e-4)):
        xs = []
        for column in (False, True):
            c = _ctx(small_ds, prec, fft=shape, graph=True)
            c.set_column_spectra(column)
            c.prepare()
            # FP64 rows of 3500 points take the one-tile kernels and column-major spectra
            want = "columns" if column or prec == "f64" else "pairs"
            assert c.spectrum_layout() == want, (prec, column)
            c.iterate(3)
            xs.append(c.estimate(f64=True))
            c.close()
        st = O.compare_stats(ref.x, xs[0])
        print(f"  {prec}: norm_linf {st['norm_linf']:.3e}", flush=True)
        assert st["norm_linf"] < tol, (prec, st)
        assert np.array_equal(xs[0], xs[1]), prec


def test_radix10_column_plan_matches_oracle(small_ds):
    """FFT 112 x 200: columns of FH = 200 = 2^3 5^2 take radix-10 edge stages
    (10 2 10, each a prime-factor 2 x 5 butterfly) in the FP32 column passes,
    and the MIXED tier's FP64 inverse of the summed adjoint spectrum runs the
    same plan with FP64 twiddles.  The transforms against numpy, every tier
    against the FP64 oracle within its usual bound, and the plain form equal
    to the batched form bit for bit in the MIXED and f32 tiers (the plain
    MIXED form inverts its summed spectrum with the same plan)."""
    from oracle import oracle as O
    from updc_amd._lib import FLAG_PER_EXPOSURE, check
    shape = (112, 200)
    print(f"case: FFT {shape}, radix-10 column edges", flush=True)
    for prec, tol in (("f32", 2e-6), ("f64", 1e-13)):
        c = _ctx(small_ds, prec, fft=shape)
        c.prepare()
        rng = np.random.default_rng(7)
        img = rng.normal(size=(shape[1], shape[0]))
        ref = np.fft.rfft2(img)
        spec = c.fft_r2c(img)
        err = np.max(np.abs(spec - ref)) / np.max(np.abs(ref))
        back = c.fft_c2r(ref)
        err2 = np.max(np.abs(back - img * shape[0] * shape[1])) / (shape[0] * shape[1] * np.max(np.abs(img)))
        print(f"  {prec}: r2c {err:.2e} c2r {err2:.2e}", flush=True)
        assert err < tol and err2 < tol, (prec, err, err2)
        c.close()
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
    for prec in ("mixed", "f32"):
        p = _ctx(small_ds, prec, fft=shape)
        check(p.lib.updc_ctx_set_flags(p.h, p.flags | FLAG_PER_EXPOSURE))
        p.prepare()
        p.iterate(3)
        assert np.array_equal(p.estimate(f64=True), xs[prec]), prec
        p.close()
