"""Every benchmark configuration at its own size on the GPU.

- Config 1 (512², 4 x 256² exposures, 63² PSF, FFT 576², 20 iterations) is
  small enough for the CPU oracle: LinkMap bytes, x0, coverage and WHT bit for
  bit (F32); 20 FP64 iterations within 1e-11 of the FP64 oracle with equal
  FreezeMu rows; F32 and MIXED within the drift of the reference's own float32
  algorithm (the f32 oracle) from FP64.
- Config 2 (8 x 2048² -> 4096², PSF 1431², FFT 5600²): the GPU FP64 tier,
  batched and per-exposure, against the CPU FP64 oracle over iterations 1-3.
  This anchors the FP64 tier that serves as the reference of the 40-iteration
  error at this size.
- Config 3 (config-2 geometry, sky 0.01) and config 4 (24 uneven exposures ->
  6144², FFT 7680²) at full size, through properties that hold at any size:
  MIXED within 1e-5 of FP64 at iterations 10, 20 and 40; batched equal to
  per-exposure launches bit for bit; FreezeMu rows equal between the tiers;
  totals accumulate (updc_core.c:1162-1164, 1425-1431).  Config 3 also forces
  FreezeMu with the reference's recipe (every background -100,
  tests/test_smoke.py:117-136) and reports the tiers' distance.
- Config 5 (4 x 1024² + a held-out fifth -> 2048², 255² two-component PSFs,
  K = 10): FP64 gradients of the held-out Poisson loss through the HIP
  operators against the FP64 torch restatement on the GPU (1e-9) and central
  finite differences (h = 1e-6) in the initial estimate, two PSF parameters
  and a background.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ctx(ds, prec, graph=True, plain=False):
    from bench import solve_ctx
    return solve_ctx(ds, prec, graph=graph and not plain, plain=plain)


def _oracle_ex(ds):
    return [dict(sci=ds.sci[k], wht=ds.wht[k], psf=ds.psfs[k], background=ds.background[k],
                 pedestal=ds.pedestal[k], linkmap=ds.linkmaps[k]) for k in range(len(ds.geoms))]


def _chunks(c, chunks):
    """Iterate in chunks; the estimate (f64) and rows after each."""
    snaps, rows, done = {}, [], 0
    for n in chunks:
        r, _ = c.iterate(n)
        rows += [(x[0] + done, x[1], x[2], x[3]) for x in r]
        done += n
        snaps[done] = c.estimate(f64=True)
    return snaps, rows


def _totals_accumulate(rows, n):
    for k in range(n):
        total = 0
        for r in rows:
            if r[1] == k + 1:
                total += r[2]
                assert r[3] == total, (k, r)


# ------------------------------------------------------------------- config 1
def test_config1_against_the_oracle():
    from bench import make_dataset
    from oracle import oracle as O
    print("case: config 1 (512^2, 4 x 256^2, PSF 63^2, 20 iterations)", flush=True)
    ds = make_dataset(1)
    n = len(ds.geoms)
    assert (ds.W, ds.H, n, ds.psfs[0].shape) == (512, 512, 4, (63, 63))
    for k in range(n):
        assert ds.linkmaps[k] == O.build_linkmap(ds.vertices(k), ds.W, ds.H)
    r32 = O.RefSolver(ds.W, ds.H, _oracle_ex(ds), "f32")
    r64 = O.RefSolver(ds.W, ds.H, _oracle_ex(ds), "f64")
    assert (r64.FW, r64.FH) == (576, 576)
    c = _ctx(ds, "f32")
    x0, cov = c.initial()
    assert np.array_equal(x0.astype(np.float32), r32.x)
    assert np.array_equal(cov.astype(np.float32), r32.coverage)
    assert np.array_equal(c.wht(), r32.coverage * np.float32(r32.weight_scale))
    c.close()
    rows64, _ = r64.iterate(20)
    r32.iterate(20)
    base = O.compare_stats(r64.x, r32.x)["norm_linf"]
    print(f"  oracle f32 vs f64 after 20: {base:.3e}", flush=True)
    for prec, factor in (("f64", None), ("f32", 5.0), ("mixed", 2.5)):
        c = _ctx(ds, prec)
        rows, _ = c.iterate(20)
        st = O.compare_stats(r64.x, c.estimate(f64=True))
        tol = 1e-11 if factor is None else factor * base + 1e-6
        print(f"  {prec}: norm_linf {st['norm_linf']:.3e} (tol {tol:.3e})", flush=True)
        assert st["norm_linf"] < tol, st
        assert [r[:4] for r in rows] == [r[:4] for r in rows64]
        c.close()


# ------------------------------------------------------------------- config 2
def test_config2_fp64_tier_against_the_cpu_oracle():
    from bench import make_dataset
    from oracle import oracle as O
    print("case: config 2 FP64 (batched and plain) vs CPU FP64 oracle, iterations 1-3", flush=True)
    ds = make_dataset(2)
    assert (ds.W, ds.H, len(ds.geoms)) == (4096, 4096, 8)
    gpu = {}
    for plain in (False, True):
        c = _ctx(ds, "f64", plain=plain)
        assert (c.scalars()["fft_width"], c.scalars()["fft_height"]) == (5600, 5600)
        gpu[plain] = _chunks(c, [1, 1, 1])
        c.close()
    assert all(np.array_equal(gpu[False][0][i], gpu[True][0][i]) for i in (1, 2, 3))
    assert gpu[False][1] == gpu[True][1]
    print("  CPU oracle set-up (f64, threaded transforms)", flush=True)
    ref = O.RefSolver(ds.W, ds.H, _oracle_ex(ds), "f64", fft_shape=(5600, 5600))
    rows_o = []
    for it in (1, 2, 3):
        r, _ = ref.iterate(1)
        rows_o += [(it, k, nn, nt) for (_, k, nn, nt, _) in r]
        st = O.compare_stats(ref.x, gpu[False][0][it])
        print(f"  iteration {it}: norm_linf {st['norm_linf']:.3e} rel_rms {st['rms'] / np.sqrt(np.mean(ref.x ** 2)):.3e}",
              flush=True)
        assert st["norm_linf"] < 1e-11, st
    assert gpu[False][1] == rows_o


# ------------------------------------------------------------------- config 3
def test_config3_full_size_properties():
    from bench import make_dataset
    from oracle import oracle as O
    print("case: config 3 (sky 0.01) at full size", flush=True)
    ds = make_dataset(3)
    n = len(ds.geoms)
    runs = {}
    for name, prec, plain in (("mixed", "mixed", False), ("mixed_plain", "mixed", True), ("f64", "f64", False)):
        c = _ctx(ds, prec, plain=plain)
        runs[name] = _chunks(c, [10, 10, 20])
        c.close()
    for it in (10, 20, 40):
        assert np.array_equal(runs["mixed"][0][it], runs["mixed_plain"][0][it])
        st = O.compare_stats(runs["f64"][0][it], runs["mixed"][0][it])
        print(f"  iteration {it}: MIXED vs FP64 norm_linf {st['norm_linf']:.3e}", flush=True)
        assert st["norm_linf"] <= 1e-5, (it, st)
    assert runs["mixed"][1] == runs["f64"][1] == runs["mixed_plain"][1]
    assert len(runs["f64"][1]) == 40 * n
    _totals_accumulate(runs["f64"][1], n)
    # FreezeMu forced by the reference's recipe: every background -100
    ds.background = [-100.0] * n
    forced = {}
    for prec in ("mixed", "f64"):
        c = _ctx(ds, prec)
        forced[prec] = _chunks(c, [2])
        c.close()
    ra, rb = forced["mixed"][1], forced["f64"][1]
    assert ra == rb and all(r[3] > 0 for r in ra[n:])
    _totals_accumulate(ra, n)
    st = O.compare_stats(forced["f64"][0][2], forced["mixed"][0][2])
    print(f"  background -100, 2 iterations: frozen {[r[3] for r in ra[n:]]}; MIXED vs FP64 norm_linf "
          f"{st['norm_linf']:.3e}", flush=True)


# ------------------------------------------------------------------- config 4
def test_config4_full_size_properties():
    from bench import make_dataset, solve_ctx
    from oracle import oracle as O
    print("case: config 4 (24 uneven exposures -> 6144^2, FFT 7680^2) at full size", flush=True)
    ds = make_dataset(4)
    n = len(ds.geoms)
    assert n == 24 and ds.W == 6144
    assert sorted({(g.iw, g.ih) for g in ds.geoms}) == [(640, 640), (1024, 2048), (2048, 2048)]
    runs = {}
    for name, prec, plain in (("mixed", "mixed", False), ("mixed_plain", "mixed", True), ("f64", "f64", False)):
        c = _ctx(ds, prec, plain=plain)
        assert (c.scalars()["fft_width"], c.scalars()["fft_height"]) == (7680, 7680)
        runs[name] = _chunks(c, [10, 10, 20])
        c.close()
    for it in (10, 20, 40):
        assert np.array_equal(runs["mixed"][0][it], runs["mixed_plain"][0][it])
        st = O.compare_stats(runs["f64"][0][it], runs["mixed"][0][it])
        print(f"  iteration {it}: MIXED vs FP64 norm_linf {st['norm_linf']:.3e}", flush=True)
        assert st["norm_linf"] <= 1e-5, (it, st)
    assert runs["mixed"][1] == runs["f64"][1] == runs["mixed_plain"][1]
    _totals_accumulate(runs["f64"][1], n)
    # the autotuned shape (hipGraph replay) against FP64 as well
    c = solve_ctx(ds, "mixed", graph=True, autotune=True)
    sc = c.scalars()
    snaps, rows = _chunks(c, [40])
    c.close()
    st = O.compare_stats(runs["f64"][0][40], snaps[40])
    print(f"  autotuned {sc['fft_width']}x{sc['fft_height']}: MIXED vs FP64 norm_linf {st['norm_linf']:.3e}",
          flush=True)
    assert st["norm_linf"] <= 1e-5, st
    assert rows == runs["f64"][1]


# ------------------------------------------------------------------- config 5
def test_config5_gradients():
    import torch
    from bench import make_dataset
    from oracle import torch_ref as R
    from updc_amd import autograd as A
    from updc_amd import synth
    from updc_amd.solver import Context, Exposure
    print("case: config 5 FP64 gradients (K=10) vs torch restatement and finite differences", flush=True)
    ds = make_dataset(5)
    n, K = len(ds.geoms), 10
    assert (ds.W, n, ds.psfs[0].shape) == (2048, 5, (255, 255))
    c = Context(ds.W, ds.H, "f64")
    for k in range(n):
        c.add_exposure(Exposure(ds.sci[k], ds.wht[k], ds.psfs[k], ds.background[k], ds.pedestal[k],
                                linkmap=ds.linkmaps[k]))
    c.prepare()
    ops = A.Operators(c)
    ex = A.exposures_from_context(c, ops, ds.sci, ds.wht, ds.pedestal)
    x0np, _ = c.initial()
    dev, size = ops.device, ds.psfs[0].shape[0]
    start = synth.config5_psf_params(n)
    start[:, :, 1] *= 1.05
    start[:, 1, 0] *= 0.8

    def loss(x0, params, bs):
        pws = [A.wrap_psf(A.gaussian_mixture(params[k], size), ops.FH, ops.FW) for k in range(n)]
        x = A.unrolled(x0, pws[:n - 1], [bs[k] for k in range(n - 1)], K, ops, ex, checkpoint=True)
        return A.poisson_nll(A.forward_model(x, pws[n - 1], bs[n - 1], n - 1, ops, ex), ex.yeff[n - 1],
                             ex.active[n - 1])

    def leaves():
        return (torch.as_tensor(x0np, device=dev).requires_grad_(True),
                torch.as_tensor(start, device=dev).requires_grad_(True),
                torch.zeros(n, dtype=torch.float64, device=dev).requires_grad_(True))

    lv = leaves()
    L = loss(*lv)
    g = [t.detach() for t in torch.autograd.grad(L, lv)]
    # the torch restatement on the GPU (torch.fft, index_add)
    m = R.RefModel(ds.W, ds.H, ops.FW, ops.FH, ds.linkmaps, device=dev)
    lr = leaves()
    pws = [R.wrap_psf(R.gaussian_mixture(lr[1][k], size), ops.FH, ops.FW) for k in range(n)]
    xk = m.unrolled(lr[0], pws[:n - 1], lr[2][:n - 1], K, ex.yeff[:n - 1], ex.w[:n - 1], ex.wmask[:n - 1],
                    ex.active[:n - 1], ex.ped[:n - 1])
    Lr = R.poisson_nll(m.forward_model(xk, pws[n - 1], lr[2][n - 1], n - 1, ex.ped[n - 1]), ex.yeff[n - 1],
                       ex.active[n - 1])
    gr = [t.detach() for t in torch.autograd.grad(Lr, lr)]
    print(f"  loss {float(L):.12g} torch {float(Lr):.12g}", flush=True)
    assert abs(float(L) - float(Lr)) <= 1e-11 * abs(float(Lr))
    for a, b in zip(g, gr):
        rel = float((a - b).norm() / b.norm())
        print(f"  gradient rel {rel:.2e}", flush=True)
        assert rel <= 1e-9
    del m, xk, Lr, gr, pws
    with torch.no_grad():
        h = 1e-6
        floor = 8 * 2.2e-16 * abs(float(L)) / h
        base = leaves()
        v = torch.randn(base[0].shape, generator=torch.Generator().manual_seed(1), dtype=torch.float64).to(dev)
        cases = [("x0", lambda s: loss(base[0] + s * v, base[1], base[2]), float((g[0] * v).sum()))]
        for idx in [(0, 0, 1), (n - 1, 1, 0)]:
            e = torch.zeros_like(base[1])
            e[idx] = 1.0
            cases.append((f"psf{idx}", lambda s, e=e: loss(base[0], base[1] + s * e, base[2]), float(g[1][idx])))
        e = torch.zeros_like(base[2])
        e[0] = 1.0
        cases.append(("background0", lambda s: loss(base[0], base[1], base[2] + s * e), float(g[2][0])))
        for name, f, an in cases:
            fd = (float(f(h)) - float(f(-h))) / (2 * h)
            print(f"  FD {name}: fd {fd:.10g} analytic {an:.10g}", flush=True)
            assert abs(fd - an) <= 1e-5 * abs(an) + floor, (name, fd, an, floor)
    del ops, ex
    c.close()
