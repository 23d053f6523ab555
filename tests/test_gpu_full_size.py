"""The headline workload at full size: 8 exposures of 2048² onto a 4096² target
grid, PSF 1431², FFT 5600².  The oracle needs seconds per iteration here, so
this test rests on properties that hold at any size:

- the FP32 tier with FP64 accumulation (MIXED) stays within 1e-5 (max |d| /
  max |x|) of FP64 after 40 iterations;
- batched FP64 and the plain per-exposure FP64 form (the reference's dataflow,
  updc_core.c:1410-1458) give the same estimate and FreezeMu rows bit for bit;
- the fused and split adjoint forms, and hipGraph replay and eager launches,
  give the same estimate bit for bit;
- FreezeMu rows number iterations x exposures, in (iteration, exposure) order
  (updc_core.c:1425-1431), and a frozen pixel stays frozen: each row's total
  is the previous total plus the newly frozen (updc_core.c:1162-1164);
- every estimate is finite and no pixel changes sign from the drizzled start
  x0, which may hold negative pixels: the update multiplies by a factor clamped
  at zero (updc_core.c:1438-1455).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ITERS = 40


@pytest.fixture(scope="module")
def full_ds():
    from bench import make_dataset
    print("case: 8 x 2048^2 -> 4096^2 synthetic workload", flush=True)
    return make_dataset(2)


def _run(ds, precision, plain=False, fused=False, graph=True):
    from bench import solve_ctx
    graph = graph and not plain
    print(f"  solve {precision} plain={plain} fused={fused} graph={graph}", flush=True)
    c = solve_ctx(ds, precision, graph=graph, plain=plain)
    if fused:
        c.set_split_adjoint(False)
    x0, _ = c.initial()
    rows, _ = c.iterate(ITERS)
    x = c.estimate(f64=True)
    c.close()
    assert np.isfinite(x).all() and (x * x0 >= 0).all()
    return x, [tuple(r[:4]) for r in rows]


def test_full_size_tiers_and_forms(full_ds):
    from oracle import oracle as O
    n = len(full_ds.geoms)
    assert (full_ds.W, full_ds.H, n) == (4096, 4096, 8)
    x64, rows64 = _run(full_ds, "f64")
    x64p, rows64p = _run(full_ds, "f64", plain=True)
    assert np.array_equal(x64, x64p)
    assert rows64 == rows64p
    xm, rowsm = _run(full_ds, "mixed")
    xmf, rowsmf = _run(full_ds, "mixed", fused=True)
    assert np.array_equal(xm, xmf)
    assert rowsm == rowsmf
    xme, rowsme = _run(full_ds, "mixed", graph=False)
    assert np.array_equal(xm, xme)
    assert rowsm == rowsme
    st = O.compare_stats(x64, xm)
    print(f"  mixed vs f64 after {ITERS}: norm_linf {st['norm_linf']:.3e} rms {st['rms']:.3e}", flush=True)
    assert st["norm_linf"] <= 1e-5, st
    for rows in (rows64, rowsm):
        assert len(rows) == ITERS * n
        assert [(r[0], r[1]) for r in rows] == [(i + 1, k + 1) for i in range(ITERS) for k in range(n)]
        for k in range(n):
            total = 0
            for r in rows:
                if r[1] == k + 1:
                    total += r[2]
                    assert r[3] == total, (k, r)
