"""FFT shape selection: the candidate rule (CPU) and the GPU autotuner.

The candidate rule follows the reference's select_fft_shape
(updc/c_core/updc_core.c:170-179, 242-291): the first five admissible sizes
per axis from the minimum, the diagonal pairs, then (w2,h0), (w0,h2),
(w4,h0), (w0,h4).  Heights use the reference's own 7-smooth rule, so the
config-2 height candidates are the reference's 5600, 5625, 5670, 5760, 5832.
Widths must also be even (the row pass is a half-length complex transform),
so they are checked against a brute-force factorisation.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O


def _candidates(min_w, min_h):
    from updc_amd import _lib
    lib = _lib.load()
    cap = 16
    w, h, n = (C.c_int * cap)(), (C.c_int * cap)(), C.c_int()
    _lib.check(lib.updc_fft_shape_candidates(min_w, min_h, w, h, cap, C.byref(n)))
    return [(w[i], h[i]) for i in range(n.value)]


def _smooth(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def _first(n0, k, even):
    out, n = [], n0
    while len(out) < k:
        if _smooth(n) and (not even or n % 2 == 0):
            out.append(n)
        n += 1
    return out


@pytest.mark.parametrize("mw,mh", [(5526, 5526), (574, 574), (2302, 2302), (7574, 7574), (1001, 777), (2, 1)])
def test_candidate_rule(mw, mh):
    shapes = _candidates(mw, mh)
    wc, hc = _first(mw, 5, True), _first(mh, 5, False)
    pattern = [(wc[i], hc[i]) for i in range(5)] + [(wc[2], hc[0]), (wc[0], hc[2]), (wc[4], hc[0]),
                                                    (wc[0], hc[4])]
    dedup = []
    for s in pattern:
        if s not in dedup:
            dedup.append(s)
    assert shapes == dedup
    assert len(shapes) <= 9
    # heights: the reference's next_fast_size rule, as restated by the oracle
    ref = O.fft_shape_list(mw, mh)
    assert [s[1] for s in shapes[:5]] == sorted({s[1] for s in ref})[:5]
    # the reference's first shape is (next_fast_size(w), next_fast_size(h)); ours
    # agrees whenever that width is even
    if O.next_fast_size(mw) % 2 == 0:
        assert shapes[0] == ref[0]


def test_config2_heights_are_the_reference_list():
    shapes = _candidates(5526, 5526)
    assert [s[1] for s in shapes[:5]] == [5600, 5625, 5670, 5760, 5832]
    assert [s[0] for s in shapes[:5]] == [5600, 5670, 5760, 5832, 5880]


@pytest.mark.gpu
def test_autotune_picks_a_candidate_and_caches(small_ds, tmp_path):
    from updc_amd.solver import Context, Exposure

    def make(**kw):
        c = Context(small_ds.W, small_ds.H, "f32", **kw)
        for k in range(len(small_ds.geoms)):
            c.add_exposure(Exposure(small_ds.sci[k], small_ds.wht[k], small_ds.psfs[k], small_ds.background[k],
                                    small_ds.pedestal[k], linkmap=small_ds.linkmaps[k]))
        return c

    print("case: autotune, cache miss then hit", flush=True)
    a = make(autotune=True, fft_cache_dir=str(tmp_path))
    a.prepare()
    rows, cached = a.autotune_report()
    assert not cached and len(rows) >= 2
    sc = a.scalars()
    shape = (sc["fft_width"], sc["fft_height"])
    pw = max(p.shape[1] for p in small_ds.psfs)
    ph = max(p.shape[0] for p in small_ds.psfs)
    cands = _candidates(small_ds.W + pw - 1, small_ds.H + ph - 1)
    assert shape in cands
    timed = [r for r in rows if r[3] > 0]
    best = min(timed, key=lambda r: r[3])
    assert (best[0], best[1]) == shape
    assert list(tmp_path.glob("fft_autotune_gpu_v2_f32_*.txt"))
    a.iterate(3)
    xa = a.estimate(f64=True)
    a.close()

    b = make(autotune=True, fft_cache_dir=str(tmp_path))
    b.prepare()
    rows_b, cached_b = b.autotune_report()
    assert cached_b and len(rows_b) == 1
    sb = b.scalars()
    assert (sb["fft_width"], sb["fft_height"]) == shape
    b.iterate(3)
    assert np.array_equal(b.estimate(f64=True), xa)
    b.close()

    # retune ignores the cache
    r = make(autotune=True, fft_cache_dir=str(tmp_path), retune=True)
    r.prepare()
    assert not r.autotune_report()[1]
    r.close()

    # a fixed shape with the same column kernel reproduces the tuned run bit for bit
    f = make(fft_shape=shape, column_kernel=best[2])
    f.prepare()
    f.iterate(3)
    assert np.array_equal(f.estimate(f64=True), xa)
    f.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("f32", 5e-6), ("f64", 1e-13)])
def test_column_kernels_agree(small_ds, prec, tol):
    """The LDS-only and the edge-stage column kernels compute the same
    transforms (different radix order, so equal to rounding)."""
    from updc_amd.solver import Context, Exposure
    print(f"case: column kernels {prec}", flush=True)
    out = []
    for kind in (0, 1):
        c = Context(small_ds.W, small_ds.H, prec, column_kernel=kind)
        for k in range(len(small_ds.geoms)):
            c.add_exposure(Exposure(small_ds.sci[k], small_ds.wht[k], small_ds.psfs[k], small_ds.background[k],
                                    small_ds.pedestal[k], linkmap=small_ds.linkmaps[k]))
        c.prepare()
        c.iterate(3)
        out.append(c.estimate(f64=True))
        c.close()
    err = np.max(np.abs(out[0] - out[1])) / np.max(np.abs(out[1]))
    assert err < tol, err
