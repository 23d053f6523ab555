"""Pins the oracle's geometry and LinkMap builder (CPU only).

Checks, each chosen so a plausible slip fails one of them:
* the reference's own overlap.c, compiled into oracle/_ref from the sources
  under the reference tree (skipped where that build is absent), on random
  and adversarial quads -- bit-identical areas;
* the reference self-test constants (updc/c_core/updc_core.c:1545-1557);
* closed forms: axis-aligned quads on the pixel lattice, a 45-degree diamond,
  a half pixel; the area of a polygon is invariant under its own rotation of
  vertex order;
* an exact rational brute force of the clipped area (fractions.Fraction);
* LinkMap structure: ABI constants (updc/cache_abi.py:9-12), header layout,
  exact file size, DELTA16/ABS32 choice, link fractions summing to ~1 for
  interior pixels, links in ascending target order.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O

REF = O.ref_lib()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref (reference overlap.c) not built here")


def _quad(rng, scale=2.0, off=(0.0, 0.0)):
    c = rng.uniform(-3, 3, 2) + np.asarray(off)
    th = rng.uniform(0, 2 * np.pi)
    sx, sy = scale * rng.uniform(0.5, 1.5, 2)
    sh = rng.uniform(-0.3, 0.3)
    base = np.array([[-.5, -.5], [-.5, .5], [.5, .5], [.5, -.5]])
    m = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]]) @ np.array([[sx, sh], [0, sy]])
    return (base @ m.T + c).astype(np.float64)


def _ref_area(q, x, y):
    import ctypes as C
    a = np.ascontiguousarray(q, np.float64)
    return REF.updc_overlap_area(a.ctypes.data_as(C.POINTER(C.c_double)), int(x), int(y))


def _ref_poly(q):
    import ctypes as C
    a = np.ascontiguousarray(q, np.float64)
    return REF.updc_polygon_area(a.ctypes.data_as(C.POINTER(C.c_double)), a.shape[0])


@needs_ref
def test_overlap_bitexact_vs_reference_overlap_c():
    rng = np.random.default_rng(1)
    n = 0
    for trial in range(400):
        q = _quad(rng, off=rng.uniform(-1000, 4000, 2) if trial % 3 == 0 else (0, 0))
        cx, cy = np.floor(q.mean(0) + 0.5).astype(int)
        for y in range(cy - 2, cy + 3):
            for x in range(cx - 2, cx + 3):
                assert np.float32(O.overlap_area(q, x, y)) == np.float32(_ref_area(q, x, y))
                n += 1
        assert np.float32(O.polygon_area(q)) == np.float32(_ref_poly(q))
    assert n == 400 * 25


@needs_ref
def test_overlap_degenerate_vs_reference():
    # vertices exactly on pixel edges, zero-length edges, vertical/horizontal edges
    cases = [
        np.array([[-0.5, -0.5], [-0.5, 0.5], [0.5, 0.5], [0.5, -0.5]]),
        np.array([[0.0, -0.5], [0.0, 0.5], [0.0, 0.5], [1.0, -0.5]]),
        np.array([[-0.5, 0.0], [0.5, 0.0], [0.5, 1e-16], [-0.5, 1e-16]]),
        np.array([[-0.25, -0.25], [-0.25, 0.75], [0.75, 0.75], [0.75, -0.25]]),
        np.array([[10.5, 20.5], [10.5, 21.5], [11.5, 21.5], [11.5, 20.5]]),
    ]
    for q in cases:
        for y in range(-1, 23):
            for x in range(-1, 13):
                assert np.float32(O.overlap_area(q, x, y)) == np.float32(_ref_area(q, x, y))


def test_self_test_constants():
    unit = np.array([[1000.0, 2000.0], [1000.0, 2001.0], [1001.0, 2001.0], [1001.0, 2000.0]])
    assert abs(O.polygon_area(unit) - 1.0) <= 2e-6
    half = np.array([[-0.5, -0.5], [-0.5, 0.5], [0.0, 0.5], [0.0, -0.5]])
    assert abs(O.overlap_area(half, 0, 0) - 0.5) <= 2e-6


def test_closed_forms():
    # axis-aligned 2x2 quad on integer offsets covers four full pixels
    q = np.array([[-0.5, -0.5], [-0.5, 1.5], [1.5, 1.5], [1.5, -0.5]])
    for (x, y) in [(0, 0), (1, 0), (0, 1), (1, 1)]:
        assert O.overlap_area(q, x, y) == 1.0
    assert O.overlap_area(q, 2, 0) == 0.0
    assert O.polygon_area(q) == 4.0
    # diamond |x|+|y| <= 1 (area 2): the unit pixel's corners lie on its edges,
    # so pixel 0 is covered fully and each edge neighbour gets the triangle
    # (0.5,-0.5),(1,0),(0.5,0.5) of area 0.25; diagonal pixels get nothing
    d = np.array([[0.0, -1.0], [-1.0, 0.0], [0.0, 1.0], [1.0, 0.0]])
    assert abs(O.polygon_area(d) - 2.0) < 1e-6
    assert abs(O.overlap_area(d, 0, 0) - 1.0) < 1e-6
    for (x, y) in [(1, 0), (-1, 0), (0, 1), (0, -1)]:
        assert abs(O.overlap_area(d, x, y) - 0.25) < 1e-6
    assert O.overlap_area(d, 1, 1) == 0.0
    # vertex order rotation leaves the area unchanged
    rng = np.random.default_rng(2)
    for _ in range(50):
        q = _quad(rng)
        a = O.polygon_area(q)
        for r in range(1, 4):
            assert abs(O.polygon_area(np.roll(q, r, axis=0)) - a) <= 4e-6 * max(a, 1.0)


def _exact_clip_area(q, x, y):
    """Sutherland-Hodgman in exact rationals, then the exact shoelace."""
    poly = [(Fraction(float(a)), Fraction(float(b))) for a, b in q]
    lims = [Fraction(x) - Fraction(1, 2), Fraction(x) + Fraction(1, 2), Fraction(y) - Fraction(1, 2),
            Fraction(y) + Fraction(1, 2)]

    def inside(p, s):
        return [p[0] >= lims[0], p[0] <= lims[1], p[1] >= lims[2], p[1] <= lims[3]][s]

    def cross(a, b, s):
        if s < 2:
            t = (lims[s] - a[0]) / (b[0] - a[0])
            return (lims[s], a[1] + t * (b[1] - a[1]))
        t = (lims[s] - a[1]) / (b[1] - a[1])
        return (a[0] + t * (b[0] - a[0]), lims[s])

    for s in range(4):
        out = []
        if not poly:
            return Fraction(0)
        prev = poly[-1]
        for cur in poly:
            if inside(cur, s):
                if not inside(prev, s):
                    out.append(cross(prev, cur, s))
                out.append(cur)
            elif inside(prev, s):
                out.append(cross(prev, cur, s))
            prev = cur
        poly = out
    if len(poly) < 3:
        return Fraction(0)
    acc = sum(poly[i][0] * poly[(i + 1) % len(poly)][1] - poly[(i + 1) % len(poly)][0] * poly[i][1]
              for i in range(len(poly)))
    return abs(acc) / 2


def test_exact_rational_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(60):
        q = _quad(rng, scale=1.7)
        cx, cy = np.floor(q.mean(0) + 0.5).astype(int)
        for y in range(cy - 2, cy + 3):
            for x in range(cx - 2, cx + 3):
                exact = float(_exact_clip_area(q, x, y))
                got = O.overlap_area(q, x, y)
                # f32 shoelace after a local-origin shift: a few f32 ulps of the quad area
                assert abs(got - exact) <= 3e-6 * 4.0, (got, exact)


def test_linkmap_abi_and_layout():
    from updc_amd import synth
    ds = synth.small_config(n_exp=1, det=20, target=40, seed=5, rotation=20.0)
    img = O.build_linkmap(ds.vertices(0), 40, 40)
    lm = O.parse_linkmap(img)
    assert lm["magic"] == 0x324D4C55 and lm["version"] == 2 and lm["header_bytes"] == 64
    assert img[:4] == b"ULM2"
    assert lm["W"] == 40 and lm["H"] == 40 and lm["nin"] == 400
    assert lm["encoding"] == 2
    # exact size rule (updc_core.c:469-483)
    nl, nin = lm["nlinks"], lm["nin"]
    pos = 64 + 4 * (nin + 1) + 4 * nin + 2 * nl
    pos = (pos + 3) & ~3
    assert len(img) == pos + 4 * nl
    assert int(lm["offset"][-1]) == nl and np.all(np.diff(lm["offset"].astype(np.int64)) >= 0)
    # within each pixel, targets ascend (ty then tx loop order)
    for p in range(nin):
        t = lm["target"][lm["offset"][p]:lm["offset"][p + 1]]
        assert np.all(np.diff(t) > 0)
    # pixels whose quad lies inside the grid: fractions sum to ~1 (only links
    # <= 1e-5 of the pixel are dropped); pixels off the grid edge sum to less
    v = ds.vertices(0)
    corners = np.stack([v[:-1, :-1], v[1:, :-1], v[1:, 1:], v[:-1, 1:]], axis=0)
    inside = np.all((corners > -0.5) & (corners < 39.5), axis=(0, 3)).ravel()
    sums = np.bincount(lm["pixel"], lm["frac"].astype(np.float64), nin)
    assert inside.sum() > 100
    assert np.all(np.abs(sums[inside] - 1.0) < 1e-4)
    assert np.all(sums[~inside] < 1.0 + 1e-4)


def test_linkmap_abs32_rule():
    """A target row wider than 65535 forces ABS32 (docs/V17_VALIDATION.md)."""
    from updc_amd import synth
    g = synth.Geometry(8, 3, 2.0, 10.0, (0.3, 0.2), None, (100.0, 3.5))
    img = O.build_linkmap(g.vertices(), 70000, 8)
    lm = O.parse_linkmap(img)
    assert lm["encoding"] == 1
    assert len(img) == 64 + 4 * (lm["nin"] + 1) + 8 * lm["nlinks"]
    img16 = O.build_linkmap(g.vertices(), 600, 8)
    assert O.parse_linkmap(img16)["encoding"] == 2


def test_linkmap_frac_recomputed_through_reference_overlap():
    if REF is None:
        pytest.skip("oracle/_ref not built")
    from updc_amd import synth
    ds = synth.small_config(n_exp=1, det=24, target=48, seed=9, rotation=30.0, distort=0.02)
    v = ds.vertices(0)
    img = O.build_linkmap(v, 48, 48)
    lm = O.parse_linkmap(img)
    for z in range(lm["nlinks"]):
        p, t = lm["pixel"][z], lm["target"][z]
        x, y = p % 24, p // 24
        q = np.array([v[y, x], v[y + 1, x], v[y + 1, x + 1], v[y, x + 1]])
        a = np.float32(_ref_area(q, t % 48, t // 48))
        full = np.float32(_ref_poly(q))
        assert a > np.float32(1e-5)
        assert np.float32(a / full) == lm["frac"][z]


def test_fft_shape_rules():
    assert O.next_fast_size(5526) == 5600
    assert O.next_fast_size(574) == 576
    assert O.next_fast_size(7574) == 7680
    assert O.next_fast_size(2302) == 2304
    shapes = O.fft_shape_list(5526, 5526)
    assert shapes[0] == (5600, 5600)
    assert (5670, 5600) in shapes and (5600, 5670) in shapes
    assert len(shapes) == len(set(shapes)) <= 9
