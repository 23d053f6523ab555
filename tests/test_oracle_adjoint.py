"""Pins the oracle's adjoint half (CPU only).

RefSolver's adjoint (blur_adjoint: sum_k crop C2R(R2C(pad img_k) . conj P^_k)
/ nf, updc_core.c:985-995, 1280-1287) is the code path of every iteration's
correction and of the denominator.  It is fixed here by mathematics, not by
its own formula:

* the dot-product identity <blur_k x, y_k> = <x, blur_adjoint(y)_k> with an
  asymmetric PSF on a non-square grid and a non-square FFT (a conjugation or
  index mistake breaks it: the planted mistake below shows by how much);
* orc_spectrum_multiply / orc_accumulate_adjoint against exact complex
  products of the same float32 inputs (a sign or conjugation error is of the
  order of the product itself, far above the rounding bound);
* Parseval for the oracle's R2C glue;
* the whole f64 iteration against the FD-pinned torch restatement
  (oracle/torch_ref.py) on the same dataset, FreezeMu firing included.
"""
import numpy as np
import pytest
import torch

from conftest import oracle_dataset, oracle_exposures
from oracle import oracle as O


def _asym_solver(prec, W=40, H=24, fft=(70, 45)):
    """Two exposures whose PSFs are asymmetric (off-centre peak, skewed
    tail) on a non-square target, with a non-square FFT shape."""
    from updc_amd import synth
    ds = synth.small_config(n_exp=2, det=12, target=48, psf=7, seed=21)
    rng = np.random.default_rng(5)
    exps = []
    for k in range(2):
        psf = rng.uniform(0, 0.05, (9, 5)).astype(np.float32)
        psf[2 + k, 4] = 1.0
        psf[7, 0] = 0.4
        v = ds.vertices(k).copy()
        v[..., 0] *= W / 48.0
        v[..., 1] *= H / 48.0
        z = np.zeros((12, 12), np.float32)
        exps.append(dict(sci=z, wht=z + 1, psf=psf, background=0.0, pedestal=0.0,
                         linkmap=O.build_linkmap(v, W, H)))
    return O.RefSolver(W, H, exps, prec, fft_shape=fft)


def _dot_identity(s, rng):
    dt = np.float32 if s.prec == "f32" else np.float64
    x = rng.normal(size=(s.H, s.W)).astype(dt)
    ys = [rng.normal(size=(s.H, s.W)).astype(dt) for _ in s.ex]
    lhs = sum(float(np.sum(b.astype(np.float64) * y)) for b, y in zip(s.blur(x), ys))
    rhs = float(np.sum(x.astype(np.float64) * s.blur_adjoint(ys).astype(np.float64)))
    scale = sum(float(np.sum(np.abs(b.astype(np.float64) * y))) for b, y in zip(s.blur(x), ys))
    return abs(lhs - rhs) / scale


@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("f32", 2e-6)])
def test_blur_adjoint_is_the_adjoint_of_blur(prec, tol):
    s = _asym_solver(prec)
    assert (s.FW, s.FH) == (70, 45) and s.W != s.H
    rng = np.random.default_rng(17)
    for _ in range(3):
        assert _dot_identity(s, rng) < tol


def test_planted_conjugation_mistake_breaks_the_identity():
    """The same identity with conj() dropped from the oracle's accumulation
    (a mistake planted in a CPU copy of the oracle) fails by orders of
    magnitude, so the tolerance above has teeth."""
    s = _asym_solver("f64")

    def wrong(acc, img, e):
        acc += s._rfft(img) * e["P"]
    s._accumulate = wrong
    assert _dot_identity(s, np.random.default_rng(17)) > 1e-3


def test_spectrum_products_against_exact_complex_products():
    rng = np.random.default_rng(9)
    n = 4096
    a = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    b = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    d0 = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    a64, b64 = a.astype(np.complex128), b.astype(np.complex128)
    eps = 2.0 ** -24
    # out = a * b: each part is two rounded products and one rounded sum
    got = O.spectrum_multiply(a, b).astype(np.complex128)
    bound = 3 * eps * (np.abs(a64.real * b64.real) + np.abs(a64.imag * b64.imag)
                       + np.abs(a64.real * b64.imag) + np.abs(a64.imag * b64.real))
    err = np.abs(got - a64 * b64)
    assert np.all(err <= bound)
    # dest += a * conj(b)
    dest = d0.copy()
    O.accumulate_adjoint(dest, a, b)
    want = d0.astype(np.complex128) + a64 * np.conj(b64)
    bound_acc = bound + 2 * eps * (np.abs(want) + np.abs(d0))
    assert np.all(np.abs(dest.astype(np.complex128) - want) <= bound_acc)
    # a conjugation error would be |2 Im(...)|-sized: far outside the bound
    wrong = d0.astype(np.complex128) + a64 * b64
    assert np.median(np.abs(dest.astype(np.complex128) - wrong) / bound_acc) > 1e5


def test_parseval_of_the_oracle_r2c():
    s = _asym_solver("f64")
    rng = np.random.default_rng(2)
    x = rng.normal(size=(s.H, s.W))
    X = s._rfft(x)
    wgt = np.full(X.shape[1], 2.0)
    wgt[0] = 1.0
    if s.FW % 2 == 0:
        wgt[-1] = 1.0
    energy = float(np.sum(np.abs(X) ** 2 * wgt)) / s.nf
    assert abs(energy - float(np.sum(x * x))) < 1e-12 * float(np.sum(x * x))


def test_refsolver_f64_iteration_matches_the_torch_reference():
    """Three f64 iterations of RefSolver against oracle/torch_ref.RefModel's
    rl_step (torch FFT and index_add, FD-pinned) on the same dataset and
    denominator.  Every background is -10 (the reference forces FreezeMu with
    -100, tests/test_smoke.py:117-121; -10 freezes pixels too while the
    iteration stays well conditioned, so rounding stays at the 1e-14 level)."""
    from oracle import torch_ref as R
    from updc_amd import synth
    ds = oracle_dataset(synth.small_config(n_exp=2, det=20, target=40, psf=7, seed=31, nsrc=8))
    ex = oracle_exposures(ds)
    for e in ex:
        e["background"] = -10.0
    s = O.RefSolver(ds.W, ds.H, ex, "f64")
    m = R.RefModel(ds.W, ds.H, s.FW, s.FH, ds.linkmaps)
    dt = torch.float64
    pws, bs, yeff, w, wmask, active, ped = [], [], [], [], [], [], []
    for k, e in enumerate(s.ex):
        pws.append(R.wrap_psf(torch.as_tensor(e["psf"], dtype=dt), s.FH, s.FW))
        bs.append(float(e["background"]))
        wk = torch.as_tensor(e["w"], dtype=dt)
        w.append(wk)
        wmask.append((wk > 0).to(dt))
        active.append((wk > 0) & m.has_links(k))
        yeff.append(torch.as_tensor(e["s"], dtype=dt) + float(e["pedestal"]))
        ped.append(float(e["pedestal"]))
    D = m.denominator(pws, w, wmask)
    assert float((D - torch.as_tensor(s.D)).abs().max()) <= 1e-13 * float(D.abs().max())
    x = torch.as_tensor(s.x.copy())
    frozen = [torch.zeros_like(a) for a in active]
    rows, _ = s.iterate(3)
    for _ in range(3):
        x, frozen = m.rl_step(x, pws, bs, D, frozen, yeff, w, wmask, active, ped)
    assert any(r[2] > 0 for r in rows), "FreezeMu must fire in this case"
    assert [int(f.sum()) for f in frozen] == [rows[-2][3], rows[-1][3]]
    err = float(np.max(np.abs(x.numpy() - s.x))) / float(np.max(np.abs(s.x)))
    assert err <= 1e-12, err
