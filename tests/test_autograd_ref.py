"""Pins the FP64 torch restatement of the differentiable model (CPU).

Central finite differences (h = 1e-6, float64) of the held-out Poisson loss
through K unrolled iterations, with respect to the initial estimate (random
directions), every PSF mixture parameter and every background; and the
adjoint identities of the blur and LinkMap operators.
"""
import numpy as np
import pytest
import torch

from conftest import oracle_dataset


def tiny_problem(K=3, seed=11):
    """Two exposures in the iteration plus a held-out third; FFT 40 x 40."""
    from oracle import torch_ref as R
    from updc_amd import synth
    ds = oracle_dataset(synth.small_config(n_exp=3, det=16, target=32, psf=7, seed=seed, rotation=4.0,
                                           dither=2.0, nsrc=6, sky=2.0, mask_fraction=0.05))
    FW = FH = 40
    model = R.RefModel(ds.W, ds.H, FW, FH, ds.linkmaps)
    dt = torch.float64
    ws = float(np.median(np.concatenate([w.ravel() for w in ds.wht])))
    yeff, w, wmask, active, ped = [], [], [], [], []
    for k in range(3):
        wk = torch.as_tensor(ds.wht[k].ravel() / ws, dtype=dt)
        yeff.append(torch.as_tensor(ds.sci[k].ravel(), dtype=dt) + ds.pedestal[k])
        w.append(wk)
        wmask.append((wk > 0).to(dt))
        active.append((wk > 0) & model.has_links(k))
        ped.append(float(ds.pedestal[k]))
    g = np.random.Generator(np.random.PCG64(seed))
    params = torch.tensor([[[1.0, 1.3, 0.1, -0.1], [0.4, 2.2, -0.3, 0.2]],
                           [[1.0, 1.1, -0.2, 0.0], [0.5, 2.0, 0.2, 0.3]],
                           [[1.0, 1.2, 0.0, 0.1], [0.3, 2.5, 0.0, 0.0]]], dtype=dt)
    bs = torch.tensor([0.05, -0.02, 0.01], dtype=dt)
    x0 = torch.as_tensor(np.clip(ds.truth, 0.1, None) + g.uniform(0, 0.5, ds.truth.shape), dtype=dt)
    return dict(ds=ds, model=model, FW=FW, FH=FH, K=K, yeff=yeff, w=w, wmask=wmask, active=active, ped=ped,
                params=params, bs=bs, x0=x0)


def ref_loss(p, x0, params, bs):
    from oracle import torch_ref as R
    m = p["model"]
    pws = [R.wrap_psf(R.gaussian_mixture(params[k], 7), p["FH"], p["FW"]) for k in range(3)]
    sl = slice(0, 2)
    x = m.unrolled(x0, pws[sl], bs[sl], p["K"], p["yeff"][sl], p["w"][sl], p["wmask"][sl], p["active"][sl],
                   p["ped"][sl])
    mu = m.forward_model(x, pws[2], bs[2], 2, p["ped"][2])
    return R.poisson_nll(mu, p["yeff"][2], p["active"][2])


@pytest.fixture(scope="module")
def prob():
    return tiny_problem()


def test_adjoint_identities(prob):
    m = prob["model"]
    from oracle import torch_ref as R
    g = torch.Generator().manual_seed(3)
    pw = R.wrap_psf(R.gaussian_mixture(prob["params"][0], 7), prob["FH"], prob["FW"])
    x = torch.rand((32, 32), generator=g, dtype=torch.float64)
    y = torch.rand((32, 32), generator=g, dtype=torch.float64)
    a = (m.blur(x, pw) * y).sum()
    b = (x * m.blur_t(y, pw)).sum()
    assert abs(a - b) <= 1e-12 * abs(a)
    u = torch.rand(16 * 16, generator=g, dtype=torch.float64)
    ones = torch.ones(16 * 16, dtype=torch.float64)
    a = (m.gather(0, x) * u).sum()
    b = (x * m.scatter(0, u, ones)).sum()
    assert abs(a - b) <= 1e-12 * abs(a)


def test_no_freezing_in_the_smooth_case(prob):
    """FD needs a differentiable point: no FreezeMu event, no clamped factor."""
    from oracle import torch_ref as R
    m = prob["model"]
    pws = [R.wrap_psf(R.gaussian_mixture(prob["params"][k], 7), prob["FH"], prob["FW"]) for k in range(2)]
    for k in range(2):
        raw = m.forward_model(prob["x0"], pws[k], prob["bs"][k], k, prob["ped"][k])
        assert bool((raw[prob["active"][k]] > 1e-3).all())


def _fd(f, h=1e-6):
    return (f(h) - f(-h)) / (2 * h)


def test_gradients_match_finite_differences(prob):
    x0 = prob["x0"].clone().requires_grad_(True)
    params = prob["params"].clone().requires_grad_(True)
    bs = prob["bs"].clone().requires_grad_(True)
    L = ref_loss(prob, x0, params, bs)
    gx, gp, gb = torch.autograd.grad(L, (x0, params, bs))
    # central differences carry the loss's own rounding, ~eps |L| / h
    floor = 8 * 2.2e-16 * abs(float(L)) / 1e-6
    with torch.no_grad():
        g = torch.Generator().manual_seed(5)
        for _ in range(3):
            v = torch.randn(x0.shape, generator=g, dtype=torch.float64)
            fd = _fd(lambda h: ref_loss(prob, prob["x0"] + h * v, prob["params"], prob["bs"]))
            an = float((gx * v).sum())
            assert abs(fd - an) <= 1e-5 * abs(an) + floor, (fd, an)
        for idx in [(0, 0, 0), (0, 0, 1), (0, 1, 2), (1, 0, 3), (1, 1, 1), (2, 0, 1), (2, 1, 2)]:
            e = torch.zeros_like(params)
            e[idx] = 1.0
            fd = _fd(lambda h: ref_loss(prob, prob["x0"], prob["params"] + h * e, prob["bs"]))
            an = float(gp[idx])
            assert abs(fd - an) <= 1e-5 * abs(an) + floor, (idx, fd, an)
        for k in range(3):
            e = torch.zeros_like(bs)
            e[k] = 1.0
            fd = _fd(lambda h: ref_loss(prob, prob["x0"], prob["params"], prob["bs"] + h * e))
            an = float(gb[k])
            assert abs(fd - an) <= 1e-5 * abs(an) + floor, (k, fd, an)
