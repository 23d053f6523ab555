"""The HIP autograd path (updc_amd.autograd) against the FP64 torch
restatement (oracle/torch_ref.py, pinned by finite differences on the CPU),
against finite differences on the GPU, and across precisions."""
import numpy as np
import pytest
import torch

from test_autograd_ref import tiny_problem

pytestmark = pytest.mark.gpu


def _setup(prec):
    from updc_amd import autograd as A
    from updc_amd.solver import Context, Exposure
    p = tiny_problem()
    ds = p["ds"]
    c = Context(ds.W, ds.H, prec, fft_shape=(p["FW"], p["FH"]))
    for k in range(3):
        c.add_exposure(Exposure(ds.sci[k], ds.wht[k], ds.psfs[k], ds.background[k], ds.pedestal[k],
                                linkmap=ds.linkmaps[k]))
    c.prepare()
    ops = A.Operators(c)
    ex = A.exposures_from_context(c, ops, ds.sci, ds.wht, ds.pedestal)
    return p, c, ops, ex


def gpu_loss(ops, ex, x0, params, bs, K, checkpoint=True):
    from updc_amd import autograd as A
    pws = [A.wrap_psf(A.gaussian_mixture(params[k], 7), ops.FH, ops.FW) for k in range(3)]
    x = A.unrolled(x0, pws[:2], [bs[0], bs[1]], K, ops, ex, checkpoint=checkpoint)
    mu = A.forward_model(x, pws[2], bs[2], 2, ops, ex)
    return A.poisson_nll(mu, ex.yeff[2], ex.active[2])


def ref_loss_with(p, ex, x0, params, bs):
    """The torch restatement fed the GPU path's normalised weights (CPU, float64)."""
    from oracle import torch_ref as R
    m = p["model"]
    cpu = lambda t: t.detach().cpu().to(torch.float64)
    yeff = [cpu(t) for t in ex.yeff]
    w = [cpu(t) for t in ex.w]
    wm = [cpu(t) for t in ex.wmask]
    act = [t.detach().cpu() for t in ex.active]
    pws = [R.wrap_psf(R.gaussian_mixture(params[k], 7), p["FH"], p["FW"]) for k in range(3)]
    x = m.unrolled(x0, pws[:2], bs[:2], p["K"], yeff[:2], w[:2], wm[:2], act[:2], ex.ped[:2])
    mu = m.forward_model(x, pws[2], bs[2], 2, ex.ped[2])
    return R.poisson_nll(mu, yeff[2], act[2])


def _grads(L, leaves):
    return [g.detach().cpu().to(torch.float64) for g in torch.autograd.grad(L, leaves)]


def _rel(a, b):
    return float((a - b).norm() / max(float(b.norm()), 1e-300))


def test_operators_against_numpy():
    print("case: autograd operators vs numpy", flush=True)
    p, c, ops, ex = _setup("f64")
    dev = ops.device
    g = np.random.Generator(np.random.PCG64(2))
    img = g.normal(size=(32, 32))
    spec = ops.r2c(torch.as_tensor(img, device=dev))
    ref = np.fft.rfft2(img, s=(p["FH"], p["FW"]))
    got = spec.cpu().numpy()[:, :p["FH"]].T
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    back = ops.c2r(spec, p["FH"], p["FW"], 1.0 / ops.nf).cpu().numpy()
    pad = np.zeros((p["FH"], p["FW"]))
    pad[:32, :32] = img
    assert np.max(np.abs(back - pad)) <= 1e-12 * np.max(np.abs(img))
    m = p["model"]
    x = torch.as_tensor(img)
    for k in range(3):
        a = ops.gather(k, x.to(dev)).cpu()
        assert float((a - m.gather(k, x)).abs().max()) <= 1e-13 * float(x.abs().max()) * 4
        v = torch.as_tensor(g.normal(size=ops.nin[k]))
        s = ops.scatter(k, v.to(dev)).cpu()
        assert float((s - m.scatter(k, v, ex.wmask[k].cpu())).abs().max()) <= 1e-13 * float(v.abs().max()) * 4
    c.close()


def test_gradients_match_torch_reference_and_fd():
    print("case: autograd f64 vs torch reference and finite differences", flush=True)
    p, c, ops, ex = _setup("f64")
    dev = ops.device
    x0 = p["x0"].to(dev).requires_grad_(True)
    params = p["params"].to(dev).requires_grad_(True)
    bs = p["bs"].to(dev).requires_grad_(True)
    L = gpu_loss(ops, ex, x0, params, bs, p["K"])
    gx, gp, gb = _grads(L, (x0, params, bs))
    x0r = p["x0"].clone().requires_grad_(True)
    pr = p["params"].clone().requires_grad_(True)
    br = p["bs"].clone().requires_grad_(True)
    Lr = ref_loss_with(p, ex, x0r, pr, br)
    rx, rp, rb = _grads(Lr, (x0r, pr, br))
    print(f"  loss {float(L):.12g} ref {float(Lr):.12g}", flush=True)
    assert abs(float(L) - float(Lr)) <= 1e-11 * abs(float(Lr))
    for a, b in ((gx, rx), (gp, rp), (gb, rb)):
        assert _rel(a, b) <= 1e-9, _rel(a, b)
    # central finite differences on the GPU path itself
    with torch.no_grad():
        h = 1e-6
        gen = torch.Generator().manual_seed(9)
        v = torch.randn(p["x0"].shape, generator=gen, dtype=torch.float64).to(dev)
        f = lambda s: float(gpu_loss(ops, ex, x0.detach() + s * v, params.detach(), bs.detach(), p["K"]))
        # central differences carry the loss's own rounding, ~eps |L| / h
        floor = 8 * 2.2e-16 * abs(float(L)) / h
        fd = (f(h) - f(-h)) / (2 * h)
        an = float((gx.to(dev) * v).sum())
        assert abs(fd - an) <= 1e-5 * abs(an) + floor, (fd, an, floor)
        for idx in [(0, 0, 1), (1, 1, 2), (2, 0, 0)]:
            e = torch.zeros_like(params)
            e[idx] = 1.0
            f = lambda s: float(gpu_loss(ops, ex, x0.detach(), params.detach() + s * e, bs.detach(), p["K"]))
            fd = (f(h) - f(-h)) / (2 * h)
            assert abs(fd - float(gp[idx])) <= 1e-5 * abs(float(gp[idx])) + floor, (idx, fd, float(gp[idx]))
    c.close()


def test_checkpointing_is_exact():
    print("case: autograd checkpointing vs stored graph", flush=True)
    p, c, ops, ex = _setup("f64")
    dev = ops.device
    out = []
    for ck in (True, False):
        x0 = p["x0"].to(dev).requires_grad_(True)
        params = p["params"].to(dev).requires_grad_(True)
        bs = p["bs"].to(dev).requires_grad_(True)
        out.append(_grads(gpu_loss(ops, ex, x0, params, bs, p["K"], checkpoint=ck), (x0, params, bs)))
    for a, b in zip(*out):
        assert torch.equal(a, b)
    c.close()


def test_fp32_gradients_close_to_fp64():
    print("case: autograd f32 vs f64", flush=True)
    res = {}
    for prec in ("f64", "mixed"):
        p, c, ops, ex = _setup(prec)
        dt, dev = ops.dtype, ops.device
        x0 = p["x0"].to(dev, dt).requires_grad_(True)
        params = p["params"].to(dev, dt).requires_grad_(True)
        bs = p["bs"].to(dev, dt).requires_grad_(True)
        res[prec] = _grads(gpu_loss(ops, ex, x0, params, bs, p["K"]), (x0, params, bs))
        c.close()
    for a, b in zip(res["mixed"], res["f64"]):
        print(f"  rel {_rel(a, b):.2e}", flush=True)
        assert _rel(a, b) <= 2e-4
