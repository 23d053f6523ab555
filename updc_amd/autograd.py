"""Differentiable UPDC forward model and reconstruction iteration on the GPU.

torch.autograd.Function wrappers over the library's device-pointer operators
(include/updc_hip.h, updc_op_*), each with a hand-written backward:

  blur(x, pw)       crop_HxW C2R(R2C(pad x) . R2C(pw)) / nf      A3, updc_core.c:753-776
  blur_t(img, pw)   crop_HxW C2R(R2C(pad img) . conj R2C(pw)) / nf  (the adjoint of blur in x, A5/A7)
  gather(img, k)    S_k img: sum_z frac_z img[t_z] per detector pixel  (updc_core.c:1154-1158)
  scatter(v, k)     S_k^T v over the pixels with w > 0           (updc_core.c:1171-1176)

pw is the PSF already normalised and wrapped into the FH x FW grid
(wrap_psf, the placement of updc_core.c:790-843), so its gradient flows to
any parameterisation built in torch.  The reconstruction iteration (A4/A8,
updc_core.c:1152-1176, 1437-1457) is composed from these operators; FreezeMu
masks and the update's guards are piecewise constants (not differentiated).
K unrolled iterations run under torch.utils.checkpoint, one segment per
iteration, so only the iterates are kept between forward and backward.

The element type is the context's: float64 for an "f64" context, float32
for "f32"/"mixed".  Tensors live on the context's device; every operator runs
on torch's current stream.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check


class Operators:
    """Device-pointer operators of a prepared Context."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.lib = ctx.lib
        fw, fh, nc, pitch, eb = (C.c_int() for _ in range(5))
        check(self.lib.updc_op_layout(ctx.h, C.byref(fw), C.byref(fh), C.byref(nc), C.byref(pitch), C.byref(eb)))
        self.FW, self.FH, self.NC, self.pitch = fw.value, fh.value, nc.value, pitch.value
        self.dtype = torch.float64 if eb.value == 8 else torch.float32
        self.cdtype = torch.complex128 if eb.value == 8 else torch.complex64
        self.W, self.H = ctx.W, ctx.H
        self.nf = float(self.FW) * float(self.FH)
        self.nin = list(ctx.nin)
        self.device = torch.device("cuda", ctx.device)

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _check(self, t: torch.Tensor, shape, dtype=None) -> torch.Tensor:
        if t.device != self.device:
            raise ValueError(f"tensor on {t.device}, operators on {self.device}")
        if t.dtype != (dtype or self.dtype):
            raise TypeError(f"tensor dtype {t.dtype}, operators need {dtype or self.dtype}")
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"tensor shape {tuple(t.shape)}, expected {tuple(shape)}")
        return t.contiguous()

    def r2c(self, img: torch.Tensor) -> torch.Tensor:
        """Half spectrum (NC, pitch) complex, column-major, of img zero-padded to FH x FW."""
        rows, cols = img.shape
        if rows > self.FH or cols > self.FW:
            raise ValueError("image larger than the FFT grid")
        img = self._check(img, (rows, cols))
        spec = torch.zeros((self.NC, self.pitch), dtype=self.cdtype, device=self.device)
        check(self.lib.updc_op_r2c(self.ctx.h, C.c_void_p(img.data_ptr()), rows, cols, cols,
                                   C.c_void_p(spec.data_ptr()), self._stream()))
        return spec

    def c2r(self, spec: torch.Tensor, rows: int, cols: int, scale: float) -> torch.Tensor:
        spec = self._check(spec, (self.NC, self.pitch), self.cdtype)
        out = torch.empty((rows, cols), dtype=self.dtype, device=self.device)
        check(self.lib.updc_op_c2r(self.ctx.h, C.c_void_p(spec.data_ptr()), C.c_void_p(out.data_ptr()), rows, cols,
                                   cols, float(scale), self._stream()))
        return out

    def gather(self, k: int, img: torch.Tensor) -> torch.Tensor:
        img = self._check(img, (self.H, self.W))
        out = torch.empty(self.nin[k], dtype=self.dtype, device=self.device)
        check(self.lib.updc_op_gather(self.ctx.h, k, C.c_void_p(img.data_ptr()), C.c_void_p(out.data_ptr()),
                                      self._stream()))
        return out

    def scatter(self, k: int, v: torch.Tensor) -> torch.Tensor:
        v = self._check(v, (self.nin[k],))
        out = torch.empty((self.H, self.W), dtype=self.dtype, device=self.device)
        check(self.lib.updc_op_scatter(self.ctx.h, k, C.c_void_p(v.data_ptr()), C.c_void_p(out.data_ptr()),
                                       self._stream()))
        return out


# ---------------------------------------------------------------- autograd Functions
class _Blur(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, pw, ops):
        X = ops.r2c(x)
        P = ops.r2c(pw)
        ctx.save_for_backward(X, P)
        ctx.ops = ops
        return ops.c2r(X * P, ops.H, ops.W, 1.0 / ops.nf)

    @staticmethod
    def backward(ctx, g):
        X, P = ctx.saved_tensors
        ops = ctx.ops
        G = ops.r2c(g)
        gx = ops.c2r(G * P.conj(), ops.H, ops.W, 1.0 / ops.nf) if ctx.needs_input_grad[0] else None
        gp = ops.c2r(G * X.conj(), ops.FH, ops.FW, 1.0 / ops.nf) if ctx.needs_input_grad[1] else None
        return gx, gp, None


class _BlurT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, img, pw, ops):
        I = ops.r2c(img)
        P = ops.r2c(pw)
        ctx.save_for_backward(I, P)
        ctx.ops = ops
        return ops.c2r(I * P.conj(), ops.H, ops.W, 1.0 / ops.nf)

    @staticmethod
    def backward(ctx, g):
        I, P = ctx.saved_tensors
        ops = ctx.ops
        G = ops.r2c(g)
        gi = ops.c2r(G * P, ops.H, ops.W, 1.0 / ops.nf) if ctx.needs_input_grad[0] else None
        gp = ops.c2r(I * G.conj(), ops.FH, ops.FW, 1.0 / ops.nf) if ctx.needs_input_grad[1] else None
        return gi, gp, None


class _Gather(torch.autograd.Function):
    @staticmethod
    def forward(ctx, img, k, ops, wmask):
        ctx.k, ctx.ops = k, ops
        ctx.save_for_backward(wmask)
        return ops.gather(k, img)

    @staticmethod
    def backward(ctx, g):
        (wmask,) = ctx.saved_tensors
        # the scatter covers the weighted pixels only; gradients arriving at
        # pixels with w <= 0 must be zero (every caller here masks them)
        return ctx.ops.scatter(ctx.k, g * wmask), None, None, None


class _Scatter(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v, k, ops, wmask):
        ctx.k, ctx.ops = k, ops
        ctx.save_for_backward(wmask)
        return ops.scatter(k, v)

    @staticmethod
    def backward(ctx, g):
        (wmask,) = ctx.saved_tensors
        return ctx.ops.gather(ctx.k, g) * wmask, None, None, None


def blur(x, pw, ops):
    return _Blur.apply(x, pw, ops)


def blur_t(img, pw, ops):
    return _BlurT.apply(img, pw, ops)


# ---------------------------------------------------------------- PSF helpers (torch)
def wrap_psf(psf: torch.Tensor, FH: int, FW: int, transposed: bool = False) -> torch.Tensor:
    """Normalised PSF placed into the FH x FW grid with its centre (h//2, w//2)
    at the origin, the placement of updc_core.c:790-843; differentiable."""
    q = psf.t() if transposed else psf
    h, w = q.shape
    yy = (torch.arange(h, device=psf.device) - h // 2 + FH) % FH
    xx = (torch.arange(w, device=psf.device) - w // 2 + FW) % FW
    out = torch.zeros((FH, FW), dtype=psf.dtype, device=psf.device)
    return out.index_put((yy[:, None].expand(h, w), xx[None, :].expand(h, w)), q / q.sum())


def gaussian_mixture(params: torch.Tensor, size: int) -> torch.Tensor:
    """Sum over components j of a_j exp(-((x-c-dx_j)^2 + (y-c-dy_j)^2) / (2 s_j^2)),
    params (ncomp, 4) = (a, s, dx, dy); unnormalised."""
    c = (size - 1) / 2.0
    ax = torch.arange(size, dtype=params.dtype, device=params.device) - c
    y, x = ax[:, None], ax[None, :]
    out = torch.zeros((size, size), dtype=params.dtype, device=params.device)
    for j in range(params.shape[0]):
        a, s, dx, dy = params[j]
        out = out + a * torch.exp(-((x - dx) ** 2 + (y - dy) ** 2) / (2.0 * s * s))
    return out


# ---------------------------------------------------------------- model and iteration
@dataclass
class Exposures:
    """Per-exposure constants of the iteration: y_eff = sci + ped, the
    normalised weights, the pedestal, and the active mask (w > 0 and at least
    one link), as the detector loop uses them (updc_core.c:1152-1170)."""
    yeff: list
    w: list
    ped: list
    active: list
    wmask: list


def exposures_from_context(ctx, ops: Operators, sci, wht, pedestal) -> Exposures:
    """Constants of the exposures added to `ctx` (raw sci/wht as given to it),
    with the context's weight normalisation and invalid-pixel rule
    (updc_core.c:378-424)."""
    sc = ctx.scalars()
    ws = float(sc["weight_scale"])
    dt, dev = ops.dtype, ops.device
    ones = torch.ones((ops.H, ops.W), dtype=dt, device=dev)
    out = Exposures([], [], [], [], [])
    for k in range(len(ops.nin)):
        s = np.asarray(sci[k], np.float32).ravel().copy()
        w = np.asarray(wht[k], np.float32).ravel().copy()
        bad = ~np.isfinite(s) | ~np.isfinite(w) | ~(w > 0)
        s[bad] = 0
        w[bad] = 0
        wt = torch.as_tensor((w / np.float32(ws)) if dt == torch.float32 else w.astype(np.float64) / ws,
                             dtype=dt, device=dev)
        st = torch.as_tensor(s, dtype=dt, device=dev)
        has = ops.gather(k, ones) > 0
        out.yeff.append(st + float(pedestal[k]))
        out.w.append(wt)
        out.ped.append(float(pedestal[k]))
        out.active.append((wt > 0) & has)
        out.wmask.append((wt > 0).to(dt))
    return out


def forward_model(x, pw, b, k, ops, ex: Exposures):
    """mu_k = S_k (P_k * x) + b_k + ped_k for every detector pixel of exposure k."""
    return _Gather.apply(blur(x, pw, ops), k, ops, ex.wmask[k]) + (b + ex.ped[k])


def denominator(pws, ops, ex: Exposures):
    """D = sum_k P_k^T S_k^T w_k (updc_core.c:997-1086), differentiable in the PSFs."""
    D = None
    for k, pw in enumerate(pws):
        d = blur_t(_Scatter.apply(ex.w[k], k, ops, ex.wmask[k]), pw, ops)
        D = d if D is None else D + d
    return D


def rl_step(x, pws, bs, D, frozen, ops, ex: Exposures):
    """One iteration: returns (x', frozen').  The FreezeMu masks and the update
    guards are taken as constants of the differentiation."""
    Cn = None
    new_frozen = []
    for k, pw in enumerate(pws):
        raw = forward_model(x, pw, bs[k], k, ops, ex)
        with torch.no_grad():
            fz = frozen[k] | (ex.active[k] & ~(raw > 0))
        live = ex.active[k] & ~fz
        ratio = torch.where(live, ex.yeff[k] / torch.where(live, raw, torch.ones_like(raw)),
                            torch.where(ex.active[k], torch.ones_like(raw), torch.zeros_like(raw)))
        c = blur_t(_Scatter.apply(ex.w[k] * ratio, k, ops, ex.wmask[k]), pw, ops)
        Cn = c if Cn is None else Cn + c
        new_frozen.append(fz)
    m = D > 1e-8
    f = torch.where(m, Cn / torch.where(m, D, torch.ones_like(D)), torch.zeros_like(D))
    f = torch.where(f < 0, torch.zeros_like(f), f)
    return torch.where(m, x * f, torch.zeros_like(x)), new_frozen


def unrolled(x0, pws, bs, K, ops, ex: Exposures, checkpoint: bool = True):
    """K unrolled iterations from x0 (FreezeMu masks start at zero); with
    checkpoint, one torch.utils.checkpoint segment per iteration."""
    from torch.utils.checkpoint import checkpoint as ckpt
    D = denominator(pws, ops, ex)
    frozen = [torch.zeros_like(a) for a in ex.active]
    x = x0
    n = len(pws)
    for _ in range(K):
        if checkpoint:
            def seg(x_, D_, *args, frozen=frozen):
                return rl_step(x_, list(args[:n]), list(args[n:]), D_, frozen, ops, ex)
            x, frozen = ckpt(seg, x, D, *pws, *bs, use_reentrant=False)
        else:
            x, frozen = rl_step(x, pws, bs, D, frozen, ops, ex)
    return x


def poisson_nll(mu, yeff, active):
    """sum over active pixels with mu > 0 of mu - y log mu."""
    m = active & (mu > 0)
    mus = torch.where(m, mu, torch.ones_like(mu))
    return torch.where(m, mus - yeff * torch.log(mus), torch.zeros_like(mu)).sum()
