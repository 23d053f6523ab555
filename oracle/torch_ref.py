"""FP64 torch restatement of the differentiable UPDC model -- TEST INFRASTRUCTURE ONLY.

Only tests/ and bench.py's checks import this module, as the checker of
updc_amd.autograd; the product never does.  It restates, with torch's own
FFT and index operations and torch's own autograd (no hand-written
backward):

  blur(x, pw)     = crop irfft2(rfft2(pad x) . rfft2(pw))            updc_core.c:753-776
  blur_t(img, pw) = crop irfft2(rfft2(pad img) . conj rfft2(pw))     updc_core.c:985-995, 1280-1287
  gather_k(img)   = sum_z frac_z img[t_z]                             updc_core.c:1154-1158
  scatter_k(v)    = S_k^T (v . [w > 0])                               updc_core.c:1171-1176
  one iteration   = FreezeMu detector loop + C/D update               updc_core.c:1152-1170, 1437-1457
  D               = sum_k blur_t(scatter_k(w_k), pw_k)                updc_core.c:997-1086

The LinkMap triples come from oracle.parse_linkmap.  Pinned by central
finite differences in float64 (tests/test_autograd_ref.py) and by the
adjoint identities <blur x, y> = <x, blur_t y>, <S u, v> = <u, S^T v>.
"""
from __future__ import annotations

import numpy as np
import torch

from . import oracle as O


class RefModel:
    def __init__(self, W: int, H: int, FW: int, FH: int, linkmaps, device="cpu", dtype=torch.float64):
        self.W, self.H, self.FW, self.FH = W, H, FW, FH
        self.device, self.dtype = torch.device(device), dtype
        self.links = []
        for img in linkmaps:
            lm = O.parse_linkmap(img)
            self.links.append(dict(
                nin=int(lm["nin"]),
                pix=torch.as_tensor(lm["pixel"], dtype=torch.long, device=self.device),
                tgt=torch.as_tensor(lm["target"], dtype=torch.long, device=self.device),
                frac=torch.as_tensor(lm["frac"].astype(np.float64), dtype=dtype, device=self.device)))

    def _pad(self, img):
        out = torch.zeros((self.FH, self.FW), dtype=img.dtype, device=img.device)
        out[:img.shape[0], :img.shape[1]] = img
        return out

    def blur(self, x, pw):
        X = torch.fft.rfft2(self._pad(x))
        P = torch.fft.rfft2(pw)
        return torch.fft.irfft2(X * P, s=(self.FH, self.FW))[:self.H, :self.W]

    def blur_t(self, img, pw):
        I = torch.fft.rfft2(self._pad(img))
        P = torch.fft.rfft2(pw)
        return torch.fft.irfft2(I * P.conj(), s=(self.FH, self.FW))[:self.H, :self.W]

    def gather(self, k, img):
        L = self.links[k]
        vals = L["frac"] * img.reshape(-1)[L["tgt"]]
        return torch.zeros(L["nin"], dtype=img.dtype, device=img.device).index_add(0, L["pix"], vals)

    def scatter(self, k, v, wmask):
        L = self.links[k]
        vals = L["frac"] * (v * wmask)[L["pix"]]
        return torch.zeros(self.H * self.W, dtype=v.dtype, device=v.device).index_add(0, L["tgt"], vals) \
            .reshape(self.H, self.W)

    def has_links(self, k):
        L = self.links[k]
        cnt = torch.zeros(L["nin"], dtype=torch.long, device=self.device)
        return cnt.index_add(0, L["pix"], torch.ones_like(L["pix"])) > 0

    # ---- the iteration
    def forward_model(self, x, pw, b, k, ped):
        return self.gather(k, self.blur(x, pw)) + b + ped

    def denominator(self, pws, w, wmask):
        return sum(self.blur_t(self.scatter(k, w[k], wmask[k]), pw) for k, pw in enumerate(pws))

    def rl_step(self, x, pws, bs, D, frozen, yeff, w, wmask, active, ped):
        Cn = torch.zeros_like(x)
        new_frozen = []
        for k, pw in enumerate(pws):
            raw = self.forward_model(x, pw, bs[k], k, ped[k])
            fz = frozen[k] | (active[k] & ~(raw.detach() > 0))
            live = active[k] & ~fz
            safe = torch.where(live, raw, torch.ones_like(raw))
            ratio = torch.where(live, yeff[k] / safe,
                                torch.where(active[k], torch.ones_like(raw), torch.zeros_like(raw)))
            Cn = Cn + self.blur_t(self.scatter(k, w[k] * ratio, wmask[k]), pw)
            new_frozen.append(fz)
        m = D > 1e-8
        f = torch.where(m, Cn / torch.where(m, D, torch.ones_like(D)), torch.zeros_like(D))
        f = torch.where(f < 0, torch.zeros_like(f), f)
        return torch.where(m, x * f, torch.zeros_like(x)), new_frozen

    def unrolled(self, x0, pws, bs, K, yeff, w, wmask, active, ped):
        D = self.denominator(pws, w, wmask)
        frozen = [torch.zeros_like(a) for a in active]
        x = x0
        for _ in range(K):
            x, frozen = self.rl_step(x, pws, bs, D, frozen, yeff, w, wmask, active, ped)
        return x


def wrap_psf(psf, FH, FW):
    """Normalised PSF with centre (h//2, w//2) at the origin of FH x FW (updc_core.c:790-843)."""
    h, w = psf.shape
    out = torch.zeros((FH, FW), dtype=psf.dtype, device=psf.device)
    for i in range(h):
        out[(i - h // 2 + FH) % FH, [(j - w // 2 + FW) % FW for j in range(w)]] = psf[i] / psf.sum()
    return out


def gaussian_mixture(params, size):
    """sum_j a_j exp(-((x-c-dx_j)^2 + (y-c-dy_j)^2) / (2 s_j^2)), params (ncomp, 4) = (a, s, dx, dy)."""
    c = (size - 1) / 2.0
    ax = torch.arange(size, dtype=params.dtype, device=params.device) - c
    y, x = ax[:, None], ax[None, :]
    return sum(params[j, 0] * torch.exp(-((x - params[j, 2]) ** 2 + (y - params[j, 3]) ** 2)
                                        / (2.0 * params[j, 1] ** 2)) for j in range(params.shape[0]))


def poisson_nll(mu, yeff, active):
    m = active & (mu > 0)
    mus = torch.where(m, mu, torch.ones_like(mu))
    return torch.where(m, mus - yeff * torch.log(mus), torch.zeros_like(mu)).sum()
