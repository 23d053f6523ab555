"""CPU oracle of the UPDC v17 iteration -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker; the product (updc_amd) never does.

Two restatements of UPDC_ALL (updc/c_core/updc_core.c:1289-1543):

* ``RefSolver(precision="f32")`` -- the reference's float32 arithmetic.  The
  byte/integer and fmaf-chain stages run in oracle/updc_oracle.c (same loop
  order, -ffp-contract=off); the FFTs use numpy's single-precision pocketfft
  in place of FFTW3f (FFTW is absent here: its outputs are the one part no
  fixture pins).
* ``RefSolver(precision="f64")`` -- the same algorithm in float64 (numpy),
  the FP64 oracle of the precision tiers.

Pinning of each piece: see oracle/updc_oracle.c's header and tests/test_oracle_*.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np
import scipy.fft as _sfft

HERE = Path(__file__).resolve().parent
_WORKERS = max(1, min(16, os.cpu_count() or 1))
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "liboverlap_ref.so"
REF_PROJECTOR = HERE / "_ref" / "updc_psf_projector"

_lib = None
_ref = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        D, F, U8, U32, U64 = (C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_uint8),
                              C.POINTER(C.c_uint32), C.POINTER(C.c_uint64))
        sig = {
            "orc_polygon_area": (C.c_float, [D, C.c_int]),
            "orc_overlap_area": (C.c_float, [D, C.c_int, C.c_int]),
            "orc_next_fast_size": (C.c_int, [C.c_int]),
            "orc_fast_candidates": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.c_int]),
            "orc_fft_shape_list": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int)]),
            "orc_linkmap_bytes": (C.c_size_t, [C.c_uint64, C.c_uint64, C.c_uint32]),
            "orc_linkmap_pass1": (C.c_int, [D, C.c_int, C.c_int, C.c_int, C.c_int, U32, U32, U64, U32]),
            "orc_linkmap_write": (C.c_int, [D, C.c_int, C.c_int, C.c_int, C.c_int, U32, U32, C.c_uint64, C.c_uint32,
                                            U8]),
            "orc_fused_detector": (C.c_int, [U8, C.c_size_t, C.c_size_t, F, F, U8, F, C.c_float, C.c_float, F, U64,
                                             U64, F, F]),
            "orc_sensitivity_scatter": (None, [U8, C.c_size_t, C.c_size_t, F, F]),
            "orc_forward_gather": (None, [U8, C.c_size_t, F, F]),
            "orc_drizzle": (C.c_int, [C.POINTER(U8), C.c_int, C.c_size_t, C.c_size_t, F, F, F, C.c_int, F, F]),
            "orc_weight_scale": (C.c_float, [F, C.c_int, C.c_size_t]),
            "orc_normalize_weights": (C.c_int, [F, F, C.c_int, C.c_size_t, C.c_float, F, U64]),
            "orc_psf_sum": (C.c_float, [F, C.c_size_t]),
            "orc_psf_wrap": (None, [F, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, F]),
            "orc_spectrum_multiply": (None, [F, F, F, C.c_size_t]),
            "orc_accumulate_adjoint": (None, [F, F, F, C.c_size_t]),
            "orc_crop_scale": (None, [F, C.c_int, C.c_int, C.c_int, C.c_int, F]),
            "orc_update": (C.c_int, [F, F, F, C.c_size_t, C.c_float, F, U64, U64]),
            "orc_psf_extent": (C.c_int, [D, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "orc_psf_project": (C.c_uint64, [F, C.c_int, C.c_int, D, C.c_int, C.c_int, F]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ref_lib() -> C.CDLL | None:
    """The reference's own overlap.c compiled into oracle/_ref (None if absent)."""
    global _ref
    if _ref is None and REF_LIB.exists():
        R = C.CDLL(str(REF_LIB))
        R.updc_polygon_area.restype = C.c_float
        R.updc_polygon_area.argtypes = [C.POINTER(C.c_double), C.c_int]
        R.updc_overlap_area.restype = C.c_float
        R.updc_overlap_area.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int]
        _ref = R
    return _ref


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


# ---------------------------------------------------------------- geometry
def polygon_area(xy: np.ndarray) -> float:
    a = np.ascontiguousarray(xy, np.float64)
    return lib().orc_polygon_area(_p(a, C.c_double), a.size // 2)


def overlap_area(quad: np.ndarray, x: int, y: int) -> float:
    a = np.ascontiguousarray(quad, np.float64)
    return lib().orc_overlap_area(_p(a, C.c_double), int(x), int(y))


def next_fast_size(n: int) -> int:
    return lib().orc_next_fast_size(int(n))


def fft_shape_list(min_w: int, min_h: int) -> list[tuple[int, int]]:
    out = (C.c_int * 32)()
    n = lib().orc_fft_shape_list(int(min_w), int(min_h), out)
    return [(out[2 * i], out[2 * i + 1]) for i in range(n)]


def psf_project(psf: np.ndarray, vertices: np.ndarray) -> tuple[np.ndarray, int]:
    """Polygon push-forward of a PSF (psf_projector.c:71-146): the projected
    (2hy+1) x (2hx+1) float32 PSF and the overlap count."""
    p = np.ascontiguousarray(psf, np.float32)
    v = np.ascontiguousarray(vertices, np.float64)
    ih, iw = p.shape
    if v.shape != (ih + 1, iw + 1, 2):
        raise ValueError("vertex lattice must be (ih+1, iw+1, 2)")
    L = lib()
    hx, hy = C.c_int(), C.c_int()
    if L.orc_psf_extent(_p(v, C.c_double), iw, ih, C.byref(hx), C.byref(hy)):
        raise FloatingPointError("non-finite PSF mapped vertex")
    out = np.empty((2 * hy.value + 1, 2 * hx.value + 1), np.float32)
    n = L.orc_psf_project(_p(p, C.c_float), iw, ih, _p(v, C.c_double), hx.value, hy.value, _p(out, C.c_float))
    return out, int(n)


def build_linkmap(vertices: np.ndarray, W: int, H: int) -> bytes:
    """ULM2 image of updc_core.c:530-609 (two passes, encoding rule)."""
    v = np.ascontiguousarray(vertices, np.float64)
    ih, iw = v.shape[0] - 1, v.shape[1] - 1
    nin = iw * ih
    off = np.zeros(nin + 1, np.uint32)
    base = np.zeros(nin, np.uint32)
    nl = C.c_uint64()
    enc = C.c_uint32()
    L = lib()
    if L.orc_linkmap_pass1(_p(v, C.c_double), iw, ih, W, H, _p(off, C.c_uint32), _p(base, C.c_uint32),
                           C.byref(nl), C.byref(enc)):
        raise ValueError("LinkMap exceeds uint32 offset capacity")
    nbytes = L.orc_linkmap_bytes(nin, nl.value, enc.value)
    out = np.zeros(nbytes, np.uint8)
    rc = L.orc_linkmap_write(_p(v, C.c_double), iw, ih, W, H, _p(off, C.c_uint32), _p(base, C.c_uint32), nl.value,
                             enc.value, _p(out, C.c_uint8))
    if rc:
        raise ValueError(f"LinkMap second pass failed ({rc})")
    return out.tobytes()


def parse_linkmap(img: bytes) -> dict:
    """Views of a ULM2 image (updc_core.c:645-662) plus expanded targets."""
    b = np.frombuffer(img, np.uint8)
    hdr = b[:64].view(np.uint32)
    magic, version, enc, hb = (int(v) for v in hdr[:4])
    nin = int(b[16:24].view(np.uint64)[0])
    nl = int(b[24:32].view(np.uint64)[0])
    W, H = int(hdr[8]), int(hdr[9])
    pos = 64
    off = b[pos:pos + 4 * (nin + 1)].view(np.uint32)
    pos += 4 * (nin + 1)
    if enc == 2:
        base = b[pos:pos + 4 * nin].view(np.uint32)
        pos += 4 * nin
        d16 = b[pos:pos + 2 * nl].view(np.uint16)
        pos = (pos + 2 * nl + 3) & ~3
        pix = np.repeat(np.arange(nin, dtype=np.int64), np.diff(off.astype(np.int64)))
        tgt = base[pix].astype(np.int64) + d16.astype(np.int64)
    else:
        t32 = b[pos:pos + 4 * nl].view(np.uint32)
        pos += 4 * nl
        tgt = t32.astype(np.int64)
        base = None
    frac = b[pos:pos + 4 * nl].view(np.float32)
    return dict(magic=magic, version=version, encoding=enc, header_bytes=hb, nin=nin, nlinks=nl, W=W, H=H,
                offset=off, base=base, target=tgt, frac=frac,
                pixel=np.repeat(np.arange(nin, dtype=np.int64), np.diff(off.astype(np.int64))))


# ---------------------------------------------------------------- stage wrappers (f32)
def fused_detector(img: bytes, nout: int, sci, w, frozen, blur, mean_offset: float, pedestal: float):
    b = np.frombuffer(img, np.uint8)
    nin = sci.size
    sc = np.zeros(nout, np.float32)
    wr = np.zeros(nin, np.float32)
    nn, nt = C.c_uint64(), C.c_uint64()
    mn = C.c_float()
    s32 = np.ascontiguousarray(sci, np.float32).ravel()
    w32 = np.ascontiguousarray(w, np.float32).ravel()
    bl = np.ascontiguousarray(blur, np.float32).ravel()
    rc = lib().orc_fused_detector(_p(b, C.c_uint8), nin, nout, _p(s32, C.c_float), _p(w32, C.c_float),
                                  _p(frozen, C.c_uint8), _p(bl, C.c_float), C.c_float(mean_offset),
                                  C.c_float(pedestal), _p(sc, C.c_float), C.byref(nn), C.byref(nt), C.byref(mn),
                                  _p(wr, C.c_float))
    if rc:
        raise FloatingPointError({-1: "non-finite forward mean", -2: "non-finite y/raw_mu ratio"}[rc])
    return wr, sc, (nn.value, nt.value, mn.value)


def sensitivity_scatter(img: bytes, nout: int, w) -> np.ndarray:
    b = np.frombuffer(img, np.uint8)
    w32 = np.ascontiguousarray(w, np.float32).ravel()
    out = np.zeros(nout, np.float32)
    lib().orc_sensitivity_scatter(_p(b, C.c_uint8), w32.size, nout, _p(w32, C.c_float), _p(out, C.c_float))
    return out


def forward_gather(img: bytes, nin: int, blur) -> np.ndarray:
    b = np.frombuffer(img, np.uint8)
    bl = np.ascontiguousarray(blur, np.float32).ravel()
    out = np.zeros(nin, np.float32)
    lib().orc_forward_gather(_p(b, C.c_uint8), nin, _p(bl, C.c_float), _p(out, C.c_float))
    return out


def weight_scale(w_all: list[np.ndarray]) -> float:
    a = np.ascontiguousarray(np.concatenate([np.asarray(w, np.float32).ravel() for w in w_all]))
    return float(lib().orc_weight_scale(_p(a, C.c_float), len(w_all), w_all[0].size))


def psf_sum(psf) -> float:
    a = np.ascontiguousarray(psf, np.float32).ravel()
    return float(lib().orc_psf_sum(_p(a, C.c_float), a.size))


def psf_wrap(psf, transposed: int, FW: int, FH: int, inv_sum: float) -> np.ndarray:
    a = np.ascontiguousarray(psf, np.float32)
    ph, pw = a.shape
    out = np.zeros((FH, FW), np.float32)
    lib().orc_psf_wrap(_p(a, C.c_float), pw, ph, int(transposed), FW, FH, C.c_float(inv_sum), _p(out, C.c_float))
    return out


def update(x, Cc, D, max_sens: float):
    x = np.ascontiguousarray(x, np.float32).copy()
    c32 = np.ascontiguousarray(Cc, np.float32)
    d32 = np.ascontiguousarray(D, np.float32)
    s = C.c_float()
    nn, bad = C.c_uint64(), C.c_uint64()
    rc = lib().orc_update(_p(x, C.c_float), _p(c32, C.c_float), _p(d32, C.c_float), x.size, C.c_float(max_sens),
                          C.byref(s), C.byref(nn), C.byref(bad))
    return rc, x, s.value, nn.value, bad.value


def spectrum_multiply(src, ker) -> np.ndarray:
    s = np.ascontiguousarray(src, np.complex64)
    k = np.ascontiguousarray(ker, np.complex64)
    out = np.empty_like(s)
    lib().orc_spectrum_multiply(_p(s, C.c_float), _p(k, C.c_float), _p(out, C.c_float), s.size)
    return out


def accumulate_adjoint(dest: np.ndarray, src, ker) -> None:
    s = np.ascontiguousarray(src, np.complex64)
    k = np.ascontiguousarray(ker, np.complex64)
    lib().orc_accumulate_adjoint(_p(dest, C.c_float), _p(s, C.c_float), _p(k, C.c_float), s.size)


# ---------------------------------------------------------------- full solve
class RefSolver:
    """UPDC_ALL restated (setup + iterations) on CPU.

    exposures: list of dicts with sci, wht, psf (float32 arrays), background,
    pedestal, and either 'linkmap' (ULM2 bytes) or 'vertices'.
    """

    def __init__(self, W: int, H: int, exposures: list[dict], precision: str = "f32", fft_shape=None,
                 psf_transposed: int = 0, workers: int = 1):
        self.W, self.H, self.prec = W, H, precision
        self.N = len(exposures)
        self.workers = max(1, min(workers, self.N))
        self.ex = exposures
        self.t = psf_transposed
        self.nout = W * H
        for e in exposures:
            if "linkmap" not in e or e["linkmap"] is None:
                e["linkmap"] = build_linkmap(e["vertices"], W, H)
            e["lm"] = parse_linkmap(e["linkmap"])
        pad_w = max((e["psf"].shape[0] if psf_transposed else e["psf"].shape[1]) for e in exposures)
        pad_h = max((e["psf"].shape[1] if psf_transposed else e["psf"].shape[0]) for e in exposures)
        if fft_shape is None:
            fw = W + pad_w - 1
            while True:
                fw = next_fast_size(fw)
                if fw % 2 == 0:
                    break
                fw += 1
            fft_shape = (fw, next_fast_size(H + pad_h - 1))
        self.FW, self.FH = fft_shape
        self.nf = self.FW * self.FH
        self._setup()

    # -- FFT helpers (pad top-left, crop, 1/nf like updc_core.c:728-744).
    # f64 transforms run on scipy's threaded pocketfft (bit-identical to
    # numpy's for float64, several times faster on many cores); f32 ones on
    # numpy's single-precision pocketfft.
    def _rfft(self, img):
        dt = np.float32 if self.prec == "f32" else np.float64
        pad = np.zeros((self.FH, self.FW), dt)
        pad[:img.shape[0], :img.shape[1]] = img
        if self.prec == "f32":
            return np.fft.rfft2(pad)
        return _sfft.rfft2(pad, workers=_WORKERS)

    def _irfft_crop(self, spec):
        if self.prec == "f32":
            full = np.fft.irfft2(spec, s=(self.FH, self.FW), norm="forward")
        else:
            full = _sfft.irfft2(spec, s=(self.FH, self.FW), norm="forward", workers=_WORKERS)
        if self.prec == "f32":
            return (full[:self.H, :self.W].astype(np.float32) * np.float32(np.float32(1.0) / np.float32(self.nf)))
        return full[:self.H, :self.W] * (1.0 / self.nf)

    def _setup(self):
        f32 = self.prec == "f32"
        # weights (updc_core.c:378-424)
        self.weight_scale = weight_scale([e["wht"] for e in self.ex])
        for e in self.ex:
            s = np.asarray(e["sci"], np.float32).ravel().copy()
            w = np.asarray(e["wht"], np.float32).ravel().copy()
            bad = ~np.isfinite(s) | ~np.isfinite(w) | ~(w > 0)
            s[bad] = 0
            w[bad] = 0
            if f32:
                w = (w / np.float32(self.weight_scale)).astype(np.float32)
                e["bg_f"] = np.float32(e["background"])
                e["ped_f"] = np.float32(e["pedestal"])
                e["mo_f"] = np.float32(e["bg_f"] + e["ped_f"])
            else:
                w = w.astype(np.float64) / float(self.weight_scale)
                s = s.astype(np.float64)
            e["s"], e["w"] = s, w
            e["frozen"] = np.zeros(s.size, np.uint8)
        # PSF spectra (updc_core.c:777-843)
        for e in self.ex:
            if f32:
                sm = psf_sum(e["psf"])
                img = psf_wrap(e["psf"], self.t, self.FW, self.FH, float(np.float32(1.0) / np.float32(sm)))
                e["P"] = np.fft.rfft2(img).astype(np.complex64)
            else:
                p = np.asarray(e["psf"], np.float64)
                img = np.zeros((self.FH, self.FW))
                q = p.T if self.t else p
                fh, fw = q.shape
                yy = (np.arange(fh) - fh // 2 + self.FH) % self.FH
                xx = (np.arange(fw) - fw // 2 + self.FW) % self.FW
                img[np.ix_(yy, xx)] = q / p.sum()
                e["P"] = _sfft.rfft2(img, workers=_WORKERS)
        # initial drizzle (updc_core.c:905-983)
        if f32:
            L = lib()
            imgs = [np.frombuffer(e["linkmap"], np.uint8) for e in self.ex]
            arr = (C.POINTER(C.c_uint8) * self.N)(*[_p(a, C.c_uint8) for a in imgs])
            s_all = np.ascontiguousarray(np.concatenate([e["s"] for e in self.ex]), np.float32)
            w_all = np.ascontiguousarray(np.concatenate([e["w"] for e in self.ex]), np.float32)
            bg = np.array([e["bg_f"] for e in self.ex], np.float32)
            est = np.zeros(self.nout, np.float32)
            cov = np.zeros(self.nout, np.float32)
            nin = self.ex[0]["s"].size
            if any(e["s"].size != nin for e in self.ex):
                raise ValueError("the f32 oracle drizzle needs equal detector sizes")
            rc = L.orc_drizzle(arr, self.N, nin, self.nout, _p(s_all, C.c_float), _p(w_all, C.c_float),
                               _p(bg, C.c_float), self.workers, _p(est, C.c_float), _p(cov, C.c_float))
            if rc:
                raise FloatingPointError("non-finite signed initial estimate")
            self.x = est.reshape(self.H, self.W)
            self.coverage = cov.reshape(self.H, self.W)
        else:
            cov = np.zeros(self.nout)
            dat = np.zeros(self.nout)
            bkg = np.zeros(self.nout)
            for e in self.ex:
                lm = e["lm"]
                q = lm["frac"].astype(np.float64) * e["w"][lm["pixel"]]
                act = e["w"][lm["pixel"]] > 0
                cov += np.bincount(lm["target"][act], q[act], self.nout)
                dat += np.bincount(lm["target"][act], (q * e["s"][lm["pixel"]])[act], self.nout)
                bkg += np.bincount(lm["target"][act], (q * float(e["background"]))[act], self.nout)
            x0 = np.zeros(self.nout)
            m = cov > 1e-12
            x0[m] = dat[m] / cov[m] - bkg[m] / cov[m]
            self.x = x0.reshape(self.H, self.W)
            self.coverage = cov.reshape(self.H, self.W)
        # denominator (updc_core.c:1026-1086)
        if f32:
            self.D = self.blur_adjoint([sensitivity_scatter(e["linkmap"], self.nout, e["w"]).reshape(self.H, self.W)
                                        for e in self.ex])
        else:
            self.D = self.blur_adjoint([self._scatter64(e, e["w"]) for e in self.ex])
        self.max_sens = float(self.D.max())
        if not (self.max_sens > 0):
            raise FloatingPointError("invalid H^T W sensitivity")

    def _scatter64(self, e, v):
        lm = e["lm"]
        act = e["w"][lm["pixel"]] > 0
        return np.bincount(lm["target"][act], (lm["frac"].astype(np.float64) * v[lm["pixel"]])[act],
                           self.nout).reshape(self.H, self.W)

    def _gather64(self, e, blur):
        lm = e["lm"]
        contrib = lm["frac"].astype(np.float64) * blur.ravel()[lm["target"]]
        return np.bincount(lm["pixel"], contrib, e["s"].size)

    def _new_acc(self) -> np.ndarray:
        return np.zeros((self.FH, self.FW // 2 + 1), np.complex64 if self.prec == "f32" else np.complex128)

    def _accumulate(self, acc: np.ndarray, img: np.ndarray, e: dict) -> None:
        """acc += R2C(pad img) . conj(P^_k): accumulate_adjoint_spectrum
        (updc_core.c:985-995), f32 without FMA in the C restatement."""
        if self.prec == "f32":
            accumulate_adjoint(acc, self._rfft(img), e["P"])
        else:
            acc += self._rfft(img) * np.conj(e["P"])

    def blur_adjoint(self, imgs) -> np.ndarray:
        """sum_k crop(C2R(R2C(pad img_k) . conj(P^_k))) / nf: the adjoint half
        of the iteration (updc_core.c:985-995, 1280-1287) and of the
        denominator (updc_core.c:1026-1086), the adjoint of blur()."""
        acc = self._new_acc()
        for e, img in zip(self.ex, imgs):
            self._accumulate(acc, np.asarray(img).reshape(self.H, self.W), e)
        return self._irfft_crop(acc)

    def blur(self, x=None) -> list[np.ndarray]:
        X = self._rfft(self.x if x is None else x)
        out = []
        for e in self.ex:
            prod = spectrum_multiply(X, e["P"]) if self.prec == "f32" else X * e["P"]
            out.append(self._irfft_crop(prod))
        return out

    def iterate(self, n: int = 1):
        rows, info = [], []
        for it in range(n):
            X = self._rfft(self.x)
            f32 = self.prec == "f32"
            acc = self._new_acc()
            for k, e in enumerate(self.ex):
                prod = spectrum_multiply(X, e["P"]) if f32 else X * e["P"]
                bl = self._irfft_crop(prod)
                if f32:
                    wr, sc, (nn, nt, mn) = fused_detector(e["linkmap"], self.nout, e["s"], e["w"], e["frozen"], bl,
                                                          float(e["mo_f"]), float(e["ped_f"]))
                    self._accumulate(acc, sc.reshape(self.H, self.W), e)
                else:
                    lm = e["lm"]
                    act = (e["w"] > 0) & (np.diff(lm["offset"].astype(np.int64)) > 0)
                    pred = self._gather64(e, bl)
                    raw = pred + (float(e["background"]) + float(e["pedestal"]))
                    if not np.all(np.isfinite(raw[act])):
                        raise FloatingPointError("non-finite forward mean")
                    newly = act & (e["frozen"] == 0) & ~(raw > 0)
                    e["frozen"][newly] = 1
                    fz = act & (e["frozen"] == 1)
                    ratio = np.ones_like(raw)
                    live = act & ~fz
                    ratio[live] = (e["s"][live] + float(e["pedestal"])) / raw[live]
                    wr = np.where(act, e["w"] * ratio, 0.0)
                    nn, nt = int(newly.sum()), int(fz.sum())
                    mn = float(raw[act].min()) if act.any() else float("inf")
                    self._accumulate(acc, self._scatter64(e, wr), e)
                rows.append((it + 1, k + 1, nn, nt, mn))
            Cc = self._irfft_crop(acc)
            if f32:
                rc, x, s, nneg, bad = update(self.x, Cc, self.D, self.max_sens)
                if rc:
                    raise FloatingPointError(f"update guard {rc} at {bad}")
                self.x = x.reshape(self.H, self.W)
            else:
                x = self.x.ravel().copy()
                Dv, Cv = self.D.ravel(), Cc.ravel()
                m = Dv > 1e-8
                f = np.zeros_like(x)
                f[m] = Cv[m] / Dv[m]
                if np.any(m & (f < 0) & (Cv < -1e-6 * self.max_sens)):
                    raise FloatingPointError("substantially negative multiplicative correction")
                f = np.where(f < 0, 0.0, f)
                x = np.where(m, x * f, 0.0)
                self.x = x.reshape(self.H, self.W)
                s, nneg = float(x.sum()), int((x < 0).sum())
            info.append(dict(iteration=it + 1, signed_sum=float(s), n_negative=int(nneg)))
        return rows, info


def compare_stats(ref: np.ndarray, new: np.ndarray) -> dict:
    """Metrics of the reference's tools/compare_fits.py:10-29 plus the
    normalised L-infinity max|d| / max|ref|."""
    a = np.asarray(ref, np.float64)
    b = np.asarray(new, np.float64)
    d = b - a
    scale = max(float(np.nanmax(np.abs(a))), 1e-30)
    mask = np.abs(a) > 1e-8 * scale
    rel = np.abs(d[mask] / a[mask]) if np.any(mask) else np.array([0.0])
    return dict(max_abs=float(np.nanmax(np.abs(d))), rms=float(np.sqrt(np.nanmean(d * d))),
                max_rel_masked=float(np.nanmax(rel)), ref_sum=float(np.nansum(a)), new_sum=float(np.nansum(b)),
                ref_negative=int(np.sum(a < 0)), new_negative=int(np.sum(b < 0)),
                nonfinite_ref=int(np.sum(~np.isfinite(a))), nonfinite_new=int(np.sum(~np.isfinite(b))),
                norm_linf=float(np.nanmax(np.abs(d)) / scale))
