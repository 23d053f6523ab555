"""Python face of the context API (one GPU solve held in HBM).

Mirrors the stages of the reference's UPDC_ALL (updc/c_core/updc_core.c:
1289-1543) as methods: add exposures (manifest rows), prepare (weights, CSC
transpose, PSF spectra, drizzle, denominator), iterate.  All compute runs in
libupdc_hip.so on the GPU; these wrappers only move numpy arrays across the
C ABI.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (FLAG_AUTOTUNE, FLAG_GRAPH, FLAG_PER_EXPOSURE, FLAG_QUIET, FLAG_FUSED_ADJOINT, FLAG_COLUMN_SPECTRA,
                   FreezeRow,
                   IterInfo, check)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _pf(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _pd(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class IterationError(_lib.UpdcError):
    """updc_iterate failed; freeze/info hold the rows written before the stop
    and completed the iterations whose update ran."""

    def __init__(self, code: int, msg: str, freeze: list, info: list, completed: int):
        super().__init__(code, msg)
        self.freeze, self.info, self.completed = freeze, info, completed


@dataclass
class Exposure:
    """One manifest row's data (updc_core.c:426-461) as arrays."""
    sci: np.ndarray            # (ih, iw) float32, raw (not normalised)
    wht: np.ndarray            # (ih, iw) float32, raw
    psf: np.ndarray            # (ph, pw) float32 as stored
    background: float = 0.0
    pedestal: float = 0.0
    vertices: np.ndarray | None = None   # (ih+1, iw+1, 2) float64 target-pixel corners
    linkmap: bytes | np.ndarray | None = None  # complete ULM2 image
    psf_transposed: bool = False


class Context:
    """A GPU solve: W x H target, precision 'f32' (reference float32), 'f64'
    or 'mixed'.  graph=True replays each iteration as a captured hipGraph."""

    def __init__(self, width: int, height: int, precision: str = "f32", fft_shape: tuple[int, int] | None = None,
                 graph: bool = False, quiet: bool = True, device: int = 0, preprocess_workers: int = 1,
                 autotune: bool = False, fft_cache_dir: str | None = None, retune: bool = False,
                 column_kernel: int = -1, setup_cache_dir: str | None = None):
        self.lib = _lib.load()
        self.W, self.H = int(width), int(height)
        self.precision = precision
        self.device = int(device)
        flags = (FLAG_GRAPH if graph else 0) | (FLAG_QUIET if quiet else 0) | (FLAG_AUTOTUNE if autotune else 0)
        fw, fh = fft_shape if fft_shape else (0, 0)
        h = C.c_void_p()
        check(self.lib.updc_ctx_create(C.byref(h), self.W, self.H, int(fw), int(fh), _lib.PREC[precision], flags,
                                       int(device)))
        self.h = h
        self.flags = flags
        self.nexp = 0
        self.nin: list[int] = []
        self.shapes: list[tuple[int, int]] = []
        self.prepared = False
        if preprocess_workers != 1:
            check(self.lib.updc_ctx_set_preprocess_workers(self.h, int(preprocess_workers)))
        if fft_cache_dir is not None or retune:
            check(self.lib.updc_ctx_set_fft_cache(self.h, (fft_cache_dir or "").encode(), int(retune)))
        if column_kernel != -1:
            check(self.lib.updc_ctx_set_column_kernel(self.h, int(column_kernel)))
        if setup_cache_dir is not None:
            check(self.lib.updc_ctx_set_setup_cache(self.h, setup_cache_dir.encode()))

    # -- lifetime
    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.updc_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_per_exposure(self, on: bool) -> None:
        """Plain form: one launch per exposure and stage (no batching)."""
        self.flags = (self.flags | FLAG_PER_EXPOSURE) if on else (self.flags & ~FLAG_PER_EXPOSURE)
        check(self.lib.updc_ctx_set_flags(self.h, self.flags))

    def setup_cache_hits(self) -> dict:
        """Which set-up caches the last prepare read."""
        h = C.c_int()
        check(self.lib.updc_get_setup_cache_hits(self.h, C.byref(h)))
        return dict(psf_spectra=bool(h.value & 1), initial=bool(h.value & 2), denominator=bool(h.value & 4))

    def autotune_report(self) -> tuple[list[tuple[int, int, int, float]], bool]:
        """(width, height, column_edge, ms) per timed candidate of the last
        autotune (ms < 0: unsupported), and whether it came from the cache."""
        cap = 32
        w, h, e = (C.c_int * cap)(), (C.c_int * cap)(), (C.c_int * cap)()
        ms = (C.c_double * cap)()
        n, cached = C.c_int(), C.c_int()
        check(self.lib.updc_get_autotune(self.h, w, h, e, ms, cap, C.byref(n), C.byref(cached)))
        return [(w[i], h[i], e[i], ms[i]) for i in range(min(n.value, cap))], bool(cached.value)

    def set_split_adjoint(self, on: bool) -> None:
        """Adjoint gather as its own launch followed by a plain row R2C (True),
        or fused into the row R2C (False)."""
        self.flags = (self.flags & ~FLAG_FUSED_ADJOINT) if on else (self.flags | FLAG_FUSED_ADJOINT)
        check(self.lib.updc_ctx_set_flags(self.h, self.flags))

    def set_column_spectra(self, on: bool) -> None:
        """Keep the batched iteration's spectra column-major (True) instead of
        the column-pair layout the FFT passes use where they can (False)."""
        self.flags = (self.flags | FLAG_COLUMN_SPECTRA) if on else (self.flags & ~FLAG_COLUMN_SPECTRA)
        check(self.lib.updc_ctx_set_flags(self.h, self.flags))

    def spectrum_layout(self) -> str:
        """'pairs' or 'columns': the layout of the batched iteration's spectra."""
        v = C.c_int()
        check(self.lib.updc_get_spectrum_layout(self.h, C.byref(v)))
        return "pairs" if v.value == 1 else "columns"

    def set_graph(self, on: bool) -> None:
        self.flags = (self.flags | FLAG_GRAPH) if on else (self.flags & ~FLAG_GRAPH)
        check(self.lib.updc_ctx_set_flags(self.h, self.flags))

    # -- setup
    def add_exposure(self, e: Exposure) -> None:
        sci, wht, psf = _f32(e.sci), _f32(e.wht), _f32(e.psf)
        ih, iw = sci.shape
        if wht.shape != sci.shape:
            raise ValueError("SCI/WHT shape mismatch")
        ph, pw = psf.shape
        if e.linkmap is not None:
            buf = np.frombuffer(bytes(e.linkmap), dtype=np.uint8) if not isinstance(e.linkmap, np.ndarray) else \
                np.ascontiguousarray(e.linkmap, dtype=np.uint8)
            check(self.lib.updc_add_exposure(self.h, _pf(sci), _pf(wht), iw, ih, buf.ctypes.data_as(C.c_void_p),
                                             buf.size, _pf(psf), pw, ph, int(e.psf_transposed),
                                             float(e.background), float(e.pedestal)))
        else:
            v = _f64(e.vertices)
            if v.shape != (ih + 1, iw + 1, 2):
                raise ValueError(f"vertices must be {(ih + 1, iw + 1, 2)}, got {v.shape}")
            check(self.lib.updc_add_exposure_vertices(self.h, _pf(sci), _pf(wht), iw, ih, _pd(v), _pf(psf), pw, ph,
                                                      int(e.psf_transposed), float(e.background),
                                                      float(e.pedestal)))
        self.nexp += 1
        self.nin.append(iw * ih)
        self.shapes.append((ih, iw))

    def linkmap(self, k: int) -> bytes:
        n = C.c_size_t()
        check(self.lib.updc_linkmap_size(self.h, k, C.byref(n)))
        buf = (C.c_uint8 * n.value)()
        check(self.lib.updc_get_linkmap(self.h, k, buf, n.value))
        return bytes(buf)

    def prepare(self) -> None:
        check(self.lib.updc_prepare(self.h))
        self.prepared = True

    # -- iteration
    def iterate(self, n: int = 1):
        """Run n iterations; returns (freeze rows, per-iteration info).  A
        numeric guard raises IterationError carrying what the reference would
        have written before stopping (updc_core.c:1410-1479)."""
        rows = (FreezeRow * (n * self.nexp))()
        info = (IterInfo * n)()
        rc = self.lib.updc_iterate(self.h, int(n), rows, info)
        msg = self.lib.updc_last_error().decode(errors="replace") if rc != _lib.UPDC_OK else ""
        done, with_rows = C.c_int(n), C.c_int(n)
        if rc != _lib.UPDC_OK:
            check(self.lib.updc_get_iterate_progress(self.h, C.byref(done), C.byref(with_rows)))
        freeze = [(r.iteration, r.exposure, r.n_new_frozen, r.n_total_frozen, r.min_raw_mu)
                  for r in rows[:with_rows.value * self.nexp]]
        its = [dict(iteration=i.iteration, signed_sum=i.signed_sum, n_negative=i.n_negative,
                    n_new_frozen=i.n_new_frozen, n_total_frozen=i.n_total_frozen, min_raw_mu=i.min_raw_mu,
                    ms=i.ms_total) for i in info[:done.value]]
        if rc != _lib.UPDC_OK:
            raise IterationError(rc, msg, freeze, its, done.value)
        return freeze, its

    # -- products
    def estimate(self, f64: bool = False) -> np.ndarray:
        if f64:
            out = np.empty((self.H, self.W), np.float64)
            check(self.lib.updc_get_estimate_f64(self.h, _pd(out)))
        else:
            out = np.empty((self.H, self.W), np.float32)
            check(self.lib.updc_get_estimate(self.h, _pf(out)))
        return out

    def set_estimate(self, x: np.ndarray) -> None:
        a = _f64(x).reshape(self.H, self.W)
        check(self.lib.updc_set_estimate_f64(self.h, _pd(a)))

    def wht(self) -> np.ndarray:
        out = np.empty((self.H, self.W), np.float32)
        check(self.lib.updc_get_wht(self.h, _pf(out)))
        return out

    def denominator(self) -> np.ndarray:
        out = np.empty((self.H, self.W), np.float64)
        check(self.lib.updc_get_denominator_f64(self.h, _pd(out)))
        return out

    def initial(self) -> tuple[np.ndarray, np.ndarray]:
        e = np.empty((self.H, self.W), np.float64)
        c = np.empty((self.H, self.W), np.float64)
        check(self.lib.updc_get_initial_f64(self.h, _pd(e), _pd(c)))
        return e, c

    def scalars(self) -> dict:
        ws, ms = C.c_double(), C.c_double()
        fw, fh = C.c_int(), C.c_int()
        masked = C.c_uint64()
        check(self.lib.updc_get_scalars(self.h, C.byref(ws), C.byref(ms), C.byref(fw), C.byref(fh),
                                        C.byref(masked)))
        return dict(weight_scale=ws.value, max_sensitivity=ms.value, fft_width=fw.value, fft_height=fh.value,
                    masked=masked.value)

    def psf_spectrum(self, k: int) -> np.ndarray:
        s = self.scalars()
        out = np.empty((s["fft_height"], s["fft_width"] // 2 + 1, 2), np.float64)
        check(self.lib.updc_get_psf_spectrum_f64(self.h, k, _pd(out)))
        return out[..., 0] + 1j * out[..., 1]

    def frozen(self, k: int) -> np.ndarray:
        out = np.empty(self.nin[k], np.uint8)
        check(self.lib.updc_get_frozen(self.h, k, out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    def exposure_info(self, k: int) -> dict:
        nin, nl, ncsc = C.c_uint64(), C.c_uint64(), C.c_uint64()
        enc = C.c_uint32()
        check(self.lib.updc_get_exposure_info(self.h, k, C.byref(nin), C.byref(nl), C.byref(enc), C.byref(ncsc)))
        n = C.c_size_t()
        check(self.lib.updc_linkmap_size(self.h, k, C.byref(n)))
        compact, cb = C.c_int(), C.c_uint64()
        check(self.lib.updc_get_csc_layout(self.h, k, C.byref(compact), C.byref(cb)))
        return dict(nin=nin.value, nlinks=nl.value, encoding=enc.value, ncsc=ncsc.value, linkmap_bytes=n.value,
                    csc_compact=bool(compact.value), csc_bytes=cb.value)

    def reset_frozen(self) -> None:
        check(self.lib.updc_reset_frozen(self.h))

    # -- stage entry points
    def forward(self, x: np.ndarray) -> np.ndarray:
        xs = _f64(x).reshape(self.H, self.W)
        out = np.empty((self.nexp, self.H, self.W), np.float64)
        check(self.lib.updc_stage_forward(self.h, _pd(xs), _pd(out)))
        return out

    def model(self, x: np.ndarray) -> list[np.ndarray]:
        """Noise-free H_k x for every exposure (list of (ih, iw) arrays)."""
        xs = _f64(x).reshape(self.H, self.W)
        out = np.empty(sum(self.nin), np.float64)
        check(self.lib.updc_stage_model(self.h, _pd(xs), _pd(out)))
        res, pos = [], 0
        for k, n in enumerate(self.nin):
            res.append(out[pos:pos + n].reshape(self.shapes[k]))
            pos += n
        return res

    def detector(self, k: int, blur: np.ndarray):
        b = _f64(blur).reshape(self.H, self.W)
        wr = np.empty(self.nin[k], np.float64)
        sc = np.empty((self.H, self.W), np.float64)
        st = np.empty(3, np.float64)
        check(self.lib.updc_stage_detector(self.h, k, _pd(b), _pd(wr), _pd(sc), _pd(st)))
        return wr, sc, (int(st[0]), int(st[1]), float(st[2]))

    def adjoint(self, v: np.ndarray) -> np.ndarray:
        vv = _f64(v).reshape(-1)
        if vv.size != sum(self.nin):
            raise ValueError("per-pixel vector has the wrong length")
        out = np.empty((self.H, self.W), np.float64)
        check(self.lib.updc_stage_adjoint(self.h, _pd(vv), _pd(out)))
        return out

    def scatter(self, k: int, v: np.ndarray) -> np.ndarray:
        vv = _f64(v).reshape(-1)
        out = np.empty((self.H, self.W), np.float64)
        check(self.lib.updc_stage_scatter(self.h, k, _pd(vv), _pd(out)))
        return out

    def adjoint_images(self, v: np.ndarray) -> np.ndarray:
        """S_k^T v_k for every exposure through the iteration's adjoint-image
        launch; v is the flattened per-pixel vector (sum of nin)."""
        vv = _f64(v).reshape(-1)
        if vv.size != sum(self.nin):
            raise ValueError("per-pixel vector has the wrong length")
        out = np.empty((self.nexp, self.H, self.W), np.float64)
        check(self.lib.updc_stage_adjoint_images(self.h, _pd(vv), _pd(out)))
        return out

    def fft_r2c(self, image: np.ndarray) -> np.ndarray:
        s = self.scalars()
        im = _f64(image).reshape(s["fft_height"], s["fft_width"])
        out = np.empty((s["fft_height"], s["fft_width"] // 2 + 1, 2), np.float64)
        check(self.lib.updc_fft_r2c_f64(self.h, _pd(im), _pd(out)))
        return out[..., 0] + 1j * out[..., 1]

    def fft_c2r(self, spec: np.ndarray) -> np.ndarray:
        s = self.scalars()
        sp = np.ascontiguousarray(np.stack([spec.real, spec.imag], axis=-1), dtype=np.float64)
        out = np.empty((s["fft_height"], s["fft_width"]), np.float64)
        check(self.lib.updc_fft_c2r_f64(self.h, _pd(sp), _pd(out)))
        return out

    def setup_times(self) -> dict:
        """Seconds per set-up stage (updc_get_setup_times)."""
        t = _lib.SetupTimes()
        check(self.lib.updc_get_setup_times(self.h, C.byref(t)))
        return {n: getattr(t, n) for n, _ in t._fields_}

    def iteration_times(self) -> list[dict]:
        """Per-iteration stage seconds of the last iterate call, named after
        the reference's timing rows (NaN where graph replay hides a stage)."""
        n = C.c_int()
        check(self.lib.updc_get_iteration_times(self.h, None, 0, C.byref(n)))
        buf = (_lib.IterTimes * max(1, n.value))()
        check(self.lib.updc_get_iteration_times(self.h, buf, n.value, C.byref(n)))
        return [{f: getattr(b, f) for f, _ in b._fields_} for b in buf[:n.value]]

    def stage_times(self) -> dict:
        cap = 16
        names = (C.c_char * 32 * cap)()
        ms = (C.c_double * cap)()
        n = (C.c_int * cap)()
        cnt = C.c_int(cap)
        check(self.lib.updc_get_stage_times(self.h, names, ms, n, C.byref(cnt)))
        return {bytes(names[i]).split(b"\0")[0].decode(): (ms[i], n[i]) for i in range(cnt.value)}


def self_test(verbose: bool = False) -> bool:
    return _lib.load().updc_self_test(int(verbose)) == 0
