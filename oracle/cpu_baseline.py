"""Threaded float32 CPU restatement of the v17 iteration -- TEST INFRASTRUCTURE ONLY.

Only bench.py's cpu_baseline leg and tests/ use this module (it builds on the
oracle); the product never does.  It is the reference's own float32
algorithm, timed on the host's cores as the baseline the GPU path is set
against:

* set-up: oracle.RefSolver(precision="f32") (weights, PSF spectra, drizzle,
  denominator; updc_core.c:1367-1398);
* per iteration (updc_core.c:1410-1458): X^ = R2C(pad x); per exposure the
  product X^ . P^_k (orc_spectrum_multiply, no FMA), blur_k = crop C2R / nf,
  the fused detector loop (orc_fused_detector: the fmaf gather, FreezeMu,
  ratio and fmaf scatter, updc_core.c:1152-1177), R2C(scatter_k) and
  acc_w += . conj(P^_k) (orc_accumulate_adjoint) into the partial spectrum
  of worker w = k mod workers; the partials summed in worker order
  (updc_core.c:1280-1285), the correction's C2R, and the update (orc_update,
  updc_core.c:1437-1457).

Threading follows the reference's exposure workers (updc_core.c:1201-1269):
the per-exposure C loops run in `workers` threads (ctypes releases the GIL),
each worker owning the exposures k = w, w + workers, ...  FFTW3f is absent
here, so the transforms are MKL's through torch.fft on the CPU, batched over
the exposures and run on `fft_threads` threads (the reference's fft_workers).
"""
from __future__ import annotations

import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import oracle as O


def host_cores(cap: int = 16) -> int:
    """Cores this process may use (its affinity set), at most `cap`."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(cap, n))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class ThreadedF32:
    """The float32 iteration on `workers` exposure threads and `fft_threads`
    MKL threads, from a prepared RefSolver(precision="f32")."""

    def __init__(self, ref: "O.RefSolver", workers: int, fft_threads: int):
        if ref.prec != "f32":
            raise ValueError("the CPU baseline restates the reference's float32 path")
        self.r = ref
        self.workers = max(1, min(workers, ref.N))
        self.fft_threads = max(1, fft_threads)
        self.pool = ThreadPoolExecutor(self.workers)
        self.scale = np.float32(np.float32(1.0) / np.float32(ref.nf))

    def close(self) -> None:
        self.pool.shutdown()

    def _owned(self, w: int) -> list[int]:
        return list(range(w, self.r.N, self.workers))

    def _rfft(self, imgs: np.ndarray) -> np.ndarray:
        """Batched R2C of (n, H, W) images zero-padded to FH x FW."""
        r = self.r
        pad = torch.zeros((imgs.shape[0], r.FH, r.FW), dtype=torch.float32)
        pad[:, :r.H, :r.W] = torch.from_numpy(imgs)
        return torch.fft.rfft2(pad).numpy()

    def _irfft_crop(self, specs: np.ndarray) -> np.ndarray:
        r = self.r
        full = torch.fft.irfft2(torch.from_numpy(specs), s=(r.FH, r.FW), norm="forward")
        return (full[:, :r.H, :r.W] * float(self.scale)).numpy()

    def iterate(self, n: int = 1) -> list[tuple]:
        r = self.r
        torch.set_num_threads(self.fft_threads)
        rows = []
        for it in range(n):
            X = self._rfft(r.x.reshape(1, r.H, r.W))[0]
            prods = np.empty((r.N,) + X.shape, np.complex64)

            def mult(w):
                for k in self._owned(w):
                    prods[k] = O.spectrum_multiply(X, r.ex[k]["P"])
            list(self.pool.map(mult, range(self.workers)))
            blurs = self._irfft_crop(prods)
            del prods
            scat = np.empty((r.N, r.H, r.W), np.float32)
            stats = [None] * r.N

            def detect(w):
                for k in self._owned(w):
                    e = r.ex[k]
                    _, sc, st = O.fused_detector(e["linkmap"], r.nout, e["s"], e["w"], e["frozen"], blurs[k],
                                                 float(e["mo_f"]), float(e["ped_f"]))
                    scat[k] = sc.reshape(r.H, r.W)
                    stats[k] = st
            list(self.pool.map(detect, range(self.workers)))
            S = self._rfft(scat)
            partial = [np.zeros(X.shape, np.complex64) for _ in range(self.workers)]

            def accumulate(w):
                for k in self._owned(w):
                    O.accumulate_adjoint(partial[w], S[k], r.ex[k]["P"])
            list(self.pool.map(accumulate, range(self.workers)))
            acc = partial[0]
            for w in range(1, self.workers):
                acc += partial[w]
            Cc = self._irfft_crop(acc.reshape((1,) + acc.shape))[0]
            rc, x, _, _, bad = O.update(r.x, Cc, r.D, r.max_sens)
            if rc:
                raise FloatingPointError(f"update guard {rc} at {bad}")
            r.x = x.reshape(r.H, r.W)
            rows += [(it + 1, k + 1) + tuple(stats[k]) for k in range(r.N)]
        return rows


def time_iterations(ref: "O.RefSolver", workers: int, fft_threads: int, iters: int) -> float:
    """Seconds per iteration of the threaded path (no untimed warm-up)."""
    path = ThreadedF32(ref, workers, fft_threads)
    try:
        t0 = time.perf_counter()
        path.iterate(iters)
        return (time.perf_counter() - t0) / iters
    finally:
        path.close()
