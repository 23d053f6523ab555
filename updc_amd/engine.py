"""Caller side of the drop-in boundary: the engine's manifest, argv, environment
and output rules for the reconstruction core, without FITS/WCS handling.

The reference engine writes a manifest and raw planes, runs the core binary
with 18 arguments and reads back the iteration images, the WHT plane and the
FreezeMu CSV (updc/_engine.py:1300-1420).  These helpers reproduce those
rules so a caller holding arrays (synthetic data, tests, notebooks) can drive
``updc_hip_core`` -- or any core with the same contract -- the way the engine
does.
"""
from __future__ import annotations

import csv
import os
import subprocess
from dataclasses import dataclass
from pathlib import Path
from typing import Sequence

import numpy as np


@dataclass
class ManifestRow:
    """One exposure of the manifest (updc_core.c:426-461)."""
    science_raw: Path
    weight_raw: Path
    psf_raw: Path
    psf_width: int
    psf_height: int
    vertices: Path
    background: float
    pedestal: float

    def line(self) -> str:
        # field order and %.17g formatting of _engine.py:1303-1308
        return (f"{self.science_raw}\t{self.weight_raw}\t{self.psf_raw}\t{self.psf_width}\t{self.psf_height}\t"
                f"{self.vertices}\t{self.background:.17g}\t{self.pedestal:.17g}\n")


def write_raw(path: Path, array: np.ndarray, dtype: str) -> Path:
    """Little-endian raw plane, C order, no header (updc_core.c:298-315)."""
    np.ascontiguousarray(array, dtype=dtype).tofile(path)
    return Path(path)


def write_exposure(work: Path, index: int, sci: np.ndarray, wht: np.ndarray, psf: np.ndarray,
                   vertices: np.ndarray, background: float, pedestal: float) -> ManifestRow:
    """Raw files of exposure `index` (1-based) under `work` and its manifest
    row.  sci/wht are (ih, iw) float32 planes, psf (ph, pw) float32, vertices
    the (ih+1, iw+1, 2) corner lattice in target pixel coordinates, float64
    (_engine.py:269-280)."""
    work = Path(work)
    ih, iw = np.shape(sci)
    if np.shape(wht) != (ih, iw):
        raise ValueError("science/weight shape mismatch")
    if np.shape(vertices) != (ih + 1, iw + 1, 2):
        raise ValueError("vertex lattice must be (ih+1, iw+1, 2)")
    ph, pw = np.shape(psf)
    s = write_raw(work / f"sci_{index}.bin", sci, "<f4")
    w = write_raw(work / f"wht_{index}.bin", wht, "<f4")
    p = write_raw(work / f"psf_{index}.bin", psf, "<f4")
    v = write_raw(work / f"vertices_{index}.bin", vertices, "<f8")
    return ManifestRow(s, w, p, pw, ph, v, float(background), float(pedestal))


def write_manifest(path: Path, rows: Sequence[ManifestRow]) -> Path:
    Path(path).write_text("".join(r.line() for r in rows))
    return Path(path)


def psf_transposed(psf_frame: str, psf_orientation: str) -> bool:
    """The core's psf_transposed argument (_engine.py:1384)."""
    return psf_frame == "target" and psf_orientation == "stored-transposed"


def core_argv(binary: Path, manifest: Path, width: int, height: int, input_width: int, input_height: int,
              exposures: int, transpose_psf: bool, iterations: int, iter_prefix: Path, wht_path: Path,
              fft_workers: int = 1, exposure_workers: int = 1, preprocess_workers: int = 1,
              fft_plan: str = "estimate", fft_autotune: bool = True, retune_fft: bool = False,
              wisdom_path: str | Path = "", save_every: int = 0) -> list[str]:
    """The 18 arguments in the engine's order (_engine.py:1381-1389)."""
    return [str(binary), str(manifest), str(width), str(height), str(input_width), str(input_height),
            str(exposures), "1" if transpose_psf else "0", str(iterations), str(iter_prefix), str(wht_path),
            str(fft_workers), str(exposure_workers), str(preprocess_workers),
            "1" if fft_plan == "measure" else "0", "1" if fft_autotune else "0", "1" if retune_fft else "0",
            str(wisdom_path), str(save_every)]


def solver_env(persistent_cache_safe: bool, base: dict | None = None) -> dict:
    """UPDC_PERSIST_CACHE is "1" only for dataset-keyed caches (_engine.py:1369-1379)."""
    env = dict(os.environ if base is None else base)
    env["UPDC_PERSIST_CACHE"] = "1" if persistent_cache_safe else "0"
    return env


def output_iterations(iterations: int, save_every: int) -> list[int]:
    """Iterations the core saves and the engine reads back (_engine.py:1409-1415)."""
    if save_every > 0:
        out = list(range(save_every, iterations + 1, save_every))
        if not out or out[-1] != iterations:
            out.append(iterations)
        return out
    return [iterations]


def read_freeze_csv(path: Path, iterations: int, exposures: int) -> list[dict]:
    """Parse <prefix>_mufreeze.csv and check its row count (_engine.py:1392-1407)."""
    with Path(path).open() as f:
        rows = [{"iteration": int(r["iteration"]), "exposure": int(r["exposure"]),
                 "n_new_frozen": int(r["n_new_frozen"]), "n_total_frozen": int(r["n_total_frozen"]),
                 "min_raw_mu": float(r["min_raw_mu"]), "mu_marker": float(r["mu_marker"]),
                 "frozen_ratio": float(r["frozen_ratio"])} for r in csv.DictReader(f)]
    if len(rows) != iterations * exposures:
        raise ValueError("Solver mu-freeze CSV row mismatch")
    return rows


def read_plane(path: Path, width: int, height: int) -> np.ndarray:
    """A saved iteration or the WHT plane: <f4 (height, width) (_engine.py:1417-1420)."""
    return np.fromfile(path, dtype="<f4", count=width * height).reshape(height, width)


@dataclass
class CoreResult:
    iterations: dict  # iteration -> (H, W) float32
    wht: np.ndarray
    freeze_rows: list
    timing: dict


def run_core(argv: Sequence[str], env: dict | None = None, timeout: float | None = None) -> CoreResult:
    """Run the core with `argv` (from core_argv) and read its outputs the way
    the engine does.  A failing core raises subprocess.CalledProcessError, as
    in the engine."""
    subprocess.run(list(argv), check=True, env=env, capture_output=True, text=True, timeout=timeout)
    width, height = int(argv[2]), int(argv[3])
    n, iterations = int(argv[6]), int(argv[8])
    prefix, wht_path, save_every = argv[9], argv[10], int(argv[18])
    rows = read_freeze_csv(Path(f"{prefix}_mufreeze.csv"), iterations, n)
    its = {i: read_plane(Path(f"{prefix}_{i:02d}.bin"), width, height)
           for i in output_iterations(iterations, save_every)}
    timing = {}
    tp = Path(f"{prefix}_timing.csv")
    if tp.exists():
        with tp.open() as f:
            timing = {r["stage"]: float(r["wall_seconds"]) for r in csv.DictReader(f)}
    return CoreResult(its, read_plane(Path(wht_path), width, height), rows, timing)


def write_iteration_products(result: CoreResult, out_dir: Path, stem: str, target_cards, metadata,
                             input_metrics: Sequence[dict], weight_mode: str, psf_records: Sequence[dict],
                             psf_names: Sequence[str]) -> list[Path]:
    """The engine's "write iteration FITS" and "validate output FITS" steps
    (_engine.py:1416-1446): one <stem>_iterNN_i2d.fits per saved iteration,
    with that iteration's FreezeMu rows, then every file re-read and checked.
    FITS handling is updc_amd.fitsout (no astropy)."""
    from . import fitsout
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    files = []
    for it in sorted(result.iterations):
        name = f"{stem}_iter{it:02d}_i2d.fits"
        rows = [r for r in result.freeze_rows if r["iteration"] == it]
        files.append(fitsout.write_iteration_product(out_dir / name, result.iterations[it], result.wht, target_cards,
                                                     metadata, input_metrics, it, name, weight_mode, rows,
                                                     psf_records, psf_names))
    if len(files) != len(result.iterations):
        raise ValueError("Missing requested iteration FITS files")
    for f in files:
        fitsout.validate_product(f, len(input_metrics))
    return files
