"""Build the gfx950 HIP library and the drop-in executable in-tree.

    python -m updc_amd.build          # or __graft_entry__.build()

Outputs (git-ignored, shipped with the tree):
    updc_amd/lib/libupdc_hip.so   -- C-ABI library (include/updc_hip.h)
    updc_amd/lib/updc_hip_core    -- drop-in for the reference's updc_jwst_c
    updc_amd/lib/updc_hip_psf_projector -- drop-in for the reference's updc_psf_projector
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libupdc_hip.so"
EXE = LIBDIR / "updc_hip_core"
PROJECTOR = LIBDIR / "updc_hip_psf_projector"

HIP_SOURCES = ["fft.hip", "stages.hip", "solver.hip", "driver.hip"]
ARCH = os.environ.get("UPDC_OFFLOAD_ARCH", "gfx950")


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def common_flags() -> list[str]:
    # -ffp-contract=off: only explicit fma()/fmaf() fuse (bit parity with the
    # reference's C11 build).  Packed FP32 instructions are disabled for all
    # device code.
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-I{ROOT / 'include'}", f"-I{CSRC}", "-Wno-unused-result",
    ]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{msg}")


def build(verbose: bool = False) -> Path:
    LIBDIR.mkdir(exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = LIBDIR / (Path(src).stem + ".o")
        cmd = [_hipcc(), *common_flags(), "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    errors = []
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            msg = "\n".join(l for l in (out + err).splitlines() if "packed-fp32-ops" not in l)
            errors.append(f"{' '.join(cmd)}\n{msg}")
    if errors:
        raise RuntimeError("build failed:\n" + "\n\n".join(errors))
    tmp = LIB.with_suffix(".so.tmp")
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    for exe, src in ((EXE, "updc_hip_core.cpp"), (PROJECTOR, "psf_projector_main.cpp")):
        _run([_hipcc(), *common_flags(), "-o", str(exe), str(CSRC / src), f"-L{LIBDIR}", "-lupdc_hip",
              f"-Wl,-rpath,$ORIGIN"])
    for o in objs:
        o.unlink(missing_ok=True)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(f"built {LIB}, {EXE} and {PROJECTOR}")
