#!/usr/bin/env python3
"""Benchmark of the UPDC v17 reconstruction iteration on one MI355X.

Workload (default): 8 exposures of 2048^2 -> 4096^2 target, Airy+wing PSFs of
1431^2, FFT 5600^2, 40 iterations (synthetic, seeded; see updc_amd/synth.py).
A "step" is one reconstruction iteration over all exposures.  --config 3 runs
the low-sky case (errors at iterations 10, 20, 40) and the reference's
FreezeMu forcing recipe; --config 4 the uneven batch of 24 exposures against
per-exposure launches; --config 5 the differentiation workload (gradients
through 10 unrolled iterations).

Reported line: iterations/s of the FP32/batched form (FP32 FFTs, PSF and S_k
products, FP64 spectrum accumulation over exposures, denominator, raw_mu test
and ratio; all exposures batched in every launch; each iteration replayed as
a hipGraph), next to the plain FP64 GPU form and the error of the FP32/batched
estimate against the FP64 one after the same number of iterations.  The
roofline block gives the dominant kernel against the HBM peak (its bytes on
this library's data layout and on the LinkMap unit) and the whole iteration's
algorithmic bytes against it.  The CPU baseline is the reference's float32
iteration on the host's cores (all of them, then one), MKL transforms in place
of FFTW3f.

Usage: python bench.py [--gpus 1] [--steps 40] [--warmup 2] [--config 2|1|3|4|5]
                       [--no-cpu] [--no-f64] [--no-autotune] [--profile]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "RL iters/s, 8x2048² exps->4096² grid: FP32/batched vs FP64, rel err vs FP64"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_dataset(config: int, stress: bool = False):
    from updc_amd import synth
    from updc_amd.solver import Context, Exposure
    t0 = time.time()
    if stress:
        ds = synth.config3_stress()
    else:
        ds = {1: synth.config1, 2: synth.config2, 3: synth.config3, 4: synth.config4, 5: synth.config5}[config]()
    log(f"synthetic {ds.name}: truth/PSFs/geometry in {time.time() - t0:.1f}s")
    t0 = time.time()
    n = len(ds.geoms)
    with Context(ds.W, ds.H, "f64") as mc:
        for k, g in enumerate(ds.geoms):
            z = np.zeros((g.ih, g.iw), np.float32)
            mc.add_exposure(Exposure(z, z + 1.0, ds.psfs[k], 0.0, 0.0, vertices=ds.vertices(k)))
        mc.prepare()
        model = mc.model(ds.truth)
        linkmaps = [mc.linkmap(k) for k in range(n)]
    synth.observe(ds, [m + b for m, b in zip(model, ds.background)], config_seed=config)
    ds.linkmaps = linkmaps
    log(f"noise-free model on GPU + observation in {time.time() - t0:.1f}s; pedestals {ds.pedestal[:2]}...")
    return ds


def solve_ctx(ds, precision: str, graph: bool, plain: bool = False, autotune: bool = False, fft=None):
    from updc_amd.solver import Context, Exposure
    c = Context(ds.W, ds.H, precision, fft_shape=fft, graph=graph, autotune=autotune)
    if plain:
        c.set_per_exposure(True)
    for k in range(len(ds.geoms)):
        c.add_exposure(Exposure(ds.sci[k], ds.wht[k], ds.psfs[k], ds.background[k], ds.pedestal[k],
                                linkmap=ds.linkmaps[k]))
    t0 = time.time()
    c.prepare()
    log(f"{precision}{' graph' if graph else ''}: prepare {time.time() - t0:.2f}s {c.scalars()}")
    if autotune:
        rows, _ = c.autotune_report()
        log("autotune: " + ", ".join(f"{w}x{h}{' edge' if e else ''} {ms:.2f}ms" for w, h, e, ms in rows))
    return c


def timed_run(c, steps: int, warmup: int):
    """W untimed iterations (graph capture happens there), reset to x0 and the
    zero FreezeMu state, then exactly K timed iterations."""
    x0, _ = c.initial()
    if warmup:
        c.iterate(warmup)
    c.set_estimate(x0)
    c.reset_frozen()
    t0 = time.perf_counter()
    rows, info = c.iterate(steps)    # synchronises before returning
    dt = time.perf_counter() - t0
    return dt, rows, info


def timed_chunks(c, chunks: list[int], warmup: int):
    """Like timed_run, in chunks: the state after each chunk is read back
    outside the timed region.  Returns (seconds, {iteration: x}, rows)."""
    x0, _ = c.initial()
    if warmup:
        c.iterate(warmup)
    c.set_estimate(x0)
    c.reset_frozen()
    dt, done, snaps, rows = 0.0, 0, {}, []
    for n in chunks:
        t0 = time.perf_counter()
        r, _ = c.iterate(n)
        dt += time.perf_counter() - t0
        done += n
        rows += r
        snaps[done] = c.estimate(f64=True)
    return dt, snaps, rows


def frozen_totals(rows, nexp: int) -> int:
    """Frozen pixels over all exposures after the last iteration in rows."""
    return int(sum(r[3] for r in rows[-nexp:]))


def run_stress(args, ds, metrics) -> dict:
    """Config 3, low sky (0.01) on the config-2 geometry:
    (a) FP32/batched against plain FP64 over the timed iterations, errors at
        iterations 10, 20 and 40;
    (b) FreezeMu forcing by the reference's own recipe (every background
        -100, tests/test_smoke.py:117-136): 2 iterations in each form, the
        freeze rows compared and their invariants checked."""
    from updc_amd._lib import UpdcError
    K = args.steps
    chunks = [10, 10, K - 20] if K > 20 else [K]
    nexp = len(ds.geoms)
    fast = solve_ctx(ds, "mixed", graph=True, autotune=not args.no_autotune)
    sc = fast.scalars()
    dt32, s32, rows32 = timed_chunks(fast, chunks, args.warmup)
    fast.close()
    its32 = K / dt32
    log(f"FP32/batched (mixed, graph): {K} iterations -> {its32:.1f} it/s")
    slow = solve_ctx(ds, "f64", graph=False, plain=True)
    dt64, s64, rows64 = timed_chunks(slow, chunks, args.warmup)
    slow.close()
    its64 = K / dt64
    log(f"plain FP64: {its64:.1f} it/s")
    errs = {}
    for it in sorted(s32):
        st = metrics.stats(s64[it], s32[it])
        errs[str(it)] = {k: st[k] for k in ("norm_linf", "rel_rms", "max_rel_masked", "max_abs", "rms")}
        log(f"iteration {it}: FP32/batched vs FP64 {errs[str(it)]}")

    # (b) FreezeMu forcing by the reference's recipe (every background -100),
    # and (c) the zero-read-noise stress (synth.config3_stress): each tier
    # against FP64 after 2, 10, 20 and 40 iterations
    def tiers_vs_fp64(d, label):
        snaps, rows = {}, {}
        for prec in ("f64", "mixed", "f32"):
            c = solve_ctx(d, prec, graph=True)
            snaps[prec], rows[prec], done = {}, [], 0
            try:
                for n in (2, 8, 10, 20):
                    r, _ = c.iterate(n)
                    rows[prec] += [(int(q[0]) + done, int(q[1]), int(q[2]), int(q[3])) for q in r]
                    done += n
                    snaps[prec][done] = c.estimate(f64=True)
            except UpdcError as e:
                log(f"{label} {prec}: {e}")
            c.close()
        out = {}
        for prec in ("mixed", "f32"):
            out[prec] = {}
            for it in sorted(snaps[prec]):
                if it not in snaps["f64"]:
                    continue
                st = metrics.stats(snaps["f64"][it], snaps[prec][it])
                out[prec][str(it)] = {k: st[k] for k in ("norm_linf", "rel_rms", "max_rel_masked")}
            out[prec]["freeze_rows_equal_to_fp64"] = rows[prec] == rows["f64"][:len(rows[prec])]
        out["frozen_total_fp64"] = {str(it): sum(r[3] for r in rows["f64"] if r[0] == it) for it in (2, 10, 20, 40)}
        log(f"{label}: {out}")
        return out

    saved_bg = list(ds.background)
    ds.background = [-100.0] * nexp
    fz = tiers_vs_fp64(ds, "FreezeMu forcing (background -100)")
    ds.background = saved_bg
    sd = make_dataset(3, stress=True)
    x0 = None
    with solve_ctx(sd, "f64", graph=False) as c:
        x0, _ = c.initial()
        pred = c.model(x0)
    raw = np.concatenate([p.ravel() + sd.background[k] + sd.pedestal[k] for k, p in enumerate(pred)])
    act = np.concatenate([w.ravel() for w in sd.wht]) > 0
    frac_small = float(np.mean(np.abs(raw[act]) < 1.0))
    stress = tiers_vs_fp64(sd, "zero-read-noise stress")
    stress["raw_mu_at_x0_below_1_fraction"] = frac_small
    stress["pedestal"] = sd.pedestal[0]
    return {
        "metric": METRIC, "value": round(its32, 3), "unit": "iters/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": round(dt32 * 1e3 / K, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 accumulation/denominator/ratio)",
        "data": "synthetic",
        "config": {"workload": f"config3 (low sky 0.01): config-2 geometry and PSFs, "
                               f"FFT {sc['fft_width']}x{sc['fft_height']}, {K} iterations",
                   "fp64_plain_iters_per_s": round(its64, 3), "fp32_over_fp64_plain": round(its32 / its64, 3),
                   "rel_err_vs_fp64_at": errs,
                   "freeze_row_mismatches": sum(1 for a, b in zip(rows32, rows64) if a[2:4] != b[2:4]),
                   "freezemu_forcing_background_minus_100": fz,
                   "zero_read_noise_stress": stress},
        "roofline": None, "cpu_baseline": None,
    }


def run_uneven(args, ds, metrics) -> dict:
    """Config 4: 24 exposures of three sizes in one launch per stage (graph,
    autotuned shape) against per-exposure launches, and the error vs FP64."""
    K = args.steps
    fast = solve_ctx(ds, "mixed", graph=True, autotune=not args.no_autotune)
    sc = fast.scalars()
    dt32, _, _ = timed_run(fast, K, args.warmup)
    x32 = fast.estimate(f64=True)
    fast.close()
    its32 = K / dt32
    log(f"FP32/batched (mixed, graph, autotuned): {its32:.2f} it/s")
    plain = solve_ctx(ds, "mixed", graph=False, plain=True)
    dtp, _, _ = timed_run(plain, K, args.warmup)
    xp = plain.estimate(f64=True)
    plain.close()
    itsp = K / dtp
    log(f"per-exposure launches (mixed, eager, default shape): {itsp:.2f} it/s")
    errs = None
    its64 = None
    if not args.no_f64:
        b64 = solve_ctx(ds, "f64", graph=True)
        dt64, _, _ = timed_run(b64, K, args.warmup)
        x64 = b64.estimate(f64=True)
        b64.close()
        its64 = K / dt64
        e = metrics.stats(x64, x32)
        ep = metrics.stats(x64, xp)
        errs = {"batched": {k: e[k] for k in ("norm_linf", "rel_rms", "max_rel_masked")},
                "per_exposure": {k: ep[k] for k in ("norm_linf", "rel_rms", "max_rel_masked")}}
        log(f"batched FP64: {its64:.2f} it/s; errors vs FP64 {errs}")
    sizes = {}
    for g in ds.geoms:
        sizes[f"{g.iw}x{g.ih}"] = sizes.get(f"{g.iw}x{g.ih}", 0) + 1
    return {
        "metric": METRIC, "value": round(its32, 3), "unit": "iters/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": round(dt32 * 1e3 / K, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 accumulation/denominator/ratio)",
        "data": "synthetic",
        "config": {"workload": f"config4 (uneven batching): {len(ds.geoms)} exposures {sizes} -> {ds.W}^2 target, "
                               f"PSFs 1431^2/511^2, FFT {sc['fft_width']}x{sc['fft_height']}, {K} iterations",
                   "per_exposure_launch_iters_per_s": round(itsp, 3),
                   "batched_over_per_exposure": round(its32 / itsp, 3),
                   "fp64_batched_iters_per_s": None if its64 is None else round(its64, 3),
                   "rel_err_vs_fp64": errs},
        "roofline": None, "cpu_baseline": None,
    }


def run_autograd(args, ds, metrics) -> dict:
    """Config 5: loss = Poisson NLL of the held-out fifth exposure after K=10
    unrolled iterations over the other four; gradients with respect to the
    initial estimate, the PSF mixture parameters and the backgrounds through
    the HIP operators (updc_amd.autograd), per-iteration checkpointing.
    Reports forward+backward time (FP32 and FP64), peak memory with and
    without checkpointing, central finite differences (h = 1e-6, FP64), the
    FP64 torch restatement (oracle/torch_ref.py, torch.fft on the GPU) and
    FP32 against FP64 gradients."""
    import torch
    from updc_amd import autograd as A
    from updc_amd import synth
    from updc_amd.solver import Context, Exposure
    K = 10
    n = len(ds.geoms)
    true_p = synth.config5_psf_params(n)
    start_p = true_p.copy()
    start_p[:, :, 1] *= 1.05          # widths 5% off
    start_p[:, 1, 0] *= 0.8           # wing amplitude 20% off
    res = {}
    grads = {}
    for prec in ("mixed", "f64"):
        c = Context(ds.W, ds.H, prec)
        for k in range(n):
            c.add_exposure(Exposure(ds.sci[k], ds.wht[k], ds.psfs[k], ds.background[k], ds.pedestal[k],
                                    linkmap=ds.linkmaps[k]))
        c.prepare()
        ops = A.Operators(c)
        ex = A.exposures_from_context(c, ops, ds.sci, ds.wht, ds.pedestal)
        x0np, _ = c.initial()
        dt, dev = ops.dtype, ops.device
        size = ds.psfs[0].shape[0]

        def loss_fn(x0, params, bs, checkpoint=True):
            pws = [A.wrap_psf(A.gaussian_mixture(params[k], size), ops.FH, ops.FW) for k in range(n)]
            x = A.unrolled(x0, pws[:n - 1], [bs[k] for k in range(n - 1)], K, ops, ex, checkpoint=checkpoint)
            mu = A.forward_model(x, pws[n - 1], bs[n - 1], n - 1, ops, ex)
            return A.poisson_nll(mu, ex.yeff[n - 1], ex.active[n - 1])

        def leaves():
            return (torch.as_tensor(x0np, dtype=dt, device=dev).requires_grad_(True),
                    torch.as_tensor(start_p, dtype=dt, device=dev).requires_grad_(True),
                    torch.zeros(n, dtype=dt, device=dev).requires_grad_(True))

        times = []
        for rep in range(3):
            lv = leaves()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            L = loss_fn(*lv)
            g = torch.autograd.grad(L, lv)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        peak = {}
        for ck in (True, False):
            torch.cuda.empty_cache()
            torch.cuda.reset_peak_memory_stats(dev)
            base = torch.cuda.memory_allocated(dev)
            lv = leaves()
            L2 = loss_fn(*lv, checkpoint=ck)
            torch.autograd.grad(L2, lv)
            torch.cuda.synchronize()
            peak["checkpointed" if ck else "stored"] = int(torch.cuda.max_memory_allocated(dev) - base)
        grads[prec] = [t.detach().to("cpu", torch.float64) for t in g]
        r = dict(loss=float(L.detach()), fwd_bwd_s=round(min(times), 4), iters_per_s=round(K / min(times), 3),
                 peak_bytes=peak, fft_shape=[ops.FW, ops.FH])
        log(f"autograd {prec}: loss {float(L.detach()):.10g}, forward+backward {min(times) * 1e3:.1f} ms "
            f"({K / min(times):.1f} unrolled it/s), peak memory {peak}")
        if prec == "f64":
            with torch.no_grad():
                h = 1e-6
                floor = 8 * 2.2e-16 * abs(float(L.detach())) / h
                lv = leaves()
                gen = torch.Generator().manual_seed(1)
                v = torch.randn(lv[0].shape, generator=gen, dtype=torch.float64).to(dev)
                fd_rows = []
                f = lambda s_: float(loss_fn(lv[0] + s_ * v, lv[1], lv[2]))
                fd = (f(h) - f(-h)) / (2 * h)
                an = float((g[0] * v).sum())
                fd_rows.append(("x0 direction", fd, an))
                for idx in [(0, 0, 1), (1, 1, 0), (n - 1, 0, 2)]:
                    e = torch.zeros_like(lv[1])
                    e[idx] = 1.0
                    f = lambda s_: float(loss_fn(lv[0], lv[1] + s_ * e, lv[2]))
                    fd_rows.append((f"psf{idx}", (f(h) - f(-h)) / (2 * h), float(g[1][idx])))
                for k in (0, n - 1):
                    e = torch.zeros_like(lv[2])
                    e[k] = 1.0
                    f = lambda s_: float(loss_fn(lv[0], lv[1], lv[2] + s_ * e))
                    fd_rows.append((f"background{k}", (f(h) - f(-h)) / (2 * h), float(g[2][k])))
            r["finite_differences"] = [dict(param=nm, fd=fdv, analytic=anv, abs_err=abs(fdv - anv),
                                            tol=1e-5 * abs(anv) + floor) for nm, fdv, anv in fd_rows]
            ok = all(x["abs_err"] <= x["tol"] for x in r["finite_differences"])
            log(f"finite differences (h=1e-6): {'pass' if ok else 'FAIL'} {r['finite_differences']}")
            r["finite_differences_pass"] = ok
            # the FP64 torch restatement on the GPU (torch.fft, index_add)
            from oracle import torch_ref as R
            m = R.RefModel(ds.W, ds.H, ops.FW, ops.FH, ds.linkmaps, device=dev)
            lv = leaves()
            pws = [R.wrap_psf(R.gaussian_mixture(lv[1][k], size), ops.FH, ops.FW) for k in range(n)]
            xk = m.unrolled(lv[0], pws[:n - 1], lv[2][:n - 1], K, ex.yeff[:n - 1], ex.w[:n - 1], ex.wmask[:n - 1],
                            ex.active[:n - 1], ex.ped[:n - 1])
            Lr = R.poisson_nll(m.forward_model(xk, pws[n - 1], lv[2][n - 1], n - 1, ex.ped[n - 1]),
                               ex.yeff[n - 1], ex.active[n - 1])
            gr = [t.detach().to("cpu", torch.float64) for t in torch.autograd.grad(Lr, lv)]
            rel = [float((a - b).norm() / b.norm()) for a, b in zip(grads["f64"], gr)]
            r["torch_reference"] = dict(loss_rel=abs(float(L.detach()) - float(Lr.detach())) / abs(float(Lr.detach())), grad_rel=rel)
            log(f"FP64 torch restatement: {r['torch_reference']}")
        res[prec] = r
        del ops, ex
        c.close()
    rel32 = [float((a - b).norm() / b.norm()) for a, b in zip(grads["mixed"], grads["f64"])]
    log(f"FP32 vs FP64 gradients (x0, psf params, backgrounds): {rel32}")
    its = res["mixed"]["iters_per_s"]
    return {
        "metric": METRIC, "value": its, "unit": "iters/s", "n_gpus": 1, "steps": K, "warmup": 1,
        "ms_per_step": round(1e3 / its, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (f64 accumulation/denominator/ratio)", "data": "synthetic",
        "config": {"workload": f"config5 (differentiation): {n - 1}x{ds.geoms[0].iw}^2 exposures + 1 held out -> "
                               f"{ds.W}^2 target, PSF {ds.psfs[0].shape[0]}^2 2-component mixtures, "
                               f"K={K} unrolled iterations forward+backward, per-iteration checkpointing",
                   "fp32": res["mixed"], "fp64": res["f64"], "fp32_vs_fp64_grad_rel": rel32},
        "roofline": None, "cpu_baseline": None,
    }


def stage_profile(c, iters: int) -> dict:
    """Per-kernel average device times over a timed region of `iters` eager
    (non-graph) iterations: HIP events recorded around every launch on the
    library's own stream, averaged per stage.  The state is restored after."""
    x = c.estimate(f64=True)
    c.set_graph(False)
    c.iterate(iters)
    t = c.stage_times()
    c.set_graph(True)
    c.set_estimate(x)
    return t


def cpu_baseline(ds, fft_shape, iters_all: int = 2, iters_one: int = 1) -> dict:
    """The reference's float32 iteration on this host's CPU cores
    (oracle/cpu_baseline.py: the oracle's C detector and spectrum loops on
    exposure threads, worker w owning exposures k = w mod workers as
    updc_core.c:1201-1269 does, MKL transforms through torch.fft because
    FFTW3f is absent), after its untimed set-up: `iters_all` iterations on all
    the cores this process may use (at most 16), then `iters_one` on one core."""
    from oracle import cpu_baseline as B
    from oracle import oracle as O
    ex = [dict(sci=ds.sci[k], wht=ds.wht[k], psf=ds.psfs[k], background=ds.background[k],
               pedestal=ds.pedestal[k], linkmap=ds.linkmaps[k]) for k in range(len(ds.geoms))]
    cores = B.host_cores()
    t0 = time.perf_counter()
    s = O.RefSolver(ds.W, ds.H, ex, "f32", fft_shape=fft_shape)
    log(f"cpu baseline: oracle set-up {time.perf_counter() - t0:.1f}s")
    t_all = B.time_iterations(s, cores, cores, iters_all)
    log(f"cpu baseline: {cores} cores, {t_all:.2f}s per iteration")
    t_one = B.time_iterations(s, 1, 1, iters_one)
    log(f"cpu baseline: 1 core, {t_one:.2f}s per iteration")
    return dict(value=1.0 / t_all, unit="iters/s", cores=cores, kind="port",
                sample=f"{iters_all} iterations on {cores} cores and {iters_one} on 1 core of the reference's "
                       f"float32 iteration over all {len(ex)} exposures of this workload (FFT {fft_shape[0]}x"
                       f"{fft_shape[1]}), after the untimed set-up; {t_all:.2f}s and {t_one:.2f}s per iteration",
                single_core=dict(value=1.0 / t_one, unit="iters/s", cores=1),
                host=dict(model=B.cpu_model(), nproc=os.cpu_count(), cores_used=cores),
                fft="MKL through torch.fft on the CPU (FFTW3f, the reference's library, is absent)",
                threading="exposure workers = FFT threads = cores (updc_core.c:159-168, 1201-1269)")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f64", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="FP32/batched form on the default FFT shape instead of the autotuned one")
    ap.add_argument("--profile", action="store_true", help="only the FP32/batched timed run (for rocprofv3)")
    ap.add_argument("--calibrate-copy", type=float, default=0.0,
                    help="first run k_copy_ref over this many GiB (HBM counter scale reference)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: the multi-GPU form (exposures sharded across GPUs, one all-reduce "
              "of the correction image per iteration) is a design, not a run; use --gpus 1", file=sys.stderr)
        return 2
    from updc_amd import metrics
    if args.calibrate_copy > 0:
        import ctypes as C
        from updc_amd import _lib
        nbytes = int(args.calibrate_copy * (1 << 30))
        ms = C.c_double()
        _lib.check(_lib.load().updc_copy_calibration(nbytes, 3, 0, C.byref(ms)))
        log(f"k_copy_ref: {nbytes} B read + {nbytes} B written per launch in {ms.value:.4f} ms "
            f"-> {2 * nbytes / (ms.value * 1e-3) / 1e9:.1f} GB/s")
    ds = make_dataset(args.config)
    K, Wm = args.steps, args.warmup
    if args.config in (3, 4, 5):
        line = {3: run_stress, 4: run_uneven, 5: run_autograd}[args.config](args, ds, metrics)
        print(json.dumps(line), flush=True)
        return 0

    fast = solve_ctx(ds, "mixed", graph=True, autotune=not args.no_autotune)
    sc = fast.scalars()
    FW, FH = sc["fft_width"], sc["fft_height"]
    dt32, rows32, info32 = timed_run(fast, K, Wm)
    its32 = K / dt32
    log(f"FP32/batched (mixed, graph): {K} iterations in {dt32 * 1e3:.1f} ms -> {its32:.1f} it/s")
    x32 = fast.estimate(f64=True)
    # the adjoint gather fused into the row R2C (FLAG_FUSED_ADJOINT): timed beside the default split form
    fast.set_split_adjoint(False)
    dtf, _, _ = timed_run(fast, K, Wm)
    xf = fast.estimate(f64=True)
    fast.set_split_adjoint(True)
    log(f"fused adjoint (graph): {K / dtf:.1f} it/s; equal to split: {bool(np.array_equal(xf, x32))}")
    if args.profile:
        print(json.dumps(dict(metric=METRIC, value=its32, unit="iters/s", steps=K, warmup=Wm)))
        return 0

    # per-kernel times of the FP32/batched form (eager launches, HIP events)
    prof = stage_profile(fast, K)
    infos = [fast.exposure_info(k) for k in range(len(ds.geoms))]
    fast.close()

    ref32 = solve_ctx(ds, "f32", graph=True)
    dtr, _, _ = timed_run(ref32, K, Wm)
    xr32 = ref32.estimate(f64=True)
    ref32.close()
    log(f"reference-precision f32 (graph): {K / dtr:.1f} it/s")

    # end to end from host memory (the boundary takes host buffers): upload of
    # the exposures, LinkMaps and PSFs, set-up, K iterations (the first one
    # capturing the graph), and the estimate's download -- reported beside
    # `value`, never as it
    t0 = time.time()
    e2e = solve_ctx(ds, "mixed", graph=True, fft=(FW, FH))
    e2e.iterate(K)
    e2e.estimate()
    t_e2e = time.time() - t0
    e2e.close()
    log(f"end to end from host buffers: {K} iterations in {t_e2e:.3f} s -> {K / t_e2e:.1f} it/s")

    its64 = None
    its64b = None
    err = None
    errr = None
    if not args.no_f64:
        # plain FP64 form: per-exposure launches, eager (the reference's dataflow)
        slow = solve_ctx(ds, "f64", graph=False, plain=True)
        dt64, rows64, _ = timed_run(slow, K, Wm)
        its64 = K / dt64
        x64 = slow.estimate(f64=True)
        slow.close()
        log(f"plain FP64 (per-exposure, eager): {K} iterations in {dt64 * 1e3:.1f} ms -> {its64:.1f} it/s")
        b64 = solve_ctx(ds, "f64", graph=True)
        dt64b, _, _ = timed_run(b64, K, Wm)
        its64b = K / dt64b
        x64b = b64.estimate(f64=True)
        b64.close()
        d64 = metrics.stats(x64, x64b)
        log(f"batched FP64 (graph): {its64b:.1f} it/s; vs plain FP64 norm_linf {d64['norm_linf']:.2e}")
        err = metrics.stats(x64, x32)
        errr = metrics.stats(x64, xr32)
        frz_mismatch = sum(1 for a, b in zip(rows32, rows64) if a[2:4] != b[2:4])
        log(f"FP32/batched vs FP64 after {K}: {err}; freeze-row mismatches {frz_mismatch}")
        log(f"reference f32 vs FP64 after {K}: norm_linf={errr['norm_linf']:.3e} rel_rms={errr['rel_rms']:.3e}")

    # roofline of the dominant kernel
    e, c, ed, ea = 4, 8, 8, 8
    sb = metrics.stage_bytes(ds.W, ds.H, FW, FH, [i["nin"] for i in infos], [i["nlinks"] for i in infos],
                             [i["ncsc"] for i in infos], [i["linkmap_bytes"] for i in infos], e, c, ed, ea,
                             fused_inverse="col_inv_corr" not in prof,
                             adj_csc_bytes=[i["csc_bytes"] for i in infos])
    kern = {k: v for k, v in prof.items() if k in sb}
    dom = max(kern, key=lambda k: kern[k][0])
    ms = kern[dom][0]
    achieved = sb[dom] / (ms * 1e-3) / 1e9
    per_stage = {k: dict(ms=round(v[0], 4), GBps=round(sb[k] / (v[0] * 1e-3) / 1e9, 1)) for k, v in kern.items()}
    log(f"per-kernel: {json.dumps(per_stage)}")

    # the whole iteration against the byte model of the path (each stage's
    # operands read once, results written once; real LinkMap sizes)
    nin_all = [i["nin"] for i in infos]
    lm_all = [i["linkmap_bytes"] for i in infos]
    b_iter = metrics.iteration_bytes(ds.W, ds.H, FW, FH, nin_all, lm_all)
    it_gbs = b_iter * its32 / 1e9
    # the dominant kernel on the LinkMap unit too: the adjoint gathers read the
    # CSC transpose this library builds, the reference reads the LinkMap
    link_unit = None
    if dom in ("adjoint_image", "row_r2c_gather"):
        csc_b = sum(i["csc_bytes"] for i in infos) if dom == "adjoint_image" else \
            sum(4 * (ds.W * ds.H + 1) + 8 * i["ncsc"] for i in infos)
        lu_bytes = sb[dom] - csc_b + sum(lm_all)
        link_unit = dict(bytes=lu_bytes, frac=round(lu_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))

    # HBM traffic of the dominant kernel, from the committed FETCH_SIZE /
    # WRITE_SIZE summary of this workload (rocprofv3 --pmc runs; see
    # updc_amd/profsum.py for the copy-kernel calibration)
    traffic = None
    tpath = ROOT / "profiles" / f"iteration_kernels_config{args.config}_traffic.txt"
    if tpath.exists():
        from updc_amd.profsum import read_traffic
        traffic = read_traffic(tpath).get(dom, {}).get("total")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(ds, (FW, FH))
        except Exception as ex:  # the baseline is informative; the GPU line stands without it
            log(f"cpu baseline failed: {ex!r}")
    line = {
        "metric": METRIC, "value": round(its32, 3), "unit": "iters/s", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": round(dt32 * 1e3 / K, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (f64 accumulation/denominator/ratio)", "data": "synthetic",
        "config": {"workload": f"{ds.name}: {len(ds.geoms)}x{ds.geoms[0].iw}^2 exposures -> {ds.W}^2 target, "
                               f"PSF {ds.psfs[0].shape[0]}^2, FFT {FW}x{FH}, {K} iterations",
                   "form": "FP32/batched (all exposures per launch, hipGraph per iteration, "
                           + ("default FFT shape)" if args.no_autotune else "autotuned FFT shape)"),
                   "fft_shape": [FW, FH],
                   "fp64_plain_iters_per_s": None if its64 is None else round(its64, 3),
                   "fp32_over_fp64_plain": None if its64 is None else round(its32 / its64, 3),
                   "fp64_batched_iters_per_s": None if its64b is None else round(its64b, 3),
                   "fp32_over_fp64_batched": None if its64b is None else round(its32 / its64b, 3),
                   "ref_f32_iters_per_s": round(K / dtr, 3),
                   "rel_err_vs_fp64": None if err is None else {k: err[k] for k in (
                       "norm_linf", "rel_rms", "max_rel_masked", "max_abs", "rms")},
                   "ref_f32_rel_err_vs_fp64": None if errr is None else {k: errr[k] for k in ("norm_linf", "rel_rms")},
                   "per_kernel": per_stage,
                   "host_to_result": {"seconds": round(t_e2e, 4), "iters_per_s": round(K / t_e2e, 2),
                                      "includes": "host->device upload of exposures, LinkMaps and PSFs; set-up; "
                                                  f"{K} iterations; estimate download"}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                     "algorithmic_bytes_per_launch": sb[dom], "avg_launch_ms": round(ms, 4),
                     "traffic_source": tpath.name if traffic is not None else None,
                     "linkmap_unit": link_unit,
                     "iteration_bytes": b_iter, "iteration_achieved": round(it_gbs, 1),
                     "iteration_frac": round(it_gbs / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
