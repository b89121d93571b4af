#!/usr/bin/env python3
"""Throughput of the Fourier-projection hot path on one MI355X.

A step renders every view of the workload once (direct sum over the Hermitian
half-plane + C2R), with the particles already registered in HBM.  The metric
is particle*frequency terms per second: N * 2*cx*cy terms per view.

  python bench.py [--steps K] [--warmup W] [--workload lognormal_field] [--form auto|plain|...]

Prints one JSON line (rank 0, one GPU).  --gpus above 1 exits: this build
measures one GPU.
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

METRIC = "particle·freq terms/s and views/s at rel-L2≤1e-5: MFMA split vs plain FP32"
PEAKS = {  # MI355X published dense peaks
    "hbm_GBps": 8000.0,
    "fp32_TFLOPs": 157.3,
    "bf16_TFLOPs": 2500.0,
    "fp16_TFLOPs": 2500.0,
}
FLOPS_PER_TERM = 8.0  # one complex multiply-accumulate of w*Phi by exp(-i theta)
# Arithmetic type of each form's products (the MFMA forms accumulate in FP32; the split forms
# multiply hi/lo halves of that type; every form reduces its partial sums in FP64).
TILES = {"default": 0, "half_x": 1, "half_y": 2, "quarter": 3}  # enum VisTile
TILE_NAMES = {v: k for k, v in TILES.items()}
DTYPE = {"plain": "f32", "mfma_f16": "f16", "mfma_split_f16": "f16", "mfma_split_bf16": "bf16"}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(name: str, n: int | None):
    from libmeshlessvis_amd import workloads as W

    if name == "lognormal_field":
        return W.lognormal_field(n or 4_194_304)
    if name == "clustered_wendland":
        return W.clustered_wendland(n or 262_144)
    if name == "uniform_signed_gaussian":
        return W.uniform_signed_gaussian(n or 1_048_576)
    if name == "uniform_gaussian_4096":
        return W.uniform_gaussian_4096()
    if name in ("timing_test", "timing_test_factor1"):
        return W.timing_test(n or 1_048_576, factor=1 if name.endswith("factor1") else 0)
    raise SystemExit(f"unknown workload {name}")


def sample_indices(cutoff, rng, low=8, n_random=512):
    """Stratified frequency sample: every index with |x| <= low and y <= low, plus n_random others.
    Returns (idx, weight) so that sum(weight * |z[idx]|^2) estimates the sum over the whole grid."""
    cx, cy = cutoff
    total = 2 * cx * cy
    x0 = np.arange(2 * cx)
    x = np.where(x0 > cx, x0 - 2 * cx, x0)
    low_idx = [y * 2 * cx + i for y in range(min(cy, low + 1)) for i in range(2 * cx) if abs(x[i]) <= low]
    low_idx = np.array(low_idx, dtype=np.int64)
    rest = np.setdiff1d(np.arange(total), low_idx)
    pick = rng.choice(rest, size=min(n_random, rest.size), replace=False)
    idx = np.concatenate([low_idx, pick])
    weight = np.concatenate([np.ones(low_idx.size), np.full(pick.size, rest.size / max(1, pick.size))])
    return idx, weight


def rel_l2_sampled(F_flat, F64_at, idx, weight) -> float:
    err = np.abs(F_flat[idx].astype(np.complex128) - F64_at) ** 2
    ref = np.abs(F64_at) ** 2
    return float(np.sqrt((weight * err).sum() / (weight * ref).sum()))


def time_renders(cfgs, ds, steps, warmup):
    for _ in range(warmup):
        for c in cfgs:
            c.render(ds)
    for c in cfgs:
        c.synchronize()
        c.reset_timing()
        c.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        for c in cfgs:
            c.render(ds)
    for c in cfgs:
        c.synchronize()
    elapsed = time.perf_counter() - t0
    launches, kernel_ms = 0, 0.0
    for c in cfgs:
        n, ms = c.kernel_timing()
        launches += n
        kernel_ms += ms
        c.set_timing(False)
    return elapsed, launches, kernel_ms


def host_info() -> dict:
    """The host the CPU baseline runs on: usable cores, CPU model, OMP_NUM_THREADS."""
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    return {"nproc": nproc, "cpu_count": os.cpu_count(), "cpu_model": model,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def baseline_threads(info: dict) -> int:
    """Every core the process may use (its affinity mask, nproc), within the host's CPU share
    for this process when OMP_NUM_THREADS sets one."""
    n = info["nproc"]
    omp = info["OMP_NUM_THREADS"]
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(wl, threads: int | None = None):
    """The FP32 restatement of the reference CPU direct sum, timed on a bounded sample of view 0,
    on the host's cores (baseline_threads) unless told otherwise."""
    import oracle

    info = host_info()
    threads = threads or baseline_threads(info)
    vw = wl.views[0]
    n_sub = min(wl.n, 262_144)
    cy = wl.cutoff[1]
    row_stride = max(1, cy // 16)
    rows = len(range(0, cy, row_stride))
    sub = (wl.positions[:n_sub], wl.weights[:n_sub], None if wl.radii is None else wl.radii[:n_sub], wl.basis)
    t0 = time.perf_counter()
    oracle.port_spectrum32([sub], wl.cutoff, wl.step, vw.u, vw.v, row_stride=row_stride, nthreads=threads)
    dt = time.perf_counter() - t0
    terms = n_sub * 2 * wl.cutoff[0] * rows
    return {
        "value": terms / dt,
        "unit": "terms/s",
        "cores": threads,
        "host": info,
        "kind": "port",
        "sample": f"view 0, {rows} of {cy} v-rows x {n_sub} of {wl.n} particles ({terms:.3e} terms, {dt:.1f} s)",
    }


def kernel_name(wl, form: str) -> str:
    if form == "plain":
        return "direct_sum_kernel"
    varying = wl.radii is not None and float(wl.radii.min()) != float(wl.radii.max())
    return "mfma_binned_kernel" if varying and wl.basis != "GAUSSIAN" else "mfma_sum_kernel"


def measured_traffic(workload: str, form: str):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc summary, if any."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    entry = json.loads(f.read_text()).get(f"{workload}/{form}")
    return entry["bytes"] if entry else None


def calibrate_copy(mib: int) -> None:
    """One elementwise copy of a known size on the GPU, for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes_read": 4 * n, "calibrate_copy_bytes_written": 4 * n}), flush=True)


def batched_bench(args) -> None:
    """16 timesteps x 32 views of drifting clusters, all 512 renders per step in one batched call
    (one captured HIP graph: per-job prep, one persistent MFMA launch, reductions, C2Rs)."""
    from libmeshlessvis_amd import host
    from libmeshlessvis_amd import workloads as W
    import oracle

    t = time.perf_counter()
    wls = W.drifting_clusters()
    form = "mfma_split_f16" if args.form == "auto" else args.form
    sets = [host.Dataset([wl.group()]) for wl in wls]
    pairs, cfgs, meta = [], [], []
    for ds, wl in zip(sets, wls):
        for vw in wl.views:
            cfgs.append(host.Config(cutoff=wl.cutoff, samples=wl.samples, step=wl.step, u=vw.u, v=vw.v, form=form,
                                    phi_mode=wl.phi_mode))
            pairs.append(ds)
            meta.append((wl, vw))
    if args.h_bin_ratio > 0.0:
        for c in cfgs:
            c.cs.h_bin_ratio = args.h_bin_ratio
    for c in cfgs:
        c.cs.tile = TILES[args.tile]
    for ds in sets:
        cfgs[0].register(ds)
    tuned = None
    if args.autotune > 0 and form != "plain":
        # One setting for the whole batch (a batch is one graph only when its renders share their
        # bin settings): tuned on view 0 of the largest timestep.
        t = time.perf_counter()
        big = max(range(len(wls)), key=lambda k: wls[k].n)
        probe = cfgs[big * len(wls[0].views)]
        tuned_ms = probe.autotune(sets[big], args.autotune)
        cs = probe.cs
        tuned = {"tolerance": args.autotune, "nodes": cs.number_of_interpolation_nodes,
                 "h_bin_ratio": round(cs.h_bin_ratio, 4), "tile": TILE_NAMES[cs.tile], "spectrum_ms": tuned_ms,
                 "timestep": big}
        for c in cfgs:
            c.cs.number_of_h_bins = 0
            c.cs.number_of_interpolation_nodes = cs.number_of_interpolation_nodes
            c.cs.h_bin_ratio = cs.h_bin_ratio
            c.cs.tile = cs.tile
        log(f"autotune: {tuned} ({time.perf_counter() - t:.1f} s)")
    n_total = sum(wl.n for wl in wls)
    terms = sum(wl.terms_per_view * len(wl.views) for wl in wls)
    log(f"workload drifting_clusters: {len(wls)} timesteps, N={n_total} total, {len(cfgs)} renders/step "
        f"({time.perf_counter() - t:.1f} s to generate and register)")

    def run(k):
        for _ in range(k):
            host.render_batched(pairs, cfgs)
        cfgs[0].synchronize()

    t = time.perf_counter()
    run(max(2, args.warmup))  # a batch runs without a graph the first time, is captured the second
    log(f"warmup (graph capture, bins) {time.perf_counter() - t:.2f} s")
    cfgs[0].reset_timing()
    cfgs[0].set_timing(True)
    t0 = time.perf_counter()
    run(args.steps)
    elapsed = time.perf_counter() - t0
    launches, graph_ms = cfgs[0].kernel_timing()
    cfgs[0].set_timing(False)
    value = terms * args.steps / elapsed
    log(f"{form} batched: {elapsed:.3f} s for {args.steps} x {len(cfgs)} renders -> {value:.3e} terms/s "
        f"({graph_ms / max(1, launches):.2f} ms per batch graph)")

    plain = None
    if not args.no_plain:
        for c in cfgs:
            c.form = "plain"
        t = time.perf_counter()
        run(1)
        p_elapsed = time.perf_counter() - t
        for c in cfgs:
            c.form = form
        plain = {"value": terms / p_elapsed, "unit": "terms/s", "renders": len(cfgs),
                 "speedup": value / (terms / p_elapsed)}
        log(f"plain FP32 (same 512 renders, one at a time): {terms / p_elapsed:.3e} terms/s; {form} is "
            f"{plain['speedup']:.1f}x")

    rel = None
    if not args.no_accuracy:
        run(1)
        rng = np.random.default_rng(0)
        errs = []
        for k in rng.choice(len(cfgs), 4, replace=False):
            wl, vw = meta[k]
            idx, weight = sample_indices(wl.cutoff, rng, n_random=256)
            F = cfgs[k].spectrum().reshape(-1)
            F64 = oracle.spectrum64([wl.group()], wl.cutoff, wl.step, vw.u, vw.v, wl.phi_mode, idx=idx)
            errs.append(rel_l2_sampled(F, F64, idx, weight))
        rel = max(errs)
        log(f"rel-L2 vs FP64 oracle (4 sampled renders, stratified frequencies): max {rel:.3e}")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wls[0])
        log(f"cpu baseline: {cpu['value']:.3e} terms/s on {cpu['cores']} threads ({cpu['sample']})")

    for ds in sets:
        cfgs[0].unregister(ds)
    animation = None
    if not args.no_animation:
        animation = animation_upload(sets, wls, cfgs)
        log(f"animation with uploads ({animation['timesteps']} timesteps, {animation['upload_bytes'] / 1e6:.0f} MB): "
            f"synchronous registration {animation['sync_s']:.3f} s, asynchronous {animation['async_s']:.3f} s")

    avg_ms = graph_ms / max(1, launches)
    achieved = FLOPS_PER_TERM * terms / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else None
    for c in cfgs:
        c.close()
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[form],
        "data": "synthetic",
        "config": {"workload": "drifting_clusters_16x32", "timesteps": len(wls), "views_per_timestep": 32,
                   "particles_total": n_total, "image": "256x256", "basis": "WENDLAND_D3_C2",
                   "phi_mode": "analytic", "form": form, "batched": True, "autotune": tuned},
        "views_per_s": len(cfgs) * args.steps / elapsed,
        "rel_l2_vs_fp64": rel,
        "plain_fp32": plain,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAKS["fp16_TFLOPs"], "unit": "TFLOP/s",
                     "frac": achieved / PEAKS["fp16_TFLOPs"] if achieved else None, "traffic": None,
                     "kernel": "mfma_binned_kernel (whole batch graph)", "kernel_ms_per_launch": avg_ms},
        "cpu_baseline": cpu,
        "animation_upload": animation,
    }
    print(json.dumps(out), flush=True)


def animation_upload(sets, wls, cfgs) -> dict:
    """Timesteps rendered one after another from host data, each registered (H2D) before its renders:
    synchronous registration vs asynchronous registration of timestep t+1 while t renders.  The
    32 views of a timestep are one batched call; both loops end with every image on the host."""
    from libmeshlessvis_amd import host
    per = len(cfgs) // len(sets)

    def loop(async_upload: bool) -> float:
        t0 = time.perf_counter()
        if async_upload:
            cfgs[0].register_async(sets[0])
        for t, ds in enumerate(sets):
            if not async_upload:
                cfgs[0].register(ds)
            group = cfgs[t * per:(t + 1) * per]
            host.render_batched([ds] * per, group)
            if async_upload and t + 1 < len(sets):
                cfgs[0].register_async(sets[t + 1])
            for c in group:
                c.image()
            cfgs[0].unregister(ds)
        return time.perf_counter() - t0

    loop(True)  # warm: bins and graphs of every timestep
    sync_s, async_s = loop(False), loop(True)
    upload = sum(wl.n * (16 + (4 if wl.radii is not None else 0)) for wl in wls)
    return {"timesteps": len(sets), "renders": len(cfgs), "upload_bytes": upload, "sync_s": sync_s,
            "async_s": async_s}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="lognormal_field")
    ap.add_argument("--n", type=int, default=None, help="override the particle count")
    ap.add_argument("--form", default="auto")
    ap.add_argument("--views", type=int, default=0, help="use only the first K views (0: all)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP32 comparison run")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-call timing of the same views (one graph per step)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-animation", action="store_true",
                    help="drifting_clusters: skip the per-timestep upload (registration) measurement")
    ap.add_argument("--autotune", type=float, default=0.0, metavar="TOL",
                    help="before timing, let vis_autotune pick nodes/bin ratio within TOL of the plain kernel")
    ap.add_argument("--h-bin-ratio", type=float, default=0.0, metavar="R",
                    help="ratio of the h-bins of the MFMA forms (0: the library's default for the kind)")
    ap.add_argument("--tile", default="default", choices=list(TILES),
                    help="MFMA output tile variant (VisTile)")
    ap.add_argument("--cull", action="store_true",
                    help="render with cull_fully_aliased_terms set (drop terms wholly outside the period window)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (views would shard "
              "across GPUs with no communication, see the design notes)", file=sys.stderr)
        sys.exit(2)

    from libmeshlessvis_amd import host
    import oracle

    if args.workload == "drifting_clusters":
        batched_bench(args)
        return

    t = time.perf_counter()
    wl = make_workload(args.workload, args.n)
    log(f"workload {wl.name}: N={wl.n} cutoff={wl.cutoff} samples={wl.samples} views={len(wl.views)} "
        f"({time.perf_counter() - t:.1f} s to generate)")
    form = "mfma_split_f16" if args.form == "auto" else args.form
    if args.views:
        wl.views = wl.views[: args.views]

    ds = host.Dataset([wl.group()])
    cfgs = [host.Config(cutoff=wl.cutoff, samples=wl.samples, step=wl.step, u=vw.u, v=vw.v, form=form,
                        phi_mode=wl.phi_mode, cull=args.cull) for vw in wl.views]
    if args.h_bin_ratio > 0.0:
        for c in cfgs:
            c.cs.h_bin_ratio = args.h_bin_ratio
    for c in cfgs:
        c.cs.tile = TILES[args.tile]
    cfgs[0].register(ds)
    tuned = None
    if args.autotune > 0 and form != "plain":
        t = time.perf_counter()
        tuned_ms = cfgs[0].autotune(ds, args.autotune)
        cs = cfgs[0].cs
        tuned = {"tolerance": args.autotune, "nodes": cs.number_of_interpolation_nodes,
                 "h_bin_ratio": round(cs.h_bin_ratio, 4), "tile": TILE_NAMES[cs.tile], "spectrum_ms": tuned_ms}
        for c in cfgs[1:]:
            c.cs.number_of_h_bins = 0
            c.cs.number_of_interpolation_nodes = cs.number_of_interpolation_nodes
            c.cs.h_bin_ratio = cs.h_bin_ratio
            c.cs.tile = cs.tile
        log(f"autotune: {tuned} ({time.perf_counter() - t:.1f} s)")
    log(f"registered; timing form={form}")
    elapsed, launches, kernel_ms = time_renders(cfgs, ds, args.steps, args.warmup)
    views = args.steps * len(wl.views)
    terms = views * wl.terms_per_view
    value = terms / elapsed
    log(f"{form}: {elapsed:.3f} s for {views} views -> {value:.3e} terms/s; kernel {kernel_ms / max(1, launches):.3f} ms/launch")

    batched = None
    if form != "plain" and len(cfgs) > 1 and not args.no_batched:
        # The same views as one batched call: one captured graph, one persistent launch over all views.
        for _ in range(max(2, args.warmup)):  # the second call captures the batch's graph
            host.render_batched([ds] * len(cfgs), cfgs)
        for c in cfgs:
            c.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            host.render_batched([ds] * len(cfgs), cfgs)
        for c in cfgs:
            c.synchronize()
        b_elapsed = time.perf_counter() - t0
        batched = {"value": terms / b_elapsed, "unit": "terms/s", "ms_per_step": b_elapsed / args.steps * 1e3,
                   "views_per_step": len(cfgs)}
        log(f"{form} batched ({len(cfgs)} views per call): {terms / b_elapsed:.3e} terms/s")

    plain = None
    if form != "plain" and not args.no_plain:
        pcfg = host.Config(cutoff=wl.cutoff, samples=wl.samples, step=wl.step, u=wl.views[0].u, v=wl.views[0].v,
                           form="plain", phi_mode=wl.phi_mode)
        p_elapsed, p_launches, p_ms = time_renders([pcfg], ds, 1, 0)
        pcfg.close()
        p_value = wl.terms_per_view / p_elapsed
        plain = {"value": p_value, "unit": "terms/s", "kernel_ms_per_launch": p_ms / max(1, p_launches),
                 "views": 1, "speedup": value / p_value}
        if batched:
            batched["speedup_vs_plain"] = batched["value"] / p_value
        log(f"plain FP32: {p_value:.3e} terms/s ({p_ms / max(1, p_launches):.1f} ms/view); "
            f"{form} is {value / p_value:.1f}x")

    rel = None
    reference_parity = None
    if not args.no_accuracy:
        rng = np.random.default_rng(0)
        idx, weight = sample_indices(wl.cutoff, rng)
        vw = wl.views[0]
        cfgs[0].render(ds)
        F = cfgs[0].spectrum().reshape(-1)
        t = time.perf_counter()
        F64 = oracle.spectrum64([wl.group()], wl.cutoff, wl.step, vw.u, vw.v, wl.phi_mode, idx=idx)
        rel = rel_l2_sampled(F, F64, idx, weight)
        log(f"rel-L2 vs FP64 oracle (stratified sample of {idx.size} frequencies): {rel:.3e} "
            f"({time.perf_counter() - t:.1f} s)")
        # The reference's CPU direct sum (FP32 port: its loop, FP32 phase and sums) on the same
        # inputs and frequencies; with the workload's Phi mode (analytic: exact transforms in
        # place of the reference's formulas, whose exp(+pi r^2) overflows at these radii).
        t = time.perf_counter()
        Fport = oracle.port_spectrum32([wl.group()], wl.cutoff, wl.step, vw.u, vw.v, phi_mode=wl.phi_mode,
                                       idx=idx).astype(np.complex128)
        e_ref = rel_l2_sampled(Fport, F64, np.arange(idx.size), weight)
        e_par = rel_l2_sampled(F, Fport, idx, weight)
        reference_parity = {"e_ref": e_ref, "e_build": rel, "e_par": e_par, "phi_mode": wl.phi_mode,
                            "frequencies": int(idx.size), "holds": bool(e_par <= e_ref + rel + 1e-6)}
        log(f"reference port: e_ref {e_ref:.3e} e_par {e_par:.3e} ({time.perf_counter() - t:.1f} s)")

    avg_ms = kernel_ms / max(1, launches)
    launch_terms = wl.terms_per_view
    achieved_tflops = FLOPS_PER_TERM * launch_terms / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else None
    if form == "plain":
        bound, peak = "valu", PEAKS["fp32_TFLOPs"]
    else:
        bound, peak = "mfma", PEAKS["bf16_TFLOPs" if DTYPE[form] == "bf16" else "fp16_TFLOPs"]
    roofline = {
        "bound": bound,
        "achieved": achieved_tflops,
        "peak": peak,
        "unit": "TFLOP/s",
        "frac": (achieved_tflops / peak) if achieved_tflops else None,
        "traffic": measured_traffic(wl.name, form),
        "kernel": kernel_name(wl, form),
        "kernel_ms_per_launch": avg_ms,
    }

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl)
        log(f"cpu baseline: {cpu['value']:.3e} terms/s on {cpu['cores']} threads ({cpu['sample']})")

    for c in cfgs[1:]:
        c.close()
    cfgs[0].unregister(ds)
    cfgs[0].close()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[form],
        "data": "synthetic",
        "config": {
            "workload": wl.name,
            "particles": wl.n,
            "cutoff": list(wl.cutoff),
            "image": f"{wl.samples[0]}x{wl.samples[1]}",
            "views_per_step": len(wl.views),
            "basis": wl.basis,
            "phi_mode": wl.phi_mode,
            "form": form,
            "autotune": tuned,
            "h_bin_ratio": args.h_bin_ratio or None,
            "tile": args.tile,
            "cull_fully_aliased_terms": bool(args.cull),
        },
        "views_per_s": views / elapsed,
        "rel_l2_vs_fp64": rel,
        "reference_parity": reference_parity,
        "batched": batched,
        "plain_fp32": plain,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
