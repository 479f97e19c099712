#!/usr/bin/env python3
"""Benchmark of the MI355X bubble finder (one JSON line on stdout).

Workload (the metric's configuration): the batched exact finder at 512^3
float32, side 400 Mpc, 10 redshift slices, radii 100 Mpc * 0.9^k down to the
cell size (47 radii), 2.5e5 point sources per slice.  Slice s uses the same
Gaussian field g rescaled by its linear growth factor, g_s = (D_s/0.25) g with
D_s = linspace(0.10, 0.25, 10), 1+delta_s = exp(g_s)/<exp(g_s)>, sources drawn
with P ~ (1+delta_s)^2 and lognormal strengths, rec_s = (1+delta_s)^2; each
slice's emissivity is scaled (bisection, untimed) to a mean x_HII of ~0.5, as
the configs' recipe asks.

A step is one pass of all radii over all slices (inputs resident in HBM).
`value` = slices * N^3 * n_radii / step time.  The plain per-radius GPU path
(the reference's dataflow) is timed on all slices, one at a time.
The differentiable mode (one redshift at 512^3, same radii) reports the exact
forward's time and the relaxed backward's, and their ratio, per tau.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "voxel·radius updates/s at 512^3 (batched vs plain GPU); grad/forward time"
HBM_PEAK_GBS = 8000.0


def gaussian_field_dev(torch, gen, n, slope=-2.5, sigma=1.2):
    """Gaussian field with P(k) ~ k^slope, std sigma (float32, device)."""
    dev = "cuda"
    white = torch.randn((n, n, n), generator=gen, device=dev, dtype=torch.float32)
    spec = torch.fft.rfftn(white)
    del white
    kx = torch.fft.fftfreq(n, device=dev) * n
    kz = torch.arange(n // 2 + 1, device=dev, dtype=torch.float32)
    k2 = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2
    amp = torch.where(k2 > 0, k2.clamp_min(1e-30) ** (slope / 4.0), torch.zeros_like(k2))
    spec *= amp
    del amp, k2
    g = torch.fft.irfftn(spec, s=(n, n, n))
    del spec
    return (g - g.mean()) / g.std() * sigma


def sources_dev(torch, gen, delta1, n_src):
    """Unit-efficiency emissivity: n_src sources with P(cell) ~ (1+delta)^2,
    lognormal(0, 1) strengths (float32, device, flat)."""
    dev = "cuda"
    n3 = delta1.numel()
    w = (delta1 ** 2).reshape(-1)
    cdf = torch.cumsum(w, 0)
    cdf /= cdf[-1].clone()
    u = torch.rand(n_src, generator=gen, device=dev, dtype=torch.float64)
    cells = torch.searchsorted(cdf, u).clamp_max(n3 - 1)
    strength = torch.exp(torch.randn(n_src, generator=gen, device=dev, dtype=torch.float64))
    emis = torch.zeros(n3, device=dev, dtype=torch.float64)
    emis.index_add_(0, cells, strength)
    return emis.float().contiguous()


def make_inputs(torch, n, side, n_slices, n_src, seed=3001, slope=-2.5, sigma=1.2):
    """Synthetic slices on the device (float32 boxes, reference layout)."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    g = gaussian_field_dev(torch, gen, n, slope, sigma)
    growth = np.linspace(0.10, 0.25, n_slices)
    recs, ions = [], []
    for s in range(n_slices):
        gs = g * float(growth[s] / 0.25)
        e = torch.exp(gs.double())
        e /= e.mean()
        ions.append(sources_dev(torch, gen, e, n_src))
        recs.append((e ** 2).float().reshape(-1).contiguous())
        del e
    torch.cuda.synchronize()
    return recs, ions


def scale_slices(torch, ctx, recs, ions, target=0.5, tol=0.02):
    """Efficiency per slice: each emissivity box is multiplied in place by
    the zeta (bisection in log zeta on the fused path, untimed) that gives
    the slice a mean x_HII of target within tol -- the configs' recipe
    scales the emissivity to x_HII ~ 0.5.  Returns the zetas."""
    zetas = []
    out = torch.empty_like(recs[0])
    for rec, ion in zip(recs, ions):
        lo, hi = -6.0, 14.0
        base = ion.clone()
        zeta = 1.0
        for _ in range(30):
            mid = 0.5 * (lo + hi)
            zeta = float(np.exp(mid))
            torch.mul(base, zeta, out=ion)
            _, neu = ctx.run_slice(rec, ion, xhii=out)
            if abs(1 - neu - target) <= tol:
                break
            lo, hi = (mid, hi) if 1 - neu < target else (lo, mid)
        zetas.append(zeta)
        del base
    torch.cuda.synchronize()
    return zetas


def measured_traffic(kernel, n, n_radii, radius_batch):
    """HBM bytes per launch of a kernel as measured by the rocprofv3 counter
    runs recorded under profiles/ (null when the workload differs)."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            rec = json.load(f)
    except OSError:
        return None
    w = rec.get("workload", {})
    if (w.get("n"), w.get("radii"), w.get("radius_batch")) != (n, n_radii, radius_batch):
        return None
    return rec.get("bytes_per_launch", {}).get(kernel)


UNEVEN_DX = 300.0 / 256


def uneven_inputs(torch, synth, n_boxes=64, seed=4000):
    """Inputs of the parameter-sweep batch: n_boxes boxes with N drawn
    uniformly from {64, 96, 128, 192, 256} at a fixed cell size dx = 300/256
    Mpc (sides 75-300 Mpc), P(k) slope ~ U[-2.8, -2.2], sigma 1.2,
    4000 (N/128)^3 sources, efficiency zeta ~ 1.5 lognormal(0, 0.5); radius
    ladder per size from R_max = side/4 down by 0.9 to the cell size.
    Returns (sizes, ladders by size, rec boxes, emissivity boxes) with the
    boxes on the device."""
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    dx = UNEVEN_DX
    sizes = [int(v) for v in rng.choice([64, 96, 128, 192, 256], n_boxes)]
    ladders = {n: synth.geometric_ladder(n * dx / 4, 0.9, dx) for n in sorted(set(sizes))}
    a, b = [], []
    for n in sizes:
        slope = float(rng.uniform(-2.8, -2.2))
        zeta = 1.5 * float(np.exp(0.5 * rng.standard_normal()))
        g = gaussian_field_dev(torch, gen, n, slope, 1.2)
        e = torch.exp(g.double())
        e /= e.mean()
        b.append((sources_dev(torch, gen, e, int(round(4000 * (n / 128) ** 3))) * zeta).contiguous())
        a.append((e ** 2).float().reshape(-1).contiguous())
        del g, e
    return sizes, ladders, a, b


def bench_uneven(torch, sf, synth, n_boxes=64, steps=3, seed=4000, tune=True):
    """Parameter-sweep batch (uneven_inputs): all boxes in one hipGraph
    launch (sf21_hii_run_multi_dev: boxes of one size in order, sizes as
    parallel branches); launch choices autotuned per size on its first box
    (not timed); per-size throughput from each box run alone."""
    dx = UNEVEN_DX
    sizes, ladders, a, b = uneven_inputs(torch, synth, n_boxes, seed)
    ctxs = {n: sf.HiiContext(n, n * dx, ladders[n], sf.INPUT_DIRECT, sf.PATH_FUSED)
            for n in sorted(set(sizes))}
    boxes = [ctxs[n] for n in sizes]
    if tune:
        for n, c in ctxs.items():
            k = sizes.index(n)
            c.autotune(a[k], b[k])
    outs = [torch.empty_like(x) for x in a]
    sf.run_multi(boxes, a, b, out_list=outs)  # capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        _, neu = sf.run_multi(boxes, a, b, out_list=outs)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    units = sum(n ** 3 * len(ladders[n]) for n in sizes)
    per = {}
    for n in sorted(set(sizes)):
        idx = [k for k, v in enumerate(sizes) if v == n]
        boxes[idx[0]].run_slice(a[idx[0]], b[idx[0]], xhii=outs[idx[0]])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for k in idx:
            boxes[k].run_slice(a[k], b[k], xhii=outs[k])
        torch.cuda.synchronize()
        mb = (time.perf_counter() - t1) * 1e3 / len(idx)
        per[str(n)] = {"boxes": len(idx), "radii": len(ladders[n]), "ms_per_box_alone": round(mb, 3),
                       "voxel_radius_per_s_alone": round(n ** 3 * len(ladders[n]) / (mb * 1e-3), 1)}
    for c in ctxs.values():
        c.close()
    return {
        "workload": (f"{n_boxes} boxes, N ~ U{{64,96,128,192,256}} at dx {dx:g} Mpc, slope ~ U[-2.8,-2.2], "
                     f"zeta ~ 1.5 lognormal(0.5), ladder side/4 x 0.9^k to dx; one hipGraph launch"),
        "voxel_radius_total": units,
        "ms_per_launch": round(ms, 3),
        "voxel_radius_per_s": round(units / (ms * 1e-3), 1),
        "mean_xhii_range": [round(1 - max(neu), 4), round(1 - min(neu), 4)],
        "per_size": per,
    }


def bench_diff(torch, sf, n, side, radii, n_src, taus=(0.05, 0.02, 0.01), reps=3, seed=5005,
               tune=True):
    """Differentiable mode at one redshift: Gaussian g (sigma 1.2, P ~ k^-2.5),
    emissivity pattern from n_src sources ~ (1+delta)^2, zeta set for
    <x> ~ 0.5.  Times the exact forward (fields, bubble finder, T, P(k)) and
    the relaxed backward per tau: recomputing every per-radius field in both
    sweeps, keeping the first sweep's for the second, and with the forward
    keeping them for the backward."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    g = gaussian_field_dev(torch, gen, n).reshape(-1).contiguous()
    e = torch.exp(g.double())
    e /= e.mean()
    emis = sources_dev(torch, gen, e, n_src)
    del e
    ctx = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
    x = torch.empty_like(g)
    grad = torch.empty_like(g)
    if tune:
        e = torch.exp(g.double())
        e /= e.mean()
        ctx.autotune((e ** 2).float().contiguous(), (emis * 75.0).contiguous())
        del e
    # zeta for <x> ~ 0.5 (bisection in log zeta on the exact forward)
    lo, hi = 1e-2, 1e6
    zeta, mx = 1.0, 0.0
    for _ in range(14):
        zeta = math.sqrt(lo * hi)
        _, mx, _ = ctx.diff_forward(g, emis, zeta, 20, x_out=x)
        if abs(mx - 0.5) < 0.02:
            break
        if mx < 0.5:
            lo = zeta
        else:
            hi = zeta
    rng = np.random.default_rng(7)
    _, _, pk = ctx.diff_forward(g, emis, zeta, 20, x_out=x)
    cp = rng.standard_normal(20) / np.where(pk > 0, pk, 1.0)
    ctx.diff_store_fields(0)  # recompute in both sweeps
    ctx.diff_backward(g, emis, zeta, taus[0], 1.0, cp, 20, grad_out=grad)  # warm: ball tables
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.diff_forward(g, emis, zeta, 20, x_out=x)
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t0) * 1e3 / reps
    bwd = {}
    for tau in taus:
        t0 = time.perf_counter()
        for _ in range(reps):
            _, gz, mxt, _ = ctx.diff_backward(g, emis, zeta, tau, 1.0, cp, 20, grad_out=grad)
        torch.cuda.synchronize()
        bwd[str(tau)] = round((time.perf_counter() - t0) * 1e3 / reps, 2)
        print(f"[bench] diff tau={tau}: backward (recompute) {bwd[str(tau)]:.1f} ms, <x~> {mxt:.4f}, "
              f"dL/dzeta {gz:.4e}", file=sys.stderr, flush=True)
    # default mode: the first sweep's pairs kept for the second when they fit
    bwd_stored = None
    try:
        ctx.diff_store_fields(1)
        bwd_stored = {}
        for tau in taus:
            ctx.diff_backward(g, emis, zeta, tau, 1.0, cp, 20, grad_out=grad)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                ctx.diff_backward(g, emis, zeta, tau, 1.0, cp, 20, grad_out=grad)
            torch.cuda.synchronize()
            bwd_stored[str(tau)] = round((time.perf_counter() - t0) * 1e3 / reps, 2)
        print(f"[bench] diff backward with stored fields {bwd_stored}", file=sys.stderr, flush=True)
    except sf.Sf21Error as e:  # does not fit
        bwd_stored = {"error": str(e)}
    # forward keeps the fields (mode 3): training-loop pairs of forward and
    # backward on the same inputs; the forward writes the pairs as it finds
    # the centres, the backward's first sweep reads them
    kept = None
    try:
        ctx.diff_store_fields(3)
        ctx.diff_forward(g, emis, zeta, 20, x_out=x)
        ctx.diff_backward(g, emis, zeta, taus[0], 1.0, cp, 20, grad_out=grad)
        torch.cuda.synchronize()
        kf, kb, used = 0.0, {}, True
        for tau in taus:
            tf = tb = 0.0
            for _ in range(reps):
                t0 = time.perf_counter()
                ctx.diff_forward(g, emis, zeta, 20, x_out=x)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                ctx.diff_backward(g, emis, zeta, tau, 1.0, cp, 20, grad_out=grad)
                torch.cuda.synchronize()
                t2 = time.perf_counter()
                used = used and ctx.diff_last_kept
                tf += t1 - t0
                tb += t2 - t1
            kf += tf
            kb[str(tau)] = round(tb * 1e3 / reps, 2)
        kf_ms = kf * 1e3 / (reps * len(taus))
        ctx.reset_stats()
        ctx.profile(True)
        ctx.diff_backward(g, emis, zeta, taus[0], 1.0, cp, 20, grad_out=grad)
        kst3 = ctx.kernel_stats()
        ctx.profile(False)
        kept = {"forward_ms": round(kf_ms, 2), "backward_ms": kb, "backward_read_kept_fields": used,
                "backward_kernels_ms": {k: [c, round(t, 3)] for k, (c, t) in kst3.items()},
                "grad_over_forward": {k: round(v / kf_ms, 3) for k, v in kb.items()},
                "forward_plus_backward_ms": {k: round(kf_ms + v, 2) for k, v in kb.items()},
                "forward_plus_backward_ms_recompute": {k: round(fwd_ms + v, 2) for k, v in bwd.items()}}
        print(f"[bench] diff forward keeping fields {kf_ms:.1f} ms, backward {kb}, "
              f"grad/forward {kept['grad_over_forward']}", file=sys.stderr, flush=True)
    except sf.Sf21Error as e:  # does not fit
        kept = {"error": str(e)}
    ctx.diff_store_fields(0)
    ctx_tuning = ctx.tuning()
    # per-kernel split of one backward (recompute mode)
    ctx.reset_stats()
    ctx.profile(True)
    ctx.diff_backward(g, emis, zeta, taus[0], 1.0, cp, 20, grad_out=grad)
    kst = ctx.kernel_stats()
    ctx.profile(False)
    ctx.close()
    worst = max(bwd.values())
    return {
        "workload": (f"differentiable mode, {n}^3 float32, side {side:g} Mpc, one redshift, "
                     f"{len(radii)} radii, {n_src} sources, zeta {zeta:.4g} (<x> {mx:.3f}), "
                     f"20 P(k) bins"),
        "forward_ms": round(fwd_ms, 2),
        "backward_ms": bwd,
        "grad_over_forward": {k: round(v / fwd_ms, 3) for k, v in bwd.items()},
        "grad_over_forward_max": round(worst / fwd_ms, 3),
        "backward_ms_fields_stored": bwd_stored,
        "tuning": ctx_tuning,
        "grad_over_forward_fields_stored": ({k: round(v / fwd_ms, 3) for k, v in bwd_stored.items()}
                                            if bwd_stored and "error" not in bwd_stored else None),
        "forward_keeps_fields": kept,
        "backward_kernels_ms": {k: [c, round(t, 3)] for k, (c, t) in kst.items()},
    }


def bench_generic(torch, sf, synth, n=600, steps=2, seed=6006, tune=True):
    """One slice at a box side without a register plan (600 = 4*2*3*5*5, the
    N_smoothed of SimFast21's sample ini): the same recipe scaled to n (its
    emissivity scaled to mean x_HII ~0.5, untimed), one radius per step
    through LDS line transforms; then the differentiable mode on the same
    fields (g = log(rec)/2, the scaled emissivity, zeta 1): one exact forward
    and one relaxed backward (tau 0.02, fields recomputed) after a warm-up."""
    side = 400.0 * n / 512
    radii = synth.geometric_ladder(100.0 * n / 512, 0.9, side / n)
    n_src = int(round(2.5e5 * (n / 512) ** 3))
    recs, ions = make_inputs(torch, n, side, 3, n_src, seed=seed)
    recs, ions = recs[1:2], ions[1:2]  # the middle growth factor of the batch recipe
    torch.cuda.empty_cache()
    ctx = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
    out = torch.empty_like(recs[0])
    zeta = scale_slices(torch, ctx, recs, ions)[0]
    # the line passes' XCD block order (results identical for every choice);
    # without tuning, the choice cached in $SF21_TUNE_CACHE (or the default)
    tune_ms = ctx.autotune(recs[0], ions[0]) if tune else None
    order = ctx.line_order
    ctx.run_slice(recs[0], ions[0], xhii=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        _, neu = ctx.run_slice(recs[0], ions[0], xhii=out)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    ctx.reset_stats()
    ctx.profile(True)
    ctx.run_slice(recs[0], ions[0], xhii=out)
    torch.cuda.synchronize()
    stats = ctx.kernel_stats()
    ctx.profile(False)
    # differentiable mode on the same fields (Delta = exp(g)/<exp(g)> = the
    # slice's (1 + delta), ion = the scaled emissivity)
    g = (0.5 * torch.log(recs[0])).contiguous()
    grad = torch.empty_like(g)
    ctx.diff_forward(g, ions[0], 1.0, 20, x_out=out)
    ctx.diff_backward(g, ions[0], 1.0, 0.02, 1.0, None, 20, grad_out=grad)  # ball tables
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, dmx, _ = ctx.diff_forward(g, ions[0], 1.0, 20, x_out=out)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.diff_backward(g, ions[0], 1.0, 0.02, 1.0, None, 20, grad_out=grad)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    diff = {"forward_ms": round((t1 - t0) * 1e3, 2), "backward_ms": round((t2 - t1) * 1e3, 2),
            "grad_over_forward": round((t2 - t1) / (t1 - t0), 3), "mean_x": round(dmx, 5),
            "tau": 0.02, "fields": "recomputed"}
    del g, grad
    ctx.close()
    v = n ** 3
    algo = {"gen_filter_i": 16 * v, "gen_inv_j": 16 * v, "gen_inv_p_crit": 8 * v, "gen_fwd": 16 * v}
    kern = {}
    for k, (c, t) in stats.items():
        e = {"launches": c, "ms": round(t, 3), "avg_launch_ms": round(t / c, 4)}
        if k in algo:
            e["hbm_frac"] = round(algo[k] / (t / c * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        kern[k] = e
    return {"n": n, "side": side, "radii": len(radii), "ms_per_slice": round(ms, 2),
            "voxel_radius_per_s": v * len(radii) / (ms * 1e-3), "mean_xhii": round(1 - neu, 6),
            "line_order": order, "tune_ms_per_slice": None if tune_ms is None else round(tune_ms, 2),
            "zeta": round(zeta, 4), "diff": diff,
            "algorithmic_bytes_per_voxel_radius": 40, "kernels": kern}


def _host_cores():
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp:
        cores = min(cores, int(omp))
    return cores


def cpu_baseline(n, side, radii_full, rec, ion):
    """Oracle (CPU restatement, reference dataflow, FP32 typing, own OpenMP
    FFT) on a bounded sample of the workload: one slice of the batch (host
    copy of the timed inputs) with every fourth radius of the ladder, so the
    sample spans the main loop (ball FFT convolutions) and the small loop
    (stencil) as the full ladder does."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as orc
    sample = radii_full[::4]
    t0 = time.perf_counter()
    orc.hii_slice(n, side, sample, rec, ion, precision=32, dilation=0)
    dt = time.perf_counter() - t0
    n_main = sum(1 for R in sample if R >= 2 * side / n)
    return {
        "value": n ** 3 * len(sample) / dt,
        "unit": "voxel·radius updates/s",
        "cores": _host_cores(),
        "kind": "port",
        "sample": (f"one {n}^3 slice of the batch (x_HII ~0.5), every 4th radius of the ladder "
                   f"({len(sample)} of {len(radii_full)}: {n_main} main-loop, {len(sample) - n_main} "
                   f"small-loop); reference dataflow (2 r2c + per radius filter, 2 c2r, criterion, "
                   f"ball + 2 r2c + c2r or stencil), FP32 typing, own OpenMP CPU FFT (not FFTW); "
                   f"{dt:.1f} s wall incl. the forward transforms"),
    }


def cpu_configs(torch, sf, synth, repeats=3):
    """CPU baselines of the two smaller configs, whole slices with all radii
    (median of `repeats`): c1, 128^3 at side 200 Mpc, 4000 sources, radii
    60 Mpc x 0.9^k, x_HII ~0.5; c2, 256^3 at side 300 Mpc, 30000 sources,
    radii 80 Mpc x 0.9^k, at the recipe's floor (zeta 1e-3, x_HII ~0.31: 0.2
    is below it) and x_HII ~0.5 and ~0.8 (zeta by bisection on the GPU)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as orc
    out = {"cores": _host_cores(), "kind": "port",
           "method": ("oracle reference dataflow, FP32 typing, own OpenMP CPU FFT (not FFTW), "
                      f"median of {repeats} whole slices with all radii")}
    for name, n, side, n_src, rmax, seed, targets in (
            ("c1", 128, 200.0, 4000, 60.0, 1001, [0.5]),
            ("c2", 256, 300.0, 30000, 80.0, 2560, [None, 0.5, 0.8])):
        delta, emis, rec = synth.direct_boxes(n, side, n_src, seed)
        radii = synth.geometric_ladder(rmax, 0.9, side / n)
        ctx = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED)
        a = torch.from_numpy(rec.ravel()).cuda()
        for tgt in targets:
            e = torch.from_numpy(emis.ravel()).cuda()
            if tgt is None:
                zeta = 1e-3
                e.mul_(zeta)
            else:
                zeta = scale_slices(torch, ctx, [a], [e], target=tgt)[0]
            ion = e.cpu().numpy()
            _, neu = ctx.run_slice(a, e)
            times = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                orc.hii_slice(n, side, radii, rec, ion, precision=32, dilation=0)
                times.append(time.perf_counter() - t0)
            med = float(np.median(times))
            key = f"{name}_x{1 - neu:.2f}"
            out[key] = {"n": n, "radii": len(radii), "zeta": round(zeta, 6),
                        "mean_xhii_gpu": round(1 - neu, 4), "seconds": [round(t, 2) for t in times],
                        "voxel_radius_per_s": round(n ** 3 * len(radii) / med, 1)}
            print(f"[bench] cpu {key}: median {med:.2f} s", file=sys.stderr, flush=True)
        ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--slices", type=int, default=10)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--radius-batch", type=int, default=0)
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-generic", action="store_true")
    ap.add_argument("--no-cpu-configs", action="store_true",
                    help="skip the CPU timings of the 128^3 and 256^3 configs")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 1 GiB device copy (torch copy_) as the HBM-counter reference")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the bubble finder's slices are independent and a "
              f"multi-GPU run is not part of this build (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import simfast21_amd as sf
    from simfast21_amd import synth

    n = args.n
    side = 400.0 * n / 512
    dx = side / n
    radii = synth.geometric_ladder(100.0 * n / 512, 0.9, dx)
    n_src = int(round(2.5e5 * (n / 512) ** 3))
    S = args.slices
    print(f"[bench] generating {S} slices at {n}^3 ({len(radii)} radii)", file=sys.stderr, flush=True)
    recs, ions = make_inputs(torch, n, side, S, n_src)
    outs = [torch.empty_like(r) for r in recs]

    if args.calib_copy:
        # HBM-counter references of known size: 1 GiB read + 1 GiB written by
        # an elementwise kernel (torch mul) and by the runtime's copy
        src = torch.ones(1 << 28, device="cuda")
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    ctx = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_FUSED,
                        radius_batch=args.radius_batch)
    torch.cuda.synchronize()
    # the recipe's efficiency: every slice scaled to mean x_HII ~0.5 (untimed)
    zetas = scale_slices(torch, ctx, recs, ions)
    print(f"[bench] slice efficiencies {[round(z, 3) for z in zetas]}", file=sys.stderr, flush=True)
    tuned = None
    if args.radius_batch <= 0 and not args.no_tune:
        # launch choices picked by the autotuner on one slice (not timed)
        tune_ms = ctx.autotune(recs[S // 2], ions[S // 2])
        tuned = dict(ctx.tuning(), tune_ms_per_slice=round(tune_ms, 3))
        print(f"[bench] autotuned: {tuned}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        _, neu = ctx.run_batch(recs, ions, out_list=outs)
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, neu = ctx.run_batch(recs, ions, out_list=outs)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    units = S * n ** 3 * len(radii)
    value = units / (ms * 1e-3)
    print(f"[bench] fused {ms:.1f} ms/step, mean x_HII {[round(1 - v, 4) for v in neu]}",
          file=sys.stderr, flush=True)

    # per-kernel HIP-event timings on the context stream (profiled steps)
    ctx.reset_stats()
    ctx.profile(True)
    prof_steps = max(1, min(2, args.steps))
    tp0 = time.perf_counter()
    for _ in range(prof_steps):
        ctx.run_batch(recs, ions, out_list=outs)
    torch.cuda.synchronize()
    prof_ms = (time.perf_counter() - tp0) * 1e3 / prof_steps
    stats = ctx.kernel_stats()
    ctx.profile(False)
    K = ctx.radius_batch
    # algorithmic HBM bytes per launch of each fused kernel (complex64 = 8 B/voxel)
    v = n ** 3
    nb = math.ceil(len(radii) / K)
    avg_k = len(radii) / nb
    algo = {
        "filter_i": v * 8 * (1 + avg_k),
        "cols_inv_j": v * 8 * 2 * avg_k,
        "inv_rows_crit": v * 8 * avg_k,
        "fwd_rows": v * (8 + 8),
        "cols_fwd_j": v * 16,
    }
    dom = max(stats.items(), key=lambda kv: kv[1][1])[0]
    cnt, tot = stats[dom]
    avg_ms = tot / cnt
    roof = None
    if dom in algo:
        ach = algo[dom] / (avg_ms * 1e-3) / 1e9
        traffic = measured_traffic(dom, n, len(radii), K)
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": ("recorded: rocprofv3 FETCH_SIZE/WRITE_SIZE runs of this workload, "
                                   "profiles/hbm_traffic.json (not measured by this run)")
                if traffic is not None else None,
                "kernel": dom, "avg_launch_ms": round(avg_ms, 4),
                "algorithmic_bytes": algo[dom]}
    # whole-step algorithmic roofline (per voxel-radius 80h+4 B, per voxel-slice 80h+25 B)
    hfac = (n // 2 + 1) / n
    step_bytes = S * v * (len(radii) * (80 * hfac + 4) + (80 * hfac + 25))
    step_roof = {"bytes": step_bytes, "achieved_GBs": round(step_bytes / (ms * 1e-3) / 1e9, 1),
                 "frac": round(step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    ctx.close()
    del ctx
    torch.cuda.empty_cache()
    plain = None
    if not args.no_plain:
        # the reference's dataflow on the GPU, one slice at a time, all slices
        pctx = sf.HiiContext(n, side, radii, sf.INPUT_DIRECT, sf.PATH_PLAIN)
        pout = torch.empty_like(recs[0])
        pctx.run_slice(recs[S // 2], ions[S // 2], xhii=pout)
        torch.cuda.synchronize()
        tq0 = time.perf_counter()
        pneu = []
        for k in range(S):
            _, pn = pctx.run_slice(recs[k], ions[k], xhii=pout)
            pneu.append(pn)
        torch.cuda.synchronize()
        pms = (time.perf_counter() - tq0) * 1e3
        plain = {"ms_per_step": round(pms, 1), "ms_per_slice": round(pms / S, 2),
                 "speedup_fused_vs_plain": round(pms / ms, 2),
                 "mean_xhii": [round(1 - v, 5) for v in pneu],
                 "max_mean_diff_vs_fused": float(f"{max(abs(a - b) for a, b in zip(pneu, neu)):.3g}")}
        print(f"[bench] plain {pms:.1f} ms/step ({pms / S:.1f} ms/slice)", file=sys.stderr, flush=True)
        pctx.close()
        del pctx

    uneven = None
    if not args.no_uneven:
        try:
            uneven = bench_uneven(torch, sf, synth, tune=not args.no_tune)
        finally:
            torch.cuda.empty_cache()
        print(f"[bench] uneven batch: {uneven['ms_per_launch']:.1f} ms per launch, "
              f"{uneven['voxel_radius_per_s']:.3g} voxel-radius/s", file=sys.stderr, flush=True)

    diff = None
    if not args.no_diff:
        try:
            diff = bench_diff(torch, sf, n, side, radii, n_src, tune=not args.no_tune)
        finally:
            torch.cuda.empty_cache()
        print(f"[bench] diff forward {diff['forward_ms']:.1f} ms, grad/forward "
              f"{diff['grad_over_forward']}", file=sys.stderr, flush=True)

    generic = None
    if not args.no_generic:
        try:
            generic = bench_generic(torch, sf, synth, tune=not args.no_tune)
        finally:
            torch.cuda.empty_cache()
        print(f"[bench] generic side {generic['n']}: {generic['ms_per_slice']:.1f} ms/slice, "
              f"{generic['voxel_radius_per_s']:.3g} voxel-radius/s", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(n, side, radii, recs[S // 2].cpu().numpy(), ions[S // 2].cpu().numpy())
        except Exception as e:  # the baseline is reported, never required
            cpu = {"error": repr(e)}
    cpu_cfg = None
    if not args.no_cpu and not args.no_cpu_configs:
        try:
            cpu_cfg = cpu_configs(torch, sf, synth)
        except Exception as e:
            cpu_cfg = {"error": repr(e)}
        finally:
            torch.cuda.empty_cache()

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel·radius updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"batched exact bubble finder, {n}^3 float32, side {side:g} Mpc, {S} slices, "
                         f"{len(radii)} radii ({100.0 * n / 512:g} Mpc x 0.9^k to the cell size), "
                         f"{n_src} sources/slice, emissivity of each slice scaled to mean x_HII ~0.5"),
            "n": n, "slices": S, "radii": len(radii), "radius_batch": K, "tuning": tuned,
            "slice_zeta": [round(z, 4) for z in zetas], "mean_xhii": [round(1 - v, 4) for v in neu],
        },
        "roofline": roof,
        "step_roofline": step_roof,
        "profiled_ms_per_step": round(prof_ms, 3),
        "kernels_ms": {k: [c, round(t, 3)] for k, (c, t) in stats.items()},
        "plain": plain,
        "uneven_batch": uneven,
        "diff": diff,
        "generic_side": generic,
        "cpu_baseline": cpu,
        "cpu_baseline_configs": cpu_cfg,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
