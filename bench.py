"""FDTD benchmark: Yee cell-updates/s and HBM-roofline fraction on one MI355X.

Default workload (C3): 3D Yee 384^3, FP32, silicon slab waveguide (eps 12.1,
22 cells thick in z, 50 wide in y, along x) on SiO2 (eps 2.1), 12-cell CPML,
Ez point dipole (Gaussian pulse), 2000 Yee steps, 2 x-normal flux planes x 64
frequencies accumulated on the fly -- the largest plain-stepper
configuration that fits one GPU.  One bench "step" is one full C3 run (2000
Yee steps); `value` is whole-job cell-updates/s with inputs resident in HBM.
The roofline object is the dominant kernel's: the one-pass interior kernel of
the split 3D step (52 algorithmic bytes per cell-update: six field
components read and written and the cell-centred eps, no CPML terms inside
the interior box), timed with HIP events on its stream.

C2 (2D TE 2048^2, 10000 steps, 4 x 100-frequency flux lines) and C1 run with
--workload c2 / c1.

    python bench.py [--steps K] [--warmup W] [--workload c1|c2|c3|c4|c5] [--no-cpu-baseline]

c4: batched forward+adjoint design iterations of 64 uneven TM variants.
c5: surrogate-assisted inverse design -- FDTD dataset, BF16 matrix-core
training, then plain-FDTD and surrogate-assisted Adam loops from one start.

Prints one JSON line on stdout (progress goes to stderr).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "meep-torch_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import torch  # noqa: E402

HBM_PEAK_GBS = 8000.0
# FP32 algorithmic bytes per cell-update: whole step (CPML shells included)
ALG_BYTES = {"c1": 32.0, "c2": 29.0, "c3": 58.0}
# ... and for the dominant kernel's cells: the 2D interior tiles are counted
# at the whole-step figure (their tiles run the same fields), as are the
# two-pass 3D step (the whole grid) and the two-step 3D pass (the slab inside
# the x and y CPML, z CPML included; two cell-updates per cell and launch);
# the interior box of the split 3D step has no psi terms: 48 B of field read
# + write and 4 B of eps
KERNEL_ALG_BYTES = {"c1": 32.0, "c2": 29.0, "c3": 58.0, "c3_split": 52.0, "c3_tb2": 58.0}
# 3D stepper variant of the C3 line: the two-pass step (the two-step pass,
# variant 5, is bit-identical and measured equal at the full run length)
C3_VARIANT = 0


DOMINANT = {
    "c1": "tm_tiled interior-tile launch (temporally blocked TM steps), HIP events on its stream",
    "c2": "te_tiled interior-tile launch (temporally blocked TE steps), HIP events on its stream",
    "c3": "h3_kernel + e3_kernel (one two-pass 3D step over the whole grid), HIP events on its stream",
    "c3_split": "y3_interior (one-pass interior box of the split 3D step), HIP events on its stream",
    "c3_tb2": "y3_tb2 psi-free-column and z-edge-column launches (two Yee steps per pass over the slab "
              "inside the x and y CPML), HIP events on its stream",
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def _traffic(workload, cell_steps_per_launch, variant=None):
    """HBM bytes per launch of the dominant kernel from the committed PMC
    measurement (FETCH_SIZE x fetch scale + WRITE_SIZE x write scale, per
    cell-step), scaled to this launch size; None if not measured (or
    measured for another kernel variant)."""
    key = "c3_tb2" if workload == "c3" and variant is not None and variant >= 4 else workload
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)[key]
    except (OSError, KeyError, ValueError):
        return None
    if "variant" in t and variant is not None and t["variant"] != variant:
        return None
    return t["bytes_per_cell_step"] * cell_steps_per_launch


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_c2(steps_override=None):
    from meep_torch_amd import synth
    from meep_torch_amd.core.coeffs import build_te_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTE, FluxMonitorTEY
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_te import init_te_state
    P = synth.C2
    n = P["n"]
    g = Grid(nx=n, ny=n, dx=P["dx"], dtype=torch.float32, device="cuda")
    eps = synth.c2_eps(device="cuda")
    m = build_te_coeffs(g, eps_inf=eps, npml=P["npml"])
    src = LineSource(g, ix=P["src_ix"], j0=P["src_lo"], j1=P["src_hi"],
                     profile=GaussianPulse(P["fcen"], P["fwidth"]))
    src.mask = src.mask[:, :-1].contiguous()
    fr = synth.c2_freqs()
    mons = [FluxMonitorTE(g, fr, ix=i, lo=P["line_lo"], hi=P["line_hi"]) for i in P["lines_x"]]
    mons += [FluxMonitorTEY(g, fr, iy=j, lo=P["line_lo"], hi=P["line_hi"]) for j in P["lines_y"]]
    steps = steps_override or P["steps"]

    def run():
        from meep_torch_amd.core.run import run_te
        return run_te(init_te_state(g), m, steps, source=src, monitors=mons)

    return run, n * n, steps, eps


def build_c3(steps_override=None):
    from meep_torch_amd import synth
    from meep_torch_amd.core.monitors_3d import FluxMonitor3D, PointSource3D, run_3d
    from meep_torch_amd.core.sources import GaussianPulse
    from meep_torch_amd.core.stepper_3d import build_3d_coeffs, init_3d_state
    P = synth.C3
    n = P["n"]
    eps = synth.c3_eps(n).cuda()
    c = build_3d_coeffs(n, n, n, P["dx"], eps=eps, npml=P["npml"], dtype=torch.float32, device="cuda")
    src = PointSource3D(c.dt, "Ez", P["npml"] + 8, n // 2, n // 2, GaussianPulse(P["fcen"], P["fwidth"]))
    fr = [P["fcen"] - 0.1 + 0.2 * i / (P["nfreq"] - 1) for i in range(P["nfreq"])]
    mons = [FluxMonitor3D(c.dt, fr, normal="x", idx=i, ap=P["npml"] + 1, dtype=torch.float32)
            for i in P["planes"]]
    steps = steps_override or P["steps"]

    def run():
        st = init_3d_state(n, n, n, npml=P["npml"], dtype=torch.float32, device="cuda")
        return run_3d(st, c, steps, source=src, monitors=mons)

    return run, n ** 3, steps, eps


def build_c1(steps_override=None):
    from meep_torch_amd import synth
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.run import run_tm
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    from meep_torch_amd.core.stepper_tm import init_tm_state
    g = Grid(nx=256, ny=256, dx=0.05, dtype=torch.float32, device="cuda")
    m = build_tm_coeffs(g, eps_inf=synth.c1_eps().cuda(), npml=16)
    src = LineSource(g, ix=24, j0=128, j1=129, profile=GaussianPulse(0.15, 0.1))
    fr = [0.1 + 0.1 * i / 49 for i in range(50)]
    mon = FluxMonitorTM(g, fr, x_normal=True, idx=232, lo=96, hi=160)
    steps = steps_override or 2000

    def run():
        return run_tm(init_tm_state(g), m, steps, source=src, monitors=[mon])

    return run, 256 * 256, steps, None


def build_c4(n_var=64, steps=4000, ck=200, seed=4, variant=None):
    """C4: n_var TM design variants, N ~ U{512..1536} (splitmix), 128x128
    density (U[0,1)) -> conic filter (r = 4 cells) -> tanh projection (beta 8)
    -> eps 2.1..12.1 painted at the centre of an eps-2.1 cell, 24-cell CPML,
    line source at x = npml + 16, FOM = sum of a 20-frequency output flux line.
    One bench step = one design iteration: batched forward + batched adjoint."""
    from meep_torch_amd import synth
    from meep_torch_amd.adjoint.filters import conic_filter, tanh_projection
    from meep_torch_amd.batch import BatchedTM
    from meep_torch_amd.core.coeffs import build_tm_coeffs
    from meep_torch_amd.core.grid import Grid
    from meep_torch_amd.core.monitors import FluxMonitorTM
    from meep_torch_amd.core.sources import GaussianPulse, LineSource
    u = synth.splitmix_uniform(n_var, seed)
    sizes = [512 + int(x * 1025) for x in u]
    npml, fr = 24, [0.4 + 0.2 * i / 19 for i in range(20)]
    probs, bases, mons = [], [], []
    for v, n in enumerate(sizes):
        g = Grid(nx=n, ny=n, dx=0.05, dtype=torch.float32, device="cuda")
        base = torch.full((n, n), 2.1, dtype=torch.float32, device="cuda")
        m = build_tm_coeffs(g, eps_inf=base, npml=npml)
        src = LineSource(g, ix=npml + 16, j0=npml, j1=n - npml, profile=GaussianPulse(0.5, 0.2))
        mon = FluxMonitorTM(g, fr, x_normal=True, idx=n - npml - 16, lo=npml, hi=n - npml)
        probs.append((m, src, [mon]))
        bases.append(base)
        mons.append(mon)
    rho = [torch.tensor(synth.splitmix_uniform(128 * 128, 100 + v).reshape(128, 128),
                        dtype=torch.float32, device="cuda") for v in range(n_var)]
    bt = BatchedTM(probs, steps, checkpoint_every=ck, variant=variant)
    cells = sum(n * n for n in sizes)
    # the design iteration replays two HIP graphs: the forward (checkpoint
    # copies, batched stepping, spectra) and the reverse sweep (checkpoint
    # restores, tape recomputes, batched adjoint steps)
    bt.set_eps(bases)
    bt.capture()
    bt.capture_adjoint()

    def eps_of(v, r):
        n = sizes[v]
        e = bases[v].clone()
        o = (n - 128) // 2
        e[o:o + 128, o:o + 128] = 2.1 + 10.0 * tanh_projection(conic_filter(r, 0.2, 0.05), 8.0, 0.5)
        return e

    def iteration():
        rs = [r.clone().requires_grad_(True) for r in rho]
        out = bt.apply([eps_of(v, r) for v, r in enumerate(rs)])
        fom = sum(mons[v].flux(o[0], 0.05).sum() for v, o in enumerate(out))
        fom.backward()
        return fom, rs

    iteration.runner = bt
    return iteration, cells, steps, sizes


def cpu_baseline_c2(budget_s=15.0):
    """The oracle (reference algorithm restated, torch CPU) on a bounded sample
    of C2: full 2048x2048 grid with its 4 flux lines, as many Yee steps as fit
    in ~budget_s seconds, on all host cores."""
    from oracle import yee2d
    from meep_torch_amd import synth
    P = synth.C2
    n = P["n"]
    torch.set_num_threads(os.cpu_count() or 1)
    eps = synth.c2_eps(device="cpu")
    c = yee2d.te_coeffs(n, n, P["dx"], eps, P["npml"], dtype=torch.float32)
    mask = yee2d.line_mask(n, n, P["src_ix"], P["src_lo"], P["src_hi"], P["dx"])[:, :-1].contiguous()
    amp = yee2d.gaussian_pulse(P["fcen"], P["fwidth"])
    fr = synth.c2_freqs()
    mons = [{"type": "flux", "kind": "te_x", "idx": i, "lo": P["line_lo"], "hi": P["line_hi"],
             "freqs": fr} for i in P["lines_x"]]
    mons += [{"type": "flux", "kind": "te_y", "idx": j, "lo": P["line_lo"], "hi": P["line_hi"],
              "freqs": fr} for j in P["lines_y"]]
    st = yee2d.te_state(n, n)
    amps = [amp((t + 0.5) * c.dt) for t in range(100000)]
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < budget_s or steps < 2:
        st, _ = yee2d.run2d("te", st, c, 1, amps=amps[steps:steps + 1], mask=mask, monitors=mons)
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": n * n * steps / dt, "unit": "cell-updates/s", "cores": torch.get_num_threads(),
            "kind": "port",
            "sample": f"C2 full 2048x2048 grid + 4x100-freq flux lines, {steps} Yee steps "
                      f"({dt:.1f} s), oracle (torch CPU eager)"}


def cpu_baseline_c3(budget_s=15.0):
    """The oracle (reference algorithm restated, torch CPU) on a bounded sample
    of C3: the full 384^3 grid with its slab, CPML, dipole and both 64-frequency
    flux planes, as many Yee steps as fit in ~budget_s seconds, on all host
    cores."""
    from oracle import yee2d, yee3d
    from meep_torch_amd import synth
    P = synth.C3
    n = P["n"]
    torch.set_num_threads(os.cpu_count() or 1)
    c = yee3d.coeffs3d(n, n, n, P["dx"], synth.c3_eps(n), P["npml"], torch.float32)
    amp = yee2d.gaussian_pulse(P["fcen"], P["fwidth"])
    fr = [P["fcen"] - 0.1 + 0.2 * i / (P["nfreq"] - 1) for i in range(P["nfreq"])]
    mons = [{"type": "flux", "normal": "x", "idx": i, "freqs": fr, "ap": P["npml"] + 1} for i in P["planes"]]
    st = yee3d.state3d(n, n, n, torch.float32)
    src = ("Ez", P["npml"] + 8, n // 2, n // 2)
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < budget_s or steps < 2:
        st, _ = yee3d.run3d(st, c, 1, [amp((steps + 0.5) * c["dt"])], src, mons)
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * steps / dt, "unit": "cell-updates/s", "cores": torch.get_num_threads(),
            "kind": "port",
            "sample": f"C3 full 384^3 grid + 2 x 64-freq flux planes, {steps} Yee steps "
                      f"({dt:.1f} s), oracle (torch CPU eager)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--c5-solves", type=int, default=20000, help="FDTD solves in the surrogate dataset")
    ap.add_argument("--c5-epochs", type=int, default=20)
    ap.add_argument("--c5-iters", type=int, default=500, help="Adam iterations of each design loop")
    ap.add_argument("--c5-yee-steps", type=int, default=2500)
    ap.add_argument("--c3-variant", type=int, default=None,
                    help="3D stepper: 0 two-pass, 1-3 split (interior kernel with 6, 14, 4 tile rows + shell), "
                         f"4-5 two steps per pass (tile rows 10, 8); default {C3_VARIANT}")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 1 GiB float32 tensor three times (reference for PMC byte counters)")
    ap.add_argument("--c4-variants", type=int, default=64)
    ap.add_argument("--c4-tile", type=int, default=None, help="tile variant of the batched launch (mt_tiled_variants)")
    ap.add_argument("--yee-steps", type=int, default=0, help="override Yee steps per run (tuning)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the FDTD path would shard by spatial slabs across GPUs)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        x = torch.ones(256 * 1024 * 1024, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log("[bench] copied 3 x 1 GiB")
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(10):
            y.copy_(x)
        ev[1].record()
        torch.cuda.synchronize()
        us = ev[0].elapsed_time(ev[1]) * 1e3 / 10
        log(f"[bench] plain 1 GiB device copy: {us:.1f} us, {2 * x.numel() * 4 / us / 1e3:.0f} GB/s (read + write)")
        return
    if args.workload == "c4":
        return bench_c4(args)
    if args.workload == "c5":
        return bench_c5(args)
    if args.workload == "c3":
        from meep_torch_amd.core import engine3d
        engine3d.DEFAULT_VARIANT = C3_VARIANT if args.c3_variant is None else args.c3_variant
    builder = {"c1": build_c1, "c2": build_c2, "c3": build_c3}[args.workload]
    run, cells, ysteps, _ = builder(args.yee_steps or None)
    log(f"[bench] {args.workload}: {cells} cells x {ysteps} Yee steps per bench step")
    for i in range(args.warmup):
        run()
        torch.cuda.synchronize()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    from meep_torch_amd import tune
    if tune.tune.last:
        log("[bench] autotuner ms per variant (-1 = two-pass):",
            {k: round(v, 3) for k, v in tune.tune.last.items()}, "chosen:", dict(tune._CACHE))
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    from meep_torch_amd import _engine as E
    L = E.lib()
    L.mt_kernel_timing_read(None, None, None)
    L.mt_kernel_timing_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record()
    for i in range(args.steps):
        run()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    L.mt_kernel_timing_enable(0)
    k_ms, k_n, k_cells = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_longlong()
    E.check(L.mt_kernel_timing_read(ctypes.byref(k_ms), ctypes.byref(k_n), ctypes.byref(k_cells)),
            "mt_kernel_timing_read")
    gpu_ms = ev0.elapsed_time(ev1)
    log(f"[bench] timed {args.steps} runs: wall {wall:.3f} s, events {gpu_ms / 1e3:.3f} s; dominant kernel "
        f"{k_n.value} launches, {k_ms.value:.1f} ms, {k_cells.value} cell-steps")
    updates = cells * ysteps * args.steps
    value = updates / wall
    alg = ALG_BYTES[args.workload]
    k_alg = KERNEL_ALG_BYTES[args.workload]
    if k_n.value:
        per_launch_cells = k_cells.value / k_n.value
        per_launch_s = k_ms.value * 1e-3 / k_n.value
    else:   # no timed launches (should not happen): whole-step events
        per_launch_cells = cells
        per_launch_s = gpu_ms / 1e3 / (ysteps * args.steps)
    variant = None
    dominant = DOMINANT[args.workload]
    if args.workload == "c3":
        from meep_torch_amd.core import engine3d
        variant = engine3d.DEFAULT_VARIANT
        if 1 <= variant <= 3:
            k_alg, dominant = KERNEL_ALG_BYTES["c3_split"], DOMINANT["c3_split"]
        elif variant >= 4:
            k_alg, dominant = KERNEL_ALG_BYTES["c3_tb2"], DOMINANT["c3_tb2"]
    elif tune._CACHE and len(set(tune._CACHE.values())) == 1:
        variant = next(iter(tune._CACHE.values()))     # the tile variant the autotuner chose
    achieved = k_alg * per_launch_cells / per_launch_s / 1e9
    traffic = _traffic(args.workload, per_launch_cells, variant)
    line = {
        "metric": "Yee cell-updates/s (FP32) with % of MI355X HBM roofline",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": {"c1": "C1 2D TM 256x256 16-cell PML 2000 steps 50-freq flux line",
                                "c2": "C2 2D TE 2048x2048 32-cell CPML 10000 steps 4 flux lines x 100 freqs",
                                "c3": "C3 3D 384^3 12-cell CPML 2000 steps 2 flux planes x 64 freqs"
                                }[args.workload],
                   "yee_steps_per_step": ysteps, "cells": cells},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_gbs": traffic / per_launch_s / 1e9 if traffic else None,
                     "kernel": dominant,
                     "launch_avg_us": per_launch_s * 1e6, "cell_steps_per_launch": per_launch_cells,
                     "alg_bytes_per_cell_update": k_alg,
                     "whole_step_alg_bytes_per_cell_update": alg,
                     "whole_step_frac": alg * cells / (gpu_ms / 1e3 / (ysteps * args.steps)) / 1e9 / HBM_PEAK_GBS},
    }
    if args.workload == "c3":
        line["config"]["stepper_variant"] = variant
    if not args.no_cpu_baseline and args.workload in ("c2", "c3"):
        log("[bench] cpu baseline (oracle) ...")
        line["cpu_baseline"] = cpu_baseline_c2() if args.workload == "c2" else cpu_baseline_c3()
    print(json.dumps(line), flush=True)


def bench_c4(args):
    steps = args.yee_steps or 4000
    it, cells, ysteps, sizes = build_c4(args.c4_variants, steps, variant=args.c4_tile)
    log(f"[bench] c4: {len(sizes)} variants, {cells} cells, {ysteps} steps, sizes {min(sizes)}..{max(sizes)}")
    for i in range(args.warmup):
        it()
        torch.cuda.synchronize()
        log(f"[bench] c4 warmup {i + 1} done")
    from meep_torch_amd import _engine as E
    L = E.lib()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        fom, rs = it()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the timed iterations replay graphs (no host launches to bracket): the
    # dominant kernel is timed on one more forward sweep issued eagerly
    L.mt_kernel_timing_read(None, None, None)
    L.mt_kernel_timing_enable(1)
    it.runner.forward(use_graph=False)
    torch.cuda.synchronize()
    L.mt_kernel_timing_enable(0)
    k_ms, k_n, k_cells = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_longlong()
    E.check(L.mt_kernel_timing_read(ctypes.byref(k_ms), ctypes.byref(k_n), ctypes.byref(k_cells)),
            "mt_kernel_timing_read")
    # ... and the fused reverse step on one more adjoint sweep issued eagerly
    L.mt_kernel_timing_enable(2)
    it.runner._gradient_body()
    torch.cuda.synchronize()
    L.mt_kernel_timing_enable(0)
    a_ms, a_n, a_cells = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_longlong()
    E.check(L.mt_kernel_timing_read(ctypes.byref(a_ms), ctypes.byref(a_n), ctypes.byref(a_cells)),
            "mt_kernel_timing_read")
    upd = cells * ysteps * 3 * args.steps   # forward + recompute + adjoint sweeps
    roof = adj_roof = None
    if k_n.value:
        per_s = k_ms.value * 1e-3 / k_n.value
        per_cells = k_cells.value / k_n.value
        alg = 29.0      # FP32 TM forward: Ez, Hx, Hy read + write and eps (interior tiles: no psi)
        gbs = alg * per_cells / per_s / 1e9
        roof = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                "traffic": _traffic("c4", per_cells), "kernel": "tm_batch interior-tile launch (forward and "
                                                               "recompute sweeps), HIP events on its stream",
                "launch_avg_us": per_s * 1e6, "cell_steps_per_launch": per_cells, "alg_bytes_per_cell_update": alg,
                "interior_launches_total_ms": k_ms.value}
    if a_n.value:
        per_s = a_ms.value * 1e-3 / a_n.value
        per_cells = a_cells.value / a_n.value
        # reads eps, Ez/Dz/Hx/Hy cotangents, eps_grad and the Ez tape; writes the
        # four cotangents and eps_grad (CPML psi cotangents not counted)
        alg = 48.0
        gbs = alg * per_cells / per_s / 1e9
        adj_roof = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": gbs / HBM_PEAK_GBS, "traffic": _traffic("c4_adjoint", per_cells),
                    "kernel": "adj_batch_fused (one reverse step of every variant), HIP events on its stream",
                    "launch_avg_us": per_s * 1e6, "cell_steps_per_launch": per_cells,
                    "alg_bytes_per_cell_update": alg}
    print(json.dumps({
        "metric": "batched design iterations/s (forward + adjoint, all variants)",
        "value": args.steps / wall, "unit": "design-iterations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"C4 batched TM fwd+adjoint, {len(sizes)} variants N in "
                               f"[{min(sizes)},{max(sizes)}], {ysteps} steps, ckpt 200",
                   "cells_per_variant_sum": cells},
        "sweep_cell_updates_per_s": upd / wall, "fom": float(fom.detach()), "graphs": "forward + adjoint",
        "roofline": roof, "adjoint_roofline": adj_roof}), flush=True)


def _gemm_roofline(net, batch=64, reps=5):
    """Roofline of the surrogate's GEMM kernel over every GEMM of `reps`
    training passes (forward + backward) at `batch`: FLOPs and algorithmic
    bytes (BF16 operands read once, FP32 result written, and read when
    accumulated into) divided by the summed duration of those launches (HIP
    events around each launch on its stream).  The FNO's products are thin
    (reductions of 64-256 over 16384-wide activations, or 16384-deep over
    256 outputs), so the bound is whichever of the MFMA and HBM fractions is
    higher; both are reported."""
    from meep_torch_amd.surrogate import _nn
    rho = torch.rand(batch, net.n, net.n, device="cuda")
    dout = torch.randn(batch, net.n_out, device="cuda")
    net.forward(rho)
    net.backward(dout)
    torch.cuda.synchronize()
    _nn.PROFILE = []
    for _ in range(reps):
        net.forward(rho)
        net.backward(dout)
    torch.cuda.synchronize()
    rec, _nn.PROFILE = _nn.PROFILE, None
    flops = sum(r[0] for r in rec)
    nbytes = sum(r[1] for r in rec)
    sec = sum(a.elapsed_time(b) for _, _, a, b in rec) * 1e-3
    tf, gbs = flops / sec / 1e12, nbytes / sec / 1e9
    f_m, f_h = tf / 2500.0, gbs / HBM_PEAK_GBS
    common = {"traffic": None, "kernel": f"gemm_kernel (BF16 MFMA, FP32 acc), all GEMMs of a B={batch} training pass",
              "gemm_launches": len(rec) // reps, "gemm_flops_per_pass": flops / reps,
              "gemm_alg_bytes_per_pass": nbytes / reps, "mfma_tflops": tf, "mfma_frac": f_m, "hbm_gbs": gbs,
              "hbm_frac": f_h}
    if f_h > f_m:
        return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": f_h, **common}
    return {"bound": "mfma", "achieved": tf, "peak": 2500.0, "unit": "TFLOP/s", "frac": f_m, **common}


def bench_c5(args):
    from meep_torch_amd.surrogate import design as D
    from meep_torch_amd.surrogate.problem import TransmissionProblem
    from meep_torch_amd.surrogate.train import build_surrogate
    pb = TransmissionProblem(steps=args.c5_yee_steps)
    t0 = time.perf_counter()
    net, data, hist = build_surrogate(pb, n_train=args.c5_solves, n_val=512, epochs=args.c5_epochs, log=log)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    roof = _gemm_roofline(net)
    log(f"[bench] c5 GEMM roofline: {roof['mfma_tflops']:.1f} TFLOP/s ({100 * roof['mfma_frac']:.1f}% of BF16 dense), "
        f"{roof['hbm_gbs']:.0f} GB/s of algorithmic bytes ({100 * roof['hbm_frac']:.1f}% of HBM): {roof['bound']}-bound")
    theta0 = D.start_design(pb)
    n_train = args.c5_solves
    rf = D.fdtd_loop(pb, iters=args.c5_iters, theta0=theta0, log=log)
    rs = D.surrogate_loop(pb, net, iters=args.c5_iters, theta0=theta0,
                          data=(data[0][:n_train], data[1][:n_train]), log=log)
    gap = abs(rs["final_fom"] - rf["final_fom"]) / max(abs(rf["final_fom"]), 1e-30)
    log(f"[bench] c5 plain FDTD: {rf['iters_per_s']:.2f} it/s final FOM {rf['final_fom']:.4f}; surrogate: "
        f"{rs['iters_per_s']:.2f} it/s final FOM {rs['final_fom']:.4f} (gap {100 * gap:.2f}%), "
        f"{rs['failed']}/{rs['verifications']} verifications failed, {rs['fdtd_iters']} FDTD iterations")
    print(json.dumps({
        "metric": "surrogate-assisted design iterations/s (final FOM vs plain FDTD loop)",
        "value": rs["iters_per_s"], "unit": "design-iterations/s", "n_gpus": 1, "steps": args.c5_iters,
        "warmup": 0, "ms_per_step": rs["wall_s"] * 1e3 / args.c5_iters, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": f"C5 surrogate-assisted design: 2D TM {pb.n}x{pb.n}, 128x128 design, "
                               f"{pb.steps} Yee steps/solve, {args.c5_solves} training solves, FNO 4x64 "
                               f"16x8 modes ({net.n_params} params), {args.c5_iters} Adam iterations, "
                               f"verify every 10, fallback above 2%"},
        "plain_fdtd_iters_per_s": rf["iters_per_s"], "speedup_vs_plain": rs["iters_per_s"] / rf["iters_per_s"],
        "final_fom_plain": rf["final_fom"], "final_fom_surrogate": rs["final_fom"], "fom_gap": gap,
        "verifications": rs["verifications"], "verifications_failed": rs["failed"],
        "rollbacks": rs["rollbacks"], "verified_errors": [round(e, 4) for e in rs["errors"]],
        "median_verified_error": (sorted(rs["errors"])[len(rs["errors"]) // 2] if rs["errors"] else None),
        "fdtd_iterations_in_surrogate_loop": rs["fdtd_iters"], "setup_s": setup_s,
        "val_err_median": hist[-1].get("val_err_median"), "val_err_p90": hist[-1].get("val_err_p90"),
        "roofline": roof}), flush=True)


if __name__ == "__main__":
    main()
