"""Benchmark of the hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--rays 1e7]

A step is one full realisation at 512^3 / 1e7 rays: Philox field + inverse
3-D FFT, gradient stencil (with the unit-RMS normalisation), beam sampling,
ray ordering, the fixed-step RK4 trace (dz = 0.5 dx, 1022 steps per ray) and
the shadowgraph / knife-edge schlieren / interferogram histograms (2048^2).
Everything is resident in HBM; the timed region has no host transfers.

value = ray-steps/s of the whole step (rays x steps per ray / step time).
roofline = the trace kernel (the dominant kernel, ~93% of the step), which is
bound by VALU instruction issue: its VALU wave-instructions per launch (the
per-wave-step count measured by rocprofv3 SQ_INSTS_VALU on this binary, times
the launch's wave-steps) over its average duration from HIP events on the
launch stream, as the FP32 rate they carry at 2 FLOP per lane (64 lanes per
wave-instruction), against MI355X's 157.3 TFLOP/s FP32 vector peak.  Its HBM
side is reported beside it: the measured bytes per launch (FETCH_SIZE and
WRITE_SIZE, calibrated on a plain copy) over the same duration, against
8 TB/s, and the 512 algorithmic gather bytes per ray-step (4 RK4 stages x 8
trilinear corners x 16 B), which registers and L1 serve.
cpu_baseline = the reference's CPU path (solve_ivp RK45 through
RegularGridInterpolator on a multiprocessing pool) timed on this host on
samples of the same rays through the same field, as shipped (one RK45 per
bundle, rtol 1e-3) and at the stated tolerance (one RK45 per ray at
rtol 5e-7), each with its error against the GPU trace of the same rays.
surrogate = the learned-correction form at its stated scale: trained on 2e7
rays through 64 seeded 512^3 realisations, error heads calibrated on one more,
evaluated on 1e8 rays over 8 held-out realisations; z-block path estimate +
MLP + RK4 fallback on flagged rays, its time against the plain trace of the
same rays (the speedup), its errors against that trace (max, p99.9, RMS), the
fallback fraction, and what the network gains over the path estimate alone
(error RMS ratios).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3     # MI355X FP32 vector peak: 2 FLOP x 64 lanes per VALU wave-instruction
BF16_PEAK_TFLOPS = 2500.0    # MI355X dense BF16 MFMA peak
BYTES_PER_RAY_STEP = 512


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--rays", type=float, default=1e7)
    ap.add_argument("--cpu-bundle", type=int, default=500, help="rays per process per as-shipped CPU round")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU rounds per mode")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--variant", type=int, default=0, help="tracer kernel: 0 cell-cached, 1 per-stage gathers")
    ap.add_argument("--plane-forms", action="store_true",
                    help="build the plane-form field for the tracer's cell cache (default off at 512^3 / 1e7 rays)")
    ap.add_argument("--workload", choices=["single", "large", "ensemble"], default="single",
                    help="single: 512^3 / 1e7 rays (default); large: 1024^3 / 1e8 rays streamed in 1e7-ray waves; "
                         "ensemble: 48 uneven realisations in one captured step")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate + fallback leg")
    ap.add_argument("--sur-fields", type=int, default=64, help="training realisations for the surrogate")
    ap.add_argument("--sur-rays", type=int, default=312_500, help="training rays per realisation")
    ap.add_argument("--sur-eval-fields", type=int, default=8, help="held-out evaluation realisations")
    ap.add_argument("--sur-eval-rays", type=int, default=12_500_000, help="rays per evaluation realisation")
    ap.add_argument("--sur-epochs", type=int, default=4, help="training epochs")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 1 GiB device-to-device copy (reference kernel for HBM counters)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this path is data-parallel over rays and the machine "
              "has one GPU per job; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "large":
        return main_large(args)
    if args.workload == "ensemble":
        return main_ensemble(args)

    # the CPU pool forks before this process touches the GPU
    base = None
    if not args.no_cpu_baseline:
        from oracle.cpu_baseline import CPUBaseline
        base = CPUBaseline()

    import numpy as np
    import torch

    from turbulence_tracing_amd.pipeline import Config, Pipeline

    cfg = Config(n=args.n, n_rays=int(args.rays), variant=args.variant, plane_forms=args.plane_forms)
    p = Pipeline(cfg)
    log(f"pipeline ready: {cfg.n}^3, {cfg.n_rays} rays")
    stream = torch.cuda.current_stream()

    # field + gradient and beam + ordering on two side streams, then trace and
    # detectors; consecutive realisations overlap (Pipeline.steps)
    if args.warmup:
        p.steps([1000 + 2 * w for w in range(args.warmup)])
    torch.cuda.synchronize()
    log(f"warmup done ({args.warmup} steps)")

    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p.steps([2 * k for k in range(args.steps)], trace_events=evs)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms/step")
    trace_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))

    ray_steps = p.ray_steps
    value = ray_steps / (ms_per_step / 1e3)

    if args.calib_copy:
        calibration_copy()

    # stage breakdown (separate, un-timed-region pass with events per stage)
    stages = {}
    names = ["field", "gradient", "beam", "sort", "trace", "detect"]
    fns = [lambda: p.generate_field(7), p.gradient, lambda: p.beam(8), p.sort, p.trace, p.detect]
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(fns) + 1)]
    ev[0].record(stream)
    for i, f in enumerate(fns):
        f()
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    for i, nm in enumerate(names):
        stages[nm] = round(ev[i].elapsed_time(ev[i + 1]), 4)

    out = {
        "metric": "ray-steps/s at 512^3, 1e7 rays; surrogate vs plain speedup at fixed error",
        "value": value,
        "unit": "ray-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Philox field and beam)",
        "config": {
            "workload": f"{cfg.n}^3 FP32 k^-11/3 field (L=10 mm, ne=1e24(1+0.1 d), 1064 nm), "
                        f"{cfg.n_rays:.0e} rays (4 mm disc, 50 urad), RK4 dz=0.5dx "
                        f"({cfg.steps_per_ray} steps/ray), {cfg.pix}^2 shadowgraph+knife-edge+interferogram",
            "grid": cfg.n, "rays": cfg.n_rays, "steps_per_ray": cfg.steps_per_ray, "pix": cfg.pix,
        },
        "rays_per_s": cfg.n_rays / (ms_per_step / 1e3),
        "stage_ms": stages,
        "roofline": trace_roofline(cfg, trace_ms, "profiles/trace_rk4_512_1e7_hbm_traffic.md"),
        "cpu_baseline": None,
    }

    if not args.no_surrogate:
        out["surrogate"] = surrogate_leg(cfg, args.sur_fields, args.sur_rays, args.sur_eval_fields,
                                         args.sur_eval_rays, args.sur_epochs, stages)

    if base is not None:
        try:
            out["cpu_baseline"] = cpu_baseline(base, p, cfg, args.cpu_bundle, args.cpu_budget, out["rays_per_s"])
            out["cpu_baseline"]["field"] = cpu_field_baseline(cfg, base.P, stages["field"])
        finally:
            base.close()
    print(json.dumps(out), flush=True)


def calibration_copy():
    """One device-to-device copy of 1 GiB, the reference dispatch that
    FETCH_SIZE / WRITE_SIZE are scaled by."""
    import torch
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    log("calibration copy: 1 GiB read + 1 GiB written")


def main_large(args):
    """1024^3 FP32 field (L = 20 mm, 15% RMS, ne 5e24), 1e8 rays in an 8 mm
    disc streamed in 1e7-ray waves, 4096^2 histograms.  A step is the whole
    streamed run (all waves) through the resident field."""
    import numpy as np
    import torch
    from turbulence_tracing_amd.waves import WaveRun, large_volume_config
    steps = args.steps if args.steps != 5 else 2
    cfg = large_volume_config(variant=args.variant, plane_forms=args.plane_forms)
    run = WaveRun(cfg, 100_000_000, 10_000_000)
    run.field(7)
    log(f"large volume ready: {cfg.n}^3, {run.total:.0e} rays in {run.n_waves} waves")
    stream = torch.cuda.current_stream()
    for w in range(args.warmup):
        run.run(8 + w)
    torch.cuda.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        run.run(100 + k)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / steps
    log(f"timed {steps} streamed runs: {ms_per_step:.1f} ms each")
    # per-wave trace kernel time (events on the launch stream) for the roofline
    p = run.p
    p.beam(100, n=run.wave, offset=0)
    p.sort()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    p.trace()
    b.record(stream)
    torch.cuda.synchronize()
    trace_ms = a.elapsed_time(b)
    if args.calib_copy:
        calibration_copy()
    out = {
        "metric": "ray-steps/s at 1024^3, 1e8 rays in 1e7-ray waves; surrogate vs plain speedup at fixed error",
        "value": run.ray_steps / (ms_per_step / 1e3),
        "unit": "ray-steps/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Philox field and beam)",
        "config": {"workload": f"{cfg.n}^3 FP32 k^-11/3 field (L=20 mm, ne=5e24(1+0.15 d), 1064 nm), 1e8 rays "
                               f"(8 mm disc, 50 urad) in 10 waves of 1e7, RK4 dz=0.5dx ({cfg.steps_per_ray} "
                               f"steps/ray), {cfg.pix}^2 histograms",
                   "grid": cfg.n, "rays": run.total, "wave": run.wave, "steps_per_ray": cfg.steps_per_ray},
        "rays_per_s": run.total / (ms_per_step / 1e3),
        "roofline": trace_roofline(run.p.cfg, trace_ms, "profiles/trace_rk4_1024_1e7_hbm_traffic.md"),
        "cpu_baseline": None,
    }
    if not args.no_surrogate:
        out["surrogate"] = surrogate_large(run, cfg, ms_per_step)
    print(json.dumps(out), flush=True)


def main_ensemble(args):
    """48 realisations of uneven size (grids from {128, 192, 256, 384}^3,
    anisotropic power-law spectra, 2-20% RMS, mean ne 1e23-1e25, 2e5-2e6 rays
    each), every stage of a step (fields, gradients, beams, ray orders, one
    persistent batched trace, detectors) captured as one HIP graph with the
    autotuned launch shape.  A step is one graph replay."""
    import torch
    from turbulence_tracing_amd.ensemble import Ensemble, draw_members
    members = draw_members(48, seed=0)
    ens = Ensemble(members)
    ens.generate()
    tuning = ens.autotune()
    log(f"ensemble: {len(members)} members, {ens.n_rays:.3e} rays, {ens.ray_steps:.3e} ray-steps; tuning {tuning}")
    ens.capture()
    for _ in range(args.warmup):
        ens.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ens.replay()
    torch.cuda.synchronize()
    ms_per_step = 1e3 * (time.perf_counter() - t0) / args.steps
    log(f"timed {args.steps} replays: {ms_per_step:.1f} ms each")
    stream = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    ens.trace()
    b.record(stream)
    torch.cuda.synchronize()
    trace_ms = a.elapsed_time(b)
    out = {
        "metric": "ray-steps/s of a 48-realisation ensemble step (one captured graph, one batched trace)",
        "value": ens.ray_steps / (ms_per_step / 1e3),
        "unit": "ray-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Philox fields and beams)",
        "config": {"workload": "48 realisations, grids {128,192,256,384}^3, P ~ (kx^2+ky^2+(kz/a)^2)^(beta/2), "
                               "beta in [-11/3,-3], a in [0.5,2], RMS 2-20%, ne 1e23-1e25, 2e5-2e6 rays each "
                               "(2 mm disc, 50 urad), RK4 dz=0.5dx, 512^2 histograms per member",
                   "members": len(members), "rays": ens.n_rays, "ray_steps": ens.ray_steps, "tuning": tuning},
        "rays_per_s": ens.n_rays / (ms_per_step / 1e3),
        "batched_trace": {"ms": trace_ms, "ray_steps_per_s": ens.ray_steps / (trace_ms / 1e3)},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def surrogate_large(run, cfg, plain_ms):
    """Surrogate + fallback over the streamed run: block size chosen by the
    error budget on 1e6 probe rays, network trained on 8 other 1024^3
    realisations, one untimed run, then the mean of two timed runs over the
    beams of the timed plain runs; errors measured on the first wave against
    the plain trace."""
    import torch
    from turbulence_tracing_amd.pipeline import Config
    from turbulence_tracing_amd.surrogate import (Surrogate, block_fields, build_dataset, choose_block, errors,
                                                  make_norm, normalise_targets, path_estimate, residual, summarize)
    p = run.p
    p.beam(100, n=run.wave, offset=0)
    p.sort()
    p.trace()
    rf0, ph0 = p.rf.clone(), p.phase.clone()
    order0 = p.order.long().clone()
    probe = 1_000_000
    k, kstats = choose_block(p.field4, p.grid, p.s0[:, :probe], rf0[:, :probe], ph0[:probe])
    log(f"surrogate: block size {k} ({kstats})")
    t0 = time.perf_counter()
    feat, resid, pt = build_dataset(8, 200_000, Config(**{**cfg.__dict__, "pix": 64}), seed=60_000, k=k)
    norm = make_norm(feat, resid)
    sur = Surrogate(norm, seed=1, k=k)
    sur.train(feat, normalise_targets(resid, norm), epochs=6, batch=4096, lr=5e-4)
    del feat, resid
    # error heads calibrated on a held-out realisation
    pt.run(seed=61_001)
    bc, fc = path_estimate(pt.field4, pt.grid, pt.s0, k=k)
    cal = sur.calibrate(fc, residual(pt.rf, pt.phase, bc), 1e-6, 1e-6, margin=0.8)
    del bc, fc, pt
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    blocks = block_fields(p.field4, p.grid, k) if k > 1 else None  # k = 1 reads the node field
    first = {}

    def grab(w, n):
        if w == 0:
            first["rf"], first["ph"], first["order"] = p.rf.clone(), p.phase.clone(), p.order.long().clone()

    run.run(8, surrogate=sur, blocks=blocks)  # warm-up: buffers, side streams, packed weights
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nfb = 0
    for r in range(2):  # the same beams as the timed plain runs
        nfb += run.run(100 + r, surrogate=sur, blocks=blocks, on_wave=grab if r == 0 else None)
    torch.cuda.synchronize()
    sur_ms = 1e3 * (time.perf_counter() - t0) / 2
    nfb /= 2
    # both runs hold wave 0 in their own entry-column order: compare ray by ray
    # in the beam's order
    def unsort(rf, ph, o):
        r, q = torch.empty_like(rf), torch.empty_like(ph)
        r[:, o] = rf
        q[o] = ph
        return r, q
    e = errors(*unsort(first["rf"], first["ph"], first["order"]), *unsort(rf0, ph0, order0))
    um = lambda d: {a: round(b * 1e6, 5) for a, b in d.items()}
    res = {"speedup": plain_ms / sur_ms, "plain_ms": plain_ms, "surrogate_ms": sur_ms, "block_cells": k,
           "block_probe": {str(a): [round(x * 1e6, 4) for x in v] for a, v in kstats.items()},
           "rays_fallback": nfb, "calibration": cal, "err_vs_plain_first_wave": {"pos_um": um(summarize(e[0])),
                                                             "ang_urad": um(summarize(e[1]))},
           "training": {"fields": 8, "rays_per_field": 200_000, "seconds": round(train_s, 2)}}
    log(f"surrogate: {sur_ms:.0f} ms vs plain {plain_ms:.0f} ms (x{res['speedup']:.2f})")
    return res


TRAFFIC_PROFILES = {(512, 0): "trace_rk4_512_1e7_hbm_traffic.md", (1024, 0): "trace_rk4_1024_1e7_hbm_traffic.md"}


def profiled_traffic(cfg):
    """HBM bytes per 1e7-ray trace launch measured by rocprofv3 (FETCH_SIZE
    and WRITE_SIZE, calibrated on a plain copy) for the 512^3 and the 1024^3
    workloads, read from their committed summaries; None for others."""
    name = TRAFFIC_PROFILES.get((cfg.n, cfg.variant))
    if name is None or cfg.n_rays != 10_000_000:
        return None
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("trace_rk4_traffic_bytes:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


VALU_PROFILES = {512: "trace_rk4_512_1e7_valu.md", 1024: "trace_rk4_1024_1e7_valu.md"}


def profiled_valu_per_wave_step(cfg):
    """VALU wave-instructions per wave-step (one RK4 step of 64 rays) of the
    trace kernel at this grid size, measured by rocprofv3 (SQ_INSTS_VALU /
    (SQ_WAVES x steps per ray)) on this binary and read from its committed
    summary; None when there is none for this workload."""
    name = VALU_PROFILES.get(cfg.n)
    if name is None or cfg.variant != 0:
        return None, None
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("trace_rk4_valu_wave_instr_per_step:"):
                return float(line.split(":")[1]), "profiles/" + name
    except OSError:
        pass
    return None, None


def trace_roofline(cfg, trace_ms, traffic_source):
    """The trace kernel against the roofline that binds it (VALU issue),
    with its HBM traffic and algorithmic gather bytes beside it."""
    import math
    waves = math.ceil(cfg.n_rays / 64)
    wave_steps = waves * cfg.steps_per_ray
    per_step, src = profiled_valu_per_wave_step(cfg)
    t = trace_ms / 1e3
    traffic = profiled_traffic(cfg)
    out = {"kernel": "trace_rk4_kernel", "bound": "valu", "unit": "TFLOP/s", "peak": FP32_PEAK_TFLOPS,
           "achieved": None, "frac": None, "traffic": traffic, "trace_ms": trace_ms}
    if per_step is not None:
        instr = per_step * wave_steps
        achieved = instr * 64 * 2 / t / 1e12
        out.update(achieved=achieved, frac=achieved / FP32_PEAK_TFLOPS,
                   valu_wave_instr_per_wave_step=per_step, valu_wave_instr_per_launch=instr,
                   valu_source=src)
    out["hbm"] = {"achieved_gbs": None if traffic is None else traffic / t / 1e9, "peak_gbs": HBM_PEAK_GBS,
                  "frac": None if traffic is None else traffic / t / 1e9 / HBM_PEAK_GBS,
                  "traffic_bytes": traffic,
                  "traffic_source": traffic_source + " (FETCH_SIZE x 2 + WRITE_SIZE x 1, calibrated on a 1 GiB copy)"}
    out["algorithmic_gather"] = {"bytes_per_ray_step": BYTES_PER_RAY_STEP,
                                 "gbs": BYTES_PER_RAY_STEP * cfg.n_rays * cfg.steps_per_ray / t / 1e9,
                                 "note": "trilinear-gather bytes, served from registers (the cell cache) and L1; "
                                         "not an HBM rate"}
    return out


MLP_FLOP_PER_RAY = 2 * (48 * 256 + 256 * 256 + 256 * 256 + 256 * 8)  # the fused network's forward pass


def surrogate_leg(cfg, n_fields, rays_per_field, eval_fields, eval_rays, epochs, stages):
    """The learned-correction form at its stated scale: the network is
    trained on rays the tracer traces through ``n_fields`` seeded 512^3
    realisations (``rays_per_field`` each), its error heads are calibrated on
    rays of one more held-out realisation, and it is evaluated against the
    plain tracer on ``eval_fields`` further held-out realisations of
    ``eval_rays`` rays each: per field the plain trace and the surrogate
    trace (block fields, path estimate, network, correction, fallback on
    flagged rays) of the same rays are timed, and the surrogate's errors are
    taken against the plain trace.  The block size is the largest whose
    bare estimate meets the error budget on 1e6 probe rays."""
    import numpy as np
    import torch
    from turbulence_tracing_amd.pipeline import Config, Pipeline
    from turbulence_tracing_amd.surrogate import (Surrogate, build_dataset, choose_block, errors, make_norm,
                                                  normalise_targets, path_estimate, residual)
    bar = (1e-6, 1e-6)
    stream = torch.cuda.current_stream()
    t0 = time.perf_counter()
    # block size from a probe realisation (held out)
    pe = Pipeline(Config(**{**cfg.__dict__, "n_rays": eval_rays, "pix": 64}))
    pe.run(seed=90_001)
    probe = 1_000_000
    k, kstats = choose_block(pe.field4, pe.grid, pe.s0[:, :probe], pe.rf[:, :probe], pe.phase[:probe])
    log(f"surrogate: block size {k} ({ {a: [round(x * 1e6, 3) for x in v] for a, v in kstats.items()} } urad/um)")
    # training set
    tcfg = Config(**{**cfg.__dict__, "pix": 64})
    feat, resid, pt = build_dataset(n_fields, rays_per_field, tcfg, seed=50_000, k=k, log=log)
    norm = make_norm(feat, resid)
    sur = Surrogate(norm, seed=1, k=k, tau=0.9)
    hist = sur.train(feat, normalise_targets(resid, norm), epochs=epochs, batch=8192, lr=5e-4, log=log)
    del feat, resid
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log(f"surrogate: trained on {n_fields} x {rays_per_field} rays in {train_s:.1f} s")
    # calibration on a held-out realisation (1e6 rays)
    pe.run(seed=90_003)
    bc, fc = path_estimate(pe.field4, pe.grid, pe.s0[:, :probe], k=k)
    cal = sur.calibrate(fc, residual(pe.rf[:, :probe], pe.phase[:probe], bc), bar[0], bar[1], margin=0.8)
    del bc, fc
    log(f"surrogate: calibration {cal}")

    def timed(fn):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        r = fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b), r

    def kth(e, q):
        e = e[torch.isfinite(e)].double()
        return float(torch.kthvalue(e.cpu(), max(1, int(np.ceil(q * e.numel()))))[0]) if e.numel() else 0.0

    plain_ms = sur_ms = 0.0
    n_rays = n_fb = n_over = 0
    emax = [0.0, 0.0, 0.0]
    p999 = [0.0, 0.0, 0.0]
    sq = [0.0, 0.0, 0.0]
    gains = None
    tim = {}
    sur.trace(pe.field4, pe.grid, pe.s0)  # warm-up: buffers and packed weights
    for f in range(eval_fields):
        pe.run(seed=91_000 + 2 * f)
        torch.cuda.synchronize()
        pms, _ = timed(pe.trace)
        sms, (rf, ph, flag) = timed(lambda: sur.trace(pe.field4, pe.grid, pe.s0, t_pos=bar[0], t_ang=bar[1]))
        plain_ms += pms
        sur_ms += sms
        e = errors(rf, ph, pe.rf, pe.phase)
        n_rays += pe.s0.shape[1]
        n_fb += int(flag.sum())
        n_over += int(((e[0] > bar[0]) | (e[1] > bar[1])).sum())
        for i in range(3):
            fin = e[i][torch.isfinite(e[i])]
            emax[i] = max(emax[i], float(fin.max()))
            p999[i] = max(p999[i], kth(e[i], 0.999))
            sq[i] += float(fin.double().pow(2).sum())
        if f == 0:
            sur.trace(pe.field4, pe.grid, pe.s0, t_pos=bar[0], t_ang=bar[1], timings=tim)
            # what the network gains over the path estimate alone, on this held-out field
            base, f2 = path_estimate(pe.field4, pe.grid, pe.s0, k=k)
            Y = sur.predict(f2)
            rfn, phn, _, _, _ = sur.apply(base, Y, 1.0, 1.0)
            eb = errors(torch.stack([base[0], base[1], base[2], base[3]]), base[4], pe.rf, pe.phase)
            en = errors(rfn, phn, pe.rf, pe.phase)
            rms = lambda v: float(v.double().pow(2).mean().sqrt())
            gains = {"pos": rms(en[0]) / rms(eb[0]), "ang": rms(en[1]) / rms(eb[1]), "phase": rms(en[2]) / rms(eb[2])}
            del base, f2, Y, rfn, phn, eb, en
        log(f"surrogate field {f}: plain {pms:.1f} ms, surrogate {sms:.1f} ms, fallback {int(flag.sum())}")
    um = lambda v: round(v * 1e6, 5)
    mlp_ms = tim.get("mlp")
    res = {
        "speedup": plain_ms / sur_ms,
        "plain_trace_ms": plain_ms, "surrogate_ms": sur_ms,
        "rays": n_rays, "fields": eval_fields,
        "rays_fallback": n_fb, "fallback_frac": n_fb / n_rays,
        "rays_over_bar": n_over,
        "thresholds": {"pos_m": bar[0], "ang_rad": bar[1]},
        "err_vs_plain": {"pos_um": {"max": um(emax[0]), "p999": um(p999[0]), "rms": um(np.sqrt(sq[0] / n_rays))},
                         "ang_urad": {"max": um(emax[1]), "p999": um(p999[1]), "rms": um(np.sqrt(sq[1] / n_rays))},
                         "phase_mrad": {"max": round(emax[2] * 1e3, 5), "p999": round(p999[2] * 1e3, 5),
                                        "rms": round(np.sqrt(sq[2] / n_rays) * 1e3, 5)}},
        "stage_ms_first_field": {a: round(b, 4) for a, b in tim.items() if a not in ("n_fallback", "chunks")},
        "network_gain_rms": gains,
        "block_cells": k,
        "block_probe": {str(a): [round(x * 1e6, 4) for x in v] for a, v in kstats.items()},
        "calibration": cal,
        "mlp_roofline": None if not mlp_ms else {
            "kernel": "mlp_infer_kernel", "bound": "mfma", "unit": "TFLOP/s", "peak": BF16_PEAK_TFLOPS,
            "achieved": MLP_FLOP_PER_RAY * eval_rays / (mlp_ms / 1e3) / 1e12,
            "frac": MLP_FLOP_PER_RAY * eval_rays / (mlp_ms / 1e3) / 1e12 / BF16_PEAK_TFLOPS, "ms": mlp_ms},
        "training": {"fields": n_fields, "rays_per_field": rays_per_field, "rays": n_fields * rays_per_field,
                     "epochs": epochs, "seconds": round(train_s, 2), "final_loss": [float(v) for v in hist[-1]]},
        "step_ms_with_surrogate": sum(v for a, v in stages.items() if a != "trace") + sur_ms / eval_fields,
    }
    log(f"surrogate: {sur_ms:.1f} ms vs plain {plain_ms:.1f} ms over {n_rays:.2e} rays (x{res['speedup']:.2f}); "
        f"fallback {n_fb}; max err pos {res['err_vs_plain']['pos_um']['max']} um, "
        f"ang {res['err_vs_plain']['ang_urad']['max']} urad")
    del pe, pt
    torch.cuda.empty_cache()
    return res


def cpu_baseline(base, p, cfg, bundle, budget, gpu_rays_per_s):
    """Both CPU modes on samples of the realisation the pipeline holds; each
    mode's error is taken against the GPU trace of the same rays (itself
    within 1e-6 cells / 1e-5 of the deflection of tight solve_ivp)."""
    import numpy as np
    import torch
    c = 299792458.0
    # the reference's gradient arrays (PT:235-237) of the same field, float64
    dz = p.grid.dz
    f4 = p.field4.cpu()
    x = np.linspace(-cfg.L / 2, cfg.L / 2, cfg.n)
    grads = [f4[..., k].permute(2, 1, 0).double().numpy() * (c ** 2 / dz) for k in range(3)]
    del f4
    base.load_gradients(grads[0], grads[1], grads[2], x, x, x)
    del grads
    n_take = base.P * bundle * 64
    s0 = p.s0_raw[:, :n_take].cpu().numpy()
    # the GPU's exit rays of those rays (p.rf is in the sorted order p.order)
    rf_raw = torch.empty_like(p.rf)
    rf_raw[:, p.order.long()] = p.rf
    rf_gpu = rf_raw[:, :n_take].cpu().numpy()
    torch.cuda.synchronize()
    log("cpu baseline: gradients in shared memory")
    launch = np.arctan(s0[[3, 4]] / s0[5])

    def err(rf, n):
        g = rf_gpu[:, :n]
        defl = float(np.sqrt(np.mean((g[[1, 3]] - launch[:, :n]) ** 2)))
        return {"pos_cells_max": float(np.abs(rf[[0, 2]] - g[[0, 2]]).max() / p.grid.dx),
                "ang_max_of_rms_deflection": float(np.abs(rf[[1, 3]] - g[[1, 3]]).max() / defl)}

    modes = {}
    for mode, b in (("as_shipped", bundle), ("stated_tolerance", 1)):
        log(f"cpu baseline ({mode}): {base.P} processes, {b} rays per process per round, ~{budget:.0f} s")
        rf, nray, wall, nfev = base.run_timed(s0, b, budget, mode, log)
        rps = nray / wall
        modes[mode] = {"rays_per_s": rps, "rays": nray, "wall_s": round(wall, 2),
                       "gpu_equivalent_ray_steps_per_s": rps * cfg.steps_per_ray,
                       "own_ray_steps_per_s": nfev / 6 / wall, "own_steps_per_ray": nfev / 6 / max(nray, 1),
                       "err_vs_gpu": err(rf, nray), "gpu_over_cpu_rays_per_s": gpu_rays_per_s / rps}
    head = modes["stated_tolerance"]
    return {
        "value": head["gpu_equivalent_ray_steps_per_s"],
        "unit": "ray-steps/s",
        "rays_per_s": head["rays_per_s"],
        "cores": base.P,
        "kind": "port",
        "sample": f"rays of the same beam through the same {cfg.n}^3 field on {base.P} processes "
                  f"(os.cpu_count() = {base.cpu_count}, affinity {base.affinity}; at most {16} used, the host-core "
                  f"share of one GPU): value = the stated-tolerance mode ({head['rays']} rays, one solve_ivp RK45 "
                  f"per ray at rtol 5e-7, within 1e-3 of the RMS deflection of a tight solve), in ray-steps of the "
                  f"GPU's problem (rays x {cfg.steps_per_ray}); the as-shipped mode ({modes['as_shipped']['rays']} "
                  f"rays, one RK45 per {bundle}-ray bundle at rtol 1e-3) is in modes",
        "modes": modes,
    }


def cpu_field_baseline(cfg, workers, gpu_field_ms):
    """The field on the host cores, as the reference makes it: the oracle's
    restatement of gaussian3D_FFT (TG:488-538: MT19937 normal draws, the
    flips, k_func, numpy's single-thread ifftn) at N = 121 (243^3 = 3^15, an
    FFT-friendly size of the reference's odd 2N + 1 family), and the
    transform alone at 256^3 with numpy (one thread) and scipy.fft (workers =
    the CPU pool's size), each the better of two runs and extrapolated to
    the benchmark's grid by n^3 log n (stated as extrapolated: a reference
    field at 512^3 takes seconds to tens of seconds of one core)."""
    import numpy as np
    import scipy.fft as sfft
    from oracle import field as F

    def k_func(K):
        with np.errstate(divide="ignore"):
            P = K ** (-11.0 / 3.0)
        P[(K < 1.0 / 243) | (K > 0.5)] = 0.0
        return P

    def scale(m):
        n = cfg.n
        return (n ** 3 * np.log(n ** 3)) / (m ** 3 * np.log(m ** 3))

    def best_of_two(fn):  # the first call also pays for faulting in fresh memory
        ts = []
        for _ in range(2):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        return min(ts)

    ref_s = best_of_two(lambda: F.gaussian3D_FFT(121, k_func, rng=np.random.RandomState(0)))
    a = np.random.RandomState(1).randn(256, 256, 256) + 0j
    np_s = best_of_two(lambda: np.fft.ifftn(a))
    sp_s = best_of_two(lambda: sfft.ifftn(a, workers=workers))
    del a
    ref_ext = ref_s * scale(243)
    log(f"cpu field: gaussian3D_FFT 243^3 {ref_s:.2f} s, ifftn 256^3 numpy {np_s:.2f} s, scipy x{workers} {sp_s:.2f} s")
    return {
        "reference_243_s": round(ref_s, 3), "numpy_ifftn_256_s": round(np_s, 3),
        f"scipy_ifftn_256_workers{workers}_s": round(sp_s, 3),
        f"extrapolated_{cfg.n}": {"reference_s": round(ref_ext, 2), "numpy_ifftn_s": round(np_s * scale(256), 2),
                                  "scipy_ifftn_s": round(sp_s * scale(256), 2)},
        "gpu_field_ms": gpu_field_ms,
        "gpu_over_reference": ref_ext / (gpu_field_ms / 1e3),
        "note": "reference = the oracle's restatement of gaussian3D_FFT (MT19937 draws, flips, k_func, "
                "single-thread numpy ifftn); extrapolated by n^3 log n",
    }


if __name__ == "__main__":
    main()
