"""Benchmark of the hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--rays 1e7]

A step is one full realisation at 512^3 / 1e7 rays: Philox field + inverse
3-D FFT, gradient stencil (with the unit-RMS normalisation), beam sampling,
ray ordering, the fixed-step RK4 trace (dz = 0.5 dx, 1022 steps per ray) and
the shadowgraph / knife-edge schlieren / interferogram histograms (2048^2).
Everything is resident in HBM; the timed region has no host transfers.

value = ray-steps/s of the whole step (rays x steps per ray / step time).
roofline = the trace kernel (the dominant kernel): 512 algorithmic bytes per
ray-step (4 RK4 stages x 8 trilinear corners x 16 B float4) over its average
duration from HIP events on the launch stream, against 8 TB/s.
cpu_baseline = the reference's CPU path (as-shipped solve_ivp RK45 through
RegularGridInterpolator, multiprocessing pool) timed on this host on a sample
of the same rays through the same field.
surrogate = the learned-correction form on the same realisation: z-block path
estimate + MLP + RK4 fallback on flagged rays, its time against the plain
trace (the speedup), its error against the plain trace of the same rays, and
what the network gains over the path estimate alone (error RMS ratios).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_RAY_STEP = 512


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--rays", type=float, default=1e7)
    ap.add_argument("--cpu-bundle", type=int, default=500, help="rays per process per CPU round")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU rounds")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--variant", type=int, default=0, help="tracer kernel: 0 cell-cached, 1 per-stage gathers")
    ap.add_argument("--workload", choices=["single", "large"], default="single",
                    help="single: 512^3 / 1e7 rays (default); large: 1024^3 / 1e8 rays streamed in 1e7-ray waves")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate + fallback leg")
    ap.add_argument("--sur-fields", type=int, default=48, help="training realisations for the surrogate")
    ap.add_argument("--sur-block", type=int, default=16, help="z-block size (cells) of the path estimate")
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

    # the CPU pool forks before this process touches the GPU
    base = None
    if not args.no_cpu_baseline:
        from oracle.cpu_baseline import CPUBaseline
        base = CPUBaseline()

    import numpy as np
    import torch

    from turbulence_tracing_amd.pipeline import Config, Pipeline

    cfg = Config(n=args.n, n_rays=int(args.rays), variant=args.variant)
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
    achieved = BYTES_PER_RAY_STEP * ray_steps / (trace_ms / 1e3) / 1e9

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
        "roofline": {
            "kernel": "trace_rk4_kernel",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": profiled_traffic(cfg),
            "traffic_source": "profiles/trace_rk4_512_1e7_hbm_traffic.md (FETCH_SIZE x 2 + WRITE_SIZE x 1, "
                              "calibrated on a 1 GiB copy)",
            "trace_ms": trace_ms,
            "bytes_per_ray_step": BYTES_PER_RAY_STEP,
            "compute": profiled_valu(cfg),
        },
        "cpu_baseline": None,
    }

    if not args.no_surrogate:
        out["surrogate"] = surrogate_leg(p, cfg, args.sur_fields, args.sur_block, stages)

    if base is not None:
        log(f"cpu baseline: {base.P} processes, bundles of {args.cpu_bundle} rays, ~{args.cpu_budget:.0f} s")
        try:
            out["cpu_baseline"] = cpu_baseline(base, p, cfg, args.cpu_bundle, args.cpu_budget)
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
    cfg = large_volume_config(variant=args.variant)
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
    wave_steps = run.wave * cfg.steps_per_ray
    achieved = BYTES_PER_RAY_STEP * wave_steps / (trace_ms / 1e3) / 1e9
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
        "roofline": {"kernel": "trace_rk4_kernel", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(run.p.cfg),
                     "traffic_source": "profiles/trace_rk4_1024_1e7_hbm_traffic.md (FETCH_SIZE x 2 + WRITE_SIZE x 1, "
                                       "calibrated on a 1 GiB copy)",
                     "trace_ms": trace_ms, "bytes_per_ray_step": BYTES_PER_RAY_STEP},
        "cpu_baseline": None,
    }
    if not args.no_surrogate:
        out["surrogate"] = surrogate_large(run, cfg, ms_per_step)
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
                                                  make_norm, normalise_targets, summarize)
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
    feat, resid, _ = build_dataset(8, 200_000, Config(**{**cfg.__dict__, "pix": 64}), seed=60_000, k=k)
    norm = make_norm(feat, resid)
    sur = Surrogate(norm, seed=1, k=k)
    sur.train(feat, normalise_targets(resid, norm), epochs=6, batch=4096, lr=5e-4)
    del feat, resid
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
           "rays_fallback": nfb, "err_vs_plain_first_wave": {"pos_um": um(summarize(e[0])),
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


def profiled_valu(cfg):
    """The trace kernel's VALU issue, measured by rocprofv3 (SQ_INSTS_VALU
    over the kernel's cycles) for the 512^3 / 1e7-ray workload: the kernel
    is bound by instruction issue, not by HBM (its measured traffic is a
    few GB per launch)."""
    if (cfg.n, cfg.n_rays, cfg.variant) != (512, 10_000_000, 0):
        return None
    path = os.path.join(ROOT, "profiles", "trace_rk4_512_1e7_valu.md")
    out = {}
    try:
        for line in open(path):
            if line.startswith("trace_rk4_valu_wave_instr_per_step:"):
                out["valu_wave_instr_per_step"] = int(line.split(":")[1])
            elif line.startswith("trace_rk4_valu_frac:"):
                out["valu_frac_of_issue_peak"] = float(line.split(":")[1])
    except OSError:
        return None
    out["source"] = "profiles/trace_rk4_512_1e7_valu.md"
    return out or None


def surrogate_leg(p, cfg, n_fields, k, stages):
    """Surrogate + fallback against the plain tracer on the bench workload
    (the realisation of the stage pass): the network is trained on rays the
    tracer traces through n_fields other seeded realisations; the speedup is
    plain trace time / surrogate time (block fields, path estimate, MLP,
    apply, fallback) at the error the surrogate reaches against the plain
    trace of the same rays."""
    import numpy as np
    import torch
    from turbulence_tracing_amd.pipeline import Config
    from turbulence_tracing_amd.surrogate import (Surrogate, build_dataset, errors, make_norm, normalise_targets,
                                                  path_estimate, block_fields, summarize)
    t0 = time.perf_counter()
    tcfg = Config(**{**cfg.__dict__, "pix": 64})
    feat, resid, _ = build_dataset(n_fields, 100_000, tcfg, seed=50_000, k=k)
    norm = make_norm(feat, resid)
    sur = Surrogate(norm, seed=1, k=k)
    hist = sur.train(feat, normalise_targets(resid, norm), epochs=6, batch=4096, lr=5e-4)
    del feat, resid
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log(f"surrogate: trained on {n_fields} x 1e5 rays in {train_s:.1f} s")
    stream = torch.cuda.current_stream()
    thr = (1e-6, 1e-6)

    def timed(fn, reps=3):
        ms = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            r = fn()
            b.record(stream)
            torch.cuda.synchronize()
            ms.append(a.elapsed_time(b))
        return float(np.mean(ms[1:] if reps > 1 else ms)), r

    plain_ms, _ = timed(p.trace)
    tim = {}
    sur_ms, (rf, ph, flag) = timed(lambda: sur.trace(p.field4, p.grid, p.s0, t_pos=thr[0], t_ang=thr[1]))
    sur.trace(p.field4, p.grid, p.s0, t_pos=thr[0], t_ang=thr[1], timings=tim)
    e = errors(rf, ph, p.rf, p.phase)
    # what the network gains over the estimate alone, on this held-out realisation
    base, f2 = path_estimate(p.field4, p.grid, p.s0, k=k)
    Y = sur.predict(f2)
    rfn, phn, _, _, _ = sur.apply(base, Y, 1.0, 1.0)
    eb = errors(torch.stack([base[0], base[1], base[2], base[3]]), base[4], p.rf, p.phase)
    en = errors(rfn, phn, p.rf, p.phase)
    rms = lambda v: float(v.double().pow(2).mean().sqrt())
    um = lambda d: {a: round(b * 1e6, 5) for a, b in d.items()}
    res = {
        "speedup": plain_ms / sur_ms,
        "plain_trace_ms": plain_ms,
        "surrogate_ms": sur_ms,
        "stage_ms": {a: round(b, 4) for a, b in tim.items() if a != "n_fallback"},
        "rays_fallback": int(tim.get("n_fallback", 0)),
        "thresholds": {"pos_m": thr[0], "ang_rad": thr[1]},
        "err_vs_plain": {"pos_um": um(summarize(e[0])), "ang_urad": um(summarize(e[1])),
                         "phase_mrad": {a: round(b * 1e3, 5) for a, b in summarize(e[2]).items()}},
        "step_ms_with_surrogate": sum(v for a, v in stages.items() if a != "trace") + sur_ms,
        "network_gain_rms": {"pos": rms(en[0]) / rms(eb[0]), "ang": rms(en[1]) / rms(eb[1]),
                             "phase": rms(en[2]) / rms(eb[2])},
        "block_cells": k,
        "training": {"fields": n_fields, "rays_per_field": 100_000, "epochs": 6, "seconds": round(train_s, 2),
                     "final_loss": [float(v) for v in hist[-1]]},
    }
    log(f"surrogate: {sur_ms:.2f} ms vs plain {plain_ms:.2f} ms (x{res['speedup']:.2f}); "
        f"max err pos {res['err_vs_plain']['pos_um'].get('max')} um, ang {res['err_vs_plain']['ang_urad'].get('max')} urad")
    return res


def cpu_baseline(base, p, cfg, bundle, budget):
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
    s0 = p.s0_raw[:, :base.P * bundle * 64].cpu().numpy()
    torch.cuda.synchronize()
    log("cpu baseline: gradients in shared memory")
    nray, wall, nfev = base.run_timed(s0, bundle, budget, log)
    rays_per_s = nray / wall
    return {
        "value": rays_per_s * cfg.steps_per_ray,
        "unit": "ray-steps/s",
        "rays_per_s": rays_per_s,
        "cores": base.P,
        "kind": "port",
        "sample": f"{nray} rays of the same beam through the same {cfg.n}^3 field, in rounds of {base.P} "
                  f"processes x {bundle} rays, one as-shipped solve_ivp RK45 (rtol 1e-3) per bundle; "
                  f"ray-steps counted as rays x {cfg.steps_per_ray} (the GPU's steps per ray); wall {wall:.2f} s, "
                  f"nfev {nfev}",
    }


if __name__ == "__main__":
    main()
