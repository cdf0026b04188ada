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

    def step(seed, ev=None):
        p.generate_field(seed)
        p.gradient()
        p.beam(seed + 1)
        p.sort()
        if ev is not None:
            ev[0].record(stream)
        p.trace()
        if ev is not None:
            ev[1].record(stream)
        p.detect()

    for w in range(args.warmup):
        step(1000 + 2 * w)
    torch.cuda.synchronize()
    log(f"warmup done ({args.warmup} steps)")

    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(2 * k, evs[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms/step")
    trace_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))

    ray_steps = p.ray_steps
    value = ray_steps / (ms_per_step / 1e3)
    achieved = BYTES_PER_RAY_STEP * ray_steps / (trace_ms / 1e3) / 1e9

    if args.calib_copy:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("calibration copy: 1 GiB read + 1 GiB written")

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
            "traffic": None,
            "trace_ms": trace_ms,
            "bytes_per_ray_step": BYTES_PER_RAY_STEP,
        },
        "cpu_baseline": None,
    }

    if base is not None:
        log(f"cpu baseline: {base.P} processes, bundles of {args.cpu_bundle} rays, ~{args.cpu_budget:.0f} s")
        try:
            out["cpu_baseline"] = cpu_baseline(base, p, cfg, args.cpu_bundle, args.cpu_budget)
        finally:
            base.close()
    print(json.dumps(out), flush=True)


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
