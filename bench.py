"""Benchmark of the MI355X Gaussian ray tracer.

Default workload (one MI355X): the 1M-Gaussian training step -- 1,000,000
Gaussians on 20 random surfaces (log-scale N(-4.5, 0.4^2)), one 980x545
pinhole camera (fov 60 deg) on a ring of radius 6; per step: forward render,
L1 loss against a random target, backward, fused Adam and BVH refit.
Prints one JSON line (metric/value/unit/...) with the dominant kernel's HBM
roofline and a bounded CPU-oracle baseline timed on the host cores.

  python bench.py [--steps K] [--warmup W] [--workload train|render] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "train-step ms & rays/s at 1M Gaussians; p50/p99 frame latency vs plain launch"

# algorithmic (compulsory) bytes per launch of each native stage:
# per Gaussian / per ray terms, see DESIGN.md "Measurement"
BYTES_PER_GAUSSIAN = {
    "trace_forward": 76,     # geometry record 64 B + colour 12 B
    "trace_backward": 216,   # geometry 64 + L2W 48 + rgb 12 + live params 32 + grads out 60
    "sh_forward": 216,       # mean 12 + SH 192 in, colour 12 out
    "sh_backward": 433,      # mean, SH rest, colour, colour grad, visibility in; SH grads out, mean grad r/w
    "adam": 1905,            # 59 floats x (p,g,m,v in + p,m,v,g out) + channel-grad zeroing + pruning counter
    "bvh_refit": 300,        # snapshot 44 in / 136 out + refit boxes, scans and node writes
}
BYTES_PER_RAY = {"trace_forward": 72, "trace_backward": 60}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def setup_raytracer(scene, cam, cfg):
    import torch

    from gray_amd import Raytracer

    rt = Raytracer(cfg, scene.count, cam.image_width, cam.image_height)
    g = rt.cuda_module.get_gaussians()
    with torch.no_grad():
        g.mean.copy_(scene.mean)
        g.rotation.copy_(scene.rotation)
        g.scale.copy_(scene.scale)
        g.opacity.copy_(scene.opacity)
        g.sh_coeffs_dc.copy_(scene.sh_dc)
        g.sh_coeffs_rest.copy_(scene.sh_rest)
        g.current_sh_degree.fill_(3)
    rt.cuda_module.rebuild_bvh()
    return rt


def cpu_baseline(scene, cam, cfg, sample_rays: int, seed: int = 0):
    """Forward + backward of the CPU oracle over a seeded sample of pixels
    against all N Gaussians; returns rays/s on the host cores."""
    import numpy as np

    from oracle import oracle as O

    W, H = cam.image_width, cam.image_height
    rng = np.random.default_rng(seed)
    pix = rng.choice(W * H, size=min(sample_rays, W * H), replace=False)
    ro, rd = O.pinhole_rays(W, H, cam.fov_y, np.asarray(cam.origin, np.float32), cam.rotation_c2w_blender())
    ro, rd = ro[pix], rd[pix]
    k = O.make_knobs(exp_power=cfg.exp_power, alpha_threshold=cfg.alpha_threshold, t_threshold=cfg.t_threshold)
    mean = scene.mean.numpy()
    t0 = time.perf_counter()
    geo, l2w = O.snapshot(mean, scene.rotation.numpy(), scene.scale.numpy(), scene.opacity.numpy(),
                          cfg.alpha_threshold, cfg.exp_power)
    rgb = O.sh_forward(mean, scene.sh_dc.numpy(), scene.sh_rest.numpy(), np.asarray(cam.origin, np.float32), 3)
    t1 = time.perf_counter()
    fwd = O.render(geo, rgb, ro, rd, k, cap=256)
    gout = (np.sign(fwd["C"] - 0.5) / (3.0 * W * H)).astype(np.float32)
    O.backward(geo, l2w, rgb, scene.rotation.numpy(), scene.scale.numpy(), scene.opacity.numpy(), ro, rd, fwd,
               gout, k, float(W * H))
    t2 = time.perf_counter()
    rays = len(pix)
    return {
        "value": rays / (t2 - t1),
        "unit": "rays/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": f"{rays} seeded pixels of the {W}x{H} frame, forward+backward against all {scene.count} "
                  f"Gaussians (brute force, C oracle); per-Gaussian prep {t1 - t0:.2f}s excluded; "
                  f"extrapolated frame {W * H / (rays / (t2 - t1)) * 1e3:.0f} ms",
        "seconds": t2 - t1,
    }


def run_train(args):
    import torch

    from gray_amd import RaytracerConfig, scenes
    from gray_amd.native import PhaseTimer, _NoTimer

    n, W, H = args.gaussians, 980, 545
    log(f"[bench] generating scene: {n} Gaussians on 20 surfaces")
    scene = scenes.surface_scene(n, 20, -4.5, 0.4, seed=0)
    cam = scenes.ring_camera(6.0, 0.0, W, H, 60.0)
    cfg = RaytracerConfig(lr_mean_init=1.6e-4, lr_scale_init=5e-3, lr_sh_dc_init=2.5e-3, lr_sh_rest=2.5e-3)
    rt = setup_raytracer(scene, cam, cfg)
    target = torch.rand(3, H, W, generator=torch.Generator().manual_seed(1)).cuda()
    nm = rt.cuda_module

    def step():
        render = rt(cam)
        loss = torch.nn.functional.l1_loss(render, target)
        rt.backward(loss)
        rt.step()

    log(f"[bench] warmup {args.warmup} steps")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    nm.device_status()
    # timed region: exactly K steps
    timer = PhaseTimer()
    nm.timer = timer
    step_events = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        step()
        e1.record()
        step_events.append((e0, e1))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    nm.device_status()
    nm.timer = _NoTimer()
    phases = timer.summary()
    step_ms = sorted(a.elapsed_time(b) for a, b in step_events)
    total_s = t1 - t0
    ms_per_step = total_s * 1e3 / args.steps
    rays = W * H
    stats = nm.get_stats()
    hits = float(stats.num_gaussians_hit.float().sum()) / max(1, (args.warmup + args.steps) * rays)
    acc = float(stats.num_gaussians_accumulated.float().sum()) / max(1, (args.warmup + args.steps) * rays)

    # dominant native kernel and its roofline
    dom = max(phases, key=lambda k: phases[k][2])
    launches, mean_ms, _ = phases[dom]
    alg = BYTES_PER_GAUSSIAN.get(dom, 0) * n + BYTES_PER_RAY.get(dom, 0) * rays
    achieved = alg / (mean_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": rays * args.steps / total_s,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded surface scene and random target; no dataset)",
        "config": {"workload": "train_step_1M_980x545", "gaussians": n, "width": W, "height": H,
                   "rays_per_step": rays, "sh_degree": 3, "step": "forward+L1+backward+Adam+BVH refit",
                   "parallelism": "single GPU"},
        "step_ms_p50": step_ms[len(step_ms) // 2],
        "step_ms_p99": step_ms[min(len(step_ms) - 1, int(math.ceil(0.99 * len(step_ms))) - 1)],
        "phases_ms": {k: round(v[1], 4) for k, v in phases.items()},
        "hits_per_ray": round(hits, 2),
        "composited_per_ray": round(acc, 2),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "algorithmic_bytes": alg, "launch_ms": round(mean_ms, 4), "launches": launches},
    }
    if not args.no_cpu:
        log(f"[bench] CPU oracle baseline on {args.cpu_rays} rays")
        cb = cpu_baseline(scene, cam, cfg, args.cpu_rays)
        cb.pop("seconds", None)
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="train", choices=["train"])
    ap.add_argument("--gaussians", type=int, default=1_000_000)
    ap.add_argument("--cpu-rays", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU design is described in "
              "DESIGN.md and is not run.", file=sys.stderr)
        sys.exit(2)
    import torch

    torch.cuda.set_device(0)
    out = run_train(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
