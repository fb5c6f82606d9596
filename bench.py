#!/usr/bin/env python3
"""Benchmark of the MDWT hot path on one MI355X.

A step is one 1-spp frame of the workload `sync_4ch256`: 4 channels of 256^3
uint8 (seeded fBm + 16 Gaussian blobs each), 8-point transfer functions with
distinct hues, 16^3 macrocells, 1024x1024 pixels, plain syncDDA (-m 2), every
frame accumulated. The volume, majorants and accumulation buffer are resident
in HBM before the timed region; the timed region is `--steps` back-to-back
frames of the frame kernel.

Prints one JSON line. The roofline object prices the frame kernel against
HBM: algorithmic bytes per frame B = 8*bpv*S + 4*k*V + 32*P (S tentative
collisions, each one channel's 8-voxel trilinear footprint; V macrocell
visits, k majorant floats per visit: C for -m 2, 1 for -m 0 passes and -m 3;
P pixels, float4 accumulator read + write), S and V from the kernel's
counting pass. The CPU baseline is the oracle (a scalar C++ port of the
reference's device code) timed on this host's cores on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "ms/frame @1spp 1024²; time-to-RMSE 1e-3, denoised vs plain, 1 GPU"
HBM_PEAK_GBPS = 8000.0
# read-only bandwidth probe: 8 GiB, far beyond the 256 MB last-level cache
READ_STREAM_BYTES = 8 << 30
VOXEL_NAMES = {"uint8": "u8", "uint16": "u16", "float16": "f16", "float32": "f32"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def frame_counts(r, frames: int) -> dict:
    """Average per-frame work counters over `frames` consecutive frames."""
    tot = {"samples": 0, "visits": 0, "rejections": 0, "rays": 0}
    for _ in range(frames):
        c = r.count_frame()
        for k in tot:
            tot[k] += c[k]
        r.render(1)
    return {k: v / frames for k, v in tot.items()}


def algorithmic_bytes(counts: dict, scene, mode: int) -> float:
    bpv = scene.bytes_per_voxel()
    k = scene.channels if mode in (1, 2) else 1
    return 8.0 * bpv * counts["samples"] + 4.0 * k * counts["visits"] + 32.0 * scene.width * scene.height


def read_traffic(workload: str, mode: int):
    p = ROOT / "profiles" / f"traffic_{workload}_m{mode}.json"
    if not p.exists():
        return None, None
    d = json.loads(p.read_text())
    return d.get("bytes_per_launch"), str(p.relative_to(ROOT))


REF_SEED = 0x13579BDF  # seed of the plain reference render; other renders use other seeds


def read_valu(workload: str, mode: int, kernel_ms: float):
    """VALU issue of the frame kernel: the profiled VALU instructions per
    launch over the live launch time, against one wave64 VALU instruction
    per CU per cycle at the profiled clock."""
    p = ROOT / "profiles" / f"valu_{workload}_m{mode}.json"
    if not p.exists():
        return None
    d = json.loads(p.read_text())
    achieved = d["valu_instructions_per_launch"] / (kernel_ms * 1e-3)
    peak = d["compute_units"] * d["clock_ghz"] * 1e9
    return {"achieved": round(achieved / 1e9, 1), "peak": round(peak / 1e9, 1),
            "unit": "G wave64 VALU instructions/s", "frac": round(achieved / peak, 4),
            "instructions_per_launch": d["valu_instructions_per_launch"],
            "lane_utilisation": d.get("lane_utilisation"),
            "source": str(p.relative_to(ROOT))}


def time_to_rmse(r, frames: int, target: float, ms_per_frame: float):
    """Frames (and seconds at the measured frame time) that plain accumulation
    needs for an expected RMSE of `target` against the converged image.

    The per-pixel variance of one frame comes from the accumulated first and
    second moments of a `frames`-frame render (unbiased, averaged over pixels
    and RGB): sigma^2. The mean of k independent frames has expected RMSE
    sqrt(sigma^2 / k) against the true image, so k = ceil(sigma^2 / target^2).
    The `frames`-frame render is the reference image; its own expected RMSE
    is reported beside it. Returns (result, reference mean, variance of the
    reference mean per pixel and channel)."""
    import numpy as np
    from mdwtracker_amd import denoise

    ref, ref_var = denoise.render_mean(r, frames, REF_SEED)
    sigma2 = float(ref_var.mean() * frames)
    k = max(1, int(np.ceil(sigma2 / (target * target))))
    res = {"target_rmse": target, "per_frame_variance": sigma2,
           "plain": {"frames": k, "seconds": round(k * ms_per_frame * 1e-3, 3)},
           "denoised": None,
           "reference_frames": frames,
           "reference_rmse": round(float(np.sqrt(sigma2 / frames)), 6)}
    return res, ref, ref_var


SWEEP_LEVELS = (1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768)


def denoised_time_to_rmse(r, scene, ttr: dict, ref, ref_var, ms_per_frame: float, args) -> dict:
    """Train the denoiser on held-out scenes of the workload family, then run
    one accumulation of the workload (with the auxiliary buffers) read at
    SWEEP_LEVELS frames, denoise every read and estimate the error of the
    plain and the denoised image against the converged image with the
    reference's own variance removed (denoise.noise_corrected_mse). The
    denoised time to `target` is the interpolated frame count at which its
    error reaches it, priced at the frame time with auxiliary buffers for
    the first denoise.AOV_FRAMES frames and the plain frame time after them,
    plus one network pass."""
    import numpy as np
    from mdwtracker_amd import denoise

    target = ttr["target_rmse"]
    t0 = time.perf_counter()
    xs, ts, ks = denoise.make_training_set(n_scenes=args.dn_scenes, views=args.dn_views,
                                           size=args.dn_size, seed=args.dn_seed,
                                           workload=scene.name, log=log)
    t1 = time.perf_counter()
    small = denoise.Denoiser(args.dn_size, args.dn_size, channels=scene.channels, seed=1)
    losses = denoise.train(small, xs, ts, args.dn_steps, batch=8, lr=args.dn_lr,
                           weights=denoise.level_weights(ks), log=log)
    params = small.get_params()
    small.close()
    t2 = time.perf_counter()
    n_pairs = len(xs)
    del xs, ts
    log(f"[bench] denoiser: {n_pairs} pairs in {t1 - t0:.1f} s, {args.dn_steps} steps in {t2 - t1:.1f} s")

    dn = denoise.Denoiser(scene.width, scene.height, channels=scene.channels, seed=1)
    dn.set_params(params)
    # the denoiser reads the standard error of the pixel mean, so its
    # accumulation keeps second moments on every frame: price its plain
    # frames with them (plain accumulation alone runs without)
    r.set_moments(True)
    r.set_aov(False)
    r.render(16)
    r.reset_accumulation()
    ta = time.perf_counter()
    r.render(256)
    ms_m2 = (time.perf_counter() - ta) * 1000.0 / 256
    r.set_aov(True)
    r.set_aov_frames(0)
    r.render(16)
    r.reset_accumulation()
    ta = time.perf_counter()
    r.render(256)
    ms_aov = (time.perf_counter() - ta) * 1000.0 / 256
    n_aov = denoise.AOV_FRAMES
    sweep, infer_ms = [], []
    for k, x in zip(SWEEP_LEVELS, denoise.accumulate_inputs(r, SWEEP_LEVELS, REF_SEED ^ 0x2468ACE)):
        d, ms = dn.infer(x)
        infer_ms.append(ms)
        pm, pse = denoise.noise_corrected_mse(x[..., :3], ref, ref_var)
        dm, dse = denoise.noise_corrected_mse(d, ref, ref_var)
        sweep.append({"frames": k, "plain_mse": pm, "plain_se": pse, "denoised_mse": dm,
                      "denoised_se": dse, "plain_rmse_vs_ref": denoise.rmse(x[..., :3], ref),
                      "denoised_rmse_vs_ref": denoise.rmse(d, ref)})
        log(f"[bench] sweep {k:6d} frames: plain rmse {np.sqrt(max(pm, 0)):.3e} "
            f"(model {np.sqrt(ttr['per_frame_variance'] / k):.3e}), denoised {np.sqrt(max(dm, 0)):.3e}")
    r.set_aov_frames(0)
    r.set_aov(False)
    dn.close()
    ms_net = float(np.median(infer_ms[1:] if len(infer_ms) > 1 else infer_ms))
    kd = denoise.frames_to_target(SWEEP_LEVELS, [s["denoised_mse"] for s in sweep], target * target)
    out = {"frame_ms_with_aov": round(ms_aov, 4), "frame_ms_with_moments": round(ms_m2, 4),
           "aov_frames": n_aov,
           "network_ms": round(ms_net, 3),
           # algorithmic FLOPs of the U-Net on the workload image
           "network_gflop": round(denoise.network_gflop(scene.width, scene.height, scene.channels), 2),
           "network_tflops": round(denoise.network_gflop(scene.width, scene.height, scene.channels)
                                   / ms_net, 1),
           "network_frac_of_bf16_peak": round(denoise.network_gflop(scene.width, scene.height,
                                                                    scene.channels)
                                              / ms_net / denoise.BF16_PEAK_TFLOPS, 4),
           "train": {"pairs": n_pairs, "size": args.dn_size, "steps": args.dn_steps,
                     "scenes": args.dn_scenes, "views": args.dn_views, "levels": list(denoise.TRAIN_LEVELS),
                     "final_loss": float(np.mean(losses[-20:])),
                     "data_seconds": round(t1 - t0, 1), "train_seconds": round(t2 - t1, 1)},
           "sweep": [{k: (round(v, 9) if isinstance(v, float) else v) for k, v in s.items()}
                     for s in sweep]}
    if kd is None:
        out.update({"frames": None, "seconds": None, "speedup": None})
    else:
        # the first n_aov frames carry the auxiliary buffers, the rest are plain
        sec = (min(kd, n_aov) * ms_aov + max(0.0, kd - n_aov) * ms_m2 + ms_net) * 1e-3
        # one-off cost: rendering the training pairs and training; it pays
        # back after break_even_renders renders of RMSE 1e-3
        one_off = (t1 - t0) + (t2 - t1)
        saving = ttr["plain"]["seconds"] - sec
        out.update({"frames": int(math.ceil(kd)), "seconds": round(sec, 3),
                    "speedup": round(ttr["plain"]["seconds"] / sec, 2),
                    "one_off_seconds": round(one_off, 1),
                    "speedup_with_training_once": round(ttr["plain"]["seconds"] / (sec + one_off), 3),
                    "break_even_renders": (int(math.ceil(one_off / saving)) if saving > 0 else None)})
    return out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, mode: int, frames: int) -> dict:
    """The oracle on the host's cores (OpenMP over pixel rows, up to 16
    threads) for `frames` full frames of the workload, and on one thread for
    one full frame."""
    from oracle import pyoracle
    from mdwtracker_amd import scenes

    threads = max(1, min(16, os.cpu_count() or 1))
    o = pyoracle.OracleRenderer(threads=threads)
    scenes.load_scene(o, scene, mode)
    t0 = time.perf_counter()
    o.render(frames, threads=threads)
    dt = time.perf_counter() - t0
    o1 = pyoracle.OracleRenderer(threads=1)
    scenes.load_scene(o1, scene, mode)
    t1 = time.perf_counter()
    o1.render(1, threads=1)
    dt1 = time.perf_counter() - t1
    return {"value": round(dt * 1000.0 / frames, 3), "unit": "ms/frame", "cores": threads,
            "kind": "port",
            "sample": f"{scene.name} -m {mode}, {scene.width}x{scene.height}, {frames} full frames "
                      f"(of {scene.frames}), OpenMP over pixel rows, {dt:.1f} s",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "single_thread_ms_per_frame": round(dt1 * 1000.0, 1),
            "samples_per_s": None}


def small_case(capi, scenes) -> dict:
    """The 2-channel 64^3 u8 CPU case (256^2, -m 0, 64 frames): the oracle on
    all host cores and on one thread, and the GPU, in ms/frame."""
    from oracle import pyoracle

    sc = scenes.make_scene("ndda_cpu_2ch64")
    threads = max(1, min(16, os.cpu_count() or 1))
    out = {"workload": sc.name, "mode": 0, "image": [sc.width, sc.height], "frames": sc.frames}
    for n in (threads, 1):
        o = pyoracle.OracleRenderer(threads=n)
        scenes.load_scene(o, sc, 0)
        t = time.perf_counter()
        o.render(sc.frames, threads=n)
        out[f"cpu_{n}_threads_ms_per_frame"] = round((time.perf_counter() - t) * 1000.0 / sc.frames, 3)
    g = capi.MdwtRenderer(0)
    scenes.load_scene(g, sc, 0)
    g.render(4)
    g.reset_accumulation()
    t = time.perf_counter()
    g.render(sc.frames)
    out["gpu_ms_per_frame"] = round((time.perf_counter() - t) * 1000.0 / sc.frames, 4)
    out["gpu_kernel_ms_per_frame"] = round(g.last_kernel_ms(), 4)
    g.close()
    return out


def batch_views(n: int, seed: int) -> list:
    """n views on a seeded sphere (radius 2.2 around the volume centre):
    width and height uniform in 256..1024, spp uniform in 1..16."""
    import numpy as np
    from mdwtracker_amd import scenes

    rng = np.random.default_rng(seed)
    cams = scenes.orbit_cameras(n, rng, radius=2.2)
    views = []
    for i in range(n):
        views.append({"from": tuple(float(c) for c in cams[i]), "at": (0.5, 0.5, 0.5),
                      "up": (0.0, 1.0, 0.0) if abs(float(cams[i][1]) - 0.5) < 0.95 * 2.2 else (1.0, 0.0, 0.0),
                      "fovy": 45.0, "width": int(rng.integers(256, 1025)),
                      "height": int(rng.integers(256, 1025)), "spp": int(rng.integers(1, 17)),
                      "seed": int(rng.integers(0, 2**31))})
    return views


def batch_bench(args, scene, r) -> dict:
    """Batched views: the autotuner's sweep over macrocell size, workgroup
    and Hilbert tile on the batch, then `--steps` batched frames at the best
    setting in each launch form (a launch per view over four streams; one
    launch for all views; one persistent launch per frame replayed from a
    captured HIP graph): ms per batched frame, aggregate samples/s, per-view
    latency. `value` is the fastest form."""
    import numpy as np

    n_views = int(scene.extra.get("views", 64))
    views = batch_views(n_views, scene.seed)
    r.batch_set_views(views)
    px = sum(v["width"] * v["height"] * v["spp"] for v in views)
    log(f"[bench] batch: {n_views} views, {px / 1e6:.1f} M pixel samples per batched frame")
    t = time.perf_counter()
    table, best = r.autotune((4, 8, 16, 32), (64, 128, 256), (8, 16, 32), frames=2)
    log(f"[bench] autotune ({time.perf_counter() - t:.1f} s): best s={best[0]} threads={best[1]} "
        f"tile={best[2]}")
    forms = {}
    for launch, name in ((0, "per_view_streams"), (1, "one_launch"), (2, "persistent_graph")):
        r.batch_set_launch(launch)
        counts = r.batch_count()
        r.batch_render(args.warmup)
        r.batch_reset()
        t0 = time.perf_counter()
        gpu_ms = r.batch_render(args.steps)
        wall_ms = (time.perf_counter() - t0) * 1000.0 / args.steps
        us = r.batch_view_times()
        kernel_ms = gpu_ms / args.steps
        forms[name] = {"ms_per_frame": round(wall_ms, 3), "kernel_ms": round(kernel_ms, 3),
                       "samples_per_s": round(counts["samples"] / (kernel_ms * 1e-3)),
                       "pixel_samples_per_s": round(px / (kernel_ms * 1e-3)),
                       "view_latency_us": {"p50": round(float(np.percentile(us, 50)), 1),
                                           "p99": round(float(np.percentile(us, 99)), 1),
                                           "max": round(float(us.max()), 1)},
                       "work_per_frame": counts}
        log(f"[bench] batch launch {name}: {wall_ms:.2f} ms per batched frame, "
            f"view latency p50 {forms[name]['view_latency_us']['p50']:.0f} us")
    best_form = min(forms, key=lambda k: forms[k]["ms_per_frame"])
    # the same scene and setting as one 1024^2 view at 1 spp (a launch that
    # fills the GPU by itself): the rate the batch is measured against
    v0 = views[0]
    r.set_frame(1024, 1024, 1, True, (0.0, 0.0, 0.0))
    r.set_camera(v0["from"], v0["at"], v0["up"], v0["fovy"])
    r.render(max(2, args.warmup))
    r.reset_accumulation()
    single_ms = r.render(args.steps * 4) / (args.steps * 4)
    single_rate = 1024 * 1024 / (single_ms * 1e-3)
    log(f"[bench] single 1024^2 view of the same scene: {single_ms:.3f} ms/frame, "
        f"{single_rate / 1e9:.2f} G pixel samples/s")
    f = forms[best_form]
    rows = [[round(float(v), 4) for v in row] for row in table]
    return {"value": f["ms_per_frame"], "kernel_ms": f["kernel_ms"],
            "samples_per_s": f["samples_per_s"], "pixel_samples_per_s": f["pixel_samples_per_s"],
            "view_latency_us": f["view_latency_us"], "work_per_frame": f["work_per_frame"],
            "launch": best_form, "launch_forms": forms,
            "single_view_1024": {"ms_per_frame": round(single_ms, 4),
                                 "pixel_samples_per_s": round(single_rate),
                                 "batch_over_single": round(f["pixel_samples_per_s"] / single_rate, 3)},
            "views": n_views,
            "autotune": {"best": {"macrocell": best[0], "threads": best[1], "tile": best[2]},
                         "columns": ["s", "threads", "tile", "ms/frame", "samples/s", "visits/s"],
                         "table": rows}}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=256)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--mode", type=int, default=2, choices=range(10))
    ap.add_argument("--workload", default="sync_4ch256")
    ap.add_argument("--cpu-frames", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra-modes", action="store_true")
    ap.add_argument("--schedule", type=int, default=1, choices=(0, 1),
                    help="frame-kernel schedule: 1 wave-coherent (default), 0 thread per pixel")
    ap.add_argument("--rmse-frames", type=int, default=4096,
                    help="frames of the reference render for time-to-RMSE (0: skip)")
    ap.add_argument("--no-denoise", action="store_true", help="skip the denoised time-to-RMSE")
    ap.add_argument("--dn-scenes", type=int, default=2)
    ap.add_argument("--dn-views", type=int, default=3)
    ap.add_argument("--dn-size", type=int, default=256)
    ap.add_argument("--dn-steps", type=int, default=6000)
    ap.add_argument("--dn-lr", type=float, default=1e-3)
    ap.add_argument("--dn-seed", type=int, default=77)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    from mdwtracker_amd import capi, scenes

    t = time.perf_counter()
    scene = scenes.make_scene(args.workload)
    log(f"[bench] scene {scene.name}: {scene.channels} x {scene.volumes[0].shape} "
        f"{scene.volumes[0].dtype}, {scene.width}x{scene.height}, mc {scene.mc} "
        f"({time.perf_counter() - t:.1f} s)")
    r = capi.MdwtRenderer(0)
    r.set_schedule(args.schedule)
    scenes.load_scene(r, scene, args.mode)
    copy_bw = r.stream_copy(2 << 30, 5)
    bw = r.stream_read(READ_STREAM_BYTES, 5)
    log(f"[bench] measured bandwidth: read-only stream {bw:.0f} GB/s over "
        f"{READ_STREAM_BYTES >> 30} GiB (roofline denominator); copy {copy_bw:.0f} GB/s "
        f"(2 GiB read + 2 GiB write)")
    if "views" in scene.extra:
        b = batch_bench(args, scene, r)
        line = {"metric": "ms per batched frame (views of 256..1024 px, spp 1..16); "
                          "samples/s; per-view latency, 1 GPU",
                "value": b.pop("value"), "unit": "ms/frame", "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "ms_per_step": None, "higher_is_better": False,
                "scaling": "weak", "vs_baseline": None,
                "dtype": "f32",
                "data": "synthetic (seeded fBm + Gaussian blobs, seeded views)",
                "config": {"workload": scene.name, "channels": scene.channels,
                           "volume": list(scene.volumes[0].shape),
                           "voxel": str(scene.volumes[0].dtype), "mode": args.mode,
                           "parallelism": "single GPU"}}
        line["ms_per_step"] = line["value"]
        line.update(b)
        r.close()
        print(json.dumps(line), flush=True)
        return 0

    r.render(args.warmup)
    counts = frame_counts(r, 4)
    r.reset_accumulation()
    t0 = time.perf_counter()
    gpu_ms = r.render(args.steps)
    t1 = time.perf_counter()
    kernel_ms = r.last_kernel_ms()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    log(f"[bench] -m {args.mode}: {args.steps} frames, wall {ms_per_step:.4f} ms/frame, "
        f"kernel {kernel_ms:.4f} ms/frame, counts/frame {counts}")

    B = algorithmic_bytes(counts, scene, args.mode)
    achieved = B / (kernel_ms * 1e-3) / 1e9
    traffic, traffic_src = read_traffic(scene.name, args.mode)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 5),
                "traffic": traffic, "algorithmic_bytes": round(B),
                "measured_read_gbps": round(bw, 1),
                "measured_read_bytes": READ_STREAM_BYTES,
                "measured_copy_gbps": round(copy_bw, 1),
                "frac_of_measured": round(achieved / bw, 5) if bw > 0 else None,
                "kernel": ((f"render_wave<{args.mode},{VOXEL_NAMES[str(scene.volumes[0].dtype)]}>")
                           if args.schedule == 1 and args.mode <= 5 else
                           f"render_kernel<{args.mode},{VOXEL_NAMES[str(scene.volumes[0].dtype)]}>")}
    if traffic_src:
        roofline["traffic_source"] = traffic_src
    valu = read_valu(scene.name, args.mode, kernel_ms)
    if valu:
        roofline["valu_issue"] = valu

    ttr = None
    if args.rmse_frames > 1:
        ttr, ref, ref_var = time_to_rmse(r, args.rmse_frames, 1e-3, ms_per_step)
        log(f"[bench] plain time-to-RMSE 1e-3: {ttr}")
        if not args.no_denoise:
            ttr["denoised"] = denoised_time_to_rmse(r, scene, ttr, ref, ref_var, ms_per_step, args)
            log(f"[bench] denoised time-to-RMSE 1e-3: "
                f"{ {k: v for k, v in ttr['denoised'].items() if k != 'sweep'} }")
        del ref, ref_var
        r.set_moments(False)

    extra_modes = {}
    if not args.no_extra_modes:
        for m in (0, 1, 2, 3):
            r.set_mode(m)
            if m != args.mode:
                r.render(4)
                r.reset_accumulation()
                r.render(32)
                extra_modes[f"m{m}_ms_per_frame"] = round(r.last_kernel_ms(), 4)
            if m in (0, 2):
                # the thread-per-pixel schedule beside the default wave-coherent one
                r.set_schedule(0)
                r.render(4)
                r.reset_accumulation()
                r.render(32)
                extra_modes[f"m{m}_thread_per_pixel_ms_per_frame"] = round(r.last_kernel_ms(), 4)
                r.set_schedule(args.schedule)
        r.set_mode(args.mode)

    line = {
        "metric": METRIC,
        "value": round(ms_per_step, 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded fBm + Gaussian blobs, seeded transfer functions)",
        "config": {"workload": scene.name, "channels": scene.channels,
                   "volume": list(scene.volumes[0].shape), "voxel": str(scene.volumes[0].dtype),
                   "image": [scene.width, scene.height], "spp": scene.spp, "mode": args.mode,
                   "macrocells": list(scene.mc), "camera": "orbit r=1.8",
                   "parallelism": "single GPU"},
        "kernel_ms_per_frame": round(kernel_ms, 4),
        "samples_per_s": round(counts["samples"] / (kernel_ms * 1e-3)),
        "cell_visits_per_s": round(counts["visits"] / (kernel_ms * 1e-3)),
        "work_per_frame": {k: round(v) for k, v in counts.items()},
        "time_to_rmse": ttr,
        "roofline": roofline,
    }
    line.update(extra_modes)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(scene, args.mode, args.cpu_frames)
        cb = line["cpu_baseline"]
        cb["samples_per_s"] = round(counts["samples"] / (cb["value"] * 1e-3))
        log(f"[bench] cpu baseline {cb}")
        line["small_case"] = small_case(capi, scenes)
        log(f"[bench] small case {line['small_case']}")
    r.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
