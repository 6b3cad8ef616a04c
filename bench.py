"""Benchmark: training it/s at 1.5M Gaussians and 1600x1040 on one MI355X.

Workload (one "step" = one msplat_trainer_step of the C API, the reference's
Trainer::step): the synthetic C1 scene -- 1.5M trainable Gaussians drawn from a
64-component isotropic mixture, 64 jittered cameras on a radius-3 sphere, 1600x1040
pinhole views with fx = 1300, ground truth rendered on the GPU from a hidden 20k-Gaussian
set and quantised to uint8 -- trained with 0.8 L1 + 0.2 (1 - SSIM), SGLD noise and the
fused Adam groups, SH degree 3 active from step 3 (sh_degree_interval = 1), fp32
throughout, cap 3M. The timed window lies before step 500, so N stays 1.5M (no MCMC
relocation), as the reference's headline it/s is quoted.

Prints one JSON line (the bench contract). `roofline` covers the dominant kernel: its
per-launch time comes from HIP events on the trainer's stream, its work from the
algorithmic bytes (HBM-bound kernels) or FLOPs (raster kernels: 16 forward, 45 backward per
valid pixel-entry pair, i.e. per entry a pixel takes, counted by the forward kernel); `traffic` is read from the committed rocprofv3 PMC summary
when one exists for that kernel and workload. `cpu_baseline` times the CPU oracle on the
same state and view.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
METRIC = "train it/s @1.5M splats 1600x1040"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--n", type=int, default=1_500_000)
    p.add_argument("--width", type=int, default=1600)
    p.add_argument("--height", type=int, default=1040)
    p.add_argument("--fx", type=float, default=1300.0)
    p.add_argument("--cams", type=int, default=64)
    p.add_argument("--hidden", type=int, default=20_000)
    p.add_argument("--cap", type=int, default=3_000_000)
    p.add_argument("--cpu-steps", type=int, default=3)
    p.add_argument("--stage-steps", type=int, default=10)
    p.add_argument("--calib-copy-mb", type=int, default=0,
                   help="launch the plain copy kernel once on this many MiB (PMC traffic calibration)")
    p.add_argument("--suite", default="c1", choices=["c1", "c0", "c1_full", "c2", "c3_bf16", "c4", "n1", "bf16_error"],
                   help="c1: the headline it/s line (default); the others print their own JSON line")
    p.add_argument("--sh-bf16", action="store_true", help="bfloat16 SH storage for the c1 suite")
    p.add_argument("--bucket-waves", type=int, default=0, choices=[0, 1, 2, 4],
                   help="waves per workgroup of the tile-list bucket kernel (0: as the LDS allows)")
    p.add_argument("--raster-ppl", type=int, default=0, choices=[0, 1, 4],
                   help="pixels per lane of the persistent raster kernels (0: library default)")
    p.add_argument("--persist-grid", type=int, default=0,
                   help="resident workgroups of the persistent raster kernels (0: library default)")
    return p.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pair_counts(final_idx, bins, W, H):
    """Evaluated (pixel, entry) pairs of the raster passes: entries from the tile's list start
    through the pixel's last contributor (final_idx)."""
    tbx = (W + 15) // 16
    ys, xs = np.divmod(np.arange(W * H), W)
    tile = (ys // 16) * tbx + xs // 16
    start = bins[tile, 0].astype(np.int64)
    fi = final_idx.reshape(-1).astype(np.int64)
    return int(np.maximum(fi - start + 1, 0).sum())


def tile_work(final_idx, bins, W, H):
    """Entries each tile's raster workgroup walks (through its last contributor), as
    (largest, total): the longest tile bounds a persistent launch from below."""
    tbx = (W + 15) // 16
    T = bins.shape[0]
    ys, xs = np.divmod(np.arange(W * H), W)
    tile = (ys // 16) * tbx + xs // 16
    last = np.full(T, -1, np.int64)
    np.maximum.at(last, tile, final_idx.reshape(-1).astype(np.int64))
    work = np.maximum(last - bins[:, 0].astype(np.int64) + 1, 0)
    return int(work.max()), int(work.sum())


def load_traffic(kernel, workload):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
    except ValueError:
        return None
    e = d.get(kernel)
    if not e or e.get("workload") != workload:
        return None
    return e.get("bytes_per_launch")


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "a design note, not implemented)"}))
        sys.exit(2)
    if a.suite != "c1":
        import bench_suites
        getattr(bench_suites, a.suite)(a)
        return
    from msplat_amd import Dataset, GaussianTrainer, TrainingConfig
    from msplat_amd import _lib

    L = _lib.lib()
    W, H, N = a.width, a.height, a.n
    workload = f"c1:n={N},w={W},h={H},f={a.fx:g},cams={a.cams},hidden={a.hidden}"
    t_setup = time.time()
    spec = f"synthetic:c1:n={N},hidden={a.hidden},cams={a.cams},w={W},h={H},f={a.fx:g}"
    ds = Dataset(spec)
    cfg = TrainingConfig(iterations=30000, sh_degree=3, sh_degree_interval=1, cap_max=a.cap)
    tr = GaussianTrainer(ds, cfg)
    if a.sh_bf16:
        tr.set_option("sh_bf16", 1)
    if a.raster_ppl:
        tr.set_option("raster_ppl", a.raster_ppl)
    if a.persist_grid:
        tr.set_option("persist_grid", a.persist_grid)
    if a.bucket_waves:
        _lib.lib().msplat_set_list_bucket_waves(a.bucket_waves)
    eng = tr.engine_handle
    log(f"setup {time.time() - t_setup:.1f}s ({ds.num_train} cameras)")

    if a.calib_copy_mb > 0:
        nb = a.calib_copy_mb << 20
        src = L.msplat_dev_alloc(nb)
        dst = L.msplat_dev_alloc(nb)
        L.msplat_dev_memset(src, 1, nb)
        assert L.msplat_dev_copy_kernel(dst, src, nb) == 0
        L.msplat_dev_free(src)
        L.msplat_dev_free(dst)
        log(f"calibration copy kernel: {nb} bytes read, {nb} bytes written")

    for _ in range(a.warmup):
        tr.step()
    L.msplat_sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        tr.step()
    L.msplat_sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / max(a.steps, 1)
    its = a.steps / (t1 - t0) if a.steps else 0.0
    assert tr.splat_count == N, "the timed window must run at fixed N"
    log(f"{its:.2f} it/s, {ms_per_step:.3f} ms/step")

    stats = (C.c_int64 * 6)()
    L.msplat_engine_stats(eng, stats)
    I = int(stats[0])
    stages = {}
    if a.stage_steps > 0:
        L.msplat_engine_reset_stage_times(eng)
        L.msplat_engine_set_stage_timing(eng, 1)
        for _ in range(a.stage_steps):
            tr.step()
        L.msplat_sync()
        ms = (C.c_double * 16)()
        nl = (C.c_int64 * 16)()
        ns = L.msplat_engine_stage_times(eng, ms, 16, nl)
        L.msplat_engine_set_stage_timing(eng, 0)
        stages = {L.msplat_engine_stage_name(k).decode(): ms[k] / max(nl[k], 1) for k in range(ns)}
        log("stages (ms/launch): " + ", ".join(f"{k}={v:.3f}" for k, v in stages.items()))

    def read(name, dtype, count):
        out = np.zeros(count, dtype)
        L.msplat_engine_read(eng, name.encode(), out.ctypes.data, out.nbytes)
        return out
    radii = read("radii", np.int32, N)
    V = int((radii > 0).sum())
    T = ((W + 15) // 16) * ((H + 15) // 16)
    bins = read("tile_bins", np.int32, 2 * T).reshape(T, 2)
    final_idx = read("final_idx", np.int32, W * H)
    pairs = pair_counts(final_idx, bins, W, H)
    # valid pairs: entries each pixel took in the forward (sigma < 5.55, alpha >= 1/255, before
    # saturation) -- the work the per-pair FLOP figures describe
    valid = int(read("n_contrib", np.int32, W * H).astype(np.int64).sum())
    tw_max, tw_total = tile_work(final_idx, bins, W, H)
    L.msplat_engine_stats(eng, stats)
    I = int(stats[0])

    # algorithmic work per launch (bytes for HBM-bound kernels, FLOPs for the raster)
    work = {
        "project_sh_fwd": ("hbm", N * ((50 + 90 + 68) if a.sh_bf16 else (56 + 180 + 68))),
        "gaussian_update": ("hbm", N * (390 if a.sh_bf16 else 408)),
        "sh_rest_adam": ("hbm", V * (810 if a.sh_bf16 else 1080) + N * 28),
        "raster_fwd": ("valu", valid * 16),
        "raster_bwd": ("valu", valid * 45),
    }
    roofline = None
    if stages:
        dom = max(stages, key=lambda k: stages[k])
        rk = dom if dom in work else max(work, key=lambda k: stages.get(k, 0.0))
        bound, amount = work[rk]
        sec = stages[rk] * 1e-3
        if bound == "hbm":
            ach = amount / sec / 1e9
            roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(ach / HBM_PEAK_GBS, 4)}
        else:
            ach = amount / sec / 1e12
            roofline = {"bound": "valu", "achieved": round(ach, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(ach / FP32_PEAK_TFLOPS, 4)}
        roofline.update({"traffic": load_traffic(rk, workload), "kernel": rk, "ms_per_launch": round(stages[rk], 4),
                         "work_per_launch": amount, "dominant_kernel": dom})

    cpu = None
    if a.cpu_steps > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        from msplat_amd import scene
        # the host cores this process may run on; OMP_NUM_THREADS, where the environment sets
        # it, is the CPU share the machine gives this GPU's process and caps them
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        omp = os.environ.get("OMP_NUM_THREADS", "")
        threads = min(avail, int(omp)) if omp.isdigit() and int(omp) > 0 else avail
        threads = O.set_threads(threads)
        cpu_model = ""
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        groups = ("means", "scales", "quats", "f_dc", "f_rest", "opac")
        shapes = {"means": (N, 3), "scales": (N, 3), "quats": (N, 4), "f_dc": (N, 3), "f_rest": (N, 15, 3),
                  "opac": (N,)}
        hg = {k: tr.read_group(i, 0).reshape(shapes[k]) for i, k in enumerate(groups)}
        hg["degree"] = 3
        mom = ({k: tr.read_group(i, 1).reshape(shapes[k]) for i, k in enumerate(groups)},
               {k: tr.read_group(i, 2).reshape(shapes[k]) for i, k in enumerate(groups)})
        ost = O.HostState(hg, moments=mom)
        times = []
        step = tr.iteration
        for k in range(a.cpu_steps):
            ci = k % ds.num_train
            cam_d = ds.camera(ci)
            cam = scene.prepare_camera(cam_d["c2w"], cam_d["fx"], cam_d["fy"], cam_d["cx"], cam_d["cy"], W, H)
            hp = tr.step_params(ci, step + 1 + k, step + 1 + k)
            tc = time.perf_counter()
            O.train_step(ost, cam, cam_d["image"], hp, want_ids=False)
            times.append(time.perf_counter() - tc)
            log(f"cpu step {k}: {times[-1]:.2f}s")
        med = statistics.median(times)
        cpu = {"value": round(1.0 / med, 4), "unit": "it/s", "cores": threads, "kind": "port",
               "sample": f"{a.cpu_steps} oracle training steps (FP32, OpenMP, {threads} threads) at N={N}, "
                         f"{W}x{H}, SH degree 3, on the GPU trainer's state; median step time",
               "step_seconds": [round(t, 3) for t in times], "nproc": os.cpu_count(), "affinity_cpus": avail,
               "omp_num_threads": omp or None, "cpu_model": cpu_model}

    out = {
        "metric": METRIC,
        "value": round(its, 3),
        "unit": "it/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+bf16-sh" if a.sh_bf16 else "f32",
        "data": "synthetic",
        "config": {"workload": "single-view 3DGS training step (C1: 64-component clustered mixture)",
                   "gaussians": N, "width": W, "height": H, "fx": a.fx, "cameras": ds.num_train,
                   "hidden_gt_gaussians": a.hidden, "sh_degree": 3, "sh_degree_interval": 1,
                   "loss": "0.8 L1 + 0.2 (1 - SSIM)",
                   "cap_max": a.cap, "intersections": I, "visible": V, "raster_pairs_walked": pairs,
                   "raster_pairs_valid": valid,
                   "tile_entries_walked_max": tw_max, "tile_entries_walked_total": tw_total,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
