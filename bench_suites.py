"""Secondary benchmark suites of bench.py (`--suite ...`), each printing one JSON line:

c0          the small configuration: 50k uniform-cube Gaussians, 16 cameras at 256x256,
            200 steps of L1 + 0.2 (1 - SSIM) (SH degree 0 throughout: interval 1000), on the
            GPU trainer and on the CPU oracle from the same initial state; it/s and final
            loss of both.
c1_full     C1 for 1000 steps with the plain MCMC relocation / growth at steps 500-800
            (cap 3M, 8 of the 64 views held out): wall time, Gaussian count, held-out PSNR.
c3_bf16     3M clustered Gaussians at 1920x1080: it/s with all-fp32 storage and with
            bfloat16 SH coefficients + moments, per-stage times and HBM bytes saved.
c2          batched multi-view: 2M clustered Gaussians, 64 cameras whose sizes are drawn from
            640x480 / 1024x768 / 1600x1040 / 1920x1080, 8 views per step, 500 steps from the
            same initial state in four modes: the views one after another through the
            single-view kernels ("sequential", the baseline), one persistent raster launch
            per direction over all views ("persistent"), that recorded once as a HIP graph
            and replayed ("graph_default"), and the graph with the launch geometry autotuned
            ("graph"); views/s of each (over all steps, after step 10, and after step 60,
            which is past the autotuner's 48 tuning steps) and the batch loss / held-out PSNR
            after 500 steps (equal up to float summation order).
c4          learned relocation proposal: 2M clustered Gaussians at 1600x1040, 8 withheld
            views; the plain GPU form (one view per step, opacity-weighted relocation), 8-view
            batches (Adam steps scaled by sqrt(8)), and 8-view batches with the learned
            proposal (judged by held-out loss),
            all from the same state, evaluated about every 100 views; steps and seconds for
            each to reach the plain form's final held-out PSNR within 0.1 dB, and the
            trajectories.
n1          the learned relocation proposal over many relocation events: 8-view batched steps
            on 2M clustered Gaussians at 1600x1040 (cap 3M), relocation every 100 steps from
            step 500 to the last step (1700 by default: 13 events), three arms from the same
            initial state and camera order (plain relocation, learned forced once trained,
            learned judged by held-out loss), replicated over three initial Gaussian sets
            (paired seeds); held-out PSNR after every event, the paired differences, and the
            network's losses on unseen outcomes against the best constant predictor's.
bf16_error  error against the FP64 oracle of the fp32 and the bf16-SH GPU paths, after 1
            step (image, loss, raster gradients, updated parameters) and after 2000 steps
            (parameters, render, loss), on a small scene (20k Gaussians, 128x128, 8 views).
"""
from __future__ import annotations

import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
GROUPS = ("means", "scales", "quats", "f_dc", "f_rest", "opac")


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    return O, O.set_threads(threads)


def _state(tr):
    return {(k, w): tr.read_group(i, w) for i, k in enumerate(GROUPS) for w in range(3)}


def _host(O, st, n, degree=3, dtype=np.float32, sh_bf16=False):
    nr = (degree + 1) ** 2 - 1
    shapes = {"means": (n, 3), "scales": (n, 3), "quats": (n, 4), "f_dc": (n, 3), "f_rest": (n, nr, 3), "opac": (n,)}
    g = {k: st[(k, 0)].reshape(shapes[k]) for k in GROUPS}
    g["degree"] = degree
    mom = ({k: st[(k, 1)].reshape(shapes[k]) for k in GROUPS}, {k: st[(k, 2)].reshape(shapes[k]) for k in GROUPS})
    return O.HostState(g, dtype=dtype, moments=mom, sh_bf16=sh_bf16)


def _camera(ds, scene, i, test=False):
    c = ds.camera(i, test=test)
    cam = scene.prepare_camera(c["c2w"], c["fx"], c["fy"], c["cx"], c["cy"], c["width"], c["height"])
    return cam, c["image"]


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def c0(a):
    from msplat_amd import Dataset, GaussianTrainer, TrainingConfig, _lib, scene
    O, threads = _oracle()
    L = _lib.lib()
    ds = Dataset("synthetic:c0")
    steps = 200
    tr = GaussianTrainer(ds, TrainingConfig(iterations=steps))
    n = tr.splat_count
    host = _host(O, _state(tr), n)
    cams = [_camera(ds, scene, i) for i in range(ds.num_train)]
    plan = []
    L.msplat_sync()
    t0 = time.perf_counter()
    for s in range(1, steps + 1):
        hp = tr.step_params(0, s, s)
        tr.step()
        plan.append((tr.last_camera, hp))
    L.msplat_sync()
    gpu_s = time.perf_counter() - t0
    gpu_loss = tr.last_loss
    t0 = time.perf_counter()
    cpu_loss = None
    for ci, hp in plan:
        out = O.train_step(host, cams[ci][0], cams[ci][1], hp, want_ids=False)
        cpu_loss = out["loss"]
    cpu_s = time.perf_counter() - t0
    print(json.dumps({"suite": "c0", "gaussians": n, "steps": steps, "views": "16 x 256x256",
                      "gpu_its": round(steps / gpu_s, 2), "cpu_its": round(steps / cpu_s, 3), "cpu_threads": threads,
                      "speedup": round(cpu_s / gpu_s, 1), "gpu_final_loss": gpu_loss, "cpu_final_loss": cpu_loss,
                      "final_loss_rel_diff": abs(gpu_loss - cpu_loss) / abs(cpu_loss),
                      "note": "same initial state, same camera order and step scalars; trajectories "
                              "separate only by float summation order"}), flush=True)


def c1_full(a):
    from msplat_amd import Dataset, GaussianTrainer, TrainingConfig, _lib
    L = _lib.lib()
    ds = Dataset("synthetic:c1", eval_mode=True, test_every=8)
    tr = GaussianTrainer(ds, TrainingConfig(iterations=1000, cap_max=3_000_000))
    counts = {}
    L.msplat_sync()
    t0 = time.perf_counter()
    for s in range(1, 1001):
        st = tr.step()
        if s % 100 == 0:
            counts[s] = st.splat_count
    L.msplat_sync()
    wall = time.perf_counter() - t0
    m = tr.evaluate()
    print(json.dumps({"suite": "c1_full", "steps": 1000, "wall_s": round(wall, 2), "mean_its": round(1000 / wall, 2),
                      "gaussians_by_step": counts, "train_views": ds.num_train, "test_views": ds.num_test,
                      "heldout": m}), flush=True)


def c3_bf16(a):
    from msplat_amd import Dataset, GaussianTrainer, TrainingConfig, _lib
    L = _lib.lib()
    n = 3_000_000
    ds = Dataset(f"synthetic:c1:n={n},w=1920,h=1080,f=1300")
    tr = GaussianTrainer(ds, TrainingConfig(iterations=30000, sh_degree_interval=1, cap_max=n))
    eng = tr.engine_handle
    res = {}
    for mode in ("fp32", "bf16"):
        tr.set_option("sh_bf16", 1 if mode == "bf16" else 0)
        for _ in range(5):
            tr.step()
        L.msplat_sync()
        t0 = time.perf_counter()
        k = 30
        for _ in range(k):
            tr.step()
        L.msplat_sync()
        its = k / (time.perf_counter() - t0)
        L.msplat_engine_reset_stage_times(eng)
        L.msplat_engine_set_stage_timing(eng, 1)
        for _ in range(5):
            tr.step()
        L.msplat_sync()
        ms = (C.c_double * 16)()
        nl = (C.c_int64 * 16)()
        ns = L.msplat_engine_stage_times(eng, ms, 16, nl)
        L.msplat_engine_set_stage_timing(eng, 0)
        stages = {L.msplat_engine_stage_name(i).decode(): round(ms[i] / max(nl[i], 1), 4) for i in range(ns)}
        radii = np.zeros(n, np.int32)
        L.msplat_engine_read(eng, b"radii", radii.ctypes.data, radii.nbytes)
        res[mode] = {"its": round(its, 2), "stages_ms": stages, "visible": int((radii > 0).sum())}
        _log(f"{mode}: {its:.2f} it/s")
    V = res["fp32"]["visible"]
    per_step_saved = n * ((56 + 180) - (50 + 90)) + n * (408 - 390) + V * (1080 - 810)
    resident_saved = n * (45 + 3) * 2 * 2 - n * (45 + 3) * 2  # moments halve, shadows add
    print(json.dumps({"suite": "c3_bf16", "gaussians": n, "view": "1920x1080", "fp32": res["fp32"],
                      "bf16_sh": res["bf16"], "speedup": round(res["bf16"]["its"] / res["fp32"]["its"], 3),
                      "hbm_bytes_saved_per_step": per_step_saved,
                      "resident_bytes_saved": resident_saved}), flush=True)


def bf16_error(a):
    from msplat_amd import Dataset, GaussianTrainer, TrainingConfig, _lib, scene
    from msplat_amd.device import G2D, G2D_CONIC, G2D_OPAC, G2D_RGB, G2D_XY
    O, threads = _oracle()
    L = _lib.lib()
    n, W, H = 20000, 128, 128
    spec = f"synthetic:c1:n={n},hidden=8000,cams=8,w={W},h={H},f=128,jitter=0"
    ds = Dataset(spec)
    # 500 * 0.833 < 500: no MCMC; cap above n / 0.75: the seeds are trained as given
    cfg = TrainingConfig(iterations=500, sh_degree_interval=1, cap_max=2 * n)
    trs = {"fp32": GaussianTrainer(ds, cfg), "bf16": GaussianTrainer(ds, cfg)}
    trs["bf16"].set_option("sh_bf16", 1)
    st0 = _state(trs["fp32"])
    ref = _host(O, st0, n, dtype=np.float64)
    cams = [_camera(ds, scene, i) for i in range(ds.num_train)]
    out = {"suite": "bf16_error", "gaussians": n, "view": f"{W}x{H}", "views": ds.num_train}
    plan = []
    total = 2000
    for s in range(1, total + 1):
        hp = trs["fp32"].step_params(0, s, s)
        for tr in trs.values():
            tr.step()
        ci = trs["fp32"].last_camera
        assert trs["bf16"].last_camera == ci
        plan.append((ci, hp))
        if s == 1:
            o64 = O.train_step(ref, cams[ci][0], cams[ci][1], hp, want_ids=False)
            one = {}
            for mode, tr in trs.items():
                eng = tr.engine_handle
                e = {"loss": abs(tr.last_loss - o64["loss"]) / abs(o64["loss"])}
                img = np.zeros(W * H * 3, np.float32)
                L.msplat_engine_read(eng, b"img", img.ctypes.data, img.nbytes)
                e["img"] = _rel(img, o64["img"].reshape(-1))
                g2d = np.zeros((n, G2D), np.float32)
                L.msplat_engine_read(eng, b"grad2d", g2d.ctypes.data, g2d.nbytes)
                for name, sl in (("v_xy", G2D_XY), ("v_conic", G2D_CONIC), ("v_rgb", G2D_RGB), ("v_opac", G2D_OPAC)):
                    e[name] = _rel(np.ascontiguousarray(g2d[:, sl]).reshape(-1), o64[name].reshape(-1))
                st = _state(tr)
                for k in GROUPS:
                    e["param_" + k] = _rel(st[(k, 0)], ref.p[k][:n].reshape(-1))
                one[mode] = e
            out["after_1_step"] = one
            _log(f"after 1 step: {one}")
    for s, (ci, hp) in enumerate(plan[1:], start=2):
        O.train_step(ref, cams[ci][0], cams[ci][1], hp, want_ids=False)
        if s % 100 == 0:
            _log(f"FP64 oracle step {s} of {total}")
    two = {}
    eval_cam = cams[0][0]
    r64 = O.render(ref, eval_cam, 3, scene.DEFAULT_BG, want_ids=False)["img"]
    for mode, tr in trs.items():
        st = _state(tr)
        e = {k: _rel(st[(k, 0)], ref.p[k][:n].reshape(-1)) for k in GROUPS}
        e["render"] = _rel(tr.render(0), r64)
        two[mode] = e
    out["after_2000_steps"] = two
    out["oracle_threads"] = threads
    out["note"] = "max-norm relative error per array against the FP64 oracle from the same initial state, " \
                  "camera order and step scalars"
    print(json.dumps(out), flush=True)


def c2(a):
    from msplat_amd import Dataset, GaussianTrainer, TrainingConfig, _lib
    L = _lib.lib()
    n, views, steps = 2_000_000, 8, int(os.environ.get("MSPLAT_C2_STEPS", "500"))
    ds = Dataset(f"synthetic:c1:n={n},cams=64,mixed=1", eval_mode=True, test_every=8)
    sizes = sorted({(c["width"], c["height"]) for c in (ds.camera(i, image=False) for i in range(ds.num_train))})
    res = {}
    for mode in ("sequential", "persistent", "graph_default", "graph"):
        tr = GaussianTrainer(ds, TrainingConfig(iterations=30000, cap_max=n))
        tr.set_option("mcmc", 0)
        tr.set_option("batch_sequential", 1 if mode == "sequential" else 0)
        tr.set_option("batch_graph", 1 if mode.startswith("graph") else 0)
        if mode == "graph":
            tr.set_option("autotune", 1)
        L.msplat_sync()
        t0 = time.perf_counter()
        t10 = t60 = None
        for s in range(1, steps + 1):
            tr.step_batch(views)
            if s in (10, 60):
                L.msplat_sync()
                if s == 10:
                    t10 = time.perf_counter()
                else:
                    t60 = time.perf_counter()
        L.msplat_sync()
        t1 = time.perf_counter()
        _, losses = tr.last_batch(losses=True)
        m = tr.evaluate()
        # steps 61.. in every mode: after the autotuner's 48 tuning steps in the tuned mode
        r = {"views_per_s": round(steps * views / (t1 - t0), 2),
             "views_per_s_after_10": round((steps - 10) * views / (t1 - t10), 2),
             "views_per_s_after_60": round((steps - 60) * views / (t1 - t60), 2),
             "ms_per_step": round(1000 * (t1 - t10) / (steps - 10), 3),
             "final_batch_loss": float(np.mean(losses)), "heldout_psnr": m["psnr"]}
        if mode == "graph":
            r["graph"] = tr.batch_info()
            r["autotune"] = tr.autotune_info()
        res[mode] = r
        _log(f"c2 {mode}: {r}")
        del tr
    base = res["sequential"]["views_per_s_after_10"]
    print(json.dumps({"suite": "c2", "gaussians": n, "views_per_step": views, "steps": steps,
                      "view_sizes": [f"{w}x{h}" for w, h in sizes], "modes": res,
                      "speedup_persistent": round(res["persistent"]["views_per_s_after_10"] / base, 3),
                      "speedup_graph": round(res["graph"]["views_per_s_after_10"] / base, 3),
                      "speedup_graph_after_tuning": round(res["graph"]["views_per_s_after_60"] /
                                                          res["sequential"]["views_per_s_after_60"], 3),
                      "note": "same initial state and camera order in every mode; relocation off "
                              "(fixed N)"}), flush=True)


def n1(a):
    from msplat_amd import Dataset, GaussianTrainer, TrainingConfig, _lib
    L = _lib.lib()
    n, views = 2_000_000, 8
    steps = int(os.environ.get("MSPLAT_N1_STEPS", "1700"))
    seeds = [int(x) for x in os.environ.get("MSPLAT_N1_SEEDS", "1,2,3").split(",")]
    arms = {"plain": None, "learned": 2, "judged": 0}
    out = {}
    for seed in seeds:
        ds = Dataset(f"synthetic:c1:n={n},cams=72,seed={seed}", eval_mode=True, test_every=9)
        res = {}
        for arm, mode in arms.items():
            tr = GaussianTrainer(ds, TrainingConfig(iterations=30000, cap_max=int(1.5 * n)))
            tr.set_option("batch_graph", 1)
            tr.set_option("batch_lr_scale", float(np.sqrt(views)))
            if mode is not None:
                tr.set_option("learned_relocation", 1)
                tr.set_option("relocation_mode", mode)
            traj = []
            for s in range(1, steps + 1):
                tr.step_batch(views)
                if s >= 500 and (s % 100 == 0 or s == steps):   # just after each relocation event
                    m = tr.evaluate()
                    traj.append({"step": s, "psnr": m["psnr"], "gaussians": m["num_gaussians"]})
            r = {"trajectory": traj, "final_psnr": traj[-1]["psnr"]}
            if mode is not None:
                r["proposal"] = tr.proposal_info()
            res[arm] = r
            _log(f"n1 seed {seed} {arm}: final {r['final_psnr']:.3f} dB {r.get('proposal', '')}")
            del tr
        out[seed] = res
        del ds
    def paired(arm):
        d = [out[sd][arm]["final_psnr"] - out[sd]["plain"]["final_psnr"] for sd in seeds]
        ev = [b["psnr"] - p["psnr"] for sd in seeds
              for b, p in zip(out[sd][arm]["trajectory"], out[sd]["plain"]["trajectory"])]
        return {"final_delta_db": [round(x, 4) for x in d], "final_delta_mean_db": round(float(np.mean(d)), 4),
                "per_event_delta_mean_db": round(float(np.mean(ev)), 4),
                "events_better": int(sum(x > 0 for x in ev)), "events": len(ev)}
    print(json.dumps({"suite": "n1", "gaussians": n, "views_per_step": views, "steps": steps, "seeds": seeds,
                      "relocation_events_per_run": sum(1 for s in range(500, steps + 1, 100)),
                      "learned_vs_plain": paired("learned"), "judged_vs_plain": paired("judged"),
                      "network_vs_constant": {str(sd): {k: out[sd]["learned"]["proposal"][k]
                                                        for k in ("val_bce", "base_bce", "val_off", "base_off",
                                                                  "learned_events", "events")} for sd in seeds},
                      "judged_learned_events": {str(sd): out[sd]["judged"]["proposal"]["learned_events"]
                                                for sd in seeds},
                      "runs": out}), flush=True)


def c4(a):
    """Time to the plain GPU form's held-out PSNR: single-view steps with the opacity-weighted
    relocation (the plain form) against 8-view batches (plain relocation) and 8-view batches
    with the learned proposal (judged by held-out loss)."""
    from msplat_amd import Dataset, GaussianTrainer, TrainingConfig, _lib
    L = _lib.lib()
    n, views = 2_000_000, 8
    steps = int(os.environ.get("MSPLAT_C4_STEPS", "1200"))
    every = 100
    ds = Dataset(f"synthetic:c1:n={n},cams=72", eval_mode=True, test_every=9)
    # Adam step sizes of the batched runs scaled linearly with the views per step (measured
    # against sqrt(8), 4 and 5.66: the larger the scale the sooner the target, up to 8)
    lr_scale = float(os.environ.get("MSPLAT_C4_LR_SCALE", str(float(views))))
    runs = {}
    for mode in ("plain", "batched", "batched_learned"):
        tr = GaussianTrainer(ds, TrainingConfig(iterations=steps, cap_max=int(1.5 * n)))
        if mode != "plain":
            tr.set_option("batch_graph", 1)
            tr.set_option("batch_lr_scale", lr_scale)
        if mode == "batched_learned":
            tr.set_option("learned_relocation", 1)
        traj = []
        L.msplat_sync()
        t0 = time.perf_counter()
        busy = 0.0
        ev = every if mode == "plain" else max(1, every // views)   # about every 100 views
        for s in range(1, steps + 1):
            tr.step() if mode == "plain" else tr.step_batch(views)
            if s % ev == 0 or s == steps:
                L.msplat_sync()
                busy += time.perf_counter() - t0
                m = tr.evaluate()
                traj.append({"step": s, "s": round(busy, 2), "psnr": m["psnr"], "gaussians": m["num_gaussians"]})
                _log(f"c4 {mode} {traj[-1]}")
                t0 = time.perf_counter()
        run = {"views_per_step": 1 if mode == "plain" else views, "trajectory": traj}
        if mode == "batched_learned":
            run["proposal"] = tr.proposal_info()
        runs[mode] = run
        del tr
    target = runs["plain"]["trajectory"][-1]["psnr"] - 0.1
    reach = {m: next((p for p in r["trajectory"] if p["psnr"] >= target), None) for m, r in runs.items()}
    print(json.dumps({"suite": "c4", "gaussians": n, "view": "1600x1040", "heldout_views": ds.num_test,
                      "steps": steps, "batched_lr_scale": round(lr_scale, 4),
                      "target_psnr": round(target, 3), "reaches_target": reach,
                      "final_psnr": {m: r["trajectory"][-1]["psnr"] for m, r in runs.items()},
                      "runs": runs}), flush=True)
