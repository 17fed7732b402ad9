#!/usr/bin/env python3
"""Benchmark of the per-pixel effect renderer on one MI355X.

Headline (one JSON line on stdout): Mpixel/s of the batched form -- 512
frame requests mixing the SDF sphere tracer, the voxel landscape and the
tunnel remap+blur+blit in a seeded 40/30/30 ratio at five resolutions,
FP16 shading -- with inputs resident in HBM.  One step renders the whole
batch.  Per-effect plain-FP32 (one frame per launch) vs batched-FP16
figures, the FP16-vs-FP64 error of the 2560x1440 precision form, the
roofline of the dominant kernel and a CPU baseline ride along in the
same line.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from cookiedough_amd import workloads as wl  # noqa: E402
from cookiedough_amd._lib import (CKD_FX_SDF, CKD_FX_TUNNEL, CKD_FX_VOXEL,  # noqa: E402
                                  CKD_SHADE_FP16, CKD_SHADE_FP32, Tuning)

PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 vector (packed) peak
PEAK_HBM_GBS = 8000.0
PEAK_FP32_UNPACKED_TFLOPS = 78.6  # without packed FP32 math

# Algorithmic FP32 FLOP of the raymarched FX-map effects (code/shadertoy.cpp),
# (per march step, per map pixel outside the march): + - * / sqrt min max
# and each float<->int conversion count 1; lutcosf 6 (scale, trunc, frac,
# lerp 3), lutsinf 7; powf 40, expf and atan2f 20; the Cephes log / exp of
# GammaAdj 33 / 28, so GammaAdj of four lanes 248.
#   plasma:   fPlasma 42 + step bookkeeping 5; map pixel: UV + direction 14,
#             the second fPlasma 45, colour 14, gamma 248, pack 8
#   nautilus: fNautilus 50 + 5; pixel: UV, roll, vNorm 25, two gradients
#             337, specular 46, colour 26, gamma 248, pack 8
#   spikey (distant) and spikey_close: fSpikey 27 + 8; pixel: UV, roll,
#             vNorm 26, gradient 98, shading 83, fog 26, lerp 12, gamma
#             248, pack 8
#   sinuses:  fSinMap 81 + 8; pixel: UV, roll, vNorm 19, normal 260,
#             shading 58, fog 26, lerp 12, gamma 248, pack 8
#   laura:    fLaura 21 + 8; pixel: UV and three rotations 36, normal 80,
#             light and specular 114, fog 26, lerp 12, gamma 248, pack 8
#   fdtunnel: no march; pixel: direction 29, flower and box 39, two
#             atan2 lookups 50, shade 26, two samples 16, two lerps 24,
#             pack 8
ST_FLOP = {"plasma": (47, 329), "nautilus": (55, 691), "spikey": (35, 475),
           "spikey_close": (35, 473), "sinuses": (89, 638), "laura": (29, 530),
           "fdtunnel": (0, 196)}

# kernel timers (CKD_KT_*) of the batched call of each classic workload
CLASSIC_KERNELS = {"sdf_c0": [("k_sdf_persistent", 1)], "voxel_c1": [("k_landscape", 2)],
                   "tunnel_c2": [("k_tunnel_h", 3), ("k_tunnel_v", 4)]}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sdf_flops(total_steps: int, pixels: int, n_spheres: int) -> float:
    """Algorithmic FLOPs of the sphere tracer: per march step F_sdf+7
    (F_sdf = 11 per sphere + 2 for the plane), per pixel 6 SDF evaluations
    for the normal and ~40 for shading (FMA = 2, sqrt/exp/log = 1)."""
    f_sdf = 11 * n_spheres + 2
    return total_steps * (f_sdf + 7) + pixels * (6 * f_sdf + 40)


def sdf_flops_executed(total_steps: int, eval_steps: int, pixels: int, eval_shades: int,
                       n_spheres: int) -> float:
    """FLOPs the batched tracer executes (same conventions): every step the
    ground, the advance and the bounding-sphere test (2 + 7 + 9), the spheres
    (11 each) only on the steps that evaluated them; every pixel the ground's
    six normal terms, the test and shading (12 + 8 + 40), the spheres' six
    normal terms (66 each) only where evaluated"""
    return (total_steps * 18 + eval_steps * 11 * n_spheres + pixels * 60
            + eval_shades * 66 * n_spheres)


def host_info() -> dict:
    """the host the CPU baselines ran on: logical CPUs, the affinity the
    process got, CPU model, sockets x cores"""
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        text = Path("/proc/cpuinfo").read_text()
        models = re.findall(r"^model name\s*:\s*(.+)$", text, re.M)
        sockets = set(re.findall(r"^physical id\s*:\s*(\d+)$", text, re.M))
        cores = re.findall(r"^cpu cores\s*:\s*(\d+)$", text, re.M)
        info["cpu_model"] = models[0].strip() if models else None
        info["sockets"] = len(sockets) or None
        info["cores_per_socket"] = int(cores[0]) if cores else None
    except OSError:
        pass
    return info


def median_rate(fn, pixels: int, warmup: int = 3, runs: int = 5) -> dict:
    """Mpixel/s of fn over `pixels`: the median of `runs` timed calls after
    `warmup` untimed ones"""
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    med = statistics.median(ts)
    return {"mpx_s": round(pixels / med / 1e6, 2), "median_s": round(med, 4),
            "runs_s": [round(t, 4) for t in ts], "warmup": warmup}


def headline_sample(items) -> list[int]:
    """indices of the batch requests whose output is checked: every
    320x180 and 640x360 request, and the first request of each effect at
    each larger size"""
    pick, seen = [], set()
    for i, it in enumerate(items):
        if it.W <= 640:
            pick.append(i)
        elif (it.effect, it.W) not in seen:
            seen.add((it.effect, it.W))
            pick.append(i)
    return pick


class Timer:
    def __init__(self, ctx):
        self.ctx = ctx

    def run(self, fn, reps: int = 1) -> float:
        """ms for `reps` calls of fn, HIP events on the context stream"""
        self.ctx.sync()
        self.ctx.record(0)
        for _ in range(reps):
            fn()
        self.ctx.record(1)
        self.ctx.sync()
        return self.ctx.elapsed_ms(0, 1)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    # the first replays after the autotune and the graph capture run 5-12%
    # slower while the clocks come back up: three untimed ones by default
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--tuning", default=None,
                    help="steps,chunk,blocksPerCU,vSeg: fixed launch shapes instead of autotuning")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the batched headline (for kernel profiles)")
    ap.add_argument("--sections", default="classic,c3,shadertoy,voxel2",
                    help="per-effect sections to run: classic (SDF c0, voxel c1, tunnel c2), "
                         "c3, shadertoy, voxel2 (ball and twister)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer to another, 4 times: the known-size "
                         "traffic that scales the HBM byte counters of a profile")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(frames and requests are independent, so several GPUs would each take a share "
              "of the request queue)",
              file=sys.stderr)
        return 2

    from cookiedough_amd.device import Context
    ctx = Context(0)
    T = Timer(ctx)

    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 4
        a, b = ctx.empty(n), ctx.empty(n)
        ctx.memset32(a, 0x01020304, n)
        for _ in range(4):
            ctx.memcpy_d2d(b, a, 4 * n)
        ctx.sync()
        print(json.dumps({"copy_bytes": 4 * n, "copies": 4}))
        ctx.close()
        return 0

    # ---------------------------------------------------------------- inputs
    t_in = time.time()
    maps = wl.landscape_maps(0)
    ctx.landscape_create(maps.height, maps.color, maps.fog)
    path = wl.landscape_path(0)
    tex = wl.noise_texture(0, 9)
    d_tex = ctx.upload(tex)
    scene8 = wl.sdf_scene(0, 8, 64)
    scene32 = wl.sdf_scene(0, 32, 128)

    items = wl.batch_requests(0, args.requests)
    total_px = sum(i.W * i.H for i in items)
    tmp_px = sum(i.W * i.H for i in items if i.effect == CKD_FX_TUNNEL)
    out = ctx.empty(total_px)
    tmp = ctx.empty(max(1, tmp_px))
    reqs, off, grids, offs = [], 0, [], []
    for it in items:
        g = ctx.upload(wl.tunnel_grid(it.W, it.H, it.t)) if it.effect == CKD_FX_TUNNEL else None
        grids.append(g)
        reqs.append(wl.make_request(it, out.ptr + 4 * off, scene8, path, g.ptr if g else None))
        offs.append(off)
        off += it.W * it.H
    sdf_only = ctx.request_array([r for r, it in zip(reqs, items) if it.effect == CKD_FX_SDF])
    reqs = ctx.request_array(reqs)
    log(f"inputs ready in {time.time() - t_in:.1f}s: {len(items)} requests, {total_px / 1e6:.1f} Mpx")

    # the persistent SDF kernel is the roofline's kernel: pcount accumulates
    # its march steps on the device and psdf["px"] the pixels it shades
    pcount = ctx.upload(np.zeros(1, np.uint64))
    psdf = {"px": 0}
    sdf_px_batch = sum(i.W * i.H for i in items if i.effect == CKD_FX_SDF)
    # launch shapes tuned on this batch (untimed), then the batch captured
    # once as a hipGraph; a step is one replay
    t_tune = time.time()
    if args.tuning:
        s_, c_, b_, v_ = (int(x) for x in args.tuning.split(","))
        ctx.set_tuning(Tuning(s_, c_, b_, v_))
        tuning = ctx.get_tuning().as_dict()
        log(f"fixed tuning: {tuning}")
    else:
        tuning = ctx.autotune_batch(reqs, scene8, CKD_SHADE_FP16, d_tex, 9, tmp, tmp_px).as_dict()
        log(f"autotuned in {time.time() - t_tune:.1f}s: {tuning}")
    graph = ctx.batch_graph(reqs, scene8, CKD_SHADE_FP16, d_tex, 9, tmp, tmp_px, pcount)

    def batch_step():
        graph.launch()

    # ------------------------------------------------------------- headline
    for _ in range(args.warmup):
        batch_step()
    ctx.sync()
    st_before = int(pcount.numpy()[0])
    ms_total = T.run(batch_step, args.steps)
    steps_total = int(pcount.numpy()[0]) - st_before
    ms_step = ms_total / args.steps
    value = total_px / (ms_step * 1e-3) / 1e6
    log(f"batch: {ms_step:.2f} ms/step, {value:.0f} Mpx/s")
    # the output of the last timed replay, for the requests the CPU leg
    # checks against the oracle
    checked = headline_sample(items)
    ctx.sync()
    headline_out = {i: out.view(offs[i], (items[i].W * items[i].H,)).numpy() for i in checked}

    # ------------------------------------------- roofline: persistent tracer
    # The kernel's own launches, bracketed by HIP events on the stream it
    # runs on: eager renders of the same batch with the same launch shapes
    # (a graph replay can only be timed as a whole, and its voxel / tunnel
    # branch runs beside the SDF kernel).
    work = ctx.upload(np.zeros(2, np.uint64))
    ctx.sdf_work_counter(work)
    ctx.kernel_timing(True)
    st_zero = int(pcount.numpy()[0])
    for _ in range(args.warmup + args.steps):
        ctx.render_batch(reqs, scene8, CKD_SHADE_FP16, d_tex, 9, tmp, tmp_px, pcount)
        psdf["px"] += sdf_px_batch
    ctx.sync()
    ms_k, launches = ctx.kernel_timing_read(1)  # CKD_KT_SDF_PERSISTENT
    ctx.kernel_timing(False)
    ctx.sdf_work_counter(None)
    eval_steps, eval_shades = (int(v) for v in work.numpy())
    steps_k = int(pcount.numpy()[0]) - st_zero
    flops_k = sdf_flops(steps_k, psdf["px"], 8)
    achieved = flops_k / (ms_k * 1e-3) / 1e12
    exe_k = sdf_flops_executed(steps_k, eval_steps, psdf["px"], eval_shades, 8)
    executed = {"achieved": round(exe_k / (ms_k * 1e-3) / 1e12, 2),
                "frac": round(exe_k / (ms_k * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
                "steps_evaluating_spheres": round(eval_steps / max(1, steps_k), 4),
                "shades_evaluating_spheres": round(eval_shades / max(1, psdf["px"]), 4),
                "algorithmic": "per launch: sum(steps)*18 + sphere_steps*88 + px*60 + "
                               "sphere_shades*528 FLOP: the spheres only where the bounding-sphere "
                               "test did not rule them out"}
    # HBM bytes per launch of this kernel on the 512-request batch, measured
    # with the FETCH_SIZE / WRITE_SIZE counters scaled by a copy of known size
    traffic = None
    prof = Path(__file__).resolve().parent / "profiles" / "k_sdf_persistent_batch512_traffic.md"
    if args.requests == 512 and prof.exists():
        m = re.search(r"traffic_bytes_per_launch = (\d+)", prof.read_text())
        traffic = int(m.group(1)) if m else None
    roofline = {"bound": "valu", "kernel": "k_sdf_persistent<fp16 shading, 8 spheres>",
                "achieved": round(achieved, 2), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-scaled)",
                "launches": launches, "avg_launch_ms": round(ms_k / max(1, launches), 4),
                "algorithmic": "per launch: sum(steps)*(F_sdf+7) + px*(6*F_sdf+40), "
                               "F_sdf = 11*8+2 = 90 FLOP (every sphere at every step, including those "
                               "the bounding-sphere test skips exactly; `executed` counts only what runs)",
                "executed": executed}
    # the same kernel alone: the batch's SDF requests only, at their own
    # autotuned launch shape (in the batch it shares the GPU with the voxel
    # and tunnel branch, which is what the figures above include)
    if sdf_px_batch:
        headline_tuning = ctx.get_tuning()
        ctx.autotune_batch(sdf_only, scene8, CKD_SHADE_FP16, d_tex, 9, tmp, tmp_px)
        ctx.kernel_timing(True)
        st0 = int(pcount.numpy()[0])
        for _ in range(args.warmup + args.steps):
            ctx.render_batch(sdf_only, scene8, CKD_SHADE_FP16, d_tex, 9, tmp, tmp_px, pcount)
        ctx.sync()
        ms_a, launches_a = ctx.kernel_timing_read(1)
        ctx.kernel_timing(False)
        n_a = args.warmup + args.steps
        ach_a = sdf_flops(int(pcount.numpy()[0]) - st0, sdf_px_batch * n_a, 8) / (ms_a * 1e-3) / 1e12
        roofline["standalone"] = {"achieved": round(ach_a, 2), "frac": round(ach_a / PEAK_FP32_TFLOPS, 4),
                                  "avg_launch_ms": round(ms_a / max(1, launches_a), 4),
                                  "tuning": ctx.get_tuning().as_dict()}
        ctx.set_tuning(headline_tuning)
    log("roofline", roofline)

    per, es, cpu, headline_err = {}, None, None, None
    sections = set(args.sections.split(","))
    if not args.headline_only:
        # ----------------------------------------------------- per effect (c0-c2)
        W, H = 1280, 720
        if "classic" in sections:

            # SDF c0: 64 frames at t=k/60
            cams = [wl.sdf_camera(scene8, k / 60) for k in range(64)]
            fr_sdf = ctx.empty(64 * W * H)
            fr_sdf_p = ctx.empty(64 * W * H)

            def sdf_plain():
                for k, c in enumerate(cams):
                    ctx.sdf_render(fr_sdf_p.view(k * W * H, (W * H,)), W, H, scene8, c, CKD_SHADE_FP32)

            sdf_reqs = []
            for k in range(64):
                r = wl.make_request(wl.BatchItem(CKD_FX_SDF, W, H, k / 60), fr_sdf.ptr + 4 * k * W * H,
                                    scene8, path, None)
                sdf_reqs.append(r)
            sdf_reqs = ctx.request_array(sdf_reqs)

            def sdf_batched():
                ctx.render_batch(sdf_reqs, scene8, CKD_SHADE_FP16, None, 9, None, 0, pcount)
    
            # voxel c1: 120 frames on the Lissajous path
            lcams = [wl.landscape_cam(path, k / 60) for k in range(120)]
            lcams_arr = ctx.landscape_cams(lcams)
            fr_vox = ctx.empty(120 * W * H)

            def vox_plain():
                for k, c in enumerate(lcams):
                    ctx.landscape_draw(fr_vox.view(k * W * H, (W * H,)), W, H, [c])

            def vox_batched():
                ctx.landscape_draw(fr_vox, W, H, lcams_arr)

            # tunnel c2: 240 frames, radius U{1..32}
            radii = wl.tunnel_radii(0, 240)
            tgrids = [ctx.upload(wl.tunnel_grid(W, H, k / 60)) for k in range(240)]
            fr_tun = ctx.empty(240 * W * H)
            t0 = ctx.empty(W * H)
            tun_tmp = ctx.empty(240 * W * H)

            def tun_plain():
                for k in range(240):
                    ctx.tunnel_frame(fr_tun.view(k * W * H, (W * H,)), t0, W, H, tgrids[k], d_tex, 9,
                                     int(radii[k]))

            tun_reqs = []
            for k in range(240):
                it = wl.BatchItem(CKD_FX_TUNNEL, W, H, k / 60, int(radii[k]))
                tun_reqs.append(wl.make_request(it, fr_tun.ptr + 4 * k * W * H, scene8, path, tgrids[k].ptr))
            tun_reqs = ctx.request_array(tun_reqs)

            def tun_batched():
                ctx.render_batch(tun_reqs, scene8, CKD_SHADE_FP16, d_tex, 9, tun_tmp, 240 * W * H)

            # the batched calls that take launch shapes are autotuned on their
            # own batch, as the headline is (the voxel call has none)
            tune = {"sdf_c0": lambda: ctx.autotune_batch(sdf_reqs, scene8, CKD_SHADE_FP16, None, 9, None, 0),
                    "tunnel_c2": lambda: ctx.autotune_batch(tun_reqs, scene8, CKD_SHADE_FP16, d_tex, 9,
                                                            tun_tmp, 240 * W * H)}
            for name, frames, plain, batched in [("sdf_c0", 64, sdf_plain, sdf_batched),
                                                 ("voxel_c1", 120, vox_plain, vox_batched),
                                                 ("tunnel_c2", 240, tun_plain, tun_batched)]:
                reps = max(1, args.steps)
                saved_tuning = ctx.get_tuning()
                if name in tune and not args.tuning:
                    tune[name]()
                plain()
                batched()
                mp = T.run(plain, reps) / reps
                st0 = int(pcount.numpy()[0])
                mb = T.run(batched, reps) / reps
                steps_b = (int(pcount.numpy()[0]) - st0) // reps
                px = frames * W * H
                per[name] = {"frames": frames, "plain_fp32_mpx_s": round(px / mp / 1e3, 1),
                             "batched_fp16_mpx_s": round(px / mb / 1e3, 1),
                             "speedup": round(mp / mb, 3)}
                # the batched call's kernels alone: HIP events around each
                # launch on its own stream, averaged over `reps` calls
                kernels = CLASSIC_KERNELS[name]
                if name == "sdf_c0":
                    ctx.memset32(work, 0, 4)
                    ctx.sdf_work_counter(work)
                st1 = int(pcount.numpy()[0])
                ctx.kernel_timing(True)
                for _ in range(reps):
                    batched()
                ctx.sync()
                ctx.sdf_work_counter(None)
                steps_t = int(pcount.numpy()[0]) - st1
                kt = {}
                for kname, kid in kernels:
                    ms_k, n_k = ctx.kernel_timing_read(kid)
                    kt[kname] = {"avg_launch_ms": round(ms_k / max(1, n_k), 4), "launches_per_call": n_k // reps}
                ctx.kernel_timing(False)
                mk = sum(v["avg_launch_ms"] * v["launches_per_call"] for v in kt.values())
                # the batched call's kernels against their bound
                if name == "sdf_c0":
                    ach = sdf_flops(steps_b, px, 8) / (mk * 1e-3) / 1e12
                    rl = {"bound": "valu", "achieved": round(ach, 2), "peak": PEAK_FP32_TFLOPS,
                          "unit": "TFLOP/s", "algorithmic": "sum(steps)*97 + px*580 FLOP (every "
                          "sphere at every step; `executed` counts only the spheres evaluated)"}
                    es_, esh_ = (int(v) for v in work.numpy())
                    exe = sdf_flops_executed(steps_t, es_, px * reps, esh_, 8) / reps
                    rl["executed"] = {"achieved": round(exe / (mk * 1e-3) / 1e12, 2),
                                      "frac": round(exe / (mk * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
                                      "steps_evaluating_spheres": round(es_ / max(1, steps_t), 4),
                                      "shades_evaluating_spheres": round(esh_ / max(1, px * reps), 4)}
                elif name == "voxel_c1":
                    # 80 integer ops per column step, 10 per span pixel,
                    # counted as 2 FLOP each against the unpacked FP32 rate
                    ops = frames * W * 512 * 80 + px * 10
                    ach = 2 * ops / (mk * 1e-3) / 1e12
                    rl = {"bound": "valu", "achieved": round(ach, 2), "peak": PEAK_FP32_UNPACKED_TFLOPS,
                          "unit": "TFLOP/s (2 x int ops)",
                          "algorithmic": "512*80 ops per column + 10 per pixel"}
                else:
                    # remap 4 B/px written, blur H+V 8 B/px, BlitAdd32 12 B/px
                    ach = 24 * px / (mk * 1e-3) / 1e9
                    rl = {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "algorithmic": "24 B per pixel (remap 4, blur 8, BlitAdd32 12); the "
                                         "kernels' own: k_tunnel_h 8 written, k_tunnel_v 8 read + 4 "
                                         "written"}
                    for kname, bpp in (("k_tunnel_h", 8), ("k_tunnel_v", 12)):
                        if kname in kt:
                            kt[kname]["achieved_gbs"] = round(bpp * px / (kt[kname]["avg_launch_ms"]
                                                                          * 1e-3) / 1e9, 1)
                rl["frac"] = round(ach / rl["peak"], 4)
                rl["kernels"] = kt
                rl["kernel_ms_per_call"] = round(mk, 4)
                rl["whole_call_ms"] = round(mb, 4)
                per[name]["roofline"] = rl
                if name in tune:
                    per[name]["tuning"] = ctx.get_tuning().as_dict()
                ctx.set_tuning(saved_tuning)
                log(name, per[name])

        if "c3" in sections:
            # --------------------------------------- c3 precision form (2560x1440)
            W3, H3 = 2560, 1440
            n3 = 64
            fr3 = ctx.empty(n3 * W3 * H3)
            fr3p = ctx.empty(n3 * W3 * H3)
            cams3 = [wl.sdf_camera(scene32, k / 60) for k in range(n3)]

            def c3_plain():
                for k, c in enumerate(cams3):
                    ctx.sdf_render(fr3p.view(k * W3 * H3, (W3 * H3,)), W3, H3, scene32, c, CKD_SHADE_FP32)

            reqs3 = [wl.make_request(wl.BatchItem(CKD_FX_SDF, W3, H3, k / 60), fr3.ptr + 4 * k * W3 * H3,
                                     scene32, path, None) for k in range(n3)]
            reqs3 = ctx.request_array(reqs3)

            def c3_batched():
                ctx.render_batch(reqs3, scene32, CKD_SHADE_FP16, None, 9, None, 0, pcount)

            saved_tuning = ctx.get_tuning()
            if not args.tuning:
                ctx.autotune_batch(reqs3, scene32, CKD_SHADE_FP16, None, 9, None, 0)
            c3_plain()
            c3_batched()
            mp = T.run(c3_plain)
            st0 = int(pcount.numpy()[0])
            mb = T.run(c3_batched)
            steps3 = int(pcount.numpy()[0]) - st0
            px3 = n3 * W3 * H3
            ach3 = sdf_flops(steps3, px3, 32) / (mb * 1e-3) / 1e12
            ctx.memset32(work, 0, 4)
            ctx.sdf_work_counter(work)
            st3 = int(pcount.numpy()[0])
            c3_batched()
            ctx.sync()
            ctx.sdf_work_counter(None)
            steps3w = int(pcount.numpy()[0]) - st3
            es3, esh3 = (int(v) for v in work.numpy())
            exe3 = sdf_flops_executed(steps3w, es3, px3, esh3, 32) / (mb * 1e-3) / 1e12
            per["sdf_c3_precision"] = {"frames": n3, "plain_fp32_mpx_s": round(px3 / mp / 1e3, 1),
                                       "batched_fp16_mpx_s": round(px3 / mb / 1e3, 1),
                                       "speedup": round(mp / mb, 3),
                                       "roofline": {"bound": "valu", "achieved": round(ach3, 2),
                                                    "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                                    "algorithmic": "sum(steps)*361 + px*2164 FLOP "
                                                                   "(every sphere at every step)",
                                                    "frac": round(ach3 / PEAK_FP32_TFLOPS, 4),
                                                    "executed": {
                                                        "achieved": round(exe3, 2),
                                                        "frac": round(exe3 / PEAK_FP32_TFLOPS, 4),
                                                        "steps_evaluating_spheres":
                                                            round(es3 / max(1, steps3w), 4),
                                                        "shades_evaluating_spheres":
                                                            round(esh3 / max(1, px3), 4)}},
                                       "tuning": ctx.get_tuning().as_dict()}
            ctx.set_tuning(saved_tuning)
            log("sdf_c3", per["sdf_c3_precision"])


        if "shadertoy" in sections:
            # ------------------- the reference's raymarched effects (1280x720 Draw)
            # plain: one Draw per call; batched: all frames of an effect in one call
            # (one map launch, then the per-frame blur / blit of its Draw function)
            n_st = 64
            st_tex = np.ascontiguousarray(wl.noise_texture(11, 10).reshape(-1))
            st_tex_hi = np.ascontiguousarray(wl.noise_texture(12, 10).reshape(-1))
            ctx.fdtunnel_create(st_tex, st_tex_hi)
            st_blur = wl.st_blur_maps(W, H)
            ctx.spikey_blurmaps(st_blur, W, H)
            fr_st = ctx.empty(n_st * W * H)
            st_cnt = ctx.upload(np.zeros(1, np.uint64))
            st_per = {}
            for fx in ["plasma", "nautilus", "spikey", "spikey_close", "sinuses", "laura", "fdtunnel"]:
                name = "spikey" if fx == "spikey_close" else fx
                close = fx == "spikey_close"
                frames = wl.st_frames(fx, n_st)
                arr = ctx._frames(name, frames)[0]
                singles = [ctx._frames(name, [f])[0] for f in frames]

                def st_plain():
                    for k in range(n_st):
                        ctx.st_draw(name, fr_st.view(k * W * H, (W * H,)), W, H, singles[k], close=close)

                def st_batched():
                    ctx.st_draw(name, fr_st, W, H, arr, close=close)

                st_plain()
                st_batched()
                mp, mb = T.run(st_plain), T.run(st_batched)
                px = n_st * W * H
                st_per[fx] = {"frames": n_st, "plain_mpx_s": round(px / mp / 1e3, 1),
                              "batched_mpx_s": round(px / mb / 1e3, 1),
                              "ms_per_frame_batched": round(mb / n_st, 4)}
                # the FX-map kernel against the FP32 rate: its march steps
                # counted on the device, its time from HIP events
                ctx.memset32(st_cnt, 0, 2)
                ctx.shadertoy_step_counter(st_cnt)
                st_batched()
                ctx.sync()
                ctx.shadertoy_step_counter(None)
                steps = int(st_cnt.numpy()[0])
                ctx.kernel_timing(True)
                for _ in range(3):
                    st_batched()
                ctx.sync()
                ms_k, n_k = ctx.kernel_timing_read(5)  # CKD_KT_SHADERTOY
                ctx.kernel_timing(False)
                map_px = n_st * (W // 2 + 4) * (H // 2 + 4)
                f_step, f_px = ST_FLOP[fx]
                flop = steps * f_step + map_px * f_px
                t_k = ms_k / max(1, n_k)
                st_per[fx]["roofline"] = {
                    "bound": "valu", "kernel": "k_st_map", "avg_launch_ms": round(t_k, 4),
                    "steps_per_map_pixel": round(steps / map_px, 2),
                    "achieved": round(flop / (t_k * 1e-3) / 1e12, 2), "peak": PEAK_FP32_TFLOPS,
                    "unit": "TFLOP/s",
                    "algorithmic": f"steps*{f_step} + map_px*{f_px} FLOP",
                    "frac": round(flop / (t_k * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4)}
                log("shadertoy", fx, st_per[fx])
            per["shadertoy_1280x720"] = st_per

        if "voxel2" in sections:
            # ------------------- the voxel ball and the torus twister (1280x720 Draw)
            # plain: one Draw per call; batched: all frames in one call (the
            # per-frame tables are staged and the launches queued back to back)
            n_v2 = 32
            fr_v2 = ctx.empty(n_v2 * W * H)
            bm_v2, tm_v2 = wl.ball_maps(W, H), wl.twister_maps(W, H)
            ctx.ball_create(bm_v2, W, H)
            ctx.twister_create(tm_v2["heightMap"], tm_v2["colorMap"], tm_v2["background"], W, H)
            tgt_v2 = ctx.empty(W * H)
            ctx.memset32(tgt_v2, 0, W * H)
            v2_sets = {"ball_beams": ("ball", wl.ball_frames(n_v2, 0)),
                       "ball_envmap": ("ball", wl.ball_frames(n_v2, 1)),
                       "twister": ("twister", wl.twister_frames(n_v2))}
            v2_per = {}
            for key, (kind, frames) in v2_sets.items():
                draw = ctx.ball_draw if kind == "ball" else ctx.twister_draw

                def v2_plain():
                    for k in range(n_v2):
                        draw(fr_v2.view(k * W * H, (W * H,)), tgt_v2, W, H, frames[k])

                def v2_batched():
                    draw(fr_v2, tgt_v2, W, H, frames)

                v2_plain()
                v2_batched()
                mp, mb = T.run(v2_plain), T.run(v2_batched)
                px = n_v2 * W * H
                v2_per[key] = {"frames": n_v2, "plain_mpx_s": round(px / mp / 1e3, 1),
                               "batched_mpx_s": round(px / mb / 1e3, 1),
                               "ms_per_frame_batched": round(mb / n_v2, 4)}
                log("voxel2", key, v2_per[key])
            per["ball_twister_1280x720"] = v2_per

        # The CPU leg is the only part of the bench that runs the oracle and
        # the reference harness: the checks of the GPU frames against them,
        # and the timed CPU baselines.
        if not args.no_cpu:
            import oracles  # noqa: E402  (the checker)
            from helpers import error_stats, merge_reports, sdf_fp64, sdf_report  # noqa: E402
            from oracles import P  # noqa: E402
            orc = oracles.oracle()
            host = host_info()
            # the GPU box's CPU share for one GPU is 16 threads
            nthr = max(1, min(16, host["affinity_cpus"]))
            host["threads_used"] = nthr
            use_ref = oracles.ref_available()
            refl = oracles.ref() if use_ref else None
            if use_ref:
                refl.ref_omp_threads(nthr)

            def sdf_dense(batched_dev, plain_dev, sc, Wd, Hd, n, row_step=4):
                """every row_step-th row of every frame of the batched FP16 and
                the plain FP32 renders against one FP64 render"""
                reps = {"batched_fp16": [], "plain_fp32": []}
                for k in range(n):
                    ref, _, cond = sdf_fp64(orc, sc, Wd, Hd, k / 60, 0, row_step, nthr)
                    for key, dev in (("batched_fp16", batched_dev), ("plain_fp32", plain_dev)):
                        img = dev.view(k * Wd * Hd, (Wd * Hd,)).numpy().reshape(Hd, Wd)
                        reps[key].append(sdf_report(img[::row_step].reshape(-1), ref, cond))
                return {"rows_sampled": f"every {row_step}th row of all {n} frames",
                        **{key: merge_reports(v) for key, v in reps.items()}}

            # ---- the headline batch: its sampled requests against the oracle
            t_chk = time.time()
            sdf_reps, vox, tun = [], {"requests": 0, "mismatched_px": 0}, {"requests": 0, "mismatched_px": 0}
            for i in checked:
                it, got = items[i], headline_out[i]
                n = it.W * it.H
                if it.effect == CKD_FX_SDF:
                    ref, _, cond = sdf_fp64(orc, scene8, it.W, it.H, it.t, 0, 1, nthr)
                    sdf_reps.append(sdf_report(got, ref, cond))
                    continue
                ref = np.zeros(n, np.uint32)
                if it.effect == CKD_FX_VOXEL:
                    oc = oracles.orc_cam(wl.landscape_cam(path, it.t))
                    orc.orc_landscape_frame(P(ref), it.W, it.H, P(maps.height), P(maps.color), P(maps.fog),
                                            C.byref(oc), nthr)
                    d = vox
                else:
                    g = wl.tunnel_grid(it.W, it.H, it.t)
                    b0, b1 = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
                    orc.orc_tunnel_frame(P(ref), P(b0), P(b1), it.W, it.H, P(g), P(tex), 9, it.radius, nthr)
                    d = tun
                d["requests"] += 1
                d["mismatched_px"] += int((ref != got).sum())
            headline_err = {
                "requests_checked": len(checked), "of": len(items),
                "checked": "every 320x180 and 640x360 request, and the first of each effect at each "
                           "larger size",
                "sdf_fp16_vs_fp64": dict(merge_reports(sdf_reps), requests=len(sdf_reps)),
                "voxel_vs_oracle": dict(vox, bit_exact=vox["mismatched_px"] == 0),
                "tunnel_vs_oracle": dict(tun, bit_exact=tun["mismatched_px"] == 0)}
            log(f"headline check in {time.time() - t_chk:.1f}s", headline_err)

            if "classic" in sections:
                # c0: plain FP32 and batched FP16 frames against the FP64
                # render, every 4th row of all 64 frames
                sdf_plain()
                sdf_batched()
                per["sdf_c0"]["error_vs_fp64"] = sdf_dense(fr_sdf, fr_sdf_p, scene8, W, H, 64)
                # c1 / c2 (integer paths): four frames each against the
                # oracle, which must be bit-exact
                vox_batched()
                got, want = [], []
                for k in [0, 17, 59, 119]:
                    buf = np.zeros(W * H, np.uint32)
                    oc = oracles.orc_cam(lcams[k])
                    orc.orc_landscape_frame(P(buf), W, H, P(maps.height), P(maps.color), P(maps.fog),
                                            C.byref(oc), nthr)
                    got.append(fr_vox.view(k * W * H, (W * H,)).numpy())
                    want.append(buf)
                per["voxel_c1"]["error_vs_oracle"] = dict(error_stats(np.concatenate(got), np.concatenate(want)),
                                                         frames=[0, 17, 59, 119])
                tun_batched()
                got, want = [], []
                for k in [0, 61, 142, 239]:
                    buf = np.zeros(W * H, np.uint32)
                    b0, b1 = np.zeros(W * H, np.uint32), np.zeros(W * H, np.uint32)
                    g = wl.tunnel_grid(W, H, k / 60)
                    orc.orc_tunnel_frame(P(buf), P(b0), P(b1), W, H, P(g), P(tex), 9, int(radii[k]), nthr)
                    got.append(fr_tun.view(k * W * H, (W * H,)).numpy())
                    want.append(buf)
                per["tunnel_c2"]["error_vs_oracle"] = dict(error_stats(np.concatenate(got), np.concatenate(want)),
                                                          frames=[0, 61, 142, 239])
                log("classic errors", {k: per[k].get("error_vs_fp64", per[k].get("error_vs_oracle"))
                                       for k in ("sdf_c0", "voxel_c1", "tunnel_c2")})

                # CPU baselines of c0-c2 (4 frames per run, median of 5 runs
                # after 3 warm-ups): the oracle's FP32 render of the SDF scene
                # (the scene has no reference counterpart); the reference's
                # own column loop of vscape_ray (code/landscape.cpp:170-191)
                # for the voxel frames; the reference's HorizontalBoxBlur32,
                # VerticalBoxBlur32 and BlitAdd32 around the oracle's grid
                # remap (no reference counterpart) for the tunnel frames
                sdf_buf = np.zeros(W * H, np.uint32)

                def cpu_c0():
                    for k in range(4):
                        orc.orc_sdf_frame(P(sdf_buf), W, H, C.byref(scene8), k / 60, 32, None, nthr)
                per["sdf_c0"]["cpu"] = dict(median_rate(cpu_c0, 4 * W * H), kind="port",
                                            what="oracle FP32 render (C, OpenMP rows)", threads=nthr)
                vbuf = oracles.aligned_empty(W * H)
                vset = []
                for k in range(4):
                    oc = oracles.orc_cam(lcams[k])
                    fx, fy = C.c_int(), C.c_int()
                    dX, dY = np.zeros(W, np.int32), np.zeros(W, np.int32)
                    fish = np.zeros(W, np.float32)
                    orc.orc_vscape_setup(W, H, C.byref(oc), C.byref(fx), C.byref(fy), P(dX), P(dY), P(fish))
                    vset.append((lcams[k].mapTilt, fx.value, fy.value, dX, dY, fish))

                if use_ref:
                    def cpu_c1():
                        for (tilt, fx, fy, dX, dY, fish) in vset:
                            refl.ref_vscape_columns(P(vbuf), P(maps.height), P(maps.color), P(maps.fog),
                                                    tilt, fx, fy, P(dX), P(dY), P(fish))
                    c1_kind, c1_what = "reference", "reference vscape_ray, OpenMP over columns (SSE4.1)"
                else:
                    def cpu_c1():
                        for k in range(4):
                            oc = oracles.orc_cam(lcams[k])
                            orc.orc_landscape_frame(P(vbuf), W, H, P(maps.height), P(maps.color),
                                                    P(maps.fog), C.byref(oc), nthr)
                    c1_kind, c1_what = "port", "oracle column caster (C, OpenMP columns)"
                per["voxel_c1"]["cpu"] = dict(median_rate(cpu_c1, 4 * W * H), kind=c1_kind, what=c1_what,
                                              threads=nthr)
                rm, t0b, t1b = (oracles.aligned_empty(W * H) for _ in range(3))
                tgrids_h = [wl.tunnel_grid(W, H, k / 60) for k in range(4)]

                if use_ref:
                    def cpu_c2():
                        for k in range(4):
                            s_ = np.float32((2.0 * int(radii[k]) + 1.5) / 255.0)
                            orc.orc_grid_remap(P(rm), W, H, P(tgrids_h[k]), P(tex), 9, nthr)
                            refl.ref_blur07_h(P(t0b), P(rm), W, H, s_)
                            refl.ref_blur07_v(P(t1b), P(t0b), W, H, s_)
                            refl.ref_blit_add32(P(rm), P(t1b), W, W, H)
                    c2_kind = "reference"
                    c2_what = ("reference HorizontalBoxBlur32 + VerticalBoxBlur32 + BlitAdd32 (SSE, OpenMP), "
                               "oracle grid remap (no reference counterpart)")
                else:
                    def cpu_c2():
                        for k in range(4):
                            orc.orc_tunnel_frame(P(rm), P(t0b), P(t1b), W, H, P(tgrids_h[k]), P(tex), 9,
                                                 int(radii[k]), nthr)
                    c2_kind, c2_what = "port", "oracle tunnel frame (C, OpenMP)"
                per["tunnel_c2"]["cpu"] = dict(median_rate(cpu_c2, 4 * W * H), kind=c2_kind, what=c2_what,
                                               threads=nthr)
                if use_ref:
                    # the reference-built tunnel frame is the oracle's, bit for bit
                    want = np.zeros(W * H, np.uint32)
                    orc.orc_tunnel_frame(P(want), P(np.zeros(W * H, np.uint32)), P(np.zeros(W * H, np.uint32)),
                                         W, H, P(tgrids_h[3]), P(tex), 9, int(radii[3]), nthr)
                    assert np.array_equal(want, rm), "reference tunnel frame differs from the oracle's"
                log("classic cpu", {k: per[k]["cpu"] for k in ("sdf_c0", "voxel_c1", "tunnel_c2")})

            if "c3" in sections:
                # precision form: FP16 batched and FP32 plain against the FP64
                # render, every 4th row of all 64 frames
                c3_plain()
                c3_batched()
                es = sdf_dense(fr3, fr3p, scene32, W3, H3, n3)
                log("precision", es)
                sdf3 = np.zeros(W3 * H3, np.uint32)

                def cpu_c3():
                    orc.orc_sdf_frame(P(sdf3), W3, H3, C.byref(scene32), 0.0, 32, None, nthr)
                per["sdf_c3_precision"]["cpu"] = dict(median_rate(cpu_c3, W3 * H3), kind="port",
                                                      what="oracle FP32 render (C, OpenMP rows)",
                                                      threads=nthr)

            # bounded sample of the batched workload, rendered by the oracle
            def cpu_render(it, nthreads):
                n = it.W * it.H
                buf = np.zeros(n, np.uint32)
                if it.effect == CKD_FX_SDF:
                    orc.orc_sdf_frame(P(buf), it.W, it.H, C.byref(scene8), it.t, 32, None, nthreads)
                elif it.effect == CKD_FX_VOXEL:
                    oc = oracles.orc_cam(wl.landscape_cam(path, it.t))
                    orc.orc_landscape_frame(P(buf), it.W, it.H, P(maps.height), P(maps.color), P(maps.fog),
                                            C.byref(oc), nthreads)
                else:
                    g = wl.tunnel_grid(it.W, it.H, it.t)
                    b0, b1 = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
                    orc.orc_tunnel_frame(P(buf), P(b0), P(b1), it.W, it.H, P(g), P(tex), 9, it.radius,
                                         nthreads)
                return n

            def first_requests(mpx):
                k, px = 0, 0
                while k < len(items) and px < mpx * 1e6:
                    px += items[k].W * items[k].H
                    k += 1
                return items[:k], px

            # the reference's Draw functions of the raymarched effects, on
            # the host, where the compiled reference harness is present (its
            # parameters come from Rocket track files written for the frames);
            # else the oracle's restatement of them
            import tempfile
            lut_h = np.zeros(2049, np.float32)
            orc.orc_cos_lut(P(lut_h))
            if use_ref and "shadertoy" in sections:
                refl.ref_st_init()
                refl.ref_st_textures(P(st_tex), P(st_tex_hi))
                refl.ref_st_blurmaps(P(st_blur), P(st_blur[st_blur.size // 2:]))
            if "shadertoy" in sections:
                for fx in st_per:
                    name = "spikey" if fx == "spikey_close" else fx
                    frames = wl.st_frames(fx, 8)
                    out_h = oracles.aligned_empty(W * H)
                    with tempfile.TemporaryDirectory() as td:
                        if use_ref:
                            cls = type(frames[0])
                            oracles.write_rocket_tracks(Path(td), {
                                tr: [getattr(f, fld) for f in frames] for fld, tr in cls.TRACKS.items()})
                            refl.ref_st_rocket(td.encode())

                        def st_cpu():
                            for k, f in enumerate(frames):
                                if use_ref:
                                    refl.ref_st_row(float(k))
                                    refl.ref_st_draw(oracles.ST_IDS[fx], P(out_h), f.time, 1.0)
                                else:
                                    orc.orc_st_draw(oracles.ST_IDS[fx], P(out_h), W, H, P(f.as_array()),
                                                    P(lut_h), P(st_tex), P(st_tex_hi), P(st_blur), 1, nthr)
                        r = median_rate(st_cpu, len(frames) * W * H, warmup=1, runs=3)
                    st_per[fx]["cpu_mpx_s"] = r["mpx_s"]
                    st_per[fx]["cpu_kind"] = (f"reference Draw (SSE, OpenMP, {nthr} threads)"
                                              if use_ref else f"oracle Draw ({nthr} threads)")
                    log("shadertoy cpu", fx, st_per[fx]["cpu_mpx_s"])

            if "voxel2" in sections:
                # the ball and the twister: the reference's Draw (tracks as above)
                # where the harness is present, else the oracle's
                m0_h = np.zeros(2 * W * H, np.int32)
                orc.orc_polar_maps(P(m0_h), P(np.zeros(2 * W * H, np.int32)), W, H, W, H)
                if use_ref:
                    refl.ref_shared_init()
                    refl.ref_polar_create()
                    refl.ref_polar_pad_forward()
                for key, (kind, frames) in v2_sets.items():
                    frames = frames[:8]
                    cls = type(frames[0])
                    out_h = oracles.aligned_empty(W * (H + 16))
                    tgt_h = np.zeros(W * H, np.uint32)
                    with tempfile.TemporaryDirectory() as td:
                        if use_ref:
                            oracles.write_rocket_tracks(Path(td), {
                                tr: [getattr(f, fld) for f in frames] for fld, tr in cls.TRACKS.items()})
                            refl.ref_st_rocket(td.encode())
                            if kind == "ball":
                                refl.ref_ball_setup(td.encode(), P(bm_v2["heightMaps"]), P(bm_v2["colorMaps"]),
                                                    P(bm_v2["beamMaps"]), P(bm_v2["envMap"]),
                                                    P(bm_v2["backgrounds"]), P(bm_v2["halo"]))
                            else:
                                refl.ref_twister_setup(td.encode(), P(tm_v2["heightMap"]), P(tm_v2["colorMap"]),
                                                       P(tm_v2["background"]))

                        def v2_cpu():
                            for k, f in enumerate(frames):
                                if use_ref:
                                    refl.ref_st_row(float(k))
                                    (refl.ref_ball_draw if kind == "ball" else refl.ref_twister_draw)(
                                        P(out_h), f.time, 1.0)
                                elif kind == "ball":
                                    orc.orc_ball_draw(P(out_h), P(tgt_h), W, H, P(bm_v2["heightMaps"]),
                                                      P(bm_v2["colorMaps"]), P(bm_v2["beamMaps"]),
                                                      P(bm_v2["envMap"]), P(bm_v2["backgrounds"]),
                                                      P(bm_v2["halo"]), P(m0_h), P(f.as_array()), nthr)
                                else:
                                    orc.orc_twister_draw(P(out_h), P(tgt_h), W, H, P(tm_v2["heightMap"]),
                                                         P(tm_v2["colorMap"]), P(tm_v2["background"]), P(m0_h),
                                                         P(f.as_array()), nthr)
                        r = median_rate(v2_cpu, len(frames) * W * H, warmup=1, runs=3)
                    v2_per[key]["cpu_mpx_s"] = r["mpx_s"]
                    v2_per[key]["cpu_kind"] = (f"reference Draw (SSE, OpenMP, {nthr} threads)"
                                               if use_ref else f"oracle Draw ({nthr} threads)")
                    log("voxel2 cpu", key, v2_per[key]["cpu_mpx_s"])

            # the headline's CPU baseline: the oracle renders the first
            # batch requests (about 100 Mpx) in order, median of 5 runs after
            # 3 warm-ups; the same on one thread over about 5 Mpx
            samp, samp_px = first_requests(100)
            r = median_rate(lambda: [cpu_render(it, nthr) for it in samp], samp_px)
            samp1, samp1_px = first_requests(5)
            r1 = median_rate(lambda: [cpu_render(it, 1) for it in samp1], samp1_px, warmup=1, runs=3)
            cpu = {"value": r["mpx_s"], "unit": "Mpixel/s", "cores": nthr, "kind": "port",
                   "sample": f"first {len(samp)} of the {len(items)} batch requests ({samp_px / 1e6:.1f} Mpx) "
                             f"rendered by the oracle (C, OpenMP) on {nthr} threads; median of 5 runs "
                             f"after 3 warm-ups",
                   "runs_s": r["runs_s"],
                   "value_1_thread": r1["mpx_s"],
                   "sample_1_thread": f"first {len(samp1)} requests ({samp1_px / 1e6:.1f} Mpx), median of 3 "
                                      f"runs after 1 warm-up",
                   "host": host}
            log("cpu", cpu)

    result = {
        "metric": "Mpixel/s per effect at <=2/255 err: batched FP16 vs plain FP32, 1 MI355X",
        "value": round(value, 1),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 march / fp16 shading; int (u8/u16/24:8) voxel, remap, blur",
        "data": "synthetic (seeded)",
        "config": {"workload": f"batched form: {len(items)} requests, 40/30/30 SDF/voxel/tunnel, "
                               f"5 resolutions 320x180..2560x1440, t U[0,60], "
                               f"{total_px / 1e6:.1f} Mpx per step, one hipGraph replay per step",
                   "march_steps_per_step": steps_total // max(1, args.steps),
                   "tuning": tuning},
        "per_effect": per,
        "error_vs_oracle": headline_err,
        "precision_c3": es,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(result))
    graph.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.path.insert(0, str(ROOT / "tests"))
    sys.exit(main())
