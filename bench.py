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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sdf_flops(total_steps: int, pixels: int, n_spheres: int) -> float:
    """Algorithmic FLOPs of the sphere tracer: per march step F_sdf+7
    (F_sdf = 11 per sphere + 2 for the plane), per pixel 6 SDF evaluations
    for the normal and ~40 for shading (FMA = 2, sqrt/exp/log = 1)."""
    f_sdf = 11 * n_spheres + 2
    return total_steps * (f_sdf + 7) + pixels * (6 * f_sdf + 40)


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
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
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
    reqs, off, grids = [], 0, []
    for it in items:
        g = ctx.upload(wl.tunnel_grid(it.W, it.H, it.t)) if it.effect == CKD_FX_TUNNEL else None
        grids.append(g)
        reqs.append(wl.make_request(it, out.ptr + 4 * off, scene8, path, g.ptr if g else None))
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

    # ------------------------------------------- roofline: persistent tracer
    # The kernel's own launches, bracketed by HIP events on the stream it
    # runs on: eager renders of the same batch with the same launch shapes
    # (a graph replay can only be timed as a whole, and its voxel / tunnel
    # branch runs beside the SDF kernel).
    ctx.kernel_timing(True)
    st_zero = int(pcount.numpy()[0])
    for _ in range(args.warmup + args.steps):
        ctx.render_batch(reqs, scene8, CKD_SHADE_FP16, d_tex, 9, tmp, tmp_px, pcount)
        psdf["px"] += sdf_px_batch
    ctx.sync()
    ms_k, launches = ctx.kernel_timing_read(1)  # CKD_KT_SDF_PERSISTENT
    ctx.kernel_timing(False)
    steps_k = int(pcount.numpy()[0]) - st_zero
    flops_k = sdf_flops(steps_k, psdf["px"], 8)
    achieved = flops_k / (ms_k * 1e-3) / 1e12
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
                               "F_sdf = 11*8+2 = 90 FLOP"}
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

    per, es, cpu = {}, None, None
    sections = set(args.sections.split(","))
    if not args.headline_only:
        # ----------------------------------------------------- per effect (c0-c2)
        W, H = 1280, 720
        if "classic" in sections:

            # SDF c0: 64 frames at t=k/60
            cams = [wl.sdf_camera(scene8, k / 60) for k in range(64)]
            fr_sdf = ctx.empty(64 * W * H)

            def sdf_plain():
                for k, c in enumerate(cams):
                    ctx.sdf_render(fr_sdf.view(k * W * H, (W * H,)), W, H, scene8, c, CKD_SHADE_FP32)

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
            fr_vox = ctx.empty(120 * W * H)

            def vox_plain():
                for k, c in enumerate(lcams):
                    ctx.landscape_draw(fr_vox.view(k * W * H, (W * H,)), W, H, [c])

            def vox_batched():
                ctx.landscape_draw(fr_vox, W, H, lcams)

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

            for name, frames, plain, batched in [("sdf_c0", 64, sdf_plain, sdf_batched),
                                                 ("voxel_c1", 120, vox_plain, vox_batched),
                                                 ("tunnel_c2", 240, tun_plain, tun_batched)]:
                plain()
                batched()
                mp = T.run(plain)
                st0 = int(pcount.numpy()[0])
                mb = T.run(batched)
                steps_b = int(pcount.numpy()[0]) - st0
                px = frames * W * H
                per[name] = {"frames": frames, "plain_fp32_mpx_s": round(px / mp / 1e3, 1),
                             "batched_fp16_mpx_s": round(px / mb / 1e3, 1),
                             "speedup": round(mp / mb, 3)}
                # the batched call against its bound, whole-call time
                if name == "sdf_c0":
                    ach = sdf_flops(steps_b, px, 8) / (mb * 1e-3) / 1e12
                    rl = {"bound": "valu", "achieved": round(ach, 2), "peak": PEAK_FP32_TFLOPS,
                          "unit": "TFLOP/s", "algorithmic": "sum(steps)*97 + px*580 FLOP"}
                elif name == "voxel_c1":
                    # 80 integer ops per column step, 10 per span pixel,
                    # counted as 2 FLOP each against the unpacked FP32 rate
                    ops = frames * W * 512 * 80 + px * 10
                    ach = 2 * ops / (mb * 1e-3) / 1e12
                    rl = {"bound": "valu", "achieved": round(ach, 2), "peak": PEAK_FP32_UNPACKED_TFLOPS,
                          "unit": "TFLOP/s (2 x int ops)",
                          "algorithmic": "512*80 ops per column + 10 per pixel"}
                else:
                    # remap 4 B/px written, blur H+V 8 B/px, BlitAdd32 12 B/px
                    ach = 24 * px / (mb * 1e-3) / 1e9
                    rl = {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "algorithmic": "24 B per pixel"}
                rl["frac"] = round(ach / rl["peak"], 4)
                per[name]["roofline"] = rl
                log(name, per[name])

        if "c3" in sections:
            # --------------------------------------- c3 precision form (2560x1440)
            W3, H3 = 2560, 1440
            n3 = 64
            fr3 = ctx.empty(n3 * W3 * H3)
            cams3 = [wl.sdf_camera(scene32, k / 60) for k in range(n3)]

            def c3_plain():
                for k, c in enumerate(cams3):
                    ctx.sdf_render(fr3.view(k * W3 * H3, (W3 * H3,)), W3, H3, scene32, c, CKD_SHADE_FP32)

            reqs3 = [wl.make_request(wl.BatchItem(CKD_FX_SDF, W3, H3, k / 60), fr3.ptr + 4 * k * W3 * H3,
                                     scene32, path, None) for k in range(n3)]
            reqs3 = ctx.request_array(reqs3)

            def c3_batched():
                ctx.render_batch(reqs3, scene32, CKD_SHADE_FP16, None, 9, None, 0)

            c3_plain()
            c3_batched()
            mp, mb = T.run(c3_plain), T.run(c3_batched)
            px3 = n3 * W3 * H3
            per["sdf_c3_precision"] = {"frames": n3, "plain_fp32_mpx_s": round(px3 / mp / 1e3, 1),
                                       "batched_fp16_mpx_s": round(px3 / mb / 1e3, 1),
                                       "speedup": round(mp / mb, 3)}
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

        # The CPU leg is the only part of the bench that runs the oracle: the
        # timed CPU baseline and the FP64 reference render the precision form's
        # error is measured against.
        if not args.no_cpu:
            import oracles  # noqa: E402  (the checker)
            from helpers import error_stats  # noqa: E402
            orc = oracles.oracle()
            nthr = min(16, os.cpu_count() or 1)
            if "classic" in sections:
                # c0: the batched FP16 frames against the FP64 render on a row
                # sample; c1 / c2 (integer paths): first and last frame against
                # the oracle, which must be bit-exact
                sdf_batched()
                errs = []
                for k in [0, 63]:
                    img = fr_sdf.view(k * W * H, (W * H,)).numpy().reshape(H, W)
                    rows = np.arange(5, H, 24)
                    ref = np.zeros(len(rows) * W, np.uint32)
                    orc.orc_sdf_rows(oracles.P(ref), W, H, C.byref(scene8), k / 60, 64, 5, 24, None, nthr)
                    errs.append((img[rows].reshape(-1), ref))
                per["sdf_c0"]["error_vs_fp64"] = error_stats(np.concatenate([e[0] for e in errs]),
                                                             np.concatenate([e[1] for e in errs]))
                vox_batched()
                got, want = [], []
                for k in [0, 119]:
                    buf = np.zeros(W * H, np.uint32)
                    oc = oracles.orc_cam(lcams[k])
                    orc.orc_landscape_frame(oracles.P(buf), W, H, oracles.P(maps.height),
                                            oracles.P(maps.color), oracles.P(maps.fog), C.byref(oc), nthr)
                    got.append(fr_vox.view(k * W * H, (W * H,)).numpy())
                    want.append(buf)
                per["voxel_c1"]["error_vs_oracle"] = error_stats(np.concatenate(got), np.concatenate(want))
                tun_batched()
                got, want = [], []
                for k in [0, 239]:
                    buf = np.zeros(W * H, np.uint32)
                    b0, b1 = np.zeros(W * H, np.uint32), np.zeros(W * H, np.uint32)
                    g = wl.tunnel_grid(W, H, k / 60)
                    orc.orc_tunnel_frame(oracles.P(buf), oracles.P(b0), oracles.P(b1), W, H, oracles.P(g),
                                         oracles.P(tex), 9, int(radii[k]), nthr)
                    got.append(fr_tun.view(k * W * H, (W * H,)).numpy())
                    want.append(buf)
                per["tunnel_c2"]["error_vs_oracle"] = error_stats(np.concatenate(got), np.concatenate(want))
                log("classic errors", {k: per[k].get("error_vs_fp64", per[k].get("error_vs_oracle"))
                                       for k in ("sdf_c0", "voxel_c1", "tunnel_c2")})
            if "c3" in sections:
                # precision form: FP16 shading vs the FP64 render on a row sample
                errs = []
                c3_batched()
                for k in [0, 63]:
                    img = fr3.view(k * W3 * H3, (W3 * H3,)).numpy().reshape(H3, W3)
                    rows = np.arange(7, H3, 36)
                    ref = np.zeros(len(rows) * W3, np.uint32)
                    orc.orc_sdf_rows(oracles.P(ref), W3, H3, C.byref(scene32), k / 60, 64, 7, 36, None, nthr)
                    errs.append((img[rows].reshape(-1), ref))
                es = error_stats(np.concatenate([e[0] for e in errs]), np.concatenate([e[1] for e in errs]))
                log("precision", es)
            # bounded sample of the batched workload: its requests in order, as
            # many as fit in about 10 s of CPU time, rendered by the oracle
            def cpu_render(it, nthreads):
                n = it.W * it.H
                buf = np.zeros(n, np.uint32)
                if it.effect == CKD_FX_SDF:
                    orc.orc_sdf_frame(oracles.P(buf), it.W, it.H, C.byref(scene8), it.t, 32, None, nthreads)
                elif it.effect == CKD_FX_VOXEL:
                    oc = oracles.orc_cam(wl.landscape_cam(path, it.t))
                    orc.orc_landscape_frame(oracles.P(buf), it.W, it.H, oracles.P(maps.height),
                                            oracles.P(maps.color), oracles.P(maps.fog), C.byref(oc),
                                            nthreads)
                else:
                    g = wl.tunnel_grid(it.W, it.H, it.t)
                    b0, b1 = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
                    orc.orc_tunnel_frame(oracles.P(buf), oracles.P(b0), oracles.P(b1), it.W, it.H,
                                         oracles.P(g), oracles.P(tex), 9, it.radius, nthreads)
                return n

            def cpu_rate(nthreads, budget_s):
                px, k, t_c = 0, 0, time.perf_counter()
                while k < len(items) and time.perf_counter() - t_c < budget_s:
                    px += cpu_render(items[k], nthreads)
                    k += 1
                return px, k, time.perf_counter() - t_c

            # the reference's own Draw functions of the raymarched effects, on
            # the host, where the compiled reference harness is present (its
            # parameters come from Rocket track files written for the frames);
            # else the oracle's restatement of them
            import tempfile
            lut_h = np.zeros(2049, np.float32)
            orc.orc_cos_lut(oracles.P(lut_h))
            use_ref = oracles.ref_available()
            refl = oracles.ref() if use_ref else None
            if use_ref and "shadertoy" in sections:
                refl.ref_st_init()
                refl.ref_st_textures(oracles.P(st_tex), oracles.P(st_tex_hi))
                refl.ref_st_blurmaps(oracles.P(st_blur), oracles.P(st_blur[st_blur.size // 2:]))
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
                        t_c = time.perf_counter()
                        for k, f in enumerate(frames):
                            if use_ref:
                                refl.ref_st_row(float(k))
                                refl.ref_st_draw(oracles.ST_IDS[fx], oracles.P(out_h), f.time, 1.0)
                            else:
                                pa = f.as_array()
                                orc.orc_st_draw(oracles.ST_IDS[fx], oracles.P(out_h), W, H, oracles.P(pa),
                                                oracles.P(lut_h), oracles.P(st_tex), oracles.P(st_tex_hi),
                                                oracles.P(st_blur), 1, nthr)
                        dt_c = time.perf_counter() - t_c
                    st_per[fx]["cpu_mpx_s"] = round(8 * W * H / dt_c / 1e6, 2)
                    st_per[fx]["cpu_kind"] = ("reference Draw (OpenMP, OMP_NUM_THREADS threads)"
                                              if use_ref else f"oracle Draw ({nthr} threads)")
                    log("shadertoy cpu", fx, st_per[fx]["cpu_mpx_s"])

            if "voxel2" in sections:
                # the ball and the twister: the reference's Draw (tracks as above)
                # where the harness is present, else the oracle's
                m0_h = np.zeros(2 * W * H, np.int32)
                orc.orc_polar_maps(oracles.P(m0_h), oracles.P(np.zeros(2 * W * H, np.int32)), W, H, W, H)
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
                                refl.ref_ball_setup(td.encode(), oracles.P(bm_v2["heightMaps"]),
                                                    oracles.P(bm_v2["colorMaps"]), oracles.P(bm_v2["beamMaps"]),
                                                    oracles.P(bm_v2["envMap"]), oracles.P(bm_v2["backgrounds"]),
                                                    oracles.P(bm_v2["halo"]))
                            else:
                                refl.ref_twister_setup(td.encode(), oracles.P(tm_v2["heightMap"]),
                                                       oracles.P(tm_v2["colorMap"]),
                                                       oracles.P(tm_v2["background"]))
                        t_c = time.perf_counter()
                        for k, f in enumerate(frames):
                            if use_ref:
                                refl.ref_st_row(float(k))
                                (refl.ref_ball_draw if kind == "ball" else refl.ref_twister_draw)(
                                    oracles.P(out_h), f.time, 1.0)
                            elif kind == "ball":
                                orc.orc_ball_draw(oracles.P(out_h), oracles.P(tgt_h), W, H,
                                                  oracles.P(bm_v2["heightMaps"]), oracles.P(bm_v2["colorMaps"]),
                                                  oracles.P(bm_v2["beamMaps"]), oracles.P(bm_v2["envMap"]),
                                                  oracles.P(bm_v2["backgrounds"]), oracles.P(bm_v2["halo"]),
                                                  oracles.P(m0_h), oracles.P(f.as_array()), nthr)
                            else:
                                orc.orc_twister_draw(oracles.P(out_h), oracles.P(tgt_h), W, H,
                                                     oracles.P(tm_v2["heightMap"]), oracles.P(tm_v2["colorMap"]),
                                                     oracles.P(tm_v2["background"]), oracles.P(m0_h),
                                                     oracles.P(f.as_array()), nthr)
                        dt_c = time.perf_counter() - t_c
                    v2_per[key]["cpu_mpx_s"] = round(len(frames) * W * H / dt_c / 1e6, 2)
                    v2_per[key]["cpu_kind"] = ("reference Draw (OpenMP, OMP_NUM_THREADS threads)"
                                               if use_ref else f"oracle Draw ({nthr} threads)")
                    log("voxel2 cpu", key, v2_per[key]["cpu_mpx_s"])

            px, k, dt = cpu_rate(nthr, 10.0)
            px1, k1, dt1 = cpu_rate(1, 5.0)
            cpu = {"value": round(px / dt / 1e6, 2), "unit": "Mpixel/s", "cores": nthr, "kind": "port",
                   "sample": f"first {k} of the {len(items)} batch requests ({px / 1e6:.1f} Mpx) "
                             f"rendered by the oracle (C, OpenMP) in {dt:.1f}s",
                   "value_1_thread": round(px1 / dt1 / 1e6, 2),
                   "sample_1_thread": f"first {k1} requests in {dt1:.1f}s"}
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
        "precision_c3": None if es is None else
        {"shading": "fp16", "vs": "fp64 CPU render", "rows_sampled": "every 36th, frames 0 and 63", **es},
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
