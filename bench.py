#!/usr/bin/env python3
"""Throughput of the MI355X neo_fft3d path on its headline workload.

Workload: YUV420P8 1920x1080 (chroma 960x540), 300 seeded frames of moving
fBm-textured rectangles over a gradient with Gaussian noise (sigma 3 luma,
2 chroma), filtered with bt=3, sigma=2, bw=bh=48, ow=oh=16, sharpen=0.3, all
three planes.  The clip is resident in HBM before timing starts; one step is
one pass of the filter over all 300 frames (three per-plane engines, frames in
order, the spectra ring in HBM; the three plane engines share their
frame-loop launches through nf3d_run_device_batch).  Prints one JSON line:

  value         frames/s over the timed steps (whole job, one GPU)
  roofline      HIP-event time of the main spectral kernel against the HBM
                peak, with the algorithmic bytes of that kernel
  cpu_baseline  the FP64 C restatement of the path (oracle/) on a bounded
                sample of the same workload on this host's cores
  psnr_db       denoised vs clean (noise-free render of the same seed)
  precision_form  the DFT-as-GEMM form (FP16 split operands on the matrix
                cores) on 3840x2160 32/16 blocks: rate and error vs FP64
  multi_clip    16 concurrent clips of mixed shape, bt and block size, one
                batched call per frame step: aggregate fps, step latency
  learned_psd   the learned-PSD form on RGBPS float planes, bt=3, 48/16 with
                coloured, spatially varying noise: network trained on the GPU,
                PSNR of plain (best global sigma) vs learned, frames/s of both
                at 1920x1080, MLP inference rate

``--gpus N`` with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fps 1080p YUV420P8 bt=3 (plain vs learned-PSD/MFMA) + PSNR dB, % HBM roofline"
HBM_PEAK_GBS = 8000.0
W, H = 1920, 1080
FILTER = dict(sigma=2.0, bw=48, bh=48, ow=16, oh=16, sharpen=0.3)
REF_ARGS = "sigma=2;bw=48;bh=48;ow=16;oh=16;sharpen=0.3;planes=0,1,2"


def plane_geometry(w, h, bw, bh, ow, oh):
    nox = (w - ow + (bw - ow - 1)) // (bw - ow) + 2
    noy = (h - oh + (bh - oh - 1)) // (bh - oh) + 2
    return nox, noy


def frame_plane_bytes(pw, ph, bt, bps=1, bw=48, bh=48, ow=16, oh=16):
    """Algorithmic HBM bytes of filtering one plane of one frame:
    2*W*H*Bps (plane in, plane out) + r(bt)*S, with S = nb*bh*(bw/2+1)*8 the
    compact spectrum of the plane and r(bt) = bt for bt >= 2 (the newest
    frame's spectrum written once, bt-1 older ones read back from the ring),
    6 for Kalman (three state arrays read and written), 0 for bt = 1 / -1."""
    nox, noy = plane_geometry(pw, ph, bw, bh, ow, oh)
    nb = nox * noy
    spec = nb * bh * (bw // 2 + 1) * 8
    r = bt if bt >= 2 else (6 if bt == 0 else 0)
    return 2 * pw * ph * bps + r * spec


TRAFFIC_PROFILE = "profiles/k_run_yuv420p8_1080p_bt3_hbm_traffic.md"


def measured_traffic():
    """HBM bytes per frame of k_run + k_seam_run from the committed rocprofv3
    counter summary (a text file: `key: value` lines)."""
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    if not os.path.exists(path):
        return None
    out = {}
    for line in open(path):
        if line.startswith(("traffic_bytes_per_frame:", "algorithmic_bytes_per_frame:")):
            k, v = line.split(":")
            out[k.strip()] = float(v)
    return out if len(out) == 2 else None


def psnr(a, b, peak=255.0):
    import numpy as np
    d = a.astype(np.float64) - b.astype(np.float64)
    mse = float(np.mean(d * d))
    return 99.0 if mse == 0 else 10 * np.log10(peak * peak / mse)


def gpu_run(args):
    import numpy as np
    import torch

    from neo_fft3d_amd import Engine, default_params, make_format
    from neo_fft3d_amd.synth import SEED_BASE, rects_clip

    F = args.frames
    t = time.perf_counter()
    planes = rects_clip(W, H, F, SEED_BASE + 1)
    gen_s = time.perf_counter() - t
    dev = torch.device("cuda", 0)
    d_src = [torch.from_numpy(p).to(dev) for p in planes]
    d_dst = [torch.empty_like(x) for x in d_src]
    torch.cuda.synchronize()
    fmt = make_format("yuv", 8)
    prm = default_params(**FILTER, bt=args.bt)
    engines = [Engine(prm, fmt, W, H, F, p) for p in range(3)]
    src_ptrs = [[x[n].data_ptr() for n in range(F)] for x in d_src]
    dst_ptrs = [[x[n].data_ptr() for n in range(F)] for x in d_dst]
    strides = [x.stride(1) for x in d_src]

    from neo_fft3d_amd import run_device_batch
    items = [(engines[p], 0, F, src_ptrs[p], strides[p], dst_ptrs[p], strides[p]) for p in range(3)]

    def step():
        # the three plane engines in one batched call: every frame-loop launch
        # covers the tiles of all three planes (one stream)
        run_device_batch(items, 0)

    def sync():
        for e in engines:
            e.sync()
        torch.cuda.synchronize()

    if args.copy_calibration:
        # a plain device copy of known size, for scaling HBM traffic counters
        nbytes = args.copy_calibration << 20
        a = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    for _ in range(args.warmup):
        step()
    sync()
    for e in engines:
        e.profile(True)
        e.profile_read()
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    sync()
    t1 = time.perf_counter()
    prof = [e.profile_read() for e in engines]
    for e in engines:
        e.profile(False)
    elapsed = t1 - t0
    fps = F * args.steps / elapsed

    # roofline of the filtering kernels (HIP events on each engine's stream):
    # the frame-loop launches (k_run + k_seam_run, 16 frames each) and the
    # one-frame 2D launches of the clip ends; the engines report the frames
    # their timed launches filtered
    pdims = [(W, H), (W // 2, H // 2), (W // 2, H // 2)]
    bytes_total = 0.0
    frames_timed = 0
    edges = args.steps * (args.bt // 2 + (args.bt - 1) // 2) if args.bt >= 2 else 0
    for (pw, ph), (ms, n) in zip(pdims, prof):
        # the clip-end frames of each pass (bt // 2 + (bt - 1) // 2) run the
        # 2D gains: pixels in and out, no spectra ring
        bytes_total += frame_plane_bytes(pw, ph, args.bt) * (n - edges) + frame_plane_bytes(pw, ph, 1) * edges
        frames_timed += n
    # the planes share their launches: each engine's timed intervals are the
    # same shared launches
    ms_total = max(ms for ms, _ in prof)
    achieved = bytes_total / (ms_total * 1e-3) / 1e9 if ms_total > 0 else 0.0
    avg_us = ms_total * 1e3 / max(frames_timed, 1)

    # a plain device copy (1 GiB read + 1 GiB written) on the same box: the
    # HBM rate this box reaches, and the fraction of it the filter reaches
    cp = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
    cq = torch.empty_like(cp)
    cq.copy_(cp)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        cq.copy_(cp)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * 5 * (1 << 30) / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del cp, cq

    # host-fed: the same frames from (pageable) host memory through the
    # pinned, chunked upload / filter / download pipeline (PCIe-inclusive)
    host = None if args.no_host_fed else host_fed(planes, args)

    # quality: denoised vs a noise-free render of the same clip
    idx = [F // 2 + k for k in range(4)]
    clean = rects_clip(W, H, F, SEED_BASE + 1, noise_luma=0.0, noise_chroma=0.0, threads=16)
    out_y = d_dst[0].cpu().numpy()
    q = {
        "noisy_luma": round(float(np.mean([psnr(planes[0][n], clean[0][n]) for n in idx])), 3),
        "denoised_luma": round(float(np.mean([psnr(out_y[n], clean[0][n]) for n in idx])), 3),
    }
    # measured HBM traffic of the same launches (rocprofv3 FETCH_SIZE /
    # WRITE_SIZE with a copy calibration; summary and commands in profiles/)
    traffic = traffic_ratio = None
    if args.bt == 3:
        t = measured_traffic()
        if t:
            traffic = int(t["traffic_bytes_per_frame"] * 16)  # one full 16-frame launch
            traffic_ratio = round(t["traffic_bytes_per_frame"] / t["algorithmic_bytes_per_frame"], 3)
    return {
        "fps": fps, "elapsed": elapsed, "ms_per_step": elapsed * 1e3 / args.steps, "gen_s": gen_s,
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_over_algorithmic": traffic_ratio,
            "traffic_source": TRAFFIC_PROFILE + " (FETCH_SIZE x2.00, WRITE_SIZE x1.00 by a 1 GiB copy)",
            "kernel": f"k_run<48,48,MODE_W3D,bt={args.bt},sharpen> + k_seam_run (3 planes per launch)",
            "copy_gbs": round(copy_gbs, 1), "frac_of_copy": round(achieved / copy_gbs, 4),
            "us_per_frame_plane": round(avg_us, 2), "frame_planes": frames_timed,
            "bytes_model": "per frame-plane: 2*W*H*Bps + bt*nb*bh*(bw/2+1)*8",
        },
        "psnr_db": q,
        "host_fed": host,
    }


def host_fed(planes, args):
    """Frames/s of the whole clip handed over in host memory and returned to
    host memory by nf3d_run_host_batch (PCIe-inclusive), from pageable numpy
    planes (copied into pinned staging by host threads on the way) and from
    page-locked planes (uploaded and returned by DMA directly): chunks
    uploaded, filtered and downloaded on three streams."""
    import numpy as np
    import torch

    from neo_fft3d_amd import Engine, default_params, make_format, run_host_batch
    F = args.frames
    fmt = make_format("yuv", 8)
    prm = default_params(**FILTER, bt=args.bt)

    def rate(src_ptrs, src_strides, dst_ptrs, dst_strides):
        engines = [Engine(prm, fmt, W, H, F, p) for p in range(3)]
        items = [(engines[p], 0, F, src_ptrs[p], src_strides[p], dst_ptrs[p], dst_strides[p]) for p in range(3)]
        run_host_batch(items)  # warm-up (device slots and staging, first transforms)
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            run_host_batch(items)
        dt = time.perf_counter() - t0
        for e in engines:
            e.close()
        return round(F * reps / dt, 1)
    outs = [np.empty_like(x) for x in planes]
    pageable = rate([[x[n].ctypes.data for n in range(F)] for x in planes], [x.strides[1] for x in planes],
                    [[o[n].ctypes.data for n in range(F)] for o in outs], [o.strides[1] for o in outs])
    pin_in = [torch.from_numpy(x).pin_memory() for x in planes]
    pin_out = [torch.empty_like(x).pin_memory() for x in pin_in]
    pinned = rate([[x[n].data_ptr() for n in range(F)] for x in pin_in], [x.stride(1) for x in pin_in],
                  [[o[n].data_ptr() for n in range(F)] for o in pin_out], [o.stride(1) for o in pin_out])
    mb = sum(x[0].nbytes for x in planes) / 1e6
    return {"fps_pageable": pageable, "fps_pinned": pinned, "unit": "frames/s",
            "note": f"PCIe-inclusive: {mb:.2f} MB in and out per frame; {F} frames x 3 passes per form"}


def precision_form(args):
    """DFT-as-GEMM precision form (nf3d_dft2d_mfma): 32x32 block DFTs of a
    3840x2160 plane's block count (32/16 blocks), forward then inverse, on the
    matrix cores with FP16 hi/lo split operands; error against float64."""
    import numpy as np
    import torch

    from neo_fft3d_amd import lib
    L = lib()
    L.nf3d_dft2d_mfma.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_float, ctypes.c_void_p]
    nox, noy = plane_geometry(3840, 2160, 32, 32, 16, 16)
    nb = nox * noy
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 1023, size=(nb, 32, 32)).astype(np.float32)
    d_x = torch.from_numpy(x).cuda()
    d_z = torch.empty((nb, 32, 17, 2), dtype=torch.float32, device="cuda")
    d_y = torch.empty_like(d_x)

    def once():
        assert L.nf3d_dft2d_mfma(ctypes.c_void_p(d_x.data_ptr()), ctypes.c_void_p(d_z.data_ptr()), nb, 32, 0,
                                 ctypes.c_float(1.0), None) == 0
        assert L.nf3d_dft2d_mfma(ctypes.c_void_p(d_z.data_ptr()), ctypes.c_void_p(d_y.data_ptr()), nb, 32, 1,
                                 ctypes.c_float(1.0 / 32), None) == 0
    once()
    torch.cuda.synchronize()
    reps = 20
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        once()
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    sub = 256
    want = np.fft.rfft2(x[:sub].astype(np.float64))
    z = d_z[:sub].cpu().numpy()
    got = z[..., 0] + 1j * z[..., 1]
    fwd_err = float(np.abs(got - want).max() / np.abs(want).max())
    back = d_y[:sub].cpu().numpy() / 1024.0
    rt_err = float(np.abs(back - x[:sub]).max())
    # FP16 MACs per block and direction: 3 split products x (32*64*32 + 2*32*64*32)
    flops = 2 * 2 * 3 * (32 * 64 * 32 + 2 * 32 * 64 * 32) * nb
    tflops = flops / (ms * 1e-3) / 1e12
    filt = precision_filter(args)
    return {"filter": filt, "blocks": nb, "ms_forward_plus_inverse": round(ms, 4), "blocks_per_s": round(nb / (ms * 1e-3)),
            "mfma_tflops_fp16": round(tflops, 1), "mfma_peak_tflops_fp16": 2500.0,
            "frac": round(tflops / 2500.0, 4), "forward_max_err_rel_fp64": fwd_err,
            "round_trip_max_err_10bit": rt_err}


def precision_filter(args):
    """The precision workload's matrix on YUV420P10 3840x2160 (sigma 3 scaled
    to 12 in 10-bit units, noise 12 / 8): bt = 0 (Kalman, kratio 2) and bt = 3,
    each at 32/16, 48/16 and 64/16, all planes: frames/s of the FFT form and of
    the DFT-as-GEMM form (dft=1, matrix cores; the frame loop at 32x32, the
    per-frame kernel at 48x48 and 64x64) on a resident clip, and how often the two
    forms' outputs agree.  (Each form's per-plane error
    against FP64 is measured by the tests.)"""
    import numpy as np
    import torch

    from neo_fft3d_amd import Engine, default_params, lib, make_format, run_device_batch
    from neo_fft3d_amd.synth import SEED_BASE, rects_clip
    L = lib()
    L.nf3d_engine_set_dft.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
    F, w, h = args.precision_frames, 3840, 2160
    planes = rects_clip(w, h, F, SEED_BASE + 3, bits=10, noise_luma=12.0, noise_chroma=8.0)
    src = [torch.from_numpy(p).cuda() for p in planes]
    fmt = make_format("yuv", 10)

    def run(bt, bw, dft):
        kw = dict(bt=bt, bw=bw, bh=bw, ow=16, oh=16)
        if bt == 0:
            kw["kratio"] = 2.0
        es = [Engine(default_params(3.0, **kw), fmt, w, h, F, p) for p in range(3)]
        if dft:
            for e in es:
                assert L.nf3d_engine_set_dft(e.h, 1, None, 0) == 0
        dst = [torch.zeros_like(x) for x in src]
        items = [(e, 0, F, [src[p][k].data_ptr() for k in range(F)], src[p][0].stride(0) * 2,
                  [dst[p][k].data_ptr() for k in range(F)], src[p][0].stride(0) * 2) for p, e in enumerate(es)]
        torch.cuda.synchronize()
        run_device_batch(items)  # warm-up pass (the Kalman state then carries on: same work per pass)
        for e in es:
            e.sync()
        out = [d.cpu().numpy() for d in dst]
        t = time.perf_counter()
        run_device_batch(items)
        for e in es:
            e.sync()
        fps = F / (time.perf_counter() - t)
        for e in es:
            e.close()
        return fps, out
    res = []
    for bt in (0, 3):
        for bw in (32, 48, 64):
            fps, out_fft = run(bt, bw, 0)
            row = {"bt": bt, "bw": bw, "ow": 16, "fps_fft": round(fps, 1)}
            fps_m, out_m = run(bt, bw, 1)
            d = [np.abs(a.astype(np.int32) - b.astype(np.int32)) for a, b in zip(out_fft, out_m)]
            row.update({"fps_mfma_dft": round(fps_m, 1), "mfma_over_fft": round(fps_m / fps, 3),
                        "mfma_kernel": "frame loop" if bw in (32, 48) else "per-frame",
                        "forms_max_diff_lsb": int(max(x.max() for x in d)),
                        "forms_equal_frac": round(float(np.mean([np.mean(x == 0) for x in d])), 6)})
            res.append(row)
    return {"workload": "yuv420p10_3840x2160", "frames": F, "runs": res}


def multi_clip(args):
    """Batched multi-clip serving: 16 concurrent clips of mixed shape -- 4x
    854x480 YUV420P8, 4x 1280x720 YUV420P8, 4x 1920x1080 YUV420P10, 4x
    3840x2160 YUV444P16 -- with bt cycling over {0, 1, 3, 5}, bw = bh over
    {32, 48, 64}, ow = oh the even number nearest bw/3 (10, 16, 22), sigma
    seeded in [1.5, 4].  One frame step = one nf3d_run_device_batch call with
    the next frame of every clip (all planes; engines of the same block shape,
    bt and sharpen case share frame-loop launches).  The learned PSD is on for
    8 clips (the five 48x48 clips and the three 8-bit 32x32 ones), each with a
    network trained on the GPU for its block shape, overlap and sample scale
    (strength 3).  Reports aggregate fps and the per-step latency."""
    import numpy as np
    import torch

    from neo_fft3d_amd import Engine, default_params, make_format, run_device_batch
    from neo_fft3d_amd import learned as LRN
    from neo_fft3d_amd.synth import SEED_BASE, rects_clip
    F = args.multi_frames
    models = {}
    t_train = time.perf_counter()
    shapes = [(854, 480, 8, 1)] * 4 + [(1280, 720, 8, 1)] * 4 + [(1920, 1080, 10, 1)] * 4 + [(3840, 2160, 16, 0)] * 4
    rng = np.random.default_rng(SEED_BASE + 5)
    clips = []
    for i, (w, h, bits, sub) in enumerate(shapes):
        bt = (0, 1, 3, 5)[i % 4]
        bw = (32, 48, 64)[i % 3]
        ow = {32: 10, 48: 16, 64: 22}[bw]
        sigma = float(rng.uniform(1.5, 4.0))
        planes = rects_clip(w, h, F, SEED_BASE + 100 + i, ssw=sub, ssh=sub, bits=bits)
        src = [torch.from_numpy(p).cuda() for p in planes]
        dst = [torch.zeros_like(x) for x in src]
        fmt = make_format("yuv", bits, ssw=sub, ssh=sub)
        es = [Engine(default_params(sigma, bt=bt, bw=bw, bh=bw, ow=ow, oh=ow), fmt, w, h, F, p) for p in range(3)]
        learned = bw == 48 or (bw == 32 and bits == 8)
        if learned:
            key = (bw, ow, bits)
            if key not in models:
                models[key] = LRN.train_model(clips=16, frames=2, width=256, height=256, steps=600, bw=bw, ow=ow,
                                              scale=float((1 << bits) - 1), seed=0x4C50 + bw + bits)[0]
            for e in es:
                LRN.set_learned(e, models[key], 3.0)
        clips.append(dict(src=src, dst=dst, es=es, w=w, h=h, bits=bits, bt=bt, bw=bw, sigma=round(sigma, 3),
                          learned=learned))
    t_train = time.perf_counter() - t_train
    torch.cuda.synchronize()

    # each plane's frame tables, built once (a serving loop keeps them)
    tables = []
    for c in clips:
        for p, e in enumerate(c["es"]):
            x, y = c["src"][p], c["dst"][p]
            st = x[0].stride(0) * x.element_size()
            tables.append((e, [x[n].data_ptr() for n in range(F)], st, [y[n].data_ptr() for n in range(F)]))

    def items_for(k):
        return [(e, k, 1, sp, st, dp, st) for e, sp, st, dp in tables]
    bstream = torch.cuda.Stream()

    def step(k):
        # the batch call's work ends on the stream it is given: one wait
        run_device_batch(items_for(k), bstream.cuda_stream)
        bstream.synchronize()
    warm = min(4, F // 4)
    for k in range(warm):
        step(k)
    lat = []
    t0 = time.perf_counter()
    for k in range(warm, F):
        t = time.perf_counter()
        step(k)
        lat.append((time.perf_counter() - t) * 1e3)
    total = time.perf_counter() - t0
    steps = F - warm
    for c in clips:
        for e in c["es"]:
            e.close()
    graph = serving_graphs(clips, F, models)
    graph_ring = serving_graphs(clips, F, models, ring=True)
    graph_ring_period = serving_graphs(clips, F, models, ring=True, period=15)
    return {"workload": "16 clips: 4x 854x480 P8, 4x 1280x720 P8, 4x 1920x1080 P10, 4x 3840x2160 YUV444P16; "
                        "bt {0,1,3,5}, bw {32,48,64}, ow 10/16/22, sigma U[1.5,4]",
            "clips": [{k: c[k] for k in ("w", "h", "bits", "bt", "bw", "sigma", "learned")} for c in clips],
            "frame_steps": steps, "fps_aggregate": round(len(clips) * steps / total, 1),
            "step_latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
            "step_latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
            "graph_replay": graph, "graph_replay_ring": graph_ring,
            "graph_replay_ring_learned_period_15": graph_ring_period,
            "learned_psd": f"on for {sum(c['learned'] for c in clips)} clips, {len(models)} networks "
                           f"(setup incl. data and training {t_train:.1f} s)"}


def serving_graphs(clips, F, models, ring=False, period=0):
    """The same 16 clips served as a stream: frame j of each plane arrives in
    staging slot j % S (S = 15, a multiple of every bt) and is filtered into
    output slot j % S, so a frame step's work repeats every S steps.  After
    three eager steps, the batched call of each phase k mod S is captured as
    one HIP graph (nf3d_run_device_batch_graph; with ring=True
    nf3d_run_device_batch_graph_ring, whose steps read the spectra the
    previous step left in the rings) and replayed once to advance the
    spectra ring; the timed steps are the frame's arrival (a device copy per
    plane, issued as one multi-tensor copy) plus one graph launch per phase.
    Fresh engines, same parameters and networks; `period` > 0 predicts the
    learned PSD once per that many frames (nf3d_engine_set_learned_period)."""
    import numpy as np
    import torch

    from neo_fft3d_amd import Engine, Graph, default_params, make_format, run_device_batch
    from neo_fft3d_amd import learned as LRN
    S, warm = 15, 3
    gcl = []
    for c in clips:
        sub = 1 if c["src"][1].shape[1] != c["src"][0].shape[1] else 0
        fmt = make_format("yuv", c["bits"], ssw=sub, ssh=sub)
        ow = {32: 10, 48: 16, 64: 22}[c["bw"]]
        es = [Engine(default_params(c["sigma"], bt=c["bt"], bw=c["bw"], bh=c["bw"], ow=ow, oh=ow), fmt, c["w"],
                     c["h"], F, p) for p in range(3)]
        if c["learned"]:
            for e in es:
                LRN.set_learned(e, models[(c["bw"], ow, c["bits"])], 3.0, period=period)
        stg = [torch.zeros((S,) + tuple(x.shape[1:]), dtype=x.dtype, device="cuda") for x in c["src"]]
        out = [torch.zeros_like(x) for x in stg]
        gcl.append(dict(src=c["src"], stg=stg, out=out, es=es))

    def arrive(j):
        torch._foreach_copy_([g["stg"][p][j % S] for g in gcl for p in range(3)],
                             [g["src"][p][j] for g in gcl for p in range(3)])

    def items(k):
        res = []
        for g in gcl:
            for p, e in enumerate(g["es"]):
                st = g["stg"][p][0].stride(0) * g["stg"][p].element_size()
                res.append((e, k, 1, [g["stg"][p][j % S].data_ptr() for j in range(F)], st,
                            [g["out"][p][j % S].data_ptr() for j in range(F)], st))
        return res
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    graphs = {}
    lat = []
    with torch.cuda.stream(stream):
        for j in range(3):
            arrive(j)
        for k in range(warm + S):
            if k > 0:
                arrive(k + 2)
            if k < warm:
                run_device_batch(items(k), sp)
            else:
                graphs[k % S] = Graph(items(k), sp, ring=ring)
                graphs[k % S].launch(sp)
            stream.synchronize()
        t0 = time.perf_counter()
        for k in range(warm + S, F - 2):
            t = time.perf_counter()
            arrive(k + 2)
            graphs[k % S].launch(sp)
            stream.synchronize()
            lat.append((time.perf_counter() - t) * 1e3)
        total = time.perf_counter() - t0
    for g in graphs.values():
        g.close()
    for g in gcl:
        for e in g["es"]:
            e.close()
    steps = len(lat)
    return {"graphs": len(graphs), "frame_steps": steps, "fps_aggregate": round(len(gcl) * steps / total, 1),
            "step_latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
            "step_latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
            "step": "arrival copies + one graph launch (the phase's batched call: every engine's frame-loop "
                    "launches, seam passes and learned-PSD predictions)"}


def learned_form(args):
    """Learned-PSD form on RGBPS float planes, bt=3, 48/16 (coloured noise with
    a Perlin sigma field): two networks trained from scratch on the GPU by
    the library's BF16 matrix-core kernels -- one on the noise periodogram,
    one on the method's own per-region pattern estimate of the noisy planes
    (no knowledge of the noise) -- against the best global sigma and the
    per-region patterns used directly; every setting tuned on one held-out
    1080p clip and scored on four others; frames/s at 1920x1080 on a resident
    clip."""
    import numpy as np
    import torch

    from neo_fft3d_amd import Engine, default_params, make_format, run_device_batch
    from neo_fft3d_amd import learned as L
    t0 = time.time()
    mlp, data, losses = L.train_model(clips=96, frames=2, width=384, height=384, steps=1500)
    t_train = time.time() - t0
    mlp_r, data_r, losses_r = L.train_model(clips=96, frames=2, width=384, height=384, steps=1500,
                                            labels="regions", regions=(4, 4))
    fmt = make_format("rgb", 32, True)

    def engines(w, h, f, sigma, m=None, strength=1.0, pfactor=0.0):
        es = [Engine(default_params(sigma, bt=3, bw=48, bh=48, ow=16, oh=16, sharpen=0.0, pfactor=pfactor), fmt,
                     w, h, f, p) for p in range(3)]
        for e in es:
            if m is not None:
                L.set_learned(e, m, strength)
            if pfactor:
                e.set_pattern_regions(4, 4)
        return es

    def filt(noisy, sigma=2.0, m=None, strength=1.0, pfactor=0.0):
        f, h, w = noisy[0].shape
        es = engines(w, h, f, sigma, m, strength, pfactor)
        outs = []
        for p, e in enumerate(es):
            s_ = torch.from_numpy(np.ascontiguousarray(noisy[p])).cuda()
            d_ = torch.zeros_like(s_)
            torch.cuda.synchronize()
            e.run_device(0, f, [s_[k].data_ptr() for k in range(f)], w * 4, [d_[k].data_ptr() for k in range(f)],
                         w * 4)
            e.sync()
            outs.append(d_.cpu().numpy())
            e.close()
        return outs

    def score(outs, clean):
        mse = np.mean([np.mean((o.astype(np.float64) - c) ** 2) for o, c in zip(outs, clean)])
        return float(10 * np.log10(1.0 / mse))
    QW, QH, QF = 1920, 1080, 5
    tc, tn, _ = L.noisy_clip(QW, QH, QF, seed=4000)
    forms = {
        "plain_best_sigma": ("sigma", [7.0, 8.0, 9.0, 10.0], lambda v: dict(sigma=v)),
        "learned_noise_labels": ("strength", [2.5, 3.0, 3.5], lambda v: dict(m=mlp, strength=v)),
        "learned_region_labels": ("strength", [0.05, 0.1, 0.25, 0.5, 1.0], lambda v: dict(m=mlp_r, strength=v)),
        "region_patterns_direct": ("pfactor", [0.05, 0.1, 0.25, 0.5, 1.0], lambda v: dict(pfactor=v)),
    }
    tuned = {k: max(vals, key=lambda v: score(filt(tn, **kw(v)), tc)) for k, (_, vals, kw) in forms.items()}
    db = {k: [] for k in forms}
    for k in range(4):
        c, n, _ = L.noisy_clip(QW, QH, QF, seed=5000 + k)
        for name, (_, _, kw) in forms.items():
            db[name].append(score(filt(n, **kw(tuned[name])), c))
    plain = float(np.mean(db["plain_best_sigma"]))
    quality = {name: {forms[name][0]: tuned[name], "psnr_db": [round(v, 2) for v in vals],
                      "mean_db": round(float(np.mean(vals)), 2),
                      "gain_over_plain_db": round(float(np.mean(vals)) - plain, 2)}
               for name, vals in db.items()}
    # throughput at 1920x1080, resident clip (seeded content, white noise of
    # the same level: the timing does not depend on the noise process)
    F = args.learned_frames
    clean = L.clean_planes(1920, 1080, F, seed=6000)
    rng = np.random.default_rng(6000)
    src = [torch.from_numpy((pl + rng.normal(0, 0.02, pl.shape)).astype(np.float32)).cuda() for pl in clean]
    dst = [torch.zeros_like(s_) for s_ in src]
    stride = 1920 * 4

    def timed(es, batch, period=0):
        # both forms through nf3d_run_device_batch; the learned form predicts
        # each plane's per-block PSD once per frame-loop launch (16 frames),
        # or once per `period` frames (each timed pass predicting afresh)
        items = [(e, 0, F, [src[p][k].data_ptr() for k in range(F)], stride, [dst[p][k].data_ptr() for k in range(F)],
                  stride) for p, e in enumerate(es)]
        torch.cuda.synchronize()
        if batch:
            run_device_batch(items)
        else:
            for it in items:
                it[0].run_device(*it[1:])
        for e in es:
            e.sync()
        t = time.perf_counter()
        for _ in range(2):
            if period:
                for e in es:
                    L.set_learned_period(e, period)
            if batch:
                run_device_batch(items)
            else:
                for it in items:
                    it[0].run_device(*it[1:])
        for e in es:
            e.sync()
        torch.cuda.synchronize()
        return 2 * F / (time.perf_counter() - t)
    es = engines(1920, 1080, F, tuned["plain_best_sigma"])
    fps_plain = timed(es, True)
    for e in es:
        e.close()
    es = engines(1920, 1080, F, 2.0, mlp, tuned["learned_noise_labels"])
    fps_learned = timed(es, True)
    fps_learned_period = {str(per): round(timed(es, True, per), 1) for per in (64, F)}
    for e in es:
        e.close()
    # MLP inference rate alone (one 1080p plane's blocks)
    nox, noy = plane_geometry(1920, 1080, 48, 48, 16, 16)
    nb = nox * noy
    xb = torch.zeros((nb, mlp.n_in), dtype=torch.bfloat16, device="cuda")
    out = torch.empty((nb, mlp.n_out), dtype=torch.float32, device="cuda")
    work = torch.empty(mlp.work_bytes(nb), dtype=torch.uint8, device="cuda")
    mlp.predict(xb.data_ptr(), nb, out.data_ptr(), work.data_ptr(), exp_out=True)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(20):
        mlp.predict(xb.data_ptr(), nb, out.data_ptr(), work.data_ptr(), exp_out=True)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / 20
    flops = 2 * nb * (mlp.n_in * mlp.nh + mlp.nh * mlp.nh + mlp.nh * mlp.n_out)
    metric_cfg = learned_on_metric_config(args)
    return {
        "metric_config": metric_cfg,
        "workload": "rgbps_float_bt3_bw48_ow16, coloured noise (power law x anisotropy, Perlin sigma 0.005-0.04)",
        "train": {"blocks": data.rows, "plane_frames": 96 * 2 * 3, "steps": 1500, "batch": 4096,
                  "seconds": round(t_train, 1),
                  "whittle_loss": [round(losses[0][1], 4), round(losses[-1][1], 4)],
                  "whittle_loss_region_labels": [round(losses_r[0][1], 4), round(losses_r[-1][1], 4)]},
        "quality": {"clips": 4, "size": f"{QW}x{QH}x{QF} frames", "tuned_on": "one held-out clip", **quality},
        "psnr_gain_db": quality["learned_noise_labels"]["gain_over_plain_db"],
        "fps_1080p_plain": round(fps_plain, 1), "fps_1080p_learned": round(fps_learned, 1),
        "fps_1080p_learned_by_prediction_period": fps_learned_period,
        "timing_frames": F,
        "mlp_forward": {"blocks": nb, "ms": round(ms, 4), "tflops_bf16": round(flops / (ms * 1e-3) / 1e12, 1),
                        "frac_of_2500": round(flops / (ms * 1e-3) / 1e12 / 2500.0, 4)},
    }


def learned_on_metric_config(args):
    """The learned-PSD form on the headline workload itself (1080p YUV420P8,
    300 frames, bt=3, 48/16, sharpen 0.3, all planes): a network for 8-bit
    48x48 blocks trained on the GPU (seeded coloured-noise data in 8-bit
    units), then frames/s of plain and learned over the resident clip in one
    batched call each, and the luma PSNR of both against the clean render
    (every 15th frame, clip ends included).  The learned form is timed with
    one prediction per frame-loop launch (16 frames) and with longer
    prediction periods (one per 64 frames, one per call)."""
    import numpy as np
    import torch

    from neo_fft3d_amd import Engine, default_params, make_format, run_device_batch
    from neo_fft3d_amd import learned as L
    from neo_fft3d_amd.synth import SEED_BASE, rects_clip
    t0 = time.time()
    mlp, data, losses = L.train_model(clips=24, frames=2, width=384, height=384, steps=800, scale=255.0)
    t_train = time.time() - t0
    F = args.frames
    planes = rects_clip(W, H, F, SEED_BASE + 1)
    clean = rects_clip(W, H, F, SEED_BASE + 1, noise_luma=0.0, noise_chroma=0.0)
    src = [torch.from_numpy(p).cuda() for p in planes]
    dst = [torch.empty_like(x) for x in src]
    fmt = make_format("yuv", 8)

    def run(learned, strength=1.0, reps=3, period=0):
        prm = default_params(**FILTER, bt=3)
        es = [Engine(prm, fmt, W, H, F, p) for p in range(3)]
        if learned:
            for e in es:
                L.set_learned(e, mlp, strength, period=period)
        items = [(es[p], 0, F, [src[p][n].data_ptr() for n in range(F)], src[p].stride(1),
                  [dst[p][n].data_ptr() for n in range(F)], dst[p].stride(1)) for p in range(3)]
        torch.cuda.synchronize()
        run_device_batch(items)  # warm-up
        for e in es:
            e.sync()
        t = time.perf_counter()
        for _ in range(reps):
            if learned and period:
                for e in es:  # each pass predicts afresh (no span kept from the pass before)
                    L.set_learned_period(e, period)
            run_device_batch(items)
        for e in es:
            e.sync()
        torch.cuda.synchronize()
        fps = reps * F / (time.perf_counter() - t)
        idx = list(range(0, F, 15)) + [F - 1]
        out = dst[0].cpu().numpy()
        db = float(np.mean([psnr(out[n], clean[0][n]) for n in idx]))
        for e in es:
            e.close()
        return fps, db
    fps_plain, db_plain = run(False)
    (fps_learned, db_learned), strength = max(((run(True, a), a) for a in (1.0, 1.5)), key=lambda r: r[0][1])
    periods = {}
    for period in (64, F):
        f, d = run(True, strength, period=period)
        periods[str(period)] = {"fps": round(f, 1), "over_plain": round(f / fps_plain, 3), "psnr_luma": round(d, 2)}
    return {"workload": "yuv420p8_1920x1080_300f_bt3_bw48_ow16_sharpen0.3",
            "train": {"blocks": data.rows, "steps": 800, "seconds": round(t_train, 1), "scale": 255.0},
            "fps_plain": round(fps_plain, 1), "fps_learned": round(fps_learned, 1),
            "learned_over_plain": round(fps_learned / fps_plain, 3),
            "psnr_luma_plain": round(db_plain, 2), "psnr_luma_learned": round(db_learned, 2),
            "psnr_frames": f"every 15th of {F} and the last", "strength": strength,
            "prediction_period_frames": periods}


# ------------------------------------------------------------- CPU baseline
def cpu_baseline(args):
    """The FP64 restatement of the path (oracle/libnf3d_oracle.so, scalar C)
    on a bounded sample of the same workload: the first `threads * per` frames
    of the seeded clip, each thread with its own oracle instance filtering a
    contiguous run of `per` frames (its window's neighbours read from the
    shared input).  The reference's own build needs FFTW, which this host
    lacks, so the restatement is the CPU figure."""
    import threading

    import numpy as np

    from neo_fft3d_amd.synth import SEED_BASE, rects_clip

    threads = args.cpu_threads
    per = args.cpu_frames_per_thread
    F = threads * per
    planes = rects_clip(W, H, F, SEED_BASE + 1, threads=threads)
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "libnf3d_oracle.so"))
    lib.nf3d_oracle_create.restype = ctypes.c_void_p
    lib.nf3d_oracle_create.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 8 + [ctypes.c_void_p, ctypes.c_void_p,
                                                                                 ctypes.c_char_p, ctypes.c_int]
    lib.nf3d_oracle_get_frame.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_char_p, ctypes.c_int]
    lib.nf3d_oracle_free.argtypes = [ctypes.c_void_p]
    ptrs = (ctypes.c_void_p * (3 * F))(*[planes[p][n].ctypes.data for n in range(F) for p in range(3)])
    strides = (ctypes.c_int * 3)(*[planes[p].strides[1] for p in range(3)])
    a = f"{REF_ARGS};bt={args.bt}".encode()
    errors = []

    def work(t):
        err = ctypes.create_string_buffer(256)
        h = lib.nf3d_oracle_create(a, 1, 0, 8, 1, 1, W, H, F, ptrs, strides, err, 256)
        if not h:
            errors.append(err.value.decode())
            return
        outs = [np.zeros_like(planes[p][0]) for p in range(3)]
        optr = (ctypes.c_void_p * 3)(*[o.ctypes.data for o in outs])
        ostr = (ctypes.c_int * 3)(*[o.strides[0] for o in outs])
        for n in range(t * per, (t + 1) * per):
            if lib.nf3d_oracle_get_frame(ctypes.c_void_p(h), n, optr, ostr, err, 256) != 0:
                errors.append(err.value.decode())
                break
        lib.nf3d_oracle_free(ctypes.c_void_p(h))

    t0 = time.perf_counter()
    pool = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
    for th in pool:
        th.start()
    for th in pool:
        th.join()
    el = time.perf_counter() - t0
    if errors:
        return {"value": None, "unit": "frames/s", "cores": threads, "kind": "port",
                "sample": "failed: " + errors[0]}
    cpu = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    desc = (f"FP64 scalar C restatement of the path (oracle/libnf3d_oracle.so) on the first {F} frames of "
            f"the same 1080p YUV420P8 bt={args.bt} 48/16 sharpen 0.3 workload, all three planes: {threads} "
            f"threads x {per} consecutive frames each, {el:.1f} s on {cpu}")
    return {"value": round(F / el, 3), "unit": "frames/s", "cores": threads, "kind": "port", "sample": desc}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=300)
    ap.add_argument("--bt", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-host-fed", action="store_true",
                    help="skip the host-fed (PCIe) leg, e.g. for a kernel trace of the resident passes only")
    ap.add_argument("--no-precision-form", action="store_true",
                    help="skip the DFT-as-GEMM (MFMA, FP16 split) transform measurement")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-PSD leg")
    ap.add_argument("--learned-metric-only", action="store_true",
                    help="of the learned-PSD leg, run only the headline-workload part (training and "
                         "plain vs learned frames/s; no quality study)")
    ap.add_argument("--learned-frames", type=int, default=96)
    ap.add_argument("--precision-frames", type=int, default=120)
    ap.add_argument("--no-multi-clip", action="store_true", help="skip the 16-clip batched leg")
    ap.add_argument("--multi-frames", type=int, default=40)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run one device copy of MIB MiB first (a reference kernel for traffic counters)")
    ap.add_argument("--cpu-frames-per-thread", type=int, default=12)
    ap.add_argument("--cpu-threads", type=int, default=min(os.cpu_count() or 1, 16))
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not a code path)"}))
        return 2
    r = gpu_run(args)
    line = {
        "metric": METRIC, "value": round(r["fps"], 2), "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(r["ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded moving fBm-textured rectangles + Gaussian noise, resident in HBM)",
        "config": {"workload": "yuv420p8_1920x1080_300f_bt3_bw48_ow16_sharpen0.3", "frames": args.frames,
                   "bt": args.bt, "sigma": 2.0, "bw": 48, "bh": 48, "ow": 16, "oh": 16, "sharpen": 0.3,
                   "planes": [0, 1, 2], "frames_per_step": args.frames},
        "roofline": r["roofline"],
        "psnr_db": r["psnr_db"],
        "host_fed": r["host_fed"],
    }
    if not args.no_precision_form:
        line["precision_form"] = precision_form(args)
    if not args.no_learned:
        if args.learned_metric_only:
            line["learned_psd"] = {"metric_config": learned_on_metric_config(args)}
        else:
            line["learned_psd"] = learned_form(args)
    if not args.no_multi_clip:
        line["multi_clip"] = multi_clip(args)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args)
        cb = line["cpu_baseline"]["value"]
        if cb:
            line["gpu_over_cpu"] = round(r["fps"] / cb, 1)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
