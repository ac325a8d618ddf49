"""Benchmark: BM3D Basic+Final on 1080p RGB48 (MI355X, one GPU).

Workload (one step): 8 seeded synthetic 1920x1080 RGB48 frames through
RGB2OPP(sample=0) -> Basic(profile='np', sigma=[10,10,10], matrix=100) ->
Final(profile='np', sigma=[10,10,10], matrix=100) -> OPP2RGB(sample=0),
the chain the reference's README recommends (README.md:358-365).  Inputs and
outputs stay in HBM (torch device tensors handed to the C ABI).

Prints one JSON line.  value = whole-job Mpix/s.  roofline = the dominant
kernel's algorithmic FLOP rate (per-launch FLOPs / average launch time from
HIP events recorded on the library's stream) against the MI355X peak.
cpu_baseline = the CPU oracle (a C restatement of the reference CPU path,
one frame per thread like VapourSynth's fmParallel) on a bounded sample.
learned_bm = the secondary metric: learned-proposal vs exhaustive block
matching on a 4K Gray frame (speedup, recall, PSNR loss).  batched = Basic ->
Final over 64 Gray float frames of uneven sizes with per-frame sigma, one
launch per stage over all frames, HIP-graph replay, against the single-frame
filters.  vbm3d = the V-BM3D chain on 30 frames 1080p YUV444 float.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vapoursynth-bm3d_amd"))

import numpy as np  # noqa: E402

FRAMES, W, H, SIGMA = 8, 1920, 1080, 10.0
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 peak (vector and matrix alike; FMA = 2 flops)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_1080p_rgb48_np.json")
METRIC = "Mpix/s BM3D Basic+Final 1080p RGB; learned-BM speedup at ≤0.05 dB PSNR loss"


def make_frames(n=FRAMES, h=H, w=W):
    import synth
    return np.stack([synth.rgb16(h, w, SIGMA, seed=2000 + k)[0] for k in range(n)])


def flops_per_frame(stats, B, wiener, planes=3):
    """Algorithmic FLOPs of one frame of a stage:
    SSD 3*B^2 per evaluated candidate; per group and plane T*(4*B^3*n + 2*B^2*n^2)
    for the separable transforms (T = 2 Basic, 3 Final) plus c*B^2*n for the
    shrinkage and aggregation arithmetic (c = 4 Basic, 8 Final)."""
    T, c = (3, 8) if wiener else (2, 4)
    ssd = 3 * B * B * stats["ssd_evals"]
    dct = planes * T * (4 * B ** 3 * stats["members"] + 2 * B * B * stats["members_sq"])
    shrink = planes * c * B * B * stats["members"]
    return ssd, dct + shrink


def executed_filter_flops(stats, B, wiener, planes=3):
    """FLOPs the filter kernel executes per frame of a stage (multiply = 1,
    add = 1, fused multiply-add = 2): the 8-point x/y passes in their even/odd
    form (64 flops per forward line, 60 per inverse line, 8n lines per pass),
    the z chains (n^2 per forward column, about n^2 - n/2 per inverse column),
    the shrinkage (Wiener 6 per coefficient) and the contribution scaling.
    Computed from the stage's sum of n and n^2 over its groups."""
    n1, n2 = stats["members"], stats["members_sq"]
    fwd = 2 * 8 * 64 * n1 + B * B * n2
    inv = 2 * 8 * 60 * n1 + B * B * (n2 - n1 / 2)
    T = 2 if wiener else 1
    elem = (6 if wiener else 0) * B * B * n1 + B * B * n1
    return planes * (T * fwd + inv + elem)


def measured_traffic(timer):
    """HBM bytes per launch of the kernel behind a timer, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this workload (calibrated by a
    plain copy); None when that kernel was not profiled."""
    try:
        with open(TRAFFIC_FILE) as fh:
            d = json.load(fh)
    except OSError:
        return None
    for k in d.get("kernels", {}).values():
        if k.get("timer") == timer:
            return k["hbm_bytes"]
    return None


LEARNED_KW = dict(profile="fast", sigma=25.0, bm_range=24, group_size=16)
LEARNED_TOPK = 32


def learned_leg(reps=3):
    """Secondary metric: block matching of Basic(profile='fast', sigma=25,
    bm_range=24, group_size=16) on one seeded 3840x2160 Gray float frame,
    exhaustive vs the learned proposal (32 best even-lattice positions by
    embedding distance, their 3x3 neighbourhoods re-ranked by exact SSD,
    trained weights of lbm_train.py), and the same with an untrained network.  Matching time per frame from the library's
    event timers; recall of the exhaustive lists; PSNR of both Basic outputs
    against the clean image."""
    import bm3d
    import lbm
    import synth
    weights = os.path.join(ROOT, "vapoursynth-bm3d_amd", "lbm_weights.npz")
    if not os.path.exists(weights):
        return {"skipped": "no trained weights (run vapoursynth-bm3d_amd/lbm_train.py)"}
    noisy, clean = synth.gray_float(2160, 3840, 25.0, seed=5999)
    clip = bm3d.Clip.from_arrays(bm3d.GRAYS, [noisy[None]])
    model = lbm.Model.load(weights)
    untrained = lbm.Model(seed=1)
    res = {}
    for name, m in (("exhaustive", None), ("learned", model), ("untrained", untrained)):
        node = bm3d.Basic(clip, **LEARNED_KW)
        if m is not None:
            node.filter.set_matcher(m, LEARNED_TOPK)
        node.get_frame(0)  # warm-up
        node.filter.profile(True)
        for _ in range(reps):
            node.clear_cache()
            out = np.asarray(node.get_frame(0).planes[0])
        prof = node.filter.profile_read()
        ms, n = prof["k_match(basic)"]
        cnt, pos, _ = node.filter.match_lists()
        res[name] = dict(ms=ms / n, out=out, cnt=cnt, pos=pos, fallback=node.filter.matcher_fallbacks())
    e, l = res["exhaustive"], res["learned"]
    p_e, p_l = lbm.psnr(e["out"], clean), lbm.psnr(l["out"], clean)
    return {"workload": "Basic(fast, sigma=25, bm_range=24, group_size=16), 1 frame 3840x2160 Gray float, top-32",
            "exhaustive_match_ms": round(e["ms"], 4), "learned_match_ms": round(l["ms"], 4),
            "speedup": round(e["ms"] / l["ms"], 3), "recall": round(lbm.recall(e["cnt"], e["pos"], l["cnt"], l["pos"]), 4),
            "psnr_exhaustive_db": round(p_e, 4), "psnr_learned_db": round(p_l, 4),
            "psnr_loss_db": round(p_e - p_l, 4), "psnr_learned_vs_exhaustive_db": round(lbm.psnr(l["out"], e["out"]), 3),
            "fallback_blocks": l["fallback"], "ref_blocks": int(len(e["cnt"])),
            "untrained_network": {"match_ms": round(res["untrained"]["ms"], 4),
                                  "recall": round(lbm.recall(e["cnt"], e["pos"], res["untrained"]["cnt"],
                                                             res["untrained"]["pos"]), 4),
                                  "psnr_loss_db": round(p_e - lbm.psnr(res["untrained"]["out"], clean), 4),
                                  "fallback_blocks": res["untrained"]["fallback"]}}


V_FRAMES, V_RADIUS, V_SIGMA = 30, 3, 15.0
V_KW = dict(radius=V_RADIUS, ps_num=2, ps_range=4, sigma=[V_SIGMA])


def vbm3d_leg(steps=1):
    """V-BM3D on 30 frames 1920x1080 YUV444 float (moving textured objects,
    AWGN sigma 15): VBasic(radius=3, ps_num=2, ps_range=4) -> VAggregate ->
    VFinal(same) -> VAggregate, every output frame, inputs resident in HBM.
    Per-kernel times from the library's event timers; VAggregate's achieved
    HBM rate from its algorithmic bytes ((window frames * 2 slots * 4 B + 4 B)
    per pixel and plane) over its average launch time."""
    import torch
    import bm3d
    import synth
    noisy, _ = synth.moving_video(V_FRAMES, H, W, V_SIGMA, seed=3000, planes=3)
    frames = [torch.from_numpy(f).cuda() for f in noisy]
    del noisy
    clip = bm3d.Clip.from_arrays(bm3d.YUV444PS, frames)
    vb = bm3d.VBasic(clip, **V_KW)
    ab = bm3d.VAggregate(vb, radius=V_RADIUS, sample=1)
    vf = bm3d.VFinal(clip, ab, **V_KW)
    out = bm3d.VAggregate(vf, radius=V_RADIUS, sample=1)
    chain = (vb, ab, vf, out)

    def run():
        for c in chain:
            c.clear_cache()
        for n in range(V_FRAMES):
            out.get_frame(n)
            if n >= 2 * V_RADIUS + 1:  # stacks no later frame needs
                vb.evict(n - 2 * V_RADIUS - 1)
                vf.evict(n - V_RADIUS - 1)
                ab.evict(n - 2 * V_RADIUS - 1)
        torch.cuda.synchronize()

    run()  # warm-up: workspace allocation
    for c in chain:
        c.filter.profile(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    dt = (time.perf_counter() - t0) / steps
    prof = {}
    for c in chain:
        for name, (ms, n) in c.filter.profile_read().items():
            p = prof.setdefault(name, [0.0, 0])
            p[0] += ms
            p[1] += n
        c.filter.profile(False)
    win = sum(min(V_FRAMES - 1, n + V_RADIUS) - max(0, n - V_RADIUS) + 1 for n in range(V_FRAMES))
    agg_bytes = (win * 2 * 4 + V_FRAMES * 4) * 3 * W * H / V_FRAMES  # per launch, average
    agg_ms = prof["k_vagg"][0] / prof["k_vagg"][1]
    for c in chain:
        c.clear_cache()
    return {"workload": "VBasic->VAggregate->VFinal->VAggregate (radius 3, ps_num 2, ps_range 4, profile fast, "
                        "sigma 15), 30 frames 1920x1080 YUV444 float, moving objects, inputs resident in HBM",
            "frames": V_FRAMES, "seconds": round(dt, 4), "fps": round(V_FRAMES / dt, 3),
            "mpix_s": round(V_FRAMES * W * H / dt / 1e6, 3),
            "kernels_ms_per_frame": {k: round(v[0] / max(v[1], 1), 4) for k, v in prof.items()},
            "vaggregate_roofline": {"bound": "hbm", "achieved": round(agg_bytes / (agg_ms * 1e-3) / 1e9, 1),
                                    "peak": 8000.0, "unit": "GB/s",
                                    "frac": round(agg_bytes / (agg_ms * 1e-3) / 8e12, 4),
                                    "bytes_per_launch": int(agg_bytes), "avg_launch_ms": round(agg_ms, 4)}}


BATCH_FRAMES = 64


def batch_workload(n=BATCH_FRAMES, seed=64):
    """64 Gray float frames of uneven sizes between 256x256 and 2048x1152
    (the two extremes included) with per-frame sigma in [5, 30]."""
    rng = np.random.RandomState(seed)
    hs = rng.randint(256, 1153, n)
    ws = rng.randint(256, 2049, n)
    hs[0], ws[0], hs[1], ws[1] = 256, 256, 1152, 2048
    sig = np.round(rng.uniform(5.0, 30.0, n), 2)
    return [(int(h), int(w)) for h, w in zip(hs, ws)], [float(s) for s in sig]


def batched_leg(reps=3, tune=True):
    """Batched Basic(np) -> Final(np) over the 64-frame uneven workload: one
    launch per stage over all frames, the chain captured in a HIP
    graph after autotuning the launch shape.  Compared with the same frames
    filtered one at a time through the single-frame filters (steady state,
    each frame's filters created and warmed up before its timed call), and
    checked bit for bit against them."""
    import torch
    import bm3d
    import synth
    sizes, sigmas = batch_workload()
    src, dst = [], []
    for k, ((h, w), s) in enumerate(zip(sizes, sigmas)):
        src.append(torch.from_numpy(synth.gray_float(h, w, s, seed=6400 + k)[0]).cuda())
        dst.append(torch.empty((h, w), dtype=torch.float32, device="cuda"))
    b = bm3d.Batch(sizes, sigmas, profile="np")
    info = b.info()
    b.run(src, dst)  # warm-up
    rep = b.tune(src, dst, reps=1) if tune else ""
    eager = [b.run(src, dst) for _ in range(reps)]
    b.run(src, dst, graph=True)  # capture
    graph = [b.run(src, dst, graph=True) for _ in range(reps)]
    torch.cuda.synchronize()
    # single-frame filters, frame by frame
    single_ms, exact = 0.0, True
    for k, ((h, w), s) in enumerate(zip(sizes, sigmas)):
        clip = bm3d.Clip.from_arrays(bm3d.GRAYS, [src[k][None]])
        basic = bm3d.Basic(clip, profile="np", sigma=s)
        final = bm3d.Final(clip, basic, profile="np", sigma=s)
        final.get_frame(0)
        basic.clear_cache()
        final.clear_cache()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = final.get_frame(0).planes[0]
        torch.cuda.synchronize()
        single_ms += (time.perf_counter() - t0) * 1e3
        exact = exact and bool(torch.equal(out.view(torch.int32), dst[k].view(torch.int32)))
        del clip, basic, final, out
    ms = min(graph)
    px = info["pixels"]
    return {"workload": f"Basic(np)->Final(np), {len(sizes)} Gray float frames 256x256..2048x1152, "
                        f"sigma 5..30 per frame, inputs resident in HBM",
            "frames": len(sizes), "pixels": px, "basic_blocks": info["basic_blocks"],
            "final_blocks": info["final_blocks"],
            "graph_ms": round(ms, 3), "eager_ms": round(min(eager), 3), "mpix_s": round(px / ms / 1e3, 3),
            "single_frame_ms": round(single_ms, 3), "single_frame_mpix_s": round(px / single_ms / 1e3, 3),
            "speedup_vs_single_frame": round(single_ms / ms, 3), "bit_exact_vs_single_frame": exact,
            "shape": b.shape, "tune": rep.strip().splitlines()}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def _ssd_per_pixel(width, height, stage_kw):
    """SSD evaluations per pixel of a stage on a width x height frame (the
    search-window rules of Block.h:683-707; float only to compare crops)."""
    def grid(dim, B, s):
        out, j = [], 0
        while j < dim - B + s:
            out.append(min(j, dim - B))
            j += s
        return out

    B, step, R = 8, stage_kw["block_step"], stage_kw["bm_range"]

    def axis(pos, last):
        return min(pos, R) + min(last - pos, R) + 1

    cy = sum(axis(p, height - B) for p in grid(height, B, step))
    cx = sum(axis(p, width - B) for p in grid(width, B, step))
    return (cy * cx - len(grid(height, B, step)) * len(grid(width, B, step))) / (width * height)


def cpu_baseline(frames, rows, nthreads, share):
    """The CPU oracle (a C restatement of the reference CPU path: SSE2 SSD,
    rcpps Wiener, the project's DCT -- FFTW is not on the box) on the same
    synthetic frames cropped to W x rows, one frame per thread as VapourSynth
    runs the reference (fmParallel): once with the 8 frames (min(8, cores)
    threads) and once with `share` frames on `share` threads (the cores one GPU's
    share of the host offers).  Crops have shorter search windows near their
    edges, so the SSD evaluations per pixel of crop and full frame are reported
    beside the figure."""
    import ctypes.util
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    pb = O.resolve("Basic", width=W, height=H, color_family="yuv", profile="np", sigma=[SIGMA] * 3, matrix=100)
    pf = O.resolve("Final", width=W, height=H, color_family="yuv", profile="np", sigma=[SIGMA] * 3, matrix=100)
    fb, ff = O.make_filter(pb), O.make_filter(pf)
    sample = np.ascontiguousarray(frames[:, :, :rows, :])
    t0 = time.perf_counter()
    O.chain_rgb48(fb, ff, sample, nthreads)
    dt = time.perf_counter() - t0
    px = sample.shape[0] * rows * W
    many = np.ascontiguousarray(np.concatenate([sample] * ((share + len(sample) - 1) // len(sample)))[:share])
    t1 = time.perf_counter()
    O.chain_rgb48(fb, ff, many, share)
    dt_share = time.perf_counter() - t1
    kb, kf = dict(block_step=4, bm_range=16), dict(block_step=3, bm_range=16)
    ssd_crop = _ssd_per_pixel(W, rows, kb) + _ssd_per_pixel(W, rows, kf)
    ssd_full = _ssd_per_pixel(W, H, kb) + _ssd_per_pixel(W, H, kf)
    return {"value": round(px / dt / 1e6, 4), "unit": "Mpix/s", "cores": nthreads, "kind": "port",
            "sample": f"{sample.shape[0]} frames cropped to {W}x{rows} (same synthetic frames), "
                      f"one frame per thread, {dt:.1f} s wall",
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
            "dct_backend": "project DCT (the oracle's DCT-II/III chains; libfftw3f "
                           + ("present but not used)" if ctypes.util.find_library("fftw3f") else "not on the host)"),
            "share_value": round(many.shape[0] * rows * W / dt_share / 1e6, 4), "share_cores": share,
            "ssd_per_pixel_crop": round(ssd_crop, 2), "ssd_per_pixel_full_frame": round(ssd_full, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-rows", type=int, default=128)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned block-matching leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched uneven-frames leg")
    ap.add_argument("--only-batched", action="store_true", help="run only the batched leg (prints its JSON)")
    ap.add_argument("--no-vbm3d", action="store_true", help="skip the V-BM3D 30-frame leg")
    ap.add_argument("--only-learned", action="store_true", help="run only the learned matching leg (prints its JSON)")
    ap.add_argument("--only-vbm3d", action="store_true", help="run only the V-BM3D leg (prints its JSON)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (three times) and exit; "
                         "used to calibrate the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(frames shard across GPUs with no communication; see the sharding design)", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_probe:
        n = args.copy_probe * (1 << 20) // 4
        a = torch.arange(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4, "copies": 3}), flush=True)
        return

    import bm3d

    if args.only_batched:
        bm3d.init(0)
        print(json.dumps(batched_leg()), flush=True)
        return
    if args.only_vbm3d:
        bm3d.init(0)
        print(json.dumps(vbm3d_leg()), flush=True)
        return
    if args.only_learned:
        bm3d.init(0)
        print(json.dumps(learned_leg()), flush=True)
        return

    frames = make_frames()
    bm3d.init(0)
    dev = torch.device("cuda:0")
    dframes = [torch.from_numpy(f.view(np.int16)).to(dev).view(torch.uint16) for f in frames]
    src = bm3d.Clip.from_arrays(bm3d.RGB48, dframes)
    opp = bm3d.RGB2OPP(src)
    basic = bm3d.Basic(opp, profile="np", sigma=[SIGMA] * 3, matrix=100)
    final = bm3d.Final(opp, basic, profile="np", sigma=[SIGMA] * 3, matrix=100)
    out = bm3d.OPP2RGB(final)
    chain = [opp, basic, final, out]

    def step():
        for c in chain:
            c.clear_cache()
        for n in range(FRAMES):
            out.get_frame(n)

    for k in range(args.warmup):
        step()
        torch.cuda.synchronize()
        print(f"warmup step {k + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
        print(f"timed step {k + 1}/{args.steps} issued", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0

    # per-kernel times (HIP events on each filter's stream) and algorithmic
    # work per launch, from one more pass with the frames one after another,
    # so that no other stage's kernels share the GPU with a timed launch
    work = {"k_match(basic)": [], "k_match(final)": [], "k_filter(hard)": [], "k_filter(wiener)": []}
    executed = {"k_filter(hard)": [], "k_filter(wiener)": []}
    for c in (basic, final):
        c.filter.profile(True)
    for c in chain:
        c.clear_cache()
    for n in range(FRAMES):
        out.get_frame(n)
        for c, wiener in ((basic, False), (final, True)):
            st = c.filter.stats()
            ssd, dct = flops_per_frame(st, 8, wiener)
            work["k_match(final)" if wiener else "k_match(basic)"].append(ssd)
            work["k_filter(wiener)" if wiener else "k_filter(hard)"].append(dct)
            executed["k_filter(wiener)" if wiener else "k_filter(hard)"].append(executed_filter_flops(st, 8, wiener))
    prof = {}
    for c in (basic, final):
        for name, (ms, n) in c.filter.profile_read().items():
            prof[name] = (ms, n)
        c.filter.profile(False)
    kern = {}
    for name, (ms, n) in prof.items():
        kern[name] = (ms / max(n, 1), n)
    dom = max(prof.items(), key=lambda kv: kv[1][0])[0]
    # the Wiener stage of a three-plane frame is one launch per plane (the
    # single-plane kernel); per-launch figures divide the stage by that
    per_stage = 3 if dom == "k_filter(wiener)" else 1
    stage_ms = prof[dom][0] / max(prof[dom][1], 1)
    avg_ms = stage_ms / per_stage
    flops_launch = float(np.mean(work[dom])) / per_stage if dom in work and work[dom] else None
    achieved = flops_launch / (avg_ms * 1e-3) / 1e12 if flops_launch else None

    ms_step = dt / args.steps * 1e3
    mpix = FRAMES * W * H / (dt / args.steps) / 1e6
    result = {
        "metric": METRIC,
        "value": round(mpix, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "RGB2OPP->Basic(np)->Final(np)->OPP2RGB, 8 frames 1920x1080 RGB48, sigma 10, "
                               "matrix=100, inputs resident in HBM",
                   "frames": FRAMES, "width": W, "height": H},
        "roofline": {"bound": "valu", "kernel": dom, "achieved": round(achieved, 3) if achieved else None,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4) if achieved else None,
                     "traffic": measured_traffic(dom),
                     "avg_launch_ms": round(avg_ms, 4),
                     "launches_per_frame": per_stage,
                     "stage_ms_per_frame": round(stage_ms, 4),
                     "flops_per_launch": flops_launch,
                     "flops_count": "algorithmic: dense separable 3D DCT T*(4*B^3*n + 2*B^2*n^2) + c*B^2*n per "
                                    "group and plane, from the frame's sum of n and n^2",
                     "executed_flops_per_launch": (float(np.mean(executed[dom])) / per_stage if executed.get(dom)
                                                   else None),
                     "executed_tflops": (round(float(np.mean(executed[dom])) / (stage_ms * 1e-3) / 1e12, 3)
                                         if executed.get(dom) else None)},
        "kernels_ms_per_frame": {k: round(v[0], 4) for k, v in kern.items()},
        # the matcher against the same FP32 peak and against the ceiling of
        # its exact-order SSD (separate sub, mul and add: no FMA, so at most
        # half the FMA-counted peak)
        "match_roofline": {
            name: {"achieved": round(float(np.mean(work[name])) / (kern[name][0] * 1e-3) / 1e12, 3),
                   "unit": "TFLOP/s", "frac": round(float(np.mean(work[name])) / (kern[name][0] * 1e-3) / 1e12
                                                    / PEAK_FP32_TFLOPS, 4),
                   "frac_of_non_fma_ceiling": round(float(np.mean(work[name])) / (kern[name][0] * 1e-3) / 1e12
                                                    / (PEAK_FP32_TFLOPS / 2), 4)}
            for name in ("k_match(basic)", "k_match(final)") if name in kern and work.get(name)},
    }
    if not args.no_learned:
        print("learned block-matching leg (4K)", file=sys.stderr, flush=True)
        result["learned_bm"] = learned_leg()
    if not args.no_batched:
        print("batched leg (64 uneven Gray frames)", file=sys.stderr, flush=True)
        result["batched"] = batched_leg()
    if not args.no_vbm3d:
        print("V-BM3D leg (30 frames 1080p YUV444 float)", file=sys.stderr, flush=True)
        result["vbm3d"] = vbm3d_leg()
    if not args.no_cpu:
        cores = os.cpu_count() or 1
        result["cpu_baseline"] = cpu_baseline(frames, args.cpu_rows, min(FRAMES, cores), min(16, cores))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
