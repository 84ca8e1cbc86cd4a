"""Throughput of the QMIV metric engine on MI355X.

Default workload (one step): IV-PSNR + WS-PSNR + PSNR of 8 frame pairs, 4096x2048, yuv420p10le,
equirectangular (WS weights on), search range 2, GCD on. Frames are synthetic (ivpsnr_amd.synth)
and resident in HBM before the timed region starts.

`--workload ssim-1080p10` times SSIM + MS-SSIM + IV-SSIM of 16 frame pairs 1920x1080 yuv420p10le
instead (its dominant kernel is the fp64 windowed-moment kernel). `--workload multiview24` times
24 views of uneven size (8 x 4096x2048 ERP, 8 x 1920x1080, 8 x 2048x1088, yuv420p10le) x 4 frames,
IV-PSNR + WS-PSNR + IV-SSIM, as one captured graph replay per frame set against the plain
per-view call loop; `multiview24-ivpsnr` is the same without IV-SSIM. `--workload grad-ivssim1024`
times the differentiable IV-SSIM (forward and gradients of both pictures) on 8 x 3 x 1024 x 1024
fp32 pictures. `--workload masked-1080p8` times masked mode (PSNR + WS-PSNR + IV-PSNR over the
pels of an 8-bit object mask) on 8 frame pairs 1920x1080 yuv420p.

Prints one JSON line:
  value      whole-job Mpix/s over the timed steps
  roofline   the dominant kernel: algorithmic bytes (or flops) per launch over its average launch
             duration, measured with HIP events on the engine stream; `traffic` is the HBM bytes per
             launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of that kernel
  cpu_baseline  the reference CPU path (oracle/_ref/qmiv_ref_v3, the reference's own sources
             built for x86-64-v3) on a bounded sample of the same workload
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
FP64_PEAK_TFLOPS = 78.6

WORKLOADS = {
    "ivpsnr-erp4k10": dict(
        metric="IV-PSNR Mpix/s @4096x2048 10b 420 + %HBM roofline; batched vs plain",
        w=4096, h=2048, bd=10, frames=8, erp=True, ml="PSNR,WSPSNR,IVPSNR", kernel="cps",
        desc="IV-PSNR+WS-PSNR+PSNR, 8 frame pairs 4096x2048 yuv420p10le ERP, B=2, GCD on",
        cpu_stages=("validate", "preproc", "gcd", "psnr", "wspsnr", "ivpsnr"), dtype="int32",
        cps_kernel="k_cps420s", traffic_file="k_cps420s_4096x2048_10b_erp.txt", value_mode="pipelined"),
    "ssim-1080p10": dict(
        metric="SSIM+MS-SSIM+IV-SSIM Mpix/s @1920x1080 10b 420",
        w=1920, h=1080, bd=10, frames=16, erp=False, ml="SSIM,MSSSIM,IVSSIM", kernel="ssim",
        desc="SSIM+MS-SSIM+IV-SSIM, 16 frame pairs 1920x1080 yuv420p10le, B=2, GCD on",
        cpu_stages=("validate", "preproc", "gcd", "scp", "ssim", "msssim", "ivssim"), dtype="f64",
        traffic_file="k_ssim_1920x1080_10b.txt", value_mode="pipelined"),
}


MULTIVIEW = [(4096, 2048, True)] * 8 + [(1920, 1080, False)] * 8 + [(2048, 1088, False)] * 8


def frame_pair_bytes(w, h, bd):
    return 2 * w * h * 3 // 2 * (1 if bd <= 8 else 2)


def ssim_flops_per_frame(w, h):
    """fp64 flops the SSIM-family kernels execute for one frame pair (SSIM + MS-SSIM + IV-SSIM):
    100 taps x 2 flops per window moment per interior pel, per component (chroma on the 4:4:4 grid).
    At full resolution one fused pass per component accumulates 11 moments (the means and squares of
    P0, P1, SCP1 and SCP0 and the cross terms P0.P1, P0.SCP1, P1.SCP0) for SSIM, MS-SSIM level 0 and
    IV-SSIM; MS-SSIM levels 1-4 run 5 moments each (batched into one launch)."""
    def interior(a, b):
        return max(a - 10, 0) * max(b - 10, 0)
    per_cmp = 11 * 200 * interior(w, h) + sum(5 * 200 * interior(w >> i, h >> i) for i in range(1, 5))
    return 3 * per_cmp


def read_profile_value(name, key):
    """A figure of the dominant kernel from its committed rocprofv3 PMC summary (a line `key: N`),
    or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def read_traffic(name):
    """HBM bytes per launch of the dominant kernel (line `traffic_bytes_per_launch: N`), or None."""
    v = read_profile_value(name, "traffic_bytes_per_launch")
    return None if v is None else int(v)


def cpu_baseline(wl, pairs, sample_frames):
    """Times the reference's own CPU path on `sample_frames` frames of the workload."""
    W, H, BD = wl["w"], wl["h"], wl["bd"]
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "qmiv_ref_v3")
    cores = min(16, os.cpu_count() or 1)
    if os.path.exists(ref_bin):
        with tempfile.TemporaryDirectory() as td:
            f0, f1 = os.path.join(td, "p0.yuv"), os.path.join(td, "p1.yuv")
            with open(f0, "wb") as a, open(f1, "wb") as b:
                for i in range(sample_frames):
                    a.write(pairs[i % len(pairs)].ref.tobytes())
                    b.write(pairs[i % len(pairs)].tst.tobytes())
            cmd = [ref_bin, "-w", str(W), "-h", str(H), "-bd", str(BD), "-cf", "420", "-i0", f0, "-i1", f1,
                   "-ml", wl["ml"], "-nth", str(cores), "-time"]
            if wl["erp"]:
                cmd[-5:-5] = ["-erp", "360", "180"]
            try:
                r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
                if r.returncode == 0:
                    t = json.loads(r.stdout)["time_ms_per_frame"]
                    ms = sum(t[k] for k in wl["cpu_stages"])
                    return {"value": round(W * H / (ms * 1e3), 3), "unit": "Mpix/s", "cores": cores,
                            "kind": "reference",
                            "sample": f"{sample_frames} frames {W}x{H} {BD}-bit of the same workload ({wl['ml']}) "
                                      f"incl. validation, preprocessing and GCD, {cores} threads; {ms:.1f} ms/frame"}
            except (subprocess.TimeoutExpired, OSError, ValueError, KeyError):
                pass
    # The reference build is missing or does not run on this host: time the C restatement.
    from oracle import pyoracle
    m = 0
    for name in wl["ml"].split(","):
        m |= getattr(pyoracle, name)
    p = pyoracle.params(W, H, BD, 420, metrics=m, erp=wl["erp"])
    t0 = time.perf_counter()
    pyoracle.frame(pairs[0].ref, pairs[0].tst, p)
    ms = (time.perf_counter() - t0) * 1e3
    return {"value": round(W * H / (ms * 1e3), 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
            "sample": f"1 frame {W}x{H} {BD}-bit ({wl['ml']}), single thread; {ms:.1f} ms/frame"}


def copy_probe():
    """A plain device copy of 512 MiB, three times: calibrates the rocprof byte counters."""
    import torch
    n = 512 << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(1)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n, "launches": 3}))


def run_multiview(args, with_ssim):
    """24 uneven views x 4 frames: one captured graph replay per frame set vs the plain loop."""
    from ivpsnr_amd import IVPSNR, IVSSIM, PSNR, WSPSNR, Engine, default_params
    from ivpsnr_amd.synth import make_pair
    import numpy as np

    bd, nframes = 10, 4
    metrics = PSNR | WSPSNR | IVPSNR | (IVSSIM if with_ssim else 0)
    print(f"[bench] multiview: generating {len(MULTIVIEW)} views", file=sys.stderr, flush=True)
    base = [make_pair(w, h, bd, seed=11 + i) for i, (w, h, _) in enumerate(MULTIVIEW)]
    eng = Engine(0)
    # frame f of a view: its synthetic pair circularly shifted by 64 f samples (new content, same statistics)
    sets = []
    for f in range(nframes):
        sets.append([(eng.picture(w, h, bd, 420, data=np.roll(p.ref, 64 * f)),
                      eng.picture(w, h, bd, 420, data=np.roll(p.tst, 64 * f)))
                     for (w, h, _), p in zip(MULTIVIEW, base)])
    prms = [default_params(erp=erp) for (_, _, erp) in MULTIVIEW]
    eng.sync()
    print("[bench] multiview: capturing and autotuning one graph per frame set", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    batches = [eng.batch(sets[0], [metrics] * len(sets[0]), prms)]
    info = batches[0].info()
    # the other frame sets have the same views: they take the tuned shape from the context's cache
    # without tuning again, and their graph instances get the same check
    batches += [eng.batch(s, [metrics] * len(s), prms) for s in sets[1:]]
    tune_s = time.perf_counter() - t0

    def step_plain():
        return [[eng.frame(p0, p1, metrics, prm) for (p0, p1), prm in zip(s, prms)] for s in sets]

    def step_batched():
        # every frame set launched, then collected: the host finalises set k while the device
        # computes the sets after it (qmiv_batch_launch / qmiv_batch_collect)
        for bt in batches:
            bt.launch()
        return [bt.collect() for bt in batches]

    def timed(step, steps, warmup):
        for _ in range(warmup):
            step()
        eng.sync()
        t = time.perf_counter()
        for _ in range(steps):
            r = step()
        eng.sync()
        return time.perf_counter() - t, r

    print(f"[bench] multiview plain loop: warmup {args.warmup}, {args.steps} steps", file=sys.stderr, flush=True)
    el_plain, r_plain = timed(step_plain, args.steps, args.warmup)
    print(f"[bench] multiview batched: warmup {args.warmup}, {args.steps} steps", file=sys.stderr, flush=True)
    el_b, r_b = timed(step_batched, args.steps, args.warmup)
    for fp, fb in zip(r_plain, r_b):
        for a, b in zip(fp, fb):
            assert a.ivpsnr == b.ivpsnr and list(a.gcd) == list(b.gcd) and list(a.wspsnr) == list(b.wspsnr)
            if with_ssim:
                assert a.ivssim == b.ivssim
    mpix = nframes * sum(w * h for (w, h, _) in MULTIVIEW) / 1e6
    ms_per_step = el_b / args.steps * 1e3
    value = mpix / (ms_per_step / 1e3)
    alg = nframes * sum(frame_pair_bytes(w, h, bd) for (w, h, _) in MULTIVIEW)
    achieved = alg / (ms_per_step / 1e3) / 1e9
    out = {
        "metric": "Batched mixed-view Mpix/s (24 views x 4 frames, 10b 420) + %HBM roofline; batched vs plain",
        "value": round(value, 2), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32+f64" if with_ssim else "int32", "data": "synthetic",
        "config": {"workload": "24 views (8x4096x2048 ERP, 8x1920x1080, 8x2048x1088) x 4 frames yuv420p10le, "
                               + ("IV-PSNR+WS-PSNR+PSNR+IV-SSIM" if with_ssim else "IV-PSNR+WS-PSNR+PSNR")
                               + ", B=2, GCD on",
                   "mode": "batched: one captured HIP graph replay per frame set, all sets launched then "
                           "collected (qmiv_batch_launch / qmiv_batch_collect)",
                   "plain_loop_mpix_s": round(mpix / (el_plain / args.steps), 2),
                   "batched_vs_plain": round(el_plain / el_b, 3),
                   "autotuned": info, "instances_per_frame_set": [bt.info()["instances"] for bt in batches],
                   "graph_capture_and_tuning_s": round(tune_s, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                     "kernel": "whole frame set (graph replay)", "bytes_per_step": alg},
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    for bt in batches:
        bt.close()
    eng.close()


def run_grad(args):
    """IV-SSIM forward + gradients of 8 fp32 picture pairs 3 x 1024 x 1024 in [0, 1]: a smooth
    random field, a copy shifted by up to 2 pels per 32 x 32 block with N(0, 0.02) noise."""
    import torch  # first: the engine library then shares torch's HIP runtime (one runtime per process)
    from ivpsnr_amd import Engine
    from ivpsnr_amd.synth import grad_batch

    n, c, h, w = 8, 3, 1024, 1024
    maxv = 1023.0
    tst, ref = grad_batch(n, c, h, w, seed=5)
    gt, gr = torch.empty_like(tst), torch.empty_like(ref)
    torch.cuda.synchronize()
    eng = Engine(0)

    def step():
        # the batch's pairs together: one wait for all values, one for all gradients
        res = eng.ivssim_backward_batch(w, h, [tst[i].data_ptr() for i in range(n)],
                                        [ref[i].data_ptr() for i in range(n)], maxv, None,
                                        [gt[i].data_ptr() for i in range(n)], [gr[i].data_ptr() for i in range(n)])
        return [r["ivssim"] for r in res]

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        vals = step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms_per_step = el / args.steps * 1e3
    mpix = n * h * w / 1e6
    # fp64 flops per picture pair: 6 forward evaluations (the values of both directions; the maps of
    # the chosen direction come with them) + 3 gradient correlations (4 maps x 100 taps x 2), per
    # interior / whole pel
    flops = n * c * (6 * 1000 * (h - 10) * (w - 10) + 3 * 800 * h * w) / 3
    achieved = flops / (ms_per_step / 1e3) / 1e12
    out = {
        "metric": "Differentiable IV-SSIM forward+backward Mpix/s @3x1024x1024 fp32, batch 8",
        "value": round(mpix / (ms_per_step / 1e3), 2), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "IV-SSIM value + gradients of both pictures, 8 x 3 x 1024 x 1024 fp32 in [0,1], "
                               "max 1023, B=2", "ivssim_first": round(vals[0], 6)},
        "roofline": {"bound": "fp64-valu", "achieved": round(achieved, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None,
                     "kernel": "whole step (k_ssimf, k_ssimf_grad, CPS, gather/scatter)"},
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    eng.close()


def run_masked(args):
    """Masked PSNR + WS-PSNR + IV-PSNR of 8 frame pairs 1920x1080 yuv420p with an 8-bit object mask
    (binary ellipses at 255 over about a third of the picture): the frame set as one
    qmiv_frames_masked call (timed), and one qmiv_frame_masked per frame beside it."""
    from ivpsnr_amd import Engine, default_params
    from ivpsnr_amd.synth import make_mask, make_pair
    W, H, BD, nframes = 1920, 1080, 8, args.frames or 8
    print(f"[bench] masked-1080p8: generating {nframes} frame pairs and masks", file=sys.stderr, flush=True)
    pairs = [make_pair(W, H, BD, seed=1 + f) for f in range(nframes)]
    masks = [make_mask(W, H, 8, seed=1 + f, kind="objects") for f in range(nframes)]
    prm = default_params()
    prm.search_range = args.search_range
    eng = Engine(0)
    pics = [(eng.picture(W, H, BD, 420, data=p.ref), eng.picture(W, H, BD, 420, data=p.tst),
             eng.picture(W, H, 8, 420, data=m)) for p, m in zip(pairs, masks)]
    eng.sync()
    metrics = 1 | 2 | 4

    def step_set():
        return eng.frames_masked(pics, metrics, prm)[0]

    def step_loop():
        return [eng.frame_masked(a, b, m, metrics, prm) for a, b, m in pics]

    def timed(steps, warmup, profile, step=step_set):
        for _ in range(warmup):
            step()
        eng.profile(profile)
        eng.profile_read(reset=True)
        eng.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            r = step()
        eng.sync()
        el = time.perf_counter() - t0
        pr = eng.profile_read(reset=True)
        eng.profile(False)
        return el, pr, r

    mpix = nframes * W * H / 1e6
    # Kernel durations for the roofline come from frames run one at a time (events around launches
    # inside the frame set also hold the waits for CUs that the other frames occupy).
    prof_steps = max(1, min(args.steps, 5))
    if args.legs == "profile":
        print(f"[bench] masked per-frame profile leg: warmup 1, {prof_steps} steps", file=sys.stderr, flush=True)
        el_p, prof_alone, _ = timed(prof_steps, 1, True, step_loop)
        print(json.dumps({"legs": "profile", "steps": prof_steps, "search_range": args.search_range,
                          "mpix_s_profiled_plain_loop":
                          round(mpix / (el_p / prof_steps), 2),
                          "kernels_us_per_frame_alone": {k: round(v[0] / nframes / prof_steps * 1e3, 2)
                                                         for k, v in prof_alone.items()}}))
        eng.close()
        return
    print(f"[bench] masked: warmup {args.warmup}, {args.steps} steps", file=sys.stderr, flush=True)
    elapsed, prof, res = timed(args.steps, args.warmup, True)
    prof_in_set = prof
    plain_mpix_s = stream_mpix_s = batched_mpix_s = pipelined_mpix_s = None
    tuned = None
    mode = "stream (qmiv_frames_masked: frame set enqueued over three streams, one wait)"
    if args.legs != "stream":
        el2, _, _ = timed(args.steps, 0, False)
        elapsed = min(elapsed, el2)
        stream_mpix_s = mpix / (elapsed / args.steps)
        el_loop, _, res_loop = timed(args.steps, 1, False, step_loop)
        plain_mpix_s = mpix / (el_loop / args.steps)
        assert [r.as_dict() for r in res] == [r.as_dict() for r in res_loop], "frame set differs from the per-frame calls"
        # The same frame set captured once (qmiv_batch_create_masked, branch count tuned) and
        # replayed: one replay per step (run), and pipelined with a second batch over the same
        # pictures (launch the next replay, then collect the previous one, so the host finalises
        # one set while the device computes the next).
        print("[bench] masked: capturing and tuning the masked batches", file=sys.stderr, flush=True)
        bt_a = eng.batch_masked(pics, metrics, prm)
        bt_b = eng.batch_masked(pics, metrics, prm)
        tuned = bt_a.info()
        want = [r.as_dict() for r in res]
        assert [r.as_dict() for r in bt_a.run()] == want, "masked batch differs from the frame set"
        assert [r.as_dict() for r in bt_b.run()] == want, "masked batch differs from the frame set"
        el_b, _, _ = timed(args.steps, args.warmup, False, bt_a.run)
        batched_mpix_s = mpix / (el_b / args.steps)

        def pipelined(steps):
            bt_a.launch()
            cur, nxt = bt_a, bt_b
            eng.sync()
            out = None
            t0 = time.perf_counter()
            for _ in range(steps):
                nxt.launch()
                out = cur.collect()
                cur, nxt = nxt, cur
            last = cur.collect()
            el = time.perf_counter() - t0
            return el, out, last

        pipelined(args.warmup)
        el_p, out_p, last_p = pipelined(args.steps)
        assert [r.as_dict() for r in out_p] == want and [r.as_dict() for r in last_p] == want
        pipelined_mpix_s = mpix / (el_p / args.steps)
        elapsed = el_p
        mode = ("pipelined batches (qmiv_batch_create_masked: the frame set captured into a graph, tuned "
                "branches; two batches, the next replay launched before the previous one is collected)")
        bt_a.close()
        bt_b.close()
        print(f"[bench] masked per-frame profile leg: warmup 1, {prof_steps} steps", file=sys.stderr, flush=True)
        _, prof, _ = timed(prof_steps, 1, True, step_loop)
    else:
        prof_steps = args.steps
    ms_per_step = elapsed / args.steps * 1e3
    kkey = "cps" if args.search_range == 2 else "cps_general"
    k_ms, k_n = prof.get(kkey, (0.0, 0))
    k_avg_ms = k_ms / max(k_n, 1)
    k_set_ms, k_set_n = prof_in_set.get(kkey, (0.0, 0))
    per_launch = frame_pair_bytes(W, H, BD) + W * H  # two pictures and the mask's luma plane
    achieved = per_launch / (k_avg_ms / 1e3) / 1e9 if k_avg_ms > 0 else 0.0
    cpu = None
    if not args.no_cpu_baseline:
        ref_bin = os.path.join(ROOT, "oracle", "_ref", "qmiv_ref_v3")
        cores = min(16, os.cpu_count() or 1)
        if os.path.exists(ref_bin):
            with tempfile.TemporaryDirectory() as td:
                f0, f1, fm = (os.path.join(td, n) for n in ("p0.yuv", "p1.yuv", "m.yuv"))
                sample = 4
                with open(f0, "wb") as a, open(f1, "wb") as b, open(fm, "wb") as c:
                    for i in range(sample):
                        a.write(pairs[i].ref.tobytes())
                        b.write(pairs[i].tst.tobytes())
                        c.write(masks[i].tobytes())
                cmd = [ref_bin, "-w", str(W), "-h", str(H), "-bd", str(BD), "-cf", "420", "-i0", f0, "-i1", f1,
                       "-im", fm, "-bdm", "8", "-cfm", "420", "-ml", "PSNR,WSPSNR,IVPSNR", "-nth", str(cores),
                       "-time"]
                try:
                    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
                    if r.returncode == 0:
                        t = json.loads(r.stdout)["time_ms_per_frame"]
                        ms = sum(t[k] for k in ("validate", "preproc", "gcd", "psnr", "wspsnr", "ivpsnr"))
                        cpu = {"value": round(W * H / (ms * 1e3), 3), "unit": "Mpix/s", "cores": cores,
                               "kind": "reference",
                               "sample": f"{sample} frames {W}x{H} 8-bit with their masks (PSNR,WSPSNR,IVPSNR "
                                         f"masked) incl. validation, preprocessing and GCD, {cores} threads; "
                                         f"{ms:.1f} ms/frame"}
                except (subprocess.TimeoutExpired, OSError, ValueError, KeyError):
                    pass
    out = {
        "metric": "masked IV-PSNR Mpix/s @1920x1080 8b 420",
        "value": round(mpix / (ms_per_step / 1e3), 2),
        "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": "masked PSNR+WS-PSNR+IV-PSNR, 8 frame pairs 1920x1080 yuv420p, 8-bit object "
                               f"masks, B={args.search_range}, GCD on", "frames_per_step": nframes,
                   "mode": mode,
                   "plain_loop_mpix_s": round(plain_mpix_s, 2) if plain_mpix_s else None,
                   "stream_mpix_s": round(stream_mpix_s, 2) if stream_mpix_s else None,
                   "batched_mpix_s": round(batched_mpix_s, 2) if batched_mpix_s else None,
                   "pipelined_mpix_s": round(pipelined_mpix_s, 2) if pipelined_mpix_s else None,
                   "value_vs_plain": round(mpix / (ms_per_step / 1e3) / plain_mpix_s, 3) if plain_mpix_s else None,
                   "autotuned": tuned,
                   "ivpsnr_last_frame": round(res[-1].ivpsnr, 6), "nnm_last_frame": res[-1].num_non_masked},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": read_traffic("k_cps420s_masked_1920x1080_8b.txt") if kkey == "cps" else None,
                     "kernel": "k_cps420s (masked)" if kkey == "cps" else "k_cps_masked",
                     "kernel_avg_us": round(k_avg_ms * 1e3, 2),
                     "kernel_avg_us_in_set": round(k_set_ms / max(k_set_n, 1) * 1e3, 2),
                     "timing": "kernel_avg_us: HIP events around each launch with the frame's kernels alone on the "
                               "device (per-frame leg); kernel_avg_us_in_set: inside the timed frame-set steps",
                     "bytes_per_launch": per_launch},
        "kernels_us_per_frame": {k: round(v[0] / nframes / prof_steps * 1e3, 2) for k, v in prof.items()},
        "kernels_us_per_frame_in_set": {k: round(v[0] / nframes / args.steps * 1e3, 2) for k, v in prof_in_set.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    eng.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["multiview24", "multiview24-ivpsnr", "grad-ivssim1024",
                                                               "masked-1080p8"],
                    default="ivpsnr-erp4k10")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--frames", type=int, default=0, help="frame pairs per step (default: the workload's)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain device copy of 512 MiB (calibrates rocprof byte counters)")
    ap.add_argument("--search-range", type=int, default=2,
                    help="masked-1080p8 only: the CPS search range B (any B other than 2 runs the general masked "
                         "search kernel k_cps_masked; the line is then not the workload's figure)")
    ap.add_argument("--legs", choices=["all", "stream", "profile"], default="all",
                    help="stream: run only the timed frame-set steps (no plain loop, no upload leg); profile: run "
                         "only the per-frame leg whose kernel durations the roofline figure is taken from, so that "
                         "a kernel trace of the run holds exactly those launches")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine's workload runs on a single MI355X "
              "(multi-GPU sharding is a design note, not run here).", file=sys.stderr)
        sys.exit(2)

    if args.workload.startswith("multiview24"):
        run_multiview(args, with_ssim=args.workload == "multiview24")
        return
    if args.workload == "grad-ivssim1024":
        run_grad(args)
        return
    if args.workload == "masked-1080p8":
        run_masked(args)
        return

    from ivpsnr_amd import Engine, default_params
    import ivpsnr_amd
    from ivpsnr_amd.synth import make_pair

    wl = WORKLOADS[args.workload]
    W, H, BD = wl["w"], wl["h"], wl["bd"]
    nframes = args.frames or wl["frames"]
    print(f"[bench] {args.workload}: generating {nframes} synthetic frame pairs {W}x{H} {BD}-bit", file=sys.stderr,
          flush=True)
    pairs = [make_pair(W, H, BD, seed=1 + f) for f in range(nframes)]
    metrics = 0
    for name in wl["ml"].split(","):
        metrics |= getattr(ivpsnr_amd, name)
    prm = default_params(erp=wl["erp"])

    eng = Engine(0)
    pics = [(eng.picture(W, H, BD, 420, data=p.ref), eng.picture(W, H, BD, 420, data=p.tst)) for p in pairs]
    eng.sync()

    def step_plain():
        # The reference's frame loop shape: one call and one wait per frame.
        return [eng.frame(p0, p1, metrics, prm) for p0, p1 in pics]

    def step_stream():
        # The whole frame set enqueued on the engine stream, one wait.
        return eng.frames(pics, metrics, prm)

    def timed(step, steps, warmup, profile):
        for _ in range(warmup):
            step()
        eng.profile(profile)
        eng.profile_read(reset=True)
        eng.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            r = step()
        eng.sync()
        el = time.perf_counter() - t0
        pr = eng.profile_read(reset=True)
        eng.profile(False)
        return el, pr, r

    key = "ivpsnr" if "IVPSNR" in wl["ml"] else "ivssim"
    # Kernel durations for the roofline come from frames run one at a time (the plain loop with
    # kernel events): in the frame-set steps up to three frames share the device, and an event pair
    # around a launch there also holds the time the launch waits for CUs the other frames occupy.
    prof_steps = max(1, min(args.steps, 5))
    if args.legs == "profile":
        print(f"[bench] per-frame profile leg: warmup 1, {prof_steps} steps", file=sys.stderr, flush=True)
        el_p, prof_alone, _ = timed(step_plain, prof_steps, 1, True)
        mpix = nframes * W * H / 1e6
        print(json.dumps({"legs": "profile", "steps": prof_steps,
                          "mpix_s_profiled_plain_loop":
                          round(mpix / (el_p / prof_steps), 2),
                          "kernels_us_per_frame_alone": {k: round(v[0] / nframes / prof_steps * 1e3, 2)
                                                         for k, v in prof_alone.items()}}))
        eng.close()
        return
    only_stream = args.legs == "stream"
    el_plain = None
    if not only_stream:
        print(f"[bench] plain loop: warmup {args.warmup}, {args.steps} steps", file=sys.stderr, flush=True)
        el_plain, _, res_plain = timed(step_plain, args.steps, args.warmup, False)
    print(f"[bench] stream: warmup {args.warmup}, {args.steps} steps", file=sys.stderr, flush=True)
    elapsed, prof, res = timed(step_stream, args.steps, args.warmup, True)
    if not only_stream:
        assert [getattr(r, key) for r in res] == [getattr(r, key) for r in res_plain]
        print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s (plain loop {el_plain:.3f} s)", file=sys.stderr,
              flush=True)
        # Kernel events serialise nothing but add a few microseconds per launch: time once more without them.
        el_noprof, _, _ = timed(step_stream, args.steps, 0, False)
        elapsed = min(elapsed, el_noprof)
    el_stream = elapsed

    # Pipelined frame sets: a second set of resident frame pairs, each set captured once as an
    # autotuned graph (qmiv_batch_create; the second set takes the first one's choice from the
    # context's cache). A step replays one set; the next set is launched before the previous one is
    # collected (qmiv_batch_launch / qmiv_batch_collect), so the host finalises one set while the
    # device computes the next. The timed region launches and collects all of its steps.
    el_pipe = pipe_info = None
    if not only_stream:
        print(f"[bench] pipelined frame sets: second set, capture and autotuning", file=sys.stderr, flush=True)
        pairs2 = [make_pair(W, H, BD, seed=1 + nframes + f) for f in range(nframes)]
        pics2 = [(eng.picture(W, H, BD, 420, data=p.ref), eng.picture(W, H, BD, 420, data=p.tst)) for p in pairs2]
        eng.sync()
        t_tune = time.perf_counter()
        bts = [eng.batch(s_, [metrics] * nframes, [prm] * nframes) for s_ in (pics, pics2)]
        tune_s = time.perf_counter() - t_tune

        def pipeline(steps):
            first = None
            bts[0].launch()
            for k in range(1, steps):
                bts[k % 2].launch()
                r_ = bts[(k - 1) % 2].collect()
                first = r_ if first is None else first
            r_ = bts[(steps - 1) % 2].collect()
            return r_ if first is None else first  # the first collected set is set 0

        if args.warmup:
            pipeline(args.warmup)
        print(f"[bench] pipelined frame sets: {args.steps} steps", file=sys.stderr, flush=True)
        eng.sync()
        t0 = time.perf_counter()
        res_pipe = pipeline(args.steps)
        eng.sync()
        el_pipe = time.perf_counter() - t0
        assert [getattr(r, key) for r in res_pipe] == [getattr(r, key) for r in res], "pipelined sets differ"
        pipe_info = dict(bts[0].info(), instances_second_set=bts[1].info()["instances"],
                         capture_and_tuning_s=round(tune_s, 2))
        for bt in bts:
            bt.close()
        if wl["value_mode"] == "pipelined":
            elapsed = el_pipe

    # Host-to-device inclusive rate (not `value`): every step first copies its frame pairs from host
    # memory into the pictures, each frame's copies on the copy stream overlapping the frames before
    # it (qmiv_frames_upload): from pinned host buffers (qmiv_host_alloc, filled once before timing,
    # as a frame reader would hold them) and, beside it, from the ordinary numpy arrays.
    el_up = el_up_pageable = None
    up_steps = max(1, min(args.steps, 5))
    if not only_stream:
        import numpy as np
        src_pageable = [(np.ascontiguousarray(p.ref).view(np.uint8).reshape(-1),
                         np.ascontiguousarray(p.tst).view(np.uint8).reshape(-1)) for p in pairs]
        src_pinned = []
        for a, b in src_pageable:
            ha, hb = eng.host_array(a.nbytes), eng.host_array(b.nbytes)
            ha[:] = a
            hb[:] = b
            src_pinned.append((ha, hb))
        for src in (src_pinned, src_pageable):
            got = eng.frames_upload(pics, src, metrics, prm)
            assert [getattr(r, key) for r in got] == [getattr(r, key) for r in res]
        el_up = timed(lambda: eng.frames_upload(pics, src_pinned, metrics, prm), up_steps, 1, False)[0]
        el_up_pageable = timed(lambda: eng.frames_upload(pics, src_pageable, metrics, prm), up_steps, 1, False)[0]
    prof_in_set = prof
    if not only_stream:
        print(f"[bench] per-frame profile leg: warmup 1, {prof_steps} steps", file=sys.stderr, flush=True)
        _, prof, _ = timed(step_plain, prof_steps, 1, True)
    else:
        prof_steps = args.steps

    ms_per_step = elapsed / args.steps * 1e3
    mpix = nframes * W * H / 1e6
    value = mpix / (ms_per_step / 1e3)
    k_ms, k_n = prof.get(wl["kernel"], (0.0, 0))
    k_avg_ms = k_ms / max(k_n, 1)
    k_set_ms, k_set_n = prof_in_set.get(wl["kernel"], (0.0, 0))
    timing_note = ("kernel_avg_us: HIP events around each launch with the frame's kernels alone on the device "
                   "(per-frame leg); kernel_avg_us_in_set: the same launches inside the timed frame-set steps, "
                   "waits for CUs held by the other frames included") if not only_stream else \
        "kernel_avg_us: HIP events around each launch inside the timed frame-set steps"
    if wl["kernel"] == "cps":
        per_launch = frame_pair_bytes(W, H, BD)
        achieved = per_launch / (k_avg_ms / 1e3) / 1e9 if k_avg_ms > 0 else 0.0
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": read_traffic(wl["traffic_file"]),
                "kernel": wl["cps_kernel"], "kernel_avg_us": round(k_avg_ms * 1e3, 2),
                "kernel_avg_us_in_set": round(k_set_ms / max(k_set_n, 1) * 1e3, 2), "timing": timing_note,
                # the kernel's launches of one step, each alone on the device, summed: in pipelined frame
                # sets the launches of different frames overlap, so a step can take less than this sum
                "kernel_us_per_step": round(k_ms / prof_steps * 1e3, 1), "bytes_per_launch": per_launch,
                # the whole job's compulsory bytes over its time per step (every kernel and overlap included)
                "achieved_whole_job": round(nframes * per_launch / (ms_per_step / 1e3) / 1e9, 1),
                # The search is bound by int32 VALU issue, not HBM: 25 candidates x 2 directions per
                # pel, and the VALU busy fraction of the same kernel from its PMC summary.
                "candidate_evals_per_s": round(50 * W * H / (k_avg_ms / 1e3), -6) if k_avg_ms > 0 else 0.0,
                "valu_busy_frac": read_profile_value(wl["traffic_file"], "valu_busy_frac")}
    else:
        flops = ssim_flops_per_frame(W, H)
        frames_prof = prof_steps * nframes
        launches = k_n / max(frames_prof, 1)  # SSIM-family launches per frame (fused, batched, single)
        per_launch = flops / max(launches, 1)
        achieved = flops / (k_ms / frames_prof / 1e3) / 1e12 if k_ms > 0 else 0.0
        roof = {"bound": "fp64-valu", "achieved": round(achieved, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": read_traffic(wl["traffic_file"]),
                "kernel": "k_ssim_multi + k_ssim_batch", "kernel_avg_us": round(k_avg_ms * 1e3, 2),
                "kernel_avg_us_in_set": round(k_set_ms / max(k_set_n, 1) * 1e3, 2), "timing": timing_note,
                "launches_per_frame": round(launches, 2), "flops_per_launch": round(per_launch)}

    cpu = None if args.no_cpu_baseline else cpu_baseline(wl, pairs, 4 if wl["kernel"] == "cps" else 2)
    out = {
        "metric": wl["metric"],
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": wl["dtype"],
        "data": "synthetic",
        "config": {"workload": wl["desc"], "frames_per_step": nframes,
                   "mode": ("pipelined frame sets (two sets of resident frame pairs, each an autotuned captured "
                            "graph; per step one replay, launched before the previous set is collected: "
                            "qmiv_batch_launch / qmiv_batch_collect)") if el_pipe and wl["value_mode"] == "pipelined"
                   else "stream (qmiv_frames: frame set enqueued, one wait)",
                   "stream_mpix_s": round(mpix / (el_stream / args.steps), 2),
                   "pipelined_mpix_s": round(mpix / (el_pipe / args.steps), 2) if el_pipe else None,
                   "autotuned": pipe_info,
                   "plain_loop_mpix_s": round(mpix / (el_plain / args.steps), 2) if el_plain else None,
                   "stream_vs_plain": round(el_plain / el_stream, 3) if el_plain else None,
                   "value_vs_plain": round(el_plain / elapsed, 3) if el_plain else None,
                   "upload_inclusive_mpix_s": round(mpix / (el_up / up_steps), 2) if el_up else None,
                   "upload_inclusive_pageable_mpix_s":
                       round(mpix / (el_up_pageable / up_steps), 2) if el_up_pageable else None,
                   key + "_last_frame": round(getattr(res[-1], key), 6)},
        "roofline": roof,
        "kernels_us_per_frame": {k: round(v[0] / nframes / prof_steps * 1e3, 2) for k, v in prof.items()},
        "kernels_us_per_frame_in_set": {k: round(v[0] / nframes / args.steps * 1e3, 2) for k, v in prof_in_set.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    eng.close()


if __name__ == "__main__":
    main()
