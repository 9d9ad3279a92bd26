"""Benchmark: plain GPU decode + develop of ProRes RAW frames on one MI355X.

Workload (one step = one frame): the 4096x2160 BGGR frame of the
plain-GPU-at-scale configuration (bitstream v1, tile alignment 3, qscale 45,
seeded quant matrix and 8-point curve, 2% of tiles corrupted) is decoded with
the Apple transform, linearized in the same kernel, and developed to Rec.2020
planes. The packet is resident in device memory before timing starts. Frames
stream through the native pipeline (prr_pipeline): frame i + 1's decode runs
on one HIP stream beside frame i's develop on another. The serial step
(decode then develop, one HIP graph a frame) is reported beside it.

Prints one JSON line. `value` is Gsensels/s (visible width x height per
second). `roofline` covers the develop kernel (the dominant one) against
HBM bandwidth; `cpu_baseline` times the oracle (a C restatement of the
crate's CPU path, threaded like its rayon code) on a bounded sample.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--width", type=int, default=4096)
    ap.add_argument("--height", type=int, default=2160)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-frames", type=int, default=64)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--learned-frames", type=int, default=32)
    ap.add_argument("--no-clip", action="store_true")
    ap.add_argument("--no-dng", action="store_true")
    ap.add_argument("--no-orient", action="store_true")
    ap.add_argument("--clip-frames", type=int, default=160)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 10 times (the reference reading "
                         "of the HBM byte counters) and exit")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_plain4096x2160.txt")


def measured_traffic(kernel, width, height):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the
    counter reads for a copy of known size), when it covers this frame
    size; else None."""
    if (width, height) != (4096, 2160) or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        f = line.split()
        if not f or f[0] != kernel:
            continue
        kv = dict(x.split("=") for x in f[1:])
        return int(1024 * (float(kv["fetch_kib"]) * float(kv["fetch_scale"]) +
                           float(kv["write_kib"]) * float(kv["write_scale"])))
    return None


UTIL_FILE = os.path.join(ROOT, "profiles", "issue_utilization.txt")


def measured_valu_rate(kernel, width, height):
    """VALU instructions `kernel` issues per SIMD-cycle on this frame size,
    from the committed counter summary (SQ_INSTS_VALU over cycles x SIMDs),
    or None."""
    if (width, height) != (4096, 2160) or not os.path.exists(UTIL_FILE):
        return None
    for line in open(UTIL_FILE):
        f = line.split()
        if len(f) > 2 and f[0] == kernel and f[1] == "plain4096x2160":
            kv = dict(x.split("=") for x in f[2:])
            return float(kv["valu_per_simd_cycle"])
    return None


def calibrate_copy(mib):
    """Ten device-to-device copies of `mib` MiB: each moves mib MiB in and
    mib MiB out, the known bytes the FETCH_SIZE / WRITE_SIZE counters are
    scaled against."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(10):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_mib": mib, "copies": 10, "ok": bool((dst == src).all())}))


def main():
    args = parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are out of scope "
                                   "(frames shard across GPUs with no exchange)"}))
        sys.exit(2)
    import torch

    import proresraw_rs_amd as P
    from proresraw_rs_amd import _native as N
    from proresraw_rs_amd import synth as S
    from proresraw_rs_amd.workloads import config1

    dev = torch.device("cuda:0")
    lib = N.lib()
    ctx = P.context()
    stream = torch.cuda.Stream()
    sc = config1(args.width, args.height)
    packet = S.encode_scene(sc)
    hdr, info = P.parse_frame_info(packet)
    raw_hdr = hdr._raw
    w, h = hdr.width, hdr.height
    cw, ch = info.coded_width, info.coded_height
    npix = w * h

    padded = ((len(packet) + 3) & ~3) + 16
    d_packet = torch.zeros(padded, dtype=torch.uint8, device=dev)
    d_packet[:len(packet)] = torch.frombuffer(bytearray(packet), dtype=torch.uint8).to(dev)
    d_samples = torch.empty(cw * ch, dtype=torch.int16, device=dev)
    d_linear = torch.empty(npix, dtype=torch.int16, device=dev)
    d_flags = torch.empty(max(info.nb_tiles, 1), dtype=torch.uint8, device=dev)
    d_status = torch.zeros(4, dtype=torch.int32, device=dev)
    planes = [torch.empty(npix, dtype=torch.float32, device=dev) for _ in range(3)]
    err = N.PrrError()
    hs = ctypes.c_void_p(stream.cuda_stream)

    def decode():
        rc = lib.prr_decode_frame_device(
            ctx.handle, ctypes.byref(raw_hdr), ctypes.byref(info), d_packet.data_ptr(),
            len(packet), 0, d_samples.data_ptr(), d_linear.data_ptr(), d_flags.data_ptr(),
            d_status.data_ptr(), hs, ctypes.byref(err))
        if rc:
            raise P.Error(err)

    def develop():
        # develop(&frame, Rec2020) from the linearized picture the decode
        # wrote (the same planes as from the code values): the noise model
        # from the frame's max_qscale, read on the device.
        rc = lib.prr_develop_linear_device(ctx.handle, ctypes.byref(raw_hdr), d_linear.data_ptr(),
                                           d_status.data_ptr(), 0, 0.0, 0.0, 2,
                                           planes[0].data_ptr(), planes[1].data_ptr(),
                                           planes[2].data_ptr(), hs, ctypes.byref(err))
        if rc:
            raise P.Error(err)

    def step():
        decode()
        develop()

    step()
    stream.synchronize()
    st = d_status.cpu().numpy().view(np.uint32)
    if st[0] != 0xFFFFFFFF or st[3]:
        raise RuntimeError(f"device status {st}")

    # Per-kernel durations with HIP events on the launch stream.
    def timed(fn, n=30):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        fn()
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_time(e1) / n

    t_decode_ms = timed(decode)
    t_develop_ms = timed(develop)

    # The serial step (decode then develop on one stream, one HIP graph a
    # frame), reported beside the pipelined stream of frames below.
    graph = None
    if not args.no_graph:
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.stream(stream):
            with torch.cuda.graph(graph, stream=stream):
                step()
        stream.synchronize()

    def run_step():
        if graph is not None:
            graph.replay()
        else:
            step()

    with torch.cuda.stream(stream):
        for _ in range(args.warmup):
            run_step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run_step()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
    serial_ms = 1e3 * elapsed / args.steps
    serial_value = npix * args.steps / elapsed / 1e9

    # The pipelined stream of frames (prr_pipeline): frame i + 1's decode on
    # one HIP stream beside frame i's develop on another, four buffer sets.
    # One step = one frame submitted; the timed region is bracketed by device
    # synchronization, so it holds exactly `steps` whole frames.
    from proresraw_rs_amd.pipeline import Pipeline
    pipe = Pipeline(depth=4)
    pipe_status = torch.zeros(4, dtype=torch.int32, device=dev)

    def submit():
        pipe.submit(hdr, info, d_packet, len(packet), planes, stage=P.Stage.Rec2020,
                    status=pipe_status)

    for _ in range(args.warmup):
        submit()
    pipe.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        submit()
    pipe.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st = pipe_status.cpu().numpy().view(np.uint32)
    if st[0] != 0xFFFFFFFF or st[3]:
        raise RuntimeError(f"device status {st}")
    # Kernel durations inside the pipelined stream, with HIP events on the
    # streams the kernels run on (the develop shares the chip with the next
    # frame's decode there).
    pipe.timing(True)
    for _ in range(min(args.steps, 60)):
        submit()
    pipe_decode_ms, pipe_develop_ms, _ = pipe.timing(False)
    pipe.close()
    ms_per_step = 1e3 * elapsed / args.steps
    value = npix * args.steps / elapsed / 1e9

    # Develop roofline: algorithmic bytes = linear values read (2 B per visible
    # sensel) + three f32 planes written (12 B per visible sensel).
    dev_bytes = 14.0 * npix
    achieved = dev_bytes / (pipe_develop_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(achieved / 8000.0, 4),
                "traffic": measured_traffic("k_develop", w, h),
                "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT),
                "valu_insts_per_simd_cycle": measured_valu_rate("k_develop", w, h),
                "valu_source": os.path.relpath(UTIL_FILE, ROOT),
                "kernel": "k_develop", "kernel_ms": round(pipe_develop_ms, 4),
                "kernel_ms_alone": round(t_develop_ms, 4),
                "bytes_per_launch": int(dev_bytes),
                "decode_ms": round(pipe_decode_ms, 4),
                "decode_ms_alone": round(t_decode_ms, 4),
                "decode_tiles_per_s": round(info.nb_tiles / (t_decode_ms * 1e-3), 1)}

    batched = None if args.no_batch else batched_bench(args)
    learned = None if args.no_learned else learned_bench(args, value)
    clip = None if args.no_clip else clip_bench(args)
    dng = None if args.no_dng else dng_bench(args)
    orient = None if args.no_orient else orient_bench(args)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(packet, args.cpu_seconds)

    line = {
        "metric": "Gsensels/s decode+develop, plain vs batched+learned demosaic at >=45 dB PSNR",
        "value": round(value, 3),
        "unit": "Gsensels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16+f32",
        "data": "synthetic",
        "config": {"workload": f"plain GPU decode(Apple)+linearize+develop(Rec2020), {w}x{h} BGGR "
                               "v1 a3 q45, seeded qmat+curve, 2% corrupt tiles; a stream of "
                               "frames, decode of frame i+1 beside develop of frame i",
                   "packet_bytes": len(packet), "pipeline_depth": 4,
                   "serial_value": round(serial_value, 3), "serial_ms_per_step": round(serial_ms, 4),
                   "serial_graph": graph is not None},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "batched_learned": learned,
        "clip": clip,
        "dng_tiles": dng,
        "orient": orient,
    }
    if cpu:
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)


def batched_bench(args):
    """The mixed-batch configuration: frames of 1920x1080 / 4096x2160 /
    5760x3240 / 4032x3024, versions 0 and 1, alignments 0-4, qscale in
    [15, 45], RGGB and BGGR; half full decode + develop(WhiteBalanced), half
    reduced at Half / Quarter / Eighth (Greens::Average on a quarter of the
    Half ones); one persistent launch per batch, replayed from a HIP graph.
    16 distinct packets are encoded and cycled to fill the batch."""
    import torch

    import proresraw_rs_amd as P
    from proresraw_rs_amd import synth as S
    from proresraw_rs_amd.batch import Batch
    from proresraw_rs_amd.workloads import seeded_curve, seeded_qmat

    sizes = [(1920, 1080), (4096, 2160), (5760, 3240), (4032, 3024)]
    distinct = []
    for i in range(16):
        w, h = sizes[i % 4]
        version = (i // 4) % 2
        sc = S.scene(w, h, 3 * ((i // 2) % 2), seed=1000 + i, version=version, alignment=i % 5,
                     qscale=(15 + (i * 7) % 31, 15 + (i * 7) % 31),
                     qmat=seeded_qmat(i) if i % 3 == 0 else None,
                     curve=seeded_curve(i) if version == 1 else None)
        distinct.append(S.encode_scene(sc))
    b = Batch()
    sensels = 0
    n = args.batch_frames
    n_full = 0
    for k in range(n):
        pk = distinct[k % 16]
        hdr, _ = P.parse_frame_info(pk)
        sensels += hdr.width * hdr.height
        if k % 2 == 0:
            b.add_full(pk, P.Idct.Apple, develop=P.Stage.WhiteBalanced)
            n_full += 1
        else:
            j = k // 2
            scale = [P.Scale.Half, P.Scale.Quarter, P.Scale.Eighth][j % 3]
            greens = P.Greens.Average if scale == P.Scale.Half and j % 4 == 0 else P.Greens.First
            b.add_reduced(pk, scale, greens)
    s = torch.cuda.Stream()
    b.run(s)
    s.synchronize()
    for slot in b.slots:
        slot.check()
    tuned = b.autotune(s)
    b.run(s)
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            b.run(s)
    s.synchronize()
    reps = 5
    with torch.cuda.stream(s):
        g.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            g.replay()
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / reps
    info = b.info()
    return {"value": round(sensels / el / 1e9, 3), "unit": "Gsensels/s", "frames": n,
            "full_develop_frames": n_full, "ms_per_batch": round(el * 1e3, 3),
            "gsensels": round(sensels / 1e9, 4), "tile_groups": info["tile_groups"],
            "persistent_grid": info["grid"], "lds_bytes": info["lds_bytes"],
            "autotuned": {"spw": tuned["spw"], "chunk": tuned["chunk"], "waves": tuned["waves"],
                          "cached": tuned["cached"],
                          "trial_ms": [round(t["ms"], 3) for t in tuned.get("trials", [])]}}


def learned_bench(args, plain_value):
    """The end-to-end batched + learned configuration: 32 frames alternating
    8192x4320 (v1, alignment 4, qscale 15) and 5760x3240 (v0, alignment 3,
    qscale 45), RGGB. One persistent batched decode (entropy decode + IDCT +
    linearize) of all frames, then per frame the learned demosaic + white
    balance + Rec.2020 matrix with the on-device fallback to plain develop
    (noise outside the trained range, or any of 8 sampled 64x64 tiles below
    45 dB against plain develop). The same batch with plain develop is timed
    beside it. Two distinct packets of each size are encoded and cycled."""
    import torch

    import proresraw_rs_amd as P
    from proresraw_rs_amd import _native as N
    from proresraw_rs_amd import learned as L
    from proresraw_rs_amd import synth as S
    from proresraw_rs_amd.batch import Batch
    from proresraw_rs_amd.workloads import seeded_curve

    kinds = [dict(size=(8192, 4320), version=1, alignment=4, q=15, vendor=b"aprh"),
             dict(size=(5760, 3240), version=0, alignment=3, q=45, vendor=b"aprn")]
    distinct = []
    for i in range(4):
        k = kinds[i % 2]
        sc = S.scene(*k["size"], 0, seed=3000 + i, version=k["version"],
                     alignment=k["alignment"], qscale=(k["q"], k["q"]),
                     curve=seeded_curve(40 + i) if k["version"] == 1 else None)
        distinct.append(S.encode_scene(sc))
    b = Batch()
    frames = []
    sensels = 0
    for j in range(args.learned_frames):
        pk = distinct[j % 4]
        slot = b.add_full(pk, P.Idct.Apple, linear=True)
        h = slot.header
        sensels += h.width * h.height
        planes = [torch.empty(h.width * h.height, dtype=torch.float32, device="cuda")
                  for _ in range(3)]
        frames.append((slot, P.DemosaicParams.new(h, 0), planes))
    params, lo, hi = L.load_weights()
    net = L.Net(params)
    net.set_range(lo, hi)
    lib, ctx = N.lib(), P.context()
    s = torch.cuda.Stream()
    hs = ctypes.c_void_p(s.cuda_stream)
    err = N.PrrError()

    def learned_step():
        b.run(s)
        for slot, prm, planes in frames:
            net.develop(slot.header, slot.samples, prm, P.Stage.Rec2020, planes, status=slot.status,
                        check_tiles=8, min_db=45.0, stream=s)

    def plain_step():
        b.run(s)
        for slot, prm, planes in frames:
            rc = lib.prr_develop_frame_device(ctx.handle, ctypes.byref(slot.header._raw),
                                              slot.samples.data_ptr(), slot.status.data_ptr(), 0,
                                              0.0, 0.0, int(P.Stage.Rec2020), planes[0].data_ptr(),
                                              planes[1].data_ptr(), planes[2].data_ptr(), hs,
                                              ctypes.byref(err))
            if rc:
                raise P.Error(err)

    with torch.cuda.stream(s):
        b.run(s)
        s.synchronize()
        for slot, _, _ in frames:
            slot.check()
        b.autotune(s)
        # Fallbacks and sampled-tile PSNR, frame by frame (eager, once).
        fallbacks, worst = 0, float("inf")
        b.run(s)
        for slot, prm, planes in frames:
            net.develop(slot.header, slot.samples, prm, P.Stage.Rec2020, planes,
                        status=slot.status, check_tiles=8, min_db=45.0, stream=s)
            used, psnr = net.last_check()
            fallbacks += int(used)
            worst = min(worst, psnr)
        results = {}
        for name, fn in (("learned", learned_step), ("plain", plain_step)):
            fn()
            s.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=s):
                fn()
            g.replay()
            s.synchronize()
            reps = 3
            t0 = time.perf_counter()
            for _ in range(reps):
                g.replay()
            s.synchronize()
            results[name] = (time.perf_counter() - t0) / reps
    v_learned = sensels / results["learned"] / 1e9
    v_plain = sensels / results["plain"] / 1e9
    return {"value": round(v_learned, 3), "unit": "Gsensels/s", "frames": args.learned_frames,
            "gsensels": round(sensels / 1e9, 4), "ms_per_batch": round(results["learned"] * 1e3, 3),
            "batched_plain_value": round(v_plain, 3),
            "vs_batched_plain": round(v_learned / v_plain, 3),
            "vs_plain_gpu_path": round(v_learned / plain_value, 3),
            "fallback_frames": fallbacks, "min_sampled_tile_psnr_db": round(worst, 2),
            "workload": "32 frames alternating 8192x4320 v1 a4 q15 and 5760x3240 v0 a3 q45, "
                        "batched decode+linearize, learned demosaic+WB+Rec2020 with fallback"}


def clip_bench(args):
    """A movie of 5760x3240 RGGB frames (v0, alignment 3, qscale 45; 8
    distinct seeded scenes cycled) written to a temporary file: the
    sample table is read with the movie reader, then the clip is decoded and
    developed (WhiteBalanced) in windows of 16 frames, host reads of the next
    window overlapping the device work on the current one. Reported against
    the same window decoded with its packets already in HBM."""
    import tempfile

    import torch

    import proresraw_rs_amd as P
    from proresraw_rs_amd import mov as M
    from proresraw_rs_amd import synth as S
    from proresraw_rs_amd.clip import ClipDecoder

    distinct = [S.encode_scene(S.scene(5760, 3240, 0, seed=500 + i, version=0, alignment=3,
                                       qscale=(45, 45))) for i in range(8)]
    packets = [distinct[i % 8] for i in range(args.clip_frames)]
    fd, path = tempfile.mkstemp(suffix=".mov")
    os.close(fd)
    try:
        with open(path, "wb") as f:
            f.write(M.clip_movie(packets, fps=(24000, 1001)))
        t0 = time.perf_counter()
        track = M.read_video_track(path)
        t_table = time.perf_counter() - t0
        dec = ClipDecoder(path, window=16, develop=P.Stage.WhiteBalanced, threads=16)
        dec.run()  # allocations, first uploads
        dec.read_seconds = 0.0
        frames, el = dec.run()
        read_s = dec.read_seconds
        nbytes = sum(s.size for s in track.samples)
        sensels = 5760 * 3240 * frames
        # The same first window with its packets already resident in HBM.
        win = next(dec.frames(0, 16))
        k = 0
        batch = dec.batches[k]
        s = dec.stream
        import ctypes
        from proresraw_rs_amd import _native as N
        err = N.PrrError()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        reps = 5
        for _ in range(reps):
            if dec.lib.prr_batch_run(batch, ctypes.c_void_p(s.cuda_stream), ctypes.byref(err)):
                raise RuntimeError("batch run failed")
        e1.record(s)
        e1.synchronize()
        resident_ms = e0.elapsed_time(e1) / reps
        for f in win:
            f.check()
        dec.close()
    finally:
        os.unlink(path)
    return {"value": round(sensels / el / 1e9, 3), "unit": "Gsensels/s", "frames": frames,
            "window": 16, "fps": round(frames / el, 1), "seconds": round(el, 4),
            "sample_table_ms": round(t_table * 1e3, 3),
            "host_read_gb_per_s": round(nbytes / read_s / 1e9, 2) if read_s > 0 else None,
            "host_read_s": round(read_s, 4), "clip_bytes": nbytes,
            "resident_window_value": round(16 * 5760 * 3240 / (resident_ms * 1e-3) / 1e9, 3),
            "fill_ms": round((el - frames / 16 * resident_ms * 1e-3) * 1e3, 3),
            "workload": f"{frames}-frame .mov of 5760x3240 RGGB v0 a3 q45 (8 scenes cycled): "
                        "sample table + windowed pread into pinned memory + copy + batched "
                        "decode+develop(WhiteBalanced), 16 frames a window"}


def dng_bench(args):
    """DNG raw tiles of a decoded 5760x3240 RGGB frame (v0, alignment 3,
    qscale 45): its visible code values, read in place at the coded stride,
    coded as 299 lossless JPEG tiles on the device; checked once against
    the oracle's bytes and timed beside the oracle on the host cores."""
    import torch

    import proresraw_rs_amd as P
    from oracle import oracle as O
    from proresraw_rs_amd import dng as D
    from proresraw_rs_amd import synth as S

    packet = S.encode_scene(S.scene(5760, 3240, 0, seed=77, version=0, alignment=3,
                                    qscale=(45, 45)))
    frame = P.decode_frame(packet)
    h = frame.header
    w, ht = h.width, h.height
    stride = (w + 15) & ~15
    samples = np.ascontiguousarray(frame.samples, dtype=np.uint16).reshape(-1, stride)
    d_samples = torch.from_numpy(samples.view(np.int16).reshape(-1)).cuda()
    n = D.tile_count(w, ht)
    out = torch.empty(D.capacity(w, ht), dtype=torch.uint8, device="cuda")
    off = torch.empty(n, dtype=torch.int32, device="cuda")
    ln = torch.empty(n, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        D.lj92_tiles_device(d_samples, w, ht, stride, out, off, ln, stream=s)
        s.synchronize()
        reps = 20
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            D.lj92_tiles_device(d_samples, w, ht, stride, out, off, ln, stream=s)
        e1.record(s)
        e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    offsets = off.cpu().numpy().view(np.uint32)
    lengths = ln.cpu().numpy().view(np.uint32)
    total = int(offsets[-1]) + int(lengths[-1])
    got = out[:total].cpu().numpy().tobytes()
    visible = samples[:ht, :w]
    t0 = time.perf_counter()
    want, w_off, w_len = O.dng_lj92_tiles(visible, threads=O.THREADS)
    cpu_s = time.perf_counter() - t0
    exact = got == want and np.array_equal(offsets, w_off) and np.array_equal(lengths, w_len)
    sensels = w * ht
    return {"value": round(sensels / (ms * 1e-3) / 1e9, 3), "unit": "Gsensels/s",
            "ms_per_frame": round(ms, 4), "tiles": n, "bytes": total,
            "bits_per_sensel": round(8.0 * total / sensels, 3),
            "algorithmic_gb_per_s": round((2 * sensels + total) / (ms * 1e-3) / 1e9, 1),
            "bit_exact_vs_oracle": bool(exact),
            "cpu_baseline": {"value": round(sensels / cpu_s / 1e9, 4), "unit": "Gsensels/s",
                             "cores": O.THREADS, "kind": "port",
                             "sample": "one frame's tiles, oracle/lj92_oracle.c"},
            "workload": "DNG LosslessJpeg tiles (256x256, 2 components, predictor 1, optimal "
                        "Huffman table a tile) of a decoded 5760x3240 RGGB v0 a3 q45 frame's "
                        "code values at the coded stride"}


def orient_bench(args):
    """Crop + display orientation of three 5760x3240 f32 develop planes
    (crop 8/8/4/4, EXIF 6 = a quarter turn clockwise, and 3 = a half turn):
    HBM bytes moved = 12 B/sensel read + 12 B/sensel written."""
    import torch

    from proresraw_rs_amd import orient as R

    w, h = 5760, 3240
    margins = (8, 8, 4, 4)
    planes = [torch.rand(w * h, device="cuda") for _ in range(3)]
    res = {}
    for o in (6, 3):
        ow, oh = R.oriented_size(w, h, margins, o)
        outs = [torch.empty(ow * oh, device="cuda") for _ in range(3)]
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            R.orient_planes_device(planes, w, h, w, margins, o, outs, ow, stream=s)
            reps = 20
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(reps):
                R.orient_planes_device(planes, w, h, w, margins, o, outs, ow, stream=s)
            e1.record(s)
            e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        moved = 24.0 * ow * oh
        res[f"exif{o}"] = {"ms": round(ms, 4), "gb_per_s": round(moved / (ms * 1e-3) / 1e9, 1),
                           "frac_of_8tbs": round(moved / (ms * 1e-3) / 8e12, 3)}
    return {"workload": "crop (8, 8, 4, 4) + orientation of three 5760x3240 f32 planes", **res}


def cpu_baseline(packet, seconds):
    """The oracle's decode + linearize + develop(Rec2020) of the same packet,
    repeated for about `seconds`, on every host core (capped at 16)."""
    from oracle import oracle as O
    threads = O.THREADS
    t0 = time.perf_counter()
    n = 0
    while True:
        h, s, sk, mq = O.decode_frame(packet, threads=threads)
        O.linearize(h, s, threads=threads)
        prm = O.params(h, h.quant_matrix[0] * mq)
        O.develop_planar(h, s, prm, 2, threads=threads)
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    npix = h.width * h.height
    return {"value": round(npix * n / el / 1e9, 5), "unit": "Gsensels/s", "cores": threads,
            "kind": "port", "sample": f"{n} frames of the same {h.width}x{h.height} packet, "
                                      f"{el:.1f} s, median-free mean"}


if __name__ == "__main__":
    main()
