"""Benchmark: plain GPU decode + develop of ProRes RAW frames on one MI355X.

Workload (one step = a batch of --frames-per-step frames, 16 by default):
the 5760x3240 frame kind of the
end-to-end stream (RGGB, bitstream v0 with the default curve, tile
alignment 3, qscale 45, seeded scene of 8 gradients + 200 discs with
Poisson-Gaussian noise) is decoded with the Apple transform, linearized in
the same kernel, and developed to Rec.2020 planes. The packet is resident
in device memory before timing starts. Frames stream through the native
pipeline (prr_pipeline): frame i + 1's decode runs on one HIP stream beside
frame i's develop on another. The serial step (decode then develop, one HIP
graph a frame) is reported beside it, and the 4096x2160 frame (BGGR v1 a3
q45, seeded quant matrix and curve, 2% corrupt tiles) as a leg in both
inverse transforms.

Prints one JSON line. `value` is Gsensels/s (visible width x height per
second). `roofline` covers the kernel that takes longest inside the stream
against HBM bandwidth, with its VALU issue rate beside it; `cpu_baseline`
times the oracle (a C restatement of the crate's CPU path, threaded like its
rayon code) on the same packet.
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
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames-per-step", type=int, default=16,
                    help="frames a step submits to the pipelined stream (a frame takes about "
                         "0.3 ms, so a step of one frame would time the pipeline's fill)")
    ap.add_argument("--width", type=int, default=5760)
    ap.add_argument("--height", type=int, default=3240)
    ap.add_argument("--pipeline-depth", type=int, default=4,
                    help="frames in flight in the plain legs' frame pipeline")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-frames", type=int, default=64)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--learned-frames", type=int, default=32)
    ap.add_argument("--learned-contexts", type=int, default=4,
                    help="learned-demosaic contexts (each on its own stream) the batch's frames "
                         "are spread over")
    ap.add_argument("--learned-distinct", type=int, default=32,
                    help="distinct packets (half of each frame kind) cycled in the learned leg "
                         "(default: every frame its own scene)")
    ap.add_argument("--net-weights", default=None,
                    help="weights of the learned demosaic (default: the shipped ones)")
    ap.add_argument("--no-clip", action="store_true")
    ap.add_argument("--no-dng", action="store_true")
    ap.add_argument("--no-orient", action="store_true")
    ap.add_argument("--clip-frames", type=int, default=160)
    ap.add_argument("--no-autotune", action="store_true",
                    help="batches keep the default launch shape (counter runs: one shape)")
    ap.add_argument("--only", choices=["plain", "c1", "batch", "learned", "clip", "dng", "orient"],
                    help="run one leg only (profiling); no CPU baseline")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 10 times (the reference reading "
                         "of the HBM byte counters) and exit")
    args = ap.parse_args()
    if args.only:
        for leg in ("c1", "batch", "learned", "clip", "dng", "orient"):
            setattr(args, "no_" + leg, args.only != leg)
        args.no_cpu = True
    return args


PROFILES = os.path.join(ROOT, "profiles")
TRAFFIC_FILE = os.path.join(PROFILES, "hbm_traffic.txt")
UTIL_FILE = os.path.join(PROFILES, "issue_utilization.txt")
N_CU = 256          # MI355X compute units
CLOCK_HZ = 2.4e9    # MI355X peak engine clock
HBM_PEAK = 8000.0   # GB/s


T_START = time.perf_counter()


def log(msg):
    """A progress line on stderr as each leg ends (stdout holds the JSON
    line only)."""
    print(f"[bench {time.perf_counter() - T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)


def profile_fields(path, kernel, workload):
    """The key=value fields of `kernel` on `workload` in a committed counter
    summary (lines `kernel workload key=value ...`), or None."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) > 2 and f[0] == kernel and f[1] == workload:
            return dict(x.split("=", 1) for x in f[2:])
    return None


def measured_traffic(kernel, workload):
    """HBM bytes per launch: FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by what the counter reads for a copy of known size."""
    kv = profile_fields(TRAFFIC_FILE, kernel, workload)
    if kv is None:
        return None
    return int(1024 * (float(kv["fetch_kib"]) * float(kv["fetch_scale"]) +
                       float(kv["write_kib"]) * float(kv["write_scale"])))


def measured_valu(kernel, workload):
    """(SQ_INSTS_VALU per dispatch, VALU instructions per SIMD-cycle) of
    `kernel` on `workload`, or (None, None)."""
    kv = profile_fields(UTIL_FILE, kernel, workload)
    if kv is None:
        return None, None
    return float(kv["SQ_INSTS_VALU"]), float(kv["valu_per_simd_cycle"])


def measured_issue(kernel, workload):
    """From the counter summary: VALU instructions per SIMD-cycle, the share
    of a SIMD's quad-cycles that issue a VALU instruction (one, or two from
    two waves) and the share of those that issue two, where recorded."""
    kv = profile_fields(UTIL_FILE, kernel, workload)
    if kv is None:
        return {}
    out = {"valu_insts_per_simd_cycle": float(kv["valu_per_simd_cycle"])}
    for k in ("valu_busy", "dual_issue"):
        if k in kv:
            out[k] = float(kv[k])
    return out


def decode_bound(kernel, workload, coded_bits, tiles, kernel_ms):
    """The decode against its per-CU bit-rate bound: the kernel's own VALU
    instruction stream (wave64 instructions per dispatch, from the counter
    summary) spread over every CU at one wave instruction per CU-cycle (four
    SIMDs, four cycles a wave64 instruction) at the peak clock gives the
    shortest time those coded bits could take with this kernel's
    instructions per bit; the fraction is that time over the measured one.
    gfx950 can also issue two VALU instructions from two waves in one SIMD
    quad-cycle (the pairs counted by SQ_ACTIVE_INST_VALU2), so the same
    instructions at two a CU-cycle are the dual-issue bound, half the time."""
    insts, _ = measured_valu(kernel, workload)
    out = {"kernel": kernel, "coded_bits": int(coded_bits), "kernel_ms": round(kernel_ms, 4),
           "tiles_per_s": round(tiles / (kernel_ms * 1e-3), 1), "valu_source":
           os.path.relpath(UTIL_FILE, ROOT)}
    if insts is None:
        return out
    t_bound = insts / (N_CU * CLOCK_HZ)
    out.update({"valu_insts": insts, "lane_ops_per_bit": round(64.0 * insts / coded_bits, 2),
                "bound_ms": round(t_bound * 1e3, 4),
                "bound_gbit_per_s": round(coded_bits / t_bound / 1e9, 1),
                "bound_tiles_per_s": round(tiles / t_bound, 1),
                "frac": round(t_bound / (kernel_ms * 1e-3), 4),
                "bound_ms_dual_issue": round(t_bound * 0.5e3, 4),
                "frac_dual_issue": round(0.5 * t_bound / (kernel_ms * 1e-3), 4)})
    return out


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


def device_packet(packet):
    import torch
    padded = ((len(packet) + 3) & ~3) + 16
    d = torch.zeros(padded, dtype=torch.uint8, device="cuda")
    d[:len(packet)] = torch.frombuffer(bytearray(packet), dtype=torch.uint8).to("cuda")
    return d


def pipelined(packet, d_packet, hdr, info, planes, stage, idct, frames, warmup,
              time_kernels=True, depth=4):
    """The pipelined stream of frames (prr_pipeline, `depth` buffer sets):
    `warmup` untimed frames, then `frames` frames submitted back to back; the
    timed region is bracketed by device synchronization, so it holds exactly
    `frames` whole frames. Returns
    (seconds, mean decode ms, mean develop ms inside the stream)."""
    import torch

    from proresraw_rs_amd.pipeline import Pipeline
    pipe = Pipeline(depth=depth)
    status = torch.zeros(4, dtype=torch.int32, device="cuda")

    def submit():
        pipe.submit(hdr, info, d_packet, len(packet), planes, stage=stage, idct=idct,
                    status=status)

    for _ in range(warmup):
        submit()
    pipe.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(frames):
        submit()
    pipe.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st = status.cpu().numpy().view(np.uint32)
    if st[0] != 0xFFFFFFFF or st[3]:
        raise RuntimeError(f"device status {st}")
    dec_ms = dev_ms = None
    if time_kernels:
        # Kernel durations inside the stream, with HIP events on the streams
        # the kernels run on (each shares the chip with the other there).
        pipe.timing(True)
        for _ in range(min(frames, 60)):
            submit()
        dec_ms, dev_ms, _ = pipe.timing(False)
    pipe.close()
    return elapsed, dec_ms, dev_ms


def plain_leg(args, packet, workload):
    """The plain path on one frame kind: the serial step (one HIP graph a
    frame), per-kernel times alone, and the pipelined stream."""
    import torch

    import proresraw_rs_amd as P
    from proresraw_rs_amd import _native as N

    lib = N.lib()
    ctx = P.context()
    stream = torch.cuda.Stream()
    hdr, info = P.parse_frame_info(packet)
    raw_hdr = hdr._raw
    w, h = hdr.width, hdr.height
    cw, ch = info.coded_width, info.coded_height
    npix = w * h
    dev = torch.device("cuda:0")
    d_packet = device_packet(packet)
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

    frames, warm = args.steps * args.frames_per_step, args.warmup * args.frames_per_step
    serial_steps = max(frames // 2, 10)
    with torch.cuda.stream(stream):
        for _ in range(warm):
            run_step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(serial_steps):
            run_step()
        torch.cuda.synchronize()
        serial_s = time.perf_counter() - t0

    elapsed, pipe_dec_ms, pipe_dev_ms = pipelined(packet, d_packet, hdr, info, planes,
                                                  P.Stage.Rec2020, P.Idct.Apple, frames, warm,
                                                  depth=args.pipeline_depth)
    table_bytes = 2 * info.nb_tiles
    coded_bits = 8 * (len(packet) - info.header_end - table_bytes)
    return {
        "npix": npix, "width": w, "height": h, "packet_bytes": len(packet),
        "nb_tiles": info.nb_tiles, "coded_width": cw, "coded_height": ch,
        "value": npix * frames / elapsed / 1e9, "ms_per_step": 1e3 * elapsed / args.steps,
        "frames": frames, "ms_per_frame": 1e3 * elapsed / frames,
        "serial_value": npix * serial_steps / serial_s / 1e9,
        "serial_ms": 1e3 * serial_s / serial_steps, "serial_graph": graph is not None,
        "decode_ms_alone": t_decode_ms, "develop_ms_alone": t_develop_ms,
        "decode_ms": pipe_dec_ms, "develop_ms": pipe_dev_ms,
        "coded_bits": coded_bits, "workload_key": workload,
    }


def roofline_of(r):
    """The roofline object for the kernel that takes longest inside the
    stream. Algorithmic bytes a launch: develop reads the linear picture
    (2 B/sensel) and writes three f32 planes (12 B/sensel); the decode
    (prepare + decode) reads the packet and writes code values (2 B per
    coded sensel) and linear values (2 B per visible sensel)."""
    npix, key = r["npix"], r["workload_key"]
    dev_bytes = 14.0 * npix
    dec_bytes = float(r["packet_bytes"]) + 2.0 * r["coded_width"] * r["coded_height"] + 2.0 * npix
    develop_dominant = r["develop_ms"] >= r["decode_ms"]
    kernel = "k_develop" if develop_dominant else "k_decode"
    nbytes = dev_bytes if develop_dominant else dec_bytes
    ms = r["develop_ms"] if develop_dominant else r["decode_ms"]
    achieved = nbytes / (ms * 1e-3) / 1e9
    insts, rate = measured_valu(kernel, key)
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK, 4), "traffic": measured_traffic(kernel, key),
           "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT), "kernel": kernel,
           "kernel_ms": round(ms, 4),
           "kernel_ms_alone": round(r["develop_ms_alone" if develop_dominant else
                                      "decode_ms_alone"], 4),
           "bytes_per_launch": int(nbytes), "bytes_per_unit": "14 B/visible sensel"
           if develop_dominant else "packet + 2 B/coded sensel + 2 B/visible sensel"}
    # The kernels' limiter is their VALU issue, not HBM: from the counter
    # summary, instructions per SIMD-cycle and the share of the SIMDs'
    # quad-cycles that issue VALU work (alone: valu_busy; gfx950 issues one
    # wave64 VALU instruction a quad-cycle, or two from two waves).
    if rate is not None:
        out.update({"limiter": "valu_issue", **measured_issue(kernel, key),
                    "valu_source": os.path.relpath(UTIL_FILE, ROOT)})
    other = "k_decode" if develop_dominant else "k_develop"
    out["other_kernel"] = {"kernel": other, "kernel_ms": round(r["decode_ms" if develop_dominant
                                                                 else "develop_ms"], 4),
                           "traffic": measured_traffic(other, key), **measured_issue(other, key)}
    out["decode_bitrate"] = decode_bound("k_decode", key, r["coded_bits"], r["nb_tiles"],
                                         r["decode_ms_alone"])
    out["decode_bitrate_frac"] = out["decode_bitrate"].get("frac")
    return out


def c1_leg(args):
    """The 4096x2160 frame (BGGR v1 a3 q45, seeded quant matrix and curve, 2%
    corrupt tiles) through the pipelined stream in each inverse transform,
    developed to Rec.2020 (the transforms are bit-exact against the oracle
    in the GPU tests)."""
    import torch

    import proresraw_rs_amd as P
    from proresraw_rs_amd import synth as S
    from proresraw_rs_amd.workloads import config1

    packet = S.encode_scene(config1(4096, 2160))
    hdr, info = P.parse_frame_info(packet)
    d_packet = device_packet(packet)
    npix = hdr.width * hdr.height
    planes = [torch.empty(npix, dtype=torch.float32, device="cuda") for _ in range(3)]
    out = {"workload": "4096x2160 BGGR v1 a3 q45, seeded qmat + curve, 2% corrupt tiles; "
                       "pipelined decode+linearize+develop(Rec2020)", "packet_bytes": len(packet)}
    steps = max(args.steps * args.frames_per_step // 2, 10)
    for idct in (P.Idct.Apple, P.Idct.Ffmpeg):
        el, dec, dv = pipelined(packet, d_packet, hdr, info, planes, P.Stage.Rec2020, idct,
                                steps, args.warmup * args.frames_per_step,
                                depth=args.pipeline_depth)
        out[idct.name.lower()] = {"value": round(npix * steps / el / 1e9, 3),
                                  "ms_per_frame": round(1e3 * el / steps, 4),
                                  "decode_ms": round(dec, 4), "develop_ms": round(dv, 4)}
    return out


def main():
    args = parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are out of scope "
                                   "(frames shard across GPUs with no exchange)"}))
        sys.exit(2)
    import torch  # noqa: F401

    from proresraw_rs_amd import synth as S
    from proresraw_rs_amd.workloads import stream_frame

    packet = S.encode_scene(stream_frame(args.width, args.height))
    key = f"plain{args.width}x{args.height}"
    log(f"encoded the {args.width}x{args.height} packet ({len(packet)} bytes)")
    r = plain_leg(args, packet, key)
    value = r["value"]
    roofline = roofline_of(r)
    log(f"plain leg: {value:.2f} Gsensels/s pipelined, {r['serial_value']:.2f} serial")

    def leg(name, skip, fn, *a):
        if skip:
            return None
        out = fn(*a)
        log(f"{name} leg done")
        return out

    c1 = leg("4096x2160 both transforms", args.no_c1, c1_leg, args)
    batched = leg("batched", args.no_batch, batched_bench, args)
    learned = leg("batched + learned", args.no_learned, learned_bench, args, value)
    clip = leg("clip", args.no_clip, clip_bench, args)
    dng = leg("DNG tiles", args.no_dng, dng_bench, args)
    orient = leg("orientation", args.no_orient, orient_bench, args)
    cpu = leg("CPU baseline", args.no_cpu, cpu_baseline, packet, args.cpu_seconds)

    line = {
        "metric": "Gsensels/s decode+develop, plain vs batched+learned demosaic at >=45 dB PSNR",
        "value": round(value, 3),
        "unit": "Gsensels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16+f32",
        "data": "synthetic",
        "config": {"workload": f"plain GPU decode(Apple)+linearize+develop(Rec2020), "
                               f"{r['width']}x{r['height']} RGGB v0 a3 q45 (the end-to-end "
                               "stream's 5760x3240 frame kind); a stream of frames, decode of "
                               "frame i+1 beside develop of frame i",
                   "packet_bytes": r["packet_bytes"], "tiles": r["nb_tiles"],
                   "coded_bits_per_sensel": round(r["coded_bits"] / r["npix"], 3),
                   "frames_per_step": args.frames_per_step, "frames_timed": r["frames"],
                   "ms_per_frame": round(r["ms_per_frame"], 4),
                   "pipeline_depth": args.pipeline_depth,
                   "serial_value": round(r["serial_value"], 3),
                   "serial_ms_per_frame": round(r["serial_ms"], 4),
                   "serial_graph": r["serial_graph"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "c1_4096x2160": c1,
        "batched": batched,
        "batched_learned": learned,
        "clip": clip,
        "dng_tiles": dng,
        "orient": orient,
    }
    if cpu:
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
        line["serial_gpu_over_cpu"] = round(r["serial_value"] / cpu["value"], 1)
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
    out_pixels = 0
    full_sensels = 0
    n = args.batch_frames
    n_full = 0
    for k in range(n):
        pk = distinct[k % 16]
        hdr, info = P.parse_frame_info(pk)
        sensels += hdr.width * hdr.height
        if k % 2 == 0:
            b.add_full(pk, P.Idct.Apple, develop=P.Stage.WhiteBalanced)
            n_full += 1
            out_pixels += hdr.width * hdr.height
            full_sensels += hdr.width * hdr.height
        else:
            j = k // 2
            scale = [P.Scale.Half, P.Scale.Quarter, P.Scale.Eighth][j % 3]
            greens = P.Greens.Average if scale == P.Scale.Half and j % 4 == 0 else P.Greens.First
            b.add_reduced(pk, scale, greens)
            # A reduced frame's output: (coded w >> s) x (coded h >> s) RGB pixels.
            out_pixels += (info.coded_width >> int(scale)) * (info.coded_height >> int(scale))
    s = torch.cuda.Stream()
    b.run(s)
    s.synchronize()
    for slot in b.slots:
        slot.check()
    tuned = b.autotune(s) if not args.no_autotune else {"spw": 64, "chunk": 1, "waves": 0,
                                                         "cached": None}
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
    return {"value": round(sensels / el / 1e9, 3), "unit": "Gsensels/s",
            "value_counts": "visible sensels of every frame's packet, reduced frames included "
                            "(not comparable with full-resolution legs; see output_value)",
            "output_value": round(out_pixels / el / 1e9, 3),
            "output_unit": "Gpixels/s of output (full frames w*h developed, reduced frames "
                           "(coded w >> s)(coded h >> s) RGB)",
            "full_develop_gsensels": round(full_sensels / 1e9, 4),
            "frames": n, "full_develop_frames": n_full, "ms_per_batch": round(el * 1e3, 3),
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
    beside it. --learned-distinct packets (half of each kind; by default one a
    frame) are encoded and cycled."""
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
    nd = max(2, args.learned_distinct & ~1)
    for i in range(nd):
        k = kinds[i % 2]
        sc = S.scene(*k["size"], 0, seed=3000 + i, version=k["version"],
                     alignment=k["alignment"], qscale=(k["q"], k["q"]),
                     curve=seeded_curve(40 + i) if k["version"] == 1 else None)
        distinct.append(S.encode_scene(sc))
    b = Batch()
    frames = []
    sensels = 0
    coded_bits = 0
    tiles = 0
    for j in range(args.learned_frames):
        pk = distinct[j % nd]
        slot = b.add_full(pk, P.Idct.Apple, linear=True)
        h = slot.header
        sensels += h.width * h.height
        coded_bits += 8 * (len(pk) - slot.info.header_end - 2 * slot.info.nb_tiles)
        tiles += slot.info.nb_tiles
        planes = [torch.empty(h.width * h.height, dtype=torch.float32, device="cuda")
                  for _ in range(3)]
        frames.append((slot, P.DemosaicParams.new(h, 0), planes))
    params, lo, hi = L.load_weights(args.net_weights) if args.net_weights else L.load_weights()
    net = L.Net(params)
    net.set_range(lo, hi)
    group = L.NetGroup(params, lo, hi, n=max(1, args.learned_contexts))
    lib, ctx = N.lib(), P.context()
    s = torch.cuda.Stream()
    hs = ctypes.c_void_p(s.cuda_stream)
    err = N.PrrError()
    jobs = [(slot.header, slot.samples, prm, planes, slot.status) for slot, prm, planes in frames]

    def learned_step():
        b.run(s)
        group.develop(jobs, P.Stage.Rec2020, check_tiles=8, min_db=45.0, stream=s)

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
        tuned = b.autotune(s) if not args.no_autotune else {"spw": 64, "chunk": 1}
        # The batched decode alone (scan, queue, persistent decode with the
        # fused linearize), HIP events on its stream.
        batch_ms = b.time_run(s, reps=5)
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
            "batch_decode_ms": round(batch_ms, 3),
            "batch_decode_autotuned": {"spw": tuned["spw"], "chunk": tuned["chunk"]},
            "coded_bits_per_sensel": round(coded_bits / sensels, 3),
            "decode_bitrate": decode_bound("k_batch_decode", "learned32", coded_bits, tiles,
                                           batch_ms),
            "distinct_packets": nd, "contexts": max(1, args.learned_contexts),
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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(packet, seconds):
    """The oracle (the C restatement of the crate's rayon path: tile rows
    for decoding, 32-row bands for develop) on the same packet:
    decode_frame then develop(&frame, Rec2020), which linearizes inside as
    the crate's does. Method of the crate's examples/bench.rs:121-140: one
    warm-up run, then timed runs (at least 5, for about `seconds`),
    reporting the median and the fastest. Threads: every CPU this process
    may run on, up to 16, the host's share for one GPU; one more run on a
    single thread gives the per-core rate."""
    from oracle import oracle as O
    affinity = len(os.sched_getaffinity(0))
    threads = min(affinity, 16)

    def run(th):
        h, s, sk, mq = O.decode_frame(packet, threads=th)
        prm = O.params(h, h.quant_matrix[0] * mq)
        O.develop_planar(h, s, prm, 2, threads=th)
        return h

    h = run(threads)
    times = []
    t_start = time.perf_counter()
    while len(times) < 5 or (time.perf_counter() - t_start < seconds and len(times) < 200):
        t0 = time.perf_counter()
        run(threads)
        times.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    run(1)
    single = time.perf_counter() - t0
    npix = h.width * h.height
    med = float(np.median(times))
    fastest = min(times)
    return {"value": round(npix / med / 1e9, 5), "unit": "Gsensels/s", "cores": threads,
            "kind": "port", "median_ms": round(med * 1e3, 2), "fastest_ms": round(fastest * 1e3, 2),
            "fastest_value": round(npix / fastest / 1e9, 5), "runs": len(times),
            "single_thread_value": round(npix / single / 1e9, 5),
            "nproc": os.cpu_count(), "affinity_cpus": affinity, "cpu_model": cpu_model(),
            "sample": f"decode_frame + develop(Rec2020) of the {h.width}x{h.height} packet: one "
                      f"warm-up, then {len(times)} timed runs ({sum(times):.1f} s) on {threads} "
                      "threads, median reported; plus one run on 1 thread"}


if __name__ == "__main__":
    main()
