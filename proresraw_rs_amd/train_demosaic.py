"""Train and evaluate the learned demosaic on the GPU.

    python -m proresraw_rs_amd.train_demosaic --out DIR [--scenes 2000]
        [--steps 20000] [--eval 50] [--eval-size 5760x3240]

1. seeded synthetic training scenes (RGGB, qscale uniform in [15, 45],
   8 gradients + 200 discs, Poisson-Gaussian noise), each decoded and
   developed at Stage::Demosaiced by the plain GPU path (bit-exact to the
   crate) into (mosaic, target) pairs kept on the device;
2. Adam (lr 1e-3, constant; --lr-min decays it along a cosine over the
   run), batches of 16 random crops of 288x288 sensels (loss on the central
   256x256), one step captured in a graph and replayed (--curves: every
   other scene v1 with a seeded custom curve);
3. held-out frames of other seeds: PSNR and max abs error of the learned
   develop against plain develop, per qscale, at Stage::Demosaiced and
   Stage::WhiteBalanced (gains 2.0 / 1.6) and Stage::Rec2020 (header
   gains and matrix), next to the bilinear base alone (the method without
   the network), the time of both develops, and the sampled-tile check
   (8 tiles, 45 dB) the end-to-end path would apply.

Writes DIR/demosaic_net.npz (parameters, trained feature range, loss
history) and DIR/eval.json.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

from . import _native as N
from . import learned as L
from .api import DemosaicParams, Stage, _check, context, decode_frame


def log(*a):
    print(*a, flush=True)


def develop_plain(torch, header, samples, params, stage, planes, stream):
    err = N.PrrError()
    raw = params._raw()
    _check(N.lib().prr_develop_device(context().handle, ctypes.byref(header._raw),
                                      samples.data_ptr(), ctypes.byref(raw), int(stage),
                                      planes[0].data_ptr(), planes[1].data_ptr(),
                                      planes[2].data_ptr(), ctypes.c_void_p(stream.cuda_stream),
                                      ctypes.byref(err)), err)


def tile_psnr(a, b, h, w):
    """PSNR (peak 1) of every whole 64x64 tile between two 3 x h x w
    tensors."""
    d = (a - b).double() ** 2
    th, tw = h // 64, w // 64
    d = d[:, :th * 64, :tw * 64].reshape(3, th, 64, tw, 64).sum(dim=(0, 2, 4)) / (3 * 4096)
    return (-10.0 * np.log10(np.maximum(d.cpu().numpy(), 1e-20))).reshape(-1)


def evaluate(torch, net, zero, packets, stages, log=log):
    """Per frame and stage: PSNR / max abs of learned and bilinear against
    plain develop, the PSNR of every 64x64 tile (stages without the output
    matrix: what the sampled-tile check compares), and develop times (ms,
    events on one stream)."""
    stream = torch.cuda.Stream()
    rows = []
    for k, (packet, q) in enumerate(packets):
        fr = decode_frame(packet)
        h = fr.header
        samples = torch.from_numpy(np.ascontiguousarray(fr.samples).view(np.int16).copy()).cuda()
        n = h.width * h.height
        for stage, wb in stages:
            params = (DemosaicParams.with_white_balance(h, fr.quantizer(), *wb) if wb
                      else DemosaicParams.new(h, fr.quantizer()))
            out = {name: [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(3)]
                   for name in ("plain", "learned", "bilinear")}
            times = {}
            with torch.cuda.stream(stream):
                for name in ("plain", "learned", "bilinear"):
                    for rep in range(2):  # the second run is timed
                        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                        e0.record(stream)
                        if name == "plain":
                            develop_plain(torch, h, samples, params, stage, out[name], stream)
                        else:
                            (net if name == "learned" else zero).develop(
                                h, samples, params, stage, out[name], check_tiles=0, stream=stream)
                        e1.record(stream)
                    stream.synchronize()
                    times[name] = e0.elapsed_time(e1)
            ref = torch.stack(out["plain"]).double()
            row = {"frame": k, "qscale": q, "stage": stage.name, "width": h.width,
                   "height": h.height, "ms": times}
            for name in ("learned", "bilinear"):
                d = torch.stack(out[name]).double() - ref
                mse = float((d * d).mean())
                row[name] = {"psnr": (-10.0 * math.log10(mse)) if mse > 0 else float("inf"),
                             "max_abs": float(d.abs().max())}
            if stage != Stage.Rec2020:
                tp = tile_psnr(torch.stack(out["learned"]).reshape(3, h.height, h.width),
                               ref.float().reshape(3, h.height, h.width), h.height, h.width)
                row["tiles"] = {"n": int(tp.size), "below_45": float((tp < 45.0).mean()),
                                "p1": float(np.percentile(tp, 1)), "min": float(tp.min())}
            checked = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(3)]
            net.develop(h, samples, params, stage, checked, check_tiles=8, min_db=45.0,
                        stream=stream)
            used, worst = net.last_check()
            row["check"] = {"fallback": used, "worst_tile_psnr": worst}
            rows.append(row)
        log(f"eval frame {k} q{q}: " + ", ".join(
            f"{r['stage']} learned {r['learned']['psnr']:.2f} dB (max {r['learned']['max_abs']:.3f}) "
            f"bilinear {r['bilinear']['psnr']:.2f} dB tiles>={r['check']['worst_tile_psnr']:.1f}"
            f"{' FALLBACK' if r['check']['fallback'] else ''}" for r in rows[-len(stages):]))
    return rows


def summarize(rows):
    out = {}
    for r in rows:
        key = f"{r['stage']}/{r['width']}x{r['height']}/q{r['qscale']}"
        s = out.setdefault(key, {"frames": 0, "learned_psnr": [], "bilinear_psnr": [],
                                 "learned_max_abs": 0.0, "fallbacks": 0, "tiles_below_45": [],
                                 "tile_min": 200.0})
        s["frames"] += 1
        if "tiles" in r:
            s["tiles_below_45"].append(r["tiles"]["below_45"])
            s["tile_min"] = min(s["tile_min"], r["tiles"]["min"])
        s["fallbacks"] += int(r["check"]["fallback"])
        s["learned_psnr"].append(r["learned"]["psnr"])
        s["bilinear_psnr"].append(r["bilinear"]["psnr"])
        s["learned_max_abs"] = max(s["learned_max_abs"], r["learned"]["max_abs"])
    for s in out.values():
        for k in ("learned_psnr", "bilinear_psnr"):
            v = s.pop(k)
            s[k + "_mean"] = float(np.mean(v))
            s[k + "_min"] = float(np.min(v))
        t = s.pop("tiles_below_45")
        s["tiles_below_45_mean"] = float(np.mean(t)) if t else None
    return out


def stream_packets(n, seed):
    """n frames of each kind of the end-to-end stream, seeds of their own:
    8192x4320 v1 (seeded curve) alignment 4 qscale 15 and 5760x3240 v0
    alignment 3 qscale 45, RGGB."""
    from . import synth as S
    from .workloads import seeded_curve
    out = []
    for i in range(n):
        sc = S.scene(8192, 4320, 0, seed=seed + 2 * i, version=1, alignment=4, qscale=(15, 15),
                     curve=seeded_curve(seed + 2 * i), white_level=4095)
        out.append((S.encode_scene(sc), 15))
        sc = S.scene(5760, 3240, 0, seed=seed + 2 * i + 1, version=0, alignment=3,
                     qscale=(45, 45), white_level=4095)
        out.append((S.encode_scene(sc), 45))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--scenes", type=int, default=2000)
    ap.add_argument("--scene-size", default="512x512")
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-min", type=float, default=0.0)
    ap.add_argument("--curves", action="store_true")
    ap.add_argument("--eval", type=int, default=50)
    ap.add_argument("--eval-size", default="5760x3240")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--init", default=None,
                    help="start from these weights (an npz of save_weights) instead of new ones")
    ap.add_argument("--hard-frac", type=float, default=0.0,
                    help="fraction of crops drawn around the 64x64 tiles where the starting "
                         "network is furthest from the target")
    ap.add_argument("--tail-tau", type=float, default=0.0,
                    help="train on d^2 + d^4 / tau^2 (weighs the few large edge errors the "
                         "sampled-tile check sees); 0: the squared error")
    ap.add_argument("--eval-stream", type=int, default=0,
                    help="also evaluate on this many frames of each end-to-end stream kind")
    args = ap.parse_args(argv)
    import torch

    os.makedirs(args.out, exist_ok=True)
    sw, sh = (int(v) for v in args.scene_size.split("x"))
    t0 = time.time()
    packets = L.scene_packets(args.scenes, (sw, sh), args.seed, workers=args.workers,
                              curves=args.curves)
    log(f"{args.scenes} training scenes encoded in {time.time() - t0:.1f} s")
    t0 = time.time()
    frames = [L.frame_pair(p) for p, _ in packets]
    del packets
    log(f"decoded and developed in {time.time() - t0:.1f} s")

    lo, hi = L.noise_feature(15), L.noise_feature(45)
    net = L.Net(L.load_weights(args.init)[0] if args.init else L.init_params(args.seed))
    net.set_range(lo, hi)
    hard = None
    if args.hard_frac > 0:
        t0 = time.time()
        hard = L.tile_errors(net, frames)
        log(f"tile errors of the starting network in {time.time() - t0:.1f} s")
    t0 = time.time()
    hist = []
    if args.steps > 0:
        hist = L.train(net, frames, args.steps, batch=args.batch, cells=144, margin=8,
                       lr=args.lr, seed=args.seed, log_every=1000, log=log,
                       lr_min=args.lr_min if args.lr_min > 0 else None, hard=hard,
                       hard_frac=args.hard_frac, tau=args.tail_tau)
    train_s = time.time() - t0
    params = net.params()
    L.save_weights(os.path.join(args.out, "demosaic_net.npz"), params, lo, hi,
                   steps=args.steps, batch=args.batch, lr=args.lr, lr_min=args.lr_min,
                   curves=args.curves, scenes=args.scenes, init=str(args.init),
                   hard_frac=args.hard_frac, tail_tau=args.tail_tau,
                   loss_history=np.array(hist, np.float64))
    log(f"trained {args.steps} steps in {train_s:.1f} s")
    del frames

    net.set_range(lo, hi)
    zero = L.Net(np.zeros(L.PARAMS, np.float32))
    zero.set_range(0.0, 100.0)
    ew, eh = (int(v) for v in args.eval_size.split("x"))
    t0 = time.time()
    held = L.scene_packets(args.eval, (ew, eh), args.seed + 7919, workers=args.workers,
                           curves=args.curves)
    if args.eval_stream:
        held = held + stream_packets(args.eval_stream, args.seed * 1000 + 50000)
    log(f"{len(held)} held-out frames encoded in {time.time() - t0:.1f} s")
    rows = evaluate(torch, net, zero, held,
                    [(Stage.Demosaiced, None), (Stage.WhiteBalanced, (2.0, 1.6)),
                     (Stage.WhiteBalanced, None), (Stage.Rec2020, None)])
    report = {"train_seconds": train_s, "steps": args.steps, "batch": args.batch,
              "scenes": args.scenes, "scene_size": args.scene_size, "eval_size": args.eval_size,
              "final_loss": hist[-1][1] if hist else None, "summary": summarize(rows),
              "frames": rows}
    with open(os.path.join(args.out, "eval.json"), "w") as f:
        json.dump(report, f, indent=1)
    log(json.dumps(report["summary"], indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
