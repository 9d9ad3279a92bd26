"""The learned demosaic (include/proresraw_learned.h) from Python: the
network handle, parameter initialisation, the training set built from
seeded synthetic scenes (mosaic + develop(Stage::Demosaiced) pairs made by
this package's plain GPU path), the graph-captured training loop and the
held-out evaluation. torch provides device memory and streams only.

The network has no counterpart in the proresraw_rs crate; its targets are
the crate's develop output (demosaic.rs:190-402), which the plain path here
reproduces bit for bit.
"""
import concurrent.futures
import ctypes
import math
import os
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _native as N
from . import synth as S
from .api import (DemosaicParams, Error, FrameHeader, Stage, _check, context, decode_frame,
                  develop_planar)

PARAMS = N.NET_PARAMS
LAYER_PARAMS = N.NET_LAYER_PARAMS
LAYERS = 8
C = 32
HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "weights", "demosaic_net.npz")

X0, H, R, OUT, DOUT, GRAD, LOSS, STEP, FRAME_X0 = range(9)
# Input channels the packing fills (prr_net.hip's layout): the mosaic (4),
# the noise feature (1), the noise-level map (4), the green pass's
# curvature ratios (4) and the above-noise tests (4).
INPUT_CHANNELS = 17


def _torch():
    import torch
    return torch


def noise_feature(quantizer: int) -> float:
    """128 * sqrt(noise[0]) with noise[0] as DemosaicParams computes it from
    the quantizer (demosaic.rs:86-101), in float32."""
    n0 = np.float32((quantizer * quantizer * 5.0 + 64.0 / 3.0) / 150994944.0)
    return float(np.float32(128.0) * np.sqrt(n0, dtype=np.float32))


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights scaled down on the residual and output layers (the
    network starts close to the bilinear base); zero biases. Input channels
    past the 17 used ones (mosaic, noise feature, noise-level map, curvature
    ratios, above-noise tests) and output channels past the 12 used ones
    stay zero (their gradients are zero, so training keeps them zero)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, np.float32)
    for l in range(LAYERS):
        cin = INPUT_CHANNELS if l == 0 else C
        cout = 12 if l == LAYERS - 1 else C
        scale = 1.0 if l == 0 else (0.25 if l < LAYERS - 1 else 0.1)
        w = np.zeros((9, C, C), np.float32)
        w[:, :cout, :cin] = rng.standard_normal((9, cout, cin)) * math.sqrt(2.0 / (9 * cin)) * scale
        p[l * LAYER_PARAMS:l * LAYER_PARAMS + 9 * C * C] = w.reshape(-1)
    return p


def load_weights(path: str = WEIGHTS):
    """(params, feature_lo, feature_hi) saved by save_weights. A file records
    the input channels its network reads (9 when it predates the curvature
    ratios and above-noise tests: its weights for those are zero); one that
    reads channels the packing does not fill is refused."""
    with np.load(path, allow_pickle=False) as z:
        n_in = int(z["input_channels"]) if "input_channels" in z.files else 9
        if n_in > INPUT_CHANNELS:
            raise ValueError(f"{path}: the network reads {n_in} input channels, the packing "
                             f"fills {INPUT_CHANNELS}")
        return z["params"].astype(np.float32), float(z["feature_lo"]), float(z["feature_hi"])


def save_weights(path: str, params, lo: float, hi: float, **info):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    info.setdefault("input_channels", INPUT_CHANNELS)
    np.savez(path, params=np.asarray(params, np.float32), feature_lo=np.float32(lo),
             feature_hi=np.float32(hi), **{k: np.asarray(v) for k, v in info.items()})


@dataclass
class TrainFrame:
    """A mosaic and its develop(Stage::Demosaiced) planes on the device."""
    header: FrameHeader
    params: DemosaicParams
    quantizer: int
    samples: object      # torch int16, coded_height x coded_width
    target: object       # torch float32, 3 x height x width

    def raw(self) -> N.PrrNetFrame:
        f = N.PrrNetFrame()
        f.d_samples = self.samples.data_ptr()
        f.stride = self.header.coded_width()
        f.header = self.header._raw
        f.params = self.params._raw()
        f.d_target = self.target.data_ptr()
        return f


def frame_pair(packet: bytes, device="cuda") -> TrainFrame:
    """Decode a packet and develop it at Stage::Demosaiced with the plain
    path; keep samples and planes on the device."""
    torch = _torch()
    fr = decode_frame(packet)
    q = fr.quantizer()
    params = DemosaicParams.new(fr.header, q)
    planes = develop_planar(fr, params, Stage.Demosaiced)
    samples = torch.from_numpy(np.ascontiguousarray(fr.samples).view(np.int16).copy()).to(device)
    target = torch.from_numpy(np.stack(planes.planes).reshape(3, fr.height(), fr.width())).to(device)
    return TrainFrame(fr.header, params, q, samples, target)


def scene_packets(n: int, size, seed: int, qscales=(15, 45), workers: int = 16,
                  pattern: int = 0, curves: bool = False) -> List[tuple]:
    """(packet, qscale) of n seeded RGGB scenes (config-0 recipe: 8
    gradients, 200 discs, Poisson-Gaussian noise, v0, default curve,
    white 4095), qscale drawn uniformly per scene; with curves, every other
    scene is v1 with a seeded 8-point curve instead. Encoded on the host in
    parallel (the encoder releases the GIL)."""
    from .workloads import seeded_curve
    w, h = size
    rng = np.random.default_rng(seed)
    qs = rng.integers(qscales[0], qscales[1] + 1, size=n)

    def one(i):
        custom = curves and i % 2 == 1
        sc = S.scene(w, h, pattern, seed=seed * 1000003 + i + 1, version=1 if custom else 0,
                     alignment=2, qscale=(int(qs[i]), int(qs[i])), white_level=4095,
                     curve=seeded_curve(seed * 7919 + i) if custom else None, threads=1)
        return S.encode_scene(sc), int(qs[i])

    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(one, range(n)))


class Net:
    """The network of include/proresraw_learned.h on the calling thread's
    context."""

    def __init__(self, params: Optional[np.ndarray] = None):
        self.lib = N.lib()
        self.ctx = context()
        h = ctypes.c_void_p()
        err = N.PrrError()
        _check(self.lib.prr_net_create(self.ctx.handle, ctypes.byref(h), ctypes.byref(err)), err)
        self.handle = h
        self.batch = self.cells = 0
        self._frames = None
        if params is not None:
            self.set_params(params)

    def __del__(self):
        try:
            self.lib.prr_net_destroy(self.handle)
        except Exception:
            pass

    def _call(self, name, *args):
        err = N.PrrError()
        _check(getattr(self.lib, name)(self.handle, *args, ctypes.byref(err)), err)

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == PARAMS
        self._call("prr_net_set_params", p.ctypes.data)

    def params(self) -> np.ndarray:
        p = np.zeros(PARAMS, np.float32)
        self._call("prr_net_get_params", p.ctypes.data)
        return p

    def set_range(self, lo: float, hi: float):
        self._call("prr_net_set_range", ctypes.c_float(lo), ctypes.c_float(hi))

    def train_setup(self, frames: List[TrainFrame], crops: np.ndarray, cells: int, margin: int):
        """crops: (steps, batch, 3) uint32 of frame index, top, left."""
        crops = np.ascontiguousarray(crops, np.uint32)
        steps, batch, _ = crops.shape
        arr = (N.PrrNetFrame * len(frames))(*[f.raw() for f in frames])
        self._frames = (frames, arr)
        self._call("prr_net_train_setup", arr, len(frames), crops.ctypes.data, steps, batch,
                   cells, margin)
        self.batch, self.cells = batch, cells

    def set_schedule(self, lr_min: float, decay_steps: int):
        self._call("prr_net_set_schedule", ctypes.c_float(lr_min), decay_steps)

    def set_loss(self, tau: float):
        """Loss of later steps: mean of d^2 + d^4 / tau^2 (0: squared error)."""
        self._call("prr_net_set_loss", ctypes.c_float(tau))

    def train_step(self, lr: float, stream=None):
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        self._call("prr_net_train_step", ctypes.c_float(lr), ctypes.c_void_p(s.cuda_stream))

    def read(self, what: int, layer: int = 0) -> np.ndarray:
        act = self.batch * self.cells * self.cells * C
        if what in (X0, H, R, DOUT):
            out = np.zeros(act, np.uint16)
        elif what == OUT:
            out = np.zeros(act, np.float32)
        elif what == GRAD:
            out = np.zeros(PARAMS, np.float32)
        elif what == LOSS:
            out = np.zeros(1, np.float32)
        else:
            out = np.zeros(1, np.uint32)
        self._call("prr_net_read", what, layer, out.ctypes.data, out.nbytes)
        if what in (X0, H, R, DOUT):
            return (out.astype(np.uint32) << 16).view(np.float32).reshape(
                self.batch, self.cells, self.cells, C)
        if what == OUT:
            return out.reshape(self.batch, self.cells, self.cells, C)
        return out

    def develop(self, header: FrameHeader, samples, params: DemosaicParams, stage: Stage,
                planes, status=None, check_tiles: int = 8, min_db: float = 45.0, stream=None):
        """Learned develop of device samples into three device planes."""
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        raw = params._raw()
        self._call("prr_net_develop_device", ctypes.byref(header._raw), samples.data_ptr(),
                   ctypes.byref(raw), status.data_ptr() if status is not None else None,
                   int(stage), check_tiles, ctypes.c_float(min_db), planes[0].data_ptr(),
                   planes[1].data_ptr(), planes[2].data_ptr(), ctypes.c_void_p(s.cuda_stream))

    def set_pairs(self, on: bool):
        """The network's layers two a launch (the default) or one a launch;
        the same bits either way."""
        self._call("prr_net_set_pairs", int(bool(on)))

    def last_check(self):
        """(used_plain, lowest sampled-tile PSNR) of the last develop."""
        used, psnr = ctypes.c_uint32(), ctypes.c_float()
        self._call("prr_net_last_check", ctypes.byref(used), ctypes.byref(psnr))
        return bool(used.value), float(psnr.value)

    def frame_input(self, header: FrameHeader) -> np.ndarray:
        """The whole-frame input the last develop packed (it must have run
        the network): (height/2, width/2, 32) float32 of the bf16 values."""
        h2, w2 = header.height // 2, header.width // 2
        out = np.zeros(h2 * w2 * C, np.uint16)
        self._call("prr_net_read", FRAME_X0, 0, out.ctypes.data, out.nbytes)
        return (out.astype(np.uint32) << 16).view(np.float32).reshape(h2, w2, C)


class NetGroup:
    """Learned develop of a batch's frames over several contexts with the
    same weights, each on a stream of its own, so that one frame's network
    runs beside the next frame's check, packing and composition (kernels
    that leave most of the matrix cores, or most of the CUs, idle). Each
    frame goes to the context with the fewest sensels queued so far (the
    first on a tie); its result is that of Net.develop (the same kernels on
    the same inputs)."""

    def __init__(self, params, lo: float, hi: float, n: int = 2):
        torch = _torch()
        if n < 1:
            raise ValueError("at least one context")
        self.nets = [Net(params) for _ in range(n)]
        for net in self.nets:
            net.set_range(lo, hi)
        self.streams = [torch.cuda.Stream() for _ in range(n)]

    def develop(self, frames, stage: Stage, check_tiles: int = 8, min_db: float = 45.0,
                stream=None):
        """frames: (header, samples, params, planes, status) tuples. Work
        queued on `stream` before the call comes first for every frame, and
        `stream` waits for all of them (capturable in a graph)."""
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream()
        for side in self.streams:
            side.wait_stream(s)
        queued = [0] * len(self.nets)
        for header, samples, params, planes, status in frames:
            k = queued.index(min(queued))
            queued[k] += header.width * header.height
            self.nets[k].develop(header, samples, params, stage, planes, status=status,
                                 check_tiles=check_tiles, min_db=min_db, stream=self.streams[k])
        for side in self.streams:
            s.wait_stream(side)


def random_crops(rng, frames: List[TrainFrame], steps: int, batch: int, cells: int,
                 hard=None, hard_frac: float = 0.0) -> np.ndarray:
    """(steps, batch, 3) crops of 2*cells sensels: frame, top, left (even).
    With `hard` (per frame, the squared error of each 64x64 tile of the
    current network's output), a fraction hard_frac of the crops are drawn
    around tiles picked in proportion to their error (frames likewise), so
    training dwells where the network is furthest from plain develop."""
    crops = np.zeros((steps, batch, 3), np.uint32)
    idx = rng.integers(0, len(frames), size=(steps, batch))
    for k, f in enumerate(frames):
        sel = idx == k
        n = int(sel.sum())
        if not n:
            continue
        crops[..., 0][sel] = k
        crops[..., 1][sel] = 2 * rng.integers(0, (f.header.height - 2 * cells) // 2 + 1, size=n)
        crops[..., 2][sel] = 2 * rng.integers(0, (f.header.width - 2 * cells) // 2 + 1, size=n)
    if hard is None or hard_frac <= 0:
        return crops
    flat = crops.reshape(-1, 3)
    pick = np.nonzero(rng.random(flat.shape[0]) < hard_frac)[0]
    sizes = [t.size for t in hard]
    weights = np.concatenate([t.reshape(-1) for t in hard]).astype(np.float64)
    weights /= weights.sum()
    starts = np.cumsum([0] + sizes)
    chosen = rng.choice(weights.size, size=pick.size, p=weights)
    k = np.searchsorted(starts, chosen, side="right") - 1
    for i, c, fk in zip(pick, chosen, k):
        f = frames[fk]
        t = c - starts[fk]
        tw = hard[fk].shape[1]
        ty, tx = divmod(int(t), tw)
        side = 2 * cells
        # The tile inside the crop's loss region, at a random place.
        top = 64 * ty + 32 - side // 2 + int(rng.integers(-cells // 2, cells // 2 + 1))
        left = 64 * tx + 32 - side // 2 + int(rng.integers(-cells // 2, cells // 2 + 1))
        top = min(max(top, 0), f.header.height - side) & ~1
        left = min(max(left, 0), f.header.width - side) & ~1
        flat[i] = (fk, top, left)
    return crops


def tile_errors(net: "Net", frames: List[TrainFrame], stream=None) -> List[np.ndarray]:
    """Per frame, the squared error of the network's develop(Demosaiced)
    against the frame's target, summed over each 64x64 tile."""
    torch = _torch()
    s = stream if stream is not None else torch.cuda.Stream()
    out = []
    for f in frames:
        h, w = f.header.height, f.header.width
        planes = [torch.empty(h * w, dtype=torch.float32, device="cuda") for _ in range(3)]
        net.develop(f.header, f.samples, f.params, Stage.Demosaiced, planes, check_tiles=0,
                    stream=s)
        s.synchronize()
        d = (torch.stack(planes).reshape(3, h, w) - f.target.reshape(3, h, w)).double() ** 2
        th, tw = h // 64, w // 64
        e = d[:, :th * 64, :tw * 64].reshape(3, th, 64, tw, 64).sum(dim=(0, 2, 4))
        out.append(e.cpu().numpy())
    return out


def train(net: Net, frames: List[TrainFrame], steps: int, batch: int = 16, cells: int = 144,
          margin: int = 8, lr: float = 1e-3, seed: int = 0, log_every: int = 1000,
          log=print, use_graph: bool = True, lr_min: Optional[float] = None, hard=None,
          hard_frac: float = 0.0, tau: float = 0.0):
    """Adam over random crops of 2*cells sensels (loss on the cells at least
    `margin` from the crop edge); one step captured in a graph and replayed.
    With lr_min, the rate decays from lr to lr_min along a cosine over the
    run; with `hard` (tile_errors) a fraction hard_frac of the crops sit
    around the hardest tiles; tau > 0 trains on d^2 + d^4 / tau^2 instead of
    the squared error. Returns the list of (step, loss)."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    net.train_setup(frames, random_crops(rng, frames, steps, batch, cells, hard, hard_frac),
                    cells, margin)
    net.set_schedule(lr_min if lr_min is not None else 0.0, steps if lr_min is not None else 0)
    net.set_loss(tau)
    stream = torch.cuda.Stream()
    history = []
    t0 = time.time()
    with torch.cuda.stream(stream):
        net.train_step(lr, stream)
        done = 1
        graph = None
        if use_graph and steps > 1:
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph, stream=stream):
                net.train_step(lr, stream)
        while done < steps:
            chunk = min(log_every - done % log_every, steps - done)
            for _ in range(chunk):
                if graph is not None:
                    graph.replay()
                else:
                    net.train_step(lr, stream)
            done += chunk
            stream.synchronize()
            loss = float(net.read(LOSS)[0])
            history.append((done, loss))
            log(f"step {done} loss {loss:.3e} ({time.time() - t0:.1f} s)")
    return history


def psnr(a, b) -> float:
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0 else -10.0 * math.log10(mse)
