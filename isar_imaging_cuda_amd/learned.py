"""Learned warm start for the fast-entropy autofocus (include/isar_learned.h).

The reference's autofocus (phase_adjustment.cu:282-324) starts from the Doppler-tracking phase
phi0 and runs 120 fixed-point iterations. The learned path runs the first FRONT_ITERS iterations
(phi1, phi2, ...), gives the per-pulse phase changes of those iterations to a small network, and
continues the same iteration from phi_F times the predicted remaining change. The network has
five dense layers of width 128, BF16 weights and FP32 accumulation on the matrix cores, and is
trained on frames the pipeline synthesises and focuses itself, with the remaining change
phi120 conj(phi_F) of the plain iteration as the target. The plain iteration remains the judge:
a learned result is accepted when its image entropy is no worse than the plain result's (within
1e-5 relative).

  python -m isar_imaging_cuda_amd.learned train --frames 20000 --out weights.npz
  python -m isar_imaging_cuda_amd.learned eval --weights weights.npz --frames 1000
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import ctypes as C
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import ImagingError, Imager, Options, RadarParams, _check, lib
from . import synth

HIDDEN = 128
N_LAYERS = 5
FRONT_ITERS = 2           # plain iterations whose phase changes form the network input
ENTROPY_REL = 1e-5
WEIGHTS_DIR = Path(__file__).resolve().parent / "weights"


def widths_for(M: int, front_iters: int = FRONT_ITERS) -> list[int]:
    """Input: the per-pulse phase changes of the first `front_iters` iterations as (re, im)
    pairs (2M values each); output: the remaining change to the converged phase (2M values)."""
    return [2 * M * front_iters] + [HIDDEN] * (N_LAYERS - 1) + [2 * M]


def he_init(widths: list[int], seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases, packed per layer as W (in x out) then b."""
    rng = np.random.default_rng(seed)
    flat = []
    for i, o in zip(widths[:-1], widths[1:]):
        flat.append((rng.standard_normal((i, o)) * np.sqrt(2.0 / i)).ravel())
        flat.append(np.zeros(o))
    return np.concatenate(flat).astype(np.float32)


def _bind() -> C.CDLL:
    L = lib()
    if getattr(L, "_learned_bound", False):
        return L
    vp, i, f = C.c_void_p, C.c_int, C.c_float
    L.isar_mlp_create.argtypes = [i, C.POINTER(C.c_int), i, C.POINTER(vp)]
    L.isar_mlp_destroy.argtypes = [vp]
    L.isar_mlp_param_count.argtypes = [vp]
    L.isar_mlp_set_params.argtypes = [vp, vp]
    L.isar_mlp_get_params.argtypes = [vp, vp]
    L.isar_mlp_get_grads.argtypes = [vp, vp]
    L.isar_mlp_forward.argtypes = [vp, i, vp, vp, vp]
    L.isar_mlp_backward.argtypes = [vp, i, vp, vp]
    L.isar_mlp_adam_step.argtypes = [vp, f, f, f, f, i, vp]
    L.isar_mse.argtypes = [i, i, vp, vp, vp, vp, vp]
    L.isar_phase_change.argtypes = [i, i, vp, vp, vp, i, vp]
    L.isar_compose_phase.argtypes = [i, i, vp, vp, vp, vp]
    L.isar_image_entropy.argtypes = [i, i, vp, vp, vp]
    L.isar_select_phase.argtypes = [i, i, vp, vp, vp, vp, vp, vp, vp]
    L.isar_half_phase.argtypes = [i, i, vp, vp, vp, vp]
    L.isar_accept_image.argtypes = [i, i, vp, vp, vp, vp, C.c_double, vp, vp, vp]
    L._learned_bound = True
    return L


def _p(x) -> C.c_void_p:
    return C.c_void_p(0 if x is None else int(x))


class Mlp:
    """The network object of isar_learned.h (device buffers for batches up to max_batch)."""

    def __init__(self, widths: list[int], max_batch: int):
        L = _bind()
        self.widths = list(widths)
        self.max_batch = max_batch
        self._h = C.c_void_p()
        arr = (C.c_int * len(widths))(*widths)
        _check(L.isar_mlp_create(len(widths) - 1, arr, max_batch, C.byref(self._h)), "isar_mlp_create")
        self.n_params = int(L.isar_mlp_param_count(self._h))

    def close(self) -> None:
        if self._h:
            _bind().isar_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        _check(_bind().isar_mlp_set_params(self._h, p.ctypes.data_as(C.c_void_p)), "isar_mlp_set_params")

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        _check(_bind().isar_mlp_get_params(self._h, p.ctypes.data_as(C.c_void_p)), "isar_mlp_get_params")
        return p

    def get_grads(self) -> np.ndarray:
        g = np.zeros(self.n_params, np.float32)
        _check(_bind().isar_mlp_get_grads(self._h, g.ctypes.data_as(C.c_void_p)), "isar_mlp_get_grads")
        return g

    def forward(self, batch: int, d_in: int, d_out: int, stream: int | None = None) -> None:
        _check(_bind().isar_mlp_forward(self._h, batch, _p(d_in), _p(d_out), _p(stream)), "isar_mlp_forward")

    def backward(self, batch: int, d_grad_out: int, stream: int | None = None) -> None:
        _check(_bind().isar_mlp_backward(self._h, batch, _p(d_grad_out), _p(stream)), "isar_mlp_backward")

    def adam_step(self, lr: float, step: int, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8,
                  stream: int | None = None) -> None:
        _check(_bind().isar_mlp_adam_step(self._h, lr, beta1, beta2, eps, step, _p(stream)), "isar_mlp_adam_step")


def mse(rows: int, cols: int, d_pred: int, d_target: int, d_grad: int | None, d_loss: int,
        stream: int | None = None) -> None:
    _check(_bind().isar_mse(rows, cols, _p(d_pred), _p(d_target), _p(d_grad), _p(d_loss), _p(stream)), "isar_mse")


def phase_change(frames: int, M: int, d_a: int, d_b: int, d_out: int, stream: int | None = None,
                 stride: int | None = None) -> None:
    _check(_bind().isar_phase_change(frames, M, _p(d_a), _p(d_b), _p(d_out), stride or 2 * M, _p(stream)),
           "isar_phase_change")


def compose_phase(frames: int, M: int, d_phi_b: int, d_pred: int, d_phi_w: int, stream: int | None = None) -> None:
    _check(_bind().isar_compose_phase(frames, M, _p(d_phi_b), _p(d_pred), _p(d_phi_w), _p(stream)),
           "isar_compose_phase")


def image_entropy(frames: int, n: int, d_img: int, d_out: int, stream: int | None = None) -> None:
    _check(_bind().isar_image_entropy(frames, n, _p(d_img), _p(d_out), _p(stream)), "isar_image_entropy")


def select_phase(frames: int, M: int, d_phi_a: int, d_ent_a: int, d_phi_b: int, d_ent_b: int, d_out: int,
                 stream: int | None = None, d_ent_out: int | None = None) -> None:
    _check(_bind().isar_select_phase(frames, M, _p(d_phi_a), _p(d_ent_a), _p(d_phi_b), _p(d_ent_b), _p(d_out),
                                     _p(d_ent_out), _p(stream)), "isar_select_phase")


def half_phase(frames: int, M: int, d_phi_a: int, d_phi_b: int, d_out: int, stream: int | None = None) -> None:
    _check(_bind().isar_half_phase(frames, M, _p(d_phi_a), _p(d_phi_b), _p(d_out), _p(stream)), "isar_half_phase")


def accept_image(frames: int, n: int, d_img_plain: int, d_ent_plain: int, d_img_learned: int, d_ent_learned: int,
                 d_out: int, d_accepted: int | None = None, rel_tol: float = ENTROPY_REL,
                 stream: int | None = None) -> None:
    _check(_bind().isar_accept_image(frames, n, _p(d_img_plain), _p(d_ent_plain), _p(d_img_learned),
                                     _p(d_ent_learned), rel_tol, _p(d_out), _p(d_accepted), _p(stream)),
           "isar_accept_image")


# ------------------------------------------------------------------------------------------
# the pipeline with a learned warm start
# ------------------------------------------------------------------------------------------
class WarmStartImager:
    """An imaging context plus the network: front (FRONT_ITERS plain iterations) -> network ->
    judge (per frame, the start of lowest image entropy among the plain phase, the prediction and
    the half step toward it) -> back (the iteration from the chosen start, then the image)."""

    def __init__(self, params: RadarParams, max_frames: int, net_params: np.ndarray | None = None,
                 front_iters: int = FRONT_ITERS):
        import torch
        self.torch = torch
        self.M = params.echo_num
        self.W = params.range_num_cut
        self.K = max_frames
        self.F = front_iters
        self.imager = Imager(params, max_frames=max_frames)
        self.net = Mlp(widths_for(self.M, front_iters), max_frames)
        if net_params is not None:
            self.net.set_params(net_params)
        dev = torch.device("cuda:0")
        # phis[i]: phase after i plain iterations (phis[0] = Doppler tracking)
        self.phis = [torch.zeros((max_frames, self.M, 2), dtype=torch.float32, device=dev)
                     for _ in range(front_iters + 1)]
        self.phi_w = torch.zeros_like(self.phis[0])
        self.feat = torch.zeros((max_frames, 2 * self.M * front_iters), dtype=torch.float32, device=dev)
        self.pred = torch.zeros((max_frames, 2 * self.M), dtype=torch.float32, device=dev)
        self.phi_s = torch.zeros_like(self.phis[0])      # the judged start
        self.phi_h = torch.zeros_like(self.phis[0])      # half step toward the prediction
        self.phi_t = torch.zeros_like(self.phis[0])      # the better of the plain phase and the prediction
        self.ent_plain = torch.zeros(max_frames, dtype=torch.float64, device=dev)
        self.ent_warm = torch.zeros(max_frames, dtype=torch.float64, device=dev)
        self.ent_half = torch.zeros(max_frames, dtype=torch.float64, device=dev)
        self.ent_t = torch.zeros(max_frames, dtype=torch.float64, device=dev)
        self.ent_start = torch.zeros(max_frames, dtype=torch.float64, device=dev)

    @property
    def phi_last(self):
        return self.phis[-1]

    def front(self, K: int, echo_ptr: int, vel_ptr: int, stream: int) -> None:
        """Range alignment through selection and the first F plain autofocus iterations, one at a
        time so that every intermediate phase is kept."""
        self.imager.run_device(echo_ptr, vel_ptr, 0, K, Options(iterations=1, stage_mode=1), stream=stream)
        self.imager.phases(K, self.phis[0].data_ptr(), self.phis[1].data_ptr(), stream)
        for i in range(2, self.F + 1):
            o = Options(iterations=1, stage_mode=2, warm_start=True)
            self.imager.run_device(0, 0, 0, K, o, phi0_ptr=self.phis[i - 1].data_ptr(), stream=stream)
            self.imager.phases(K, None, self.phis[i].data_ptr(), stream)

    def features(self, K: int, stream: int) -> None:
        for i in range(1, self.F + 1):
            phase_change(K, self.M, self.phis[i - 1].data_ptr(), self.phis[i].data_ptr(),
                         self.feat.data_ptr() + (i - 1) * 2 * self.M * 4, stream, stride=2 * self.M * self.F)

    def predict(self, K: int, stream: int) -> None:
        """phi_w = phi_F * unit(net(phase changes of the first F iterations))."""
        self.features(K, stream)
        self.net.forward(K, self.feat.data_ptr(), self.pred.data_ptr(), stream)
        compose_phase(K, self.M, self.phi_last.data_ptr(), self.pred.data_ptr(), self.phi_w.data_ptr(), stream)

    def back(self, K: int, img_ptr: int, iterations: int, stream: int, warm: bool = True, phi=None) -> None:
        """Autofocus from `phi` (default: the judged start phi_s; with warm=False the plain
        Doppler start) and the image."""
        o = Options(iterations=iterations, stage_mode=2, warm_start=warm)
        start = (self.phi_s if phi is None else phi).data_ptr() if warm else None
        self.imager.run_device(0, 0, img_ptr, K, o, phi0_ptr=start, stream=stream)

    def judge(self, K: int, img_ptr: int, stream: int) -> None:
        """The plain objective decides: image entropies from phi_F, from the prediction phi_w and
        from the half step toward it (img_ptr: K x M x W floats of scratch); phi_s = the start of
        lowest entropy per frame (the first of them on ties), ent_start its entropy."""
        n = self.M * self.W
        half_phase(K, self.M, self.phi_last.data_ptr(), self.phi_w.data_ptr(), self.phi_h.data_ptr(), stream)
        for phi, ent in ((self.phi_last, self.ent_plain), (self.phi_w, self.ent_warm), (self.phi_h, self.ent_half)):
            self.back(K, img_ptr, 0, stream, phi=phi)
            image_entropy(K, n, img_ptr, ent.data_ptr(), stream)
        select_phase(K, self.M, self.phi_last.data_ptr(), self.ent_plain.data_ptr(), self.phi_w.data_ptr(),
                     self.ent_warm.data_ptr(), self.phi_t.data_ptr(), stream, self.ent_t.data_ptr())
        select_phase(K, self.M, self.phi_t.data_ptr(), self.ent_t.data_ptr(), self.phi_h.data_ptr(),
                     self.ent_half.data_ptr(), self.phi_s.data_ptr(), stream, self.ent_start.data_ptr())

    def run(self, K: int, echo_ptr: int, vel_ptr: int, img_ptr: int, iterations: int, stream: int,
            final_judge: bool = True):
        """The learned chain: F + `iterations` autofocus iterations in total. With final_judge
        (the default) the plain 120-iteration result is formed too and remains the judge: a
        frame keeps its learned image only when that image's entropy is no worse than the plain
        image's (within 1e-5 relative), else it gets the plain image. Returns the per-frame
        acceptance flags (a K-element int32 device tensor) or None without the judgement."""
        self.front(K, echo_ptr, vel_ptr, stream)
        self.predict(K, stream)
        self.judge(K, img_ptr, stream)
        self.back(K, img_ptr, iterations, stream)
        if not final_judge:
            return None
        torch = self.torch
        if getattr(self, "img_plain", None) is None:
            dev = self.phis[0].device
            self.img_plain = torch.zeros((self.K, self.M, self.W), dtype=torch.float32, device=dev)
            self.ent_final_plain = torch.zeros(self.K, dtype=torch.float64, device=dev)
            self.ent_final_learned = torch.zeros(self.K, dtype=torch.float64, device=dev)
            self.accepted = torch.zeros(self.K, dtype=torch.int32, device=dev)
        n = self.M * self.W
        image_entropy(K, n, img_ptr, self.ent_final_learned.data_ptr(), stream)
        self.back(K, self.img_plain.data_ptr(), 120, stream, warm=False)
        image_entropy(K, n, self.img_plain.data_ptr(), self.ent_final_plain.data_ptr(), stream)
        accept_image(K, n, self.img_plain.data_ptr(), self.ent_final_plain.data_ptr(), img_ptr,
                     self.ent_final_learned.data_ptr(), img_ptr, self.accepted.data_ptr(), stream=stream)
        return self.accepted[:K]


# ------------------------------------------------------------------------------------------
# data, training and evaluation
# ------------------------------------------------------------------------------------------
def synth_frames(config: int, seeds: list[int], threads: int = 16, **over) -> tuple[np.ndarray, np.ndarray, dict]:
    cfgs = [synth.preset(config, seed=s, **over) for s in seeds]
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        data = list(ex.map(synth.frame, cfgs))
    echo = np.stack([d[0] for d in data])
    vel = np.stack([d[1] for d in data])
    return echo, vel, synth.radar_header(cfgs[0])


def synth_frames_device(config: int, seeds: list[int], **over):
    """The same frames as synth_frames, generated on the GPU: (echo (K, M, N, 2) float32 and
    velocity (K, M) float64 device tensors, radar header)."""
    import torch
    cfg = synth.preset(config, **over)
    dev = torch.device("cuda:0")
    e = torch.empty((len(seeds), cfg.echo_num, cfg.range_num, 2), dtype=torch.float32, device=dev)
    v = torch.empty((len(seeds), cfg.echo_num), dtype=torch.float64, device=dev)
    synth.frames_device(cfg, seeds, e.data_ptr(), v.data_ptr())
    return e, v, synth.radar_header(cfg)


def host_echo(echo: np.ndarray) -> np.ndarray:
    """complex64 frames (K, M, N) as the float32 (K, M, N, 2) array of (re, im) pairs that the
    device echo tensors hold (a plain float32 view would be (K, M, 2N))."""
    if echo.ndim != 3 or echo.dtype != np.complex64:
        raise ImagingError(f"host_echo: expected complex64 (K, M, N) frames, got {echo.dtype} {echo.shape}")
    K, M, N = echo.shape
    return np.ascontiguousarray(echo).view(np.float32).reshape(K, M, N, 2)


def _frames(config: int, seeds: list[int], gpu_synth: bool, over: dict):
    """(echo (K, M, N, 2) float32, velocity (K, M) float64) device tensors and the radar header
    of the frames of `seeds`; the callers take K, M, N from the echo tensor's shape."""
    import torch
    if gpu_synth:
        e, v, h = synth_frames_device(config, seeds, **over)
    else:
        echo, vel, h = synth_frames(config, seeds, **over)
        dev = torch.device("cuda:0")
        e, v = torch.from_numpy(host_echo(echo)).to(dev), torch.from_numpy(vel).to(dev)
    if e.dim() != 4 or e.shape[3] != 2 or tuple(v.shape) != tuple(e.shape[:2]):
        raise ImagingError(f"frame tensors of shape {tuple(e.shape)} and {tuple(v.shape)}; expected (K, M, N, 2) and (K, M)")
    return e, v, h


def _radar(h: dict, M: int, N: int) -> RadarParams:
    return RadarParams(M, N, h["band_width"], h["fc"], h["Tp"], h["Fs"])


def make_dataset(seeds: list[int], chunk: int = 256, config: int = 4, log=print, front_iters: int = FRONT_ITERS,
                 gpu_synth: bool = False, **over):
    """Features (the phase changes of the first FRONT_ITERS plain iterations) and targets
    (phi120 conj(phi_F), the remaining change of the plain iteration) for the frames of `seeds`
    (synthesised on the host or, with gpu_synth, on the GPU; imaged on the GPU). `over` replaces
    fields of the synthetic preset (e.g. range_num)."""
    import torch
    dev = torch.device("cuda:0")
    feats, targets = [], []
    ws = None
    t0 = time.time()
    for c0 in range(0, len(seeds), chunk):
        part = seeds[c0:c0 + chunk]
        e, v, h = _frames(config, part, gpu_synth, over)
        K, M, N = e.shape[0], e.shape[1], e.shape[2]
        if ws is None:
            ws = WarmStartImager(_radar(h, M, N), chunk, front_iters=front_iters)
            phi_end = torch.zeros((chunk, M, 2), dtype=torch.float32, device=dev)
        s = torch.cuda.current_stream(dev).cuda_stream
        ws.front(K, e.data_ptr(), v.data_ptr(), s)
        ws.features(K, s)
        ws.back(K, 0, 120, s, warm=False)
        ws.imager.phases(K, None, phi_end.data_ptr(), s)
        tgt = torch.zeros((K, 2 * M), dtype=torch.float32, device=dev)
        phase_change(K, M, ws.phi_last.data_ptr(), phi_end.data_ptr(), tgt.data_ptr(), s)
        torch.cuda.synchronize()
        feats.append(ws.feat[:K].clone())
        targets.append(tgt)
        log(f"[data] {c0 + K}/{len(seeds)} frames, {time.time() - t0:.1f} s")
    return torch.cat(feats), torch.cat(targets)


def train(feats, targets, epochs: int = 30, batch: int = 256, lr: float = 1e-3, seed: int = 0, log=print):
    """Adam on the MSE between the predicted and the plain phase change; all on the GPU through
    the network's own forward, backward and update kernels."""
    import torch
    dev = feats.device
    n, d_in = feats.shape
    d = targets.shape[1]
    net = Mlp(widths_for(d // 2, d_in // d), batch)
    net.set_params(he_init(net.widths, seed))
    out = torch.zeros((batch, d), dtype=torch.float32, device=dev)
    grad = torch.zeros_like(out)
    loss_rows = torch.zeros(batch, dtype=torch.float64, device=dev)
    s = torch.cuda.current_stream(dev).cuda_stream
    rng = np.random.default_rng(seed)
    step, hist = 0, []
    for ep in range(epochs):
        order = torch.from_numpy(rng.permutation(n)).to(dev)
        tot, cnt = 0.0, 0
        for b0 in range(0, n - batch + 1, batch):
            idx = order[b0:b0 + batch]
            x = feats[idx].contiguous()
            y = targets[idx].contiguous()
            net.forward(batch, x.data_ptr(), out.data_ptr(), s)
            mse(batch, d, out.data_ptr(), y.data_ptr(), grad.data_ptr(), loss_rows.data_ptr(), s)
            net.backward(batch, grad.data_ptr(), s)
            step += 1
            net.adam_step(lr * (0.5 ** (ep // 10)), step, stream=s)
            tot += float(loss_rows.sum().item())
            cnt += 1
        hist.append(tot / max(cnt, 1))
        log(f"[train] epoch {ep + 1}/{epochs} loss {hist[-1]:.5f}")
    return net.get_params(), hist


def iterations_to_reach(ent: np.ndarray, target: np.ndarray) -> np.ndarray:
    """First column index k with ent[:, k] <= target (len(ks) when never)."""
    ok = ent <= target[:, None]
    first = np.where(ok.any(axis=1), ok.argmax(axis=1), ent.shape[1])
    return first


def evaluate(params: np.ndarray, seeds: list[int], chunk: int = 256, config: int = 4, max_iter: int = 120,
             log=print, front_iters: int = FRONT_ITERS, gpu_synth: bool = False, **over) -> dict:
    """Iterations until the image entropy is within 1e-5 relative of the plain 120-iteration
    result, plain (from phi0) versus learned (FRONT_ITERS iterations, the network, the entropy
    judgement, then the iteration from the judged start), on the frames of `seeds`. The learned
    chain shares the plain front iterations: a frame the plain iteration settles within them counts
    the same for both. A learned frame that never reaches the plain entropy within max_iter
    iterations falls back to the plain result and is charged both."""
    import torch
    dev = torch.device("cuda:0")
    plain_it, learned_it, accepted, e_plain_all, e_learned_all, kept = [], [], [], [], [], []
    ws = None
    for c0 in range(0, len(seeds), chunk):
        part = seeds[c0:c0 + chunk]
        e, v, h = _frames(config, part, gpu_synth, over)
        K, M, N = e.shape[0], e.shape[1], e.shape[2]
        if ws is None:
            ws = WarmStartImager(_radar(h, M, N), chunk, params, front_iters=front_iters)
            img = torch.zeros((chunk, M, ws.W), dtype=torch.float32, device=dev)
            ent = torch.zeros(chunk, dtype=torch.float64, device=dev)
        s = torch.cuda.current_stream(dev).cuda_stream
        ws.front(K, e.data_ptr(), v.data_ptr(), s)
        ws.predict(K, s)
        ws.judge(K, img.data_ptr(), s)
        kept.append((ws.ent_start[:K] < ws.ent_plain[:K]).cpu().numpy())
        curve_p = np.zeros((K, max_iter + 1))
        curve_l = np.zeros((K, max_iter + 1))
        for k in range(max_iter + 1):
            ws.back(K, img.data_ptr(), k, s, warm=False)
            image_entropy(K, M * ws.W, img.data_ptr(), ent.data_ptr(), s)
            curve_p[:, k] = ent[:K].cpu().numpy()
            ws.back(K, img.data_ptr(), k, s, warm=True)
            image_entropy(K, M * ws.W, img.data_ptr(), ent.data_ptr(), s)
            curve_l[:, k] = ent[:K].cpu().numpy()
        e_plain = curve_p[:, 120] if max_iter >= 120 else curve_p[:, -1]
        target = e_plain * (1.0 + ENTROPY_REL)
        ip = iterations_to_reach(curve_p, target)
        il = iterations_to_reach(curve_l, target)
        ok = il <= max_iter
        plain_it.append(ip)
        learned_it.append(np.where(ip <= ws.F, ip, np.where(ok, ws.F + il, ws.F + max_iter + ip)))
        accepted.append(ok)
        e_plain_all.append(e_plain)
        e_learned_all.append(np.where(ok, curve_l[np.arange(K), np.minimum(il, max_iter)], e_plain))
        log(f"[eval] {c0 + K}/{len(seeds)} frames")
    ip = np.concatenate(plain_it)
    il = np.concatenate(learned_it)
    acc = np.concatenate(accepted)
    return {"frames": int(ip.size), "plain_mean_iterations": float(ip.mean()),
            "learned_mean_iterations": float(il.mean()),
            "iteration_ratio": float(ip.mean() / max(il.mean(), 1e-9)),
            "accepted_fraction": float(acc.mean()),
            "prediction_kept_fraction": float(np.concatenate(kept).mean()),
            "plain_median": float(np.median(ip)), "learned_median": float(np.median(il)),
            "entropy_rel_diff_max": float(np.max((np.concatenate(e_learned_all) - np.concatenate(e_plain_all))
                                                 / np.concatenate(e_plain_all)))}


def stage_times(params: np.ndarray, seeds: list[int], config: int = 4, front_iters: int = FRONT_ITERS,
                repeats: int = 5, gpu_synth: bool = False, **over) -> dict:
    """Device time per frame (ms, HIP events on the run stream, one batch of len(seeds) frames, the
    median of `repeats`) of the pieces the two chains are made of: one autofocus iteration, the
    image formation after the autofocus (MTRC, azimuth FFT), the network (features, forward,
    composed start) and the judgement of the start (three image formations, their entropies
    and two selections). Range alignment, centring, cut and bin selection are common to both
    chains and left out."""
    import torch
    dev = torch.device("cuda:0")
    e, v, h = _frames(config, seeds, gpu_synth, over)
    K, M, N = e.shape[0], e.shape[1], e.shape[2]
    ws = WarmStartImager(_radar(h, M, N), K, params, front_iters=front_iters)
    img = torch.zeros((K, M, ws.W), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    s = stream.cuda_stream
    ws.front(K, e.data_ptr(), v.data_ptr(), s)

    def timed(fn) -> float:
        out = []
        for _ in range(repeats + 1):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            torch.cuda.synchronize()
            out.append(a.elapsed_time(b))
        return float(np.median(out[1:]))

    t0 = timed(lambda: ws.back(K, img.data_ptr(), 0, s, warm=False))
    t120 = timed(lambda: ws.back(K, img.data_ptr(), 120, s, warm=False))
    t_net = timed(lambda: ws.predict(K, s))
    t_judge = timed(lambda: ws.judge(K, img.data_ptr(), s))
    return {"iteration": (t120 - t0) / 120 / K, "image": t0 / K, "network": t_net / K, "judge": t_judge / K,
            "batch_frames": K}


def load_weights(path: Path | str | None = None, M: int = 256) -> tuple[np.ndarray, int]:
    """Trained parameters and the number of front iterations they were trained for."""
    p = Path(path) if path else WEIGHTS_DIR / f"warm_start_m{M}.npz"
    if not p.exists():
        raise ImagingError(f"{p} is missing: train the warm-start network first")
    with np.load(p, allow_pickle=False) as z:
        return np.asarray(z["params"], dtype=np.float32), int(z["front_iters"])


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--frames", type=int, default=20000)
    t.add_argument("--epochs", type=int, default=30)
    t.add_argument("--seed0", type=int, default=1)
    t.add_argument("--out", default=str(WEIGHTS_DIR / "warm_start_m256.npz"))
    t.add_argument("--eval-frames", type=int, default=256)
    t.add_argument("--front-iters", type=int, default=FRONT_ITERS)
    t.add_argument("--config", type=int, default=4, help="synthetic preset of the training frames")
    t.add_argument("--range-num", type=int, default=0, help="range cells of the training frames (0: the preset's)")
    t.add_argument("--gpu-synth", action="store_true", help="synthesise the frames on the GPU")
    t.add_argument("--chunk", type=int, default=256)
    ev = sub.add_parser("eval")
    ev.add_argument("--weights", default=None)
    ev.add_argument("--frames", type=int, default=1000)
    ev.add_argument("--seed0", type=int, default=10_000_001)
    ev.add_argument("--config", type=int, default=4)
    ev.add_argument("--range-num", type=int, default=0)
    ev.add_argument("--gpu-synth", action="store_true")
    ev.add_argument("--chunk", type=int, default=256)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    over = {"range_num": a.range_num} if a.range_num else {}
    if a.cmd == "train":
        t0 = time.time()
        feats, targets = make_dataset(list(range(a.seed0, a.seed0 + a.frames)), chunk=a.chunk, config=a.config, log=log,
                                      front_iters=a.front_iters, gpu_synth=a.gpu_synth, **over)
        params, hist = train(feats, targets, epochs=a.epochs, log=log)
        out = Path(a.out)
        out.parent.mkdir(parents=True, exist_ok=True)
        np.savez(out, params=params, loss=np.asarray(hist), frames=a.frames, epochs=a.epochs,
                 front_iters=a.front_iters, config=a.config, range_num=a.range_num)
        log(f"[train] saved {out} after {time.time() - t0:.1f} s")
        if a.eval_frames:
            res = evaluate(params, list(range(10_000_001, 10_000_001 + a.eval_frames)), chunk=min(a.chunk, a.eval_frames),
                           config=a.config, log=log, front_iters=a.front_iters, gpu_synth=a.gpu_synth, **over)
            log(json.dumps(res))
    else:
        params, fi = load_weights(a.weights)
        res = evaluate(params, list(range(a.seed0, a.seed0 + a.frames)), chunk=min(a.chunk, a.frames), config=a.config,
                       log=log, front_iters=fi, gpu_synth=a.gpu_synth, **over)
        log(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
