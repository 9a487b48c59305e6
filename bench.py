"""Throughput benchmark of the MI355X imaging pipeline on 256 x 4096 complex64 frames.

One step = the full imaging chain (HPC, range alignment, centring and cut, Doppler tracking +
120-iteration fast-entropy autofocus, MTRC and the azimuth FFT) over one batch of `--frames`
distinct synthetic frames already resident in HBM. Prints one JSON line.

  python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--frames 64] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, GB/s
M, N, W = 256, 4096, 512


def algorithmic_bytes_per_frame(M: int, N: int, W: int, U: float) -> float:
    """Minimum-pass HBM bytes of one frame: 8MN(L+2) + 76MW + 8MU, L = ceil(log2 M)."""
    L = int(np.ceil(np.log2(M)))
    return 8.0 * M * N * (L + 2) + 76.0 * M * W + 8.0 * M * U


def stage_bytes(stage: str, M: int, N: int, W: int, U: float, frames: int, launches_per_run: int) -> float:
    """Algorithmic HBM bytes of one launch of a kernel family (whole batch), see DESIGN.md."""
    per_frame = {
        # every merge level re-reads the M/2 shifted rows (8N each) and writes their energy (4N)
        "ra_upd": (M / 2) * (8.0 * N + 4.0 * N),
        # group energies read (2 groups per pair) and the first one written back: ~ M/2 rows of 4N
        "ra_corr": (M / 2) * 4.0 * N * 1.5,
        # level 0: every echo row read (8N), its spectrum stored (8N), one energy row per pair (4N)
        "ra_pair0": M * (8.0 * N + 8.0 * N) + (M / 2) * 4.0 * N,
        "autofocus": 8.0 * M * W + 8.0 * M * U * 2,
        "mtrc_czt": 2 * 8.0 * M * W,
        "mtrc_rows": 2 * 8.0 * M * W,
        "mtrc_image": 8.0 * M * W + 4.0 * M * W,
        "pa_colfft": 8.0 * M * W,
    }[stage]
    if stage in ("ra_upd", "ra_corr"):
        per_frame /= 1.0
    return per_frame * frames


def cpu_baseline(frames: int = 24) -> dict:
    """FP64 oracle with OpenMP over frames on this host (bounded sample of the same workload)."""
    from isar_imaging_cuda_amd import synth
    from oracle import oracle as O

    cfgs = [synth.preset(2, seed=0xC0FFEE + k) for k in range(frames)]
    data = [synth.frame(c) for c in cfgs]
    echo = np.stack([d[0] for d in data])
    vel = np.stack([d[1] for d in data])
    h = synth.radar_header(cfgs[0])
    p = O.params(M, N, W, h["band_width"], h["fc"], h["Tp"], h["Fs"])
    O.isar_batch(echo[:1], vel[:1], p)                      # warm-up (plan caches)
    t0 = time.perf_counter()
    _, threads = O.isar_batch(echo, vel, p)
    dt = time.perf_counter() - t0
    return {"value": frames / dt, "unit": "frames/s", "cores": int(threads), "kind": "port",
            "sample": f"{frames} distinct 256x4096 frames, FP64 oracle, one OpenMP thread per frame, "
                      f"{dt:.1f} s wall"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=64, help="frames per batch (one step)")
    ap.add_argument("--distinct", type=int, default=64, help="distinct synthetic frames in the batch")
    ap.add_argument("--probe", default="auto", help="kernel family timed for the roofline figure")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-frames", type=int, default=24)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the imaging path shards "
              f"frames across GPUs with no exchange, see DESIGN.md).", file=sys.stderr)
        return 2

    import torch

    import isar_imaging_cuda_amd as isar
    from isar_imaging_cuda_amd import synth

    dev = torch.device("cuda:0")
    K = args.frames
    D = min(args.distinct, K)
    cfgs = [synth.preset(2, seed=0xBEEF + k) for k in range(D)]
    frames = [synth.frame(c) for c in cfgs]
    h = synth.radar_header(cfgs[0])
    echo_h = np.stack([frames[k % D][0] for k in range(K)])
    vel_h = np.stack([frames[k % D][1] for k in range(K)])
    echo = torch.from_numpy(echo_h.view(np.float32)).to(dev)          # K x M x 2N float32
    vel = torch.from_numpy(vel_h).to(dev)
    img = torch.empty((K, M, W), dtype=torch.float32, device=dev)
    params = isar.RadarParams(M, N, h["band_width"], h["fc"], h["Tp"], h["Fs"])
    opts = isar.Options()
    im = isar.Imager(params, max_frames=K)
    stream = torch.cuda.current_stream(dev)

    def step():
        im.run_device(echo.data_ptr(), vel.data_ptr(), img.data_ptr(), K, opts, stream=stream.cuda_stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # mean number of selected bins (needed for the algorithmic byte count)
    U = float(np.mean([im.trace(k).unit2_num for k in range(min(K, D))]))

    probe = "autofocus" if args.probe == "auto" else args.probe
    im.probe_enable(isar.STAGES[probe])
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    total_ms = ev0.elapsed_time(ev1)
    probe_ms, probe_n = im.probe_read()
    im.probe_enable(-1)
    ms_per_step = total_ms / args.steps
    fps = K * args.steps / (total_ms / 1e3)

    launches_per_run = probe_n / max(args.steps, 1)
    avg_launch_ms = probe_ms / max(probe_n, 1)
    alg = stage_bytes(probe, M, N, W, U, K, launches_per_run)
    achieved_gbs = alg / (avg_launch_ms / 1e3) / 1e9
    frame_bytes = algorithmic_bytes_per_frame(M, N, W, U)
    out = {
        "metric": "frames/s @256x4096 cf32, % HBM roofline",
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 complex; fp64 phase arguments)",
        "data": "synthetic",
        "config": {"workload": "config 2: 256 pulses x 4096 range cells complex64, full plain pipeline "
                               "(HPC, RA, centring, 120-iteration fast entropy, MTRC, azimuth FFT)",
                   "frames_per_step": K, "distinct_frames": D, "mean_selected_bins": U},
        "roofline": {"bound": "hbm", "kernel": probe, "achieved": round(achieved_gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved_gbs / HBM_PEAK_GBS, 5), "traffic": None,
                     "avg_launch_ms": round(avg_launch_ms, 5), "launches_per_step": launches_per_run,
                     "algorithmic_bytes_per_launch": alg},
        "pipeline_roofline": {"algorithmic_bytes_per_frame": frame_bytes,
                              "achieved_GBs": round(frame_bytes * fps / 1e9, 2),
                              "frac": round(frame_bytes * fps / 1e9 / HBM_PEAK_GBS, 5)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_frames)
        out["gpu_over_cpu"] = round(fps / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out))
    im.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
