"""Throughput benchmark of the MI355X imaging pipeline on 256 x 4096 complex64 frames.

One step = the full imaging chain (HPC, range alignment, centring and cut, Doppler tracking +
120-iteration fast-entropy autofocus, MTRC and the azimuth FFT) over one batch of `--frames`
distinct synthetic frames already resident in HBM. Prints one JSON line.

  python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--frames 256] [--probe ra_upd] [--no-cpu]
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


def _ra_levels(M: int) -> list[int]:
    """Merge strides s >= 2 of range alignment (level 0 is the fused pair kernel)."""
    out, s = [], 2
    while s < M:
        out.append(s)
        s *= 2
    return out


def stage_bytes_per_frame(stage: str, M: int, N: int, W: int, U: float, chunk: int = 16) -> float:
    """Algorithmic HBM bytes of one frame for one launch of a kernel family, averaged over the
    launches of that family in one run (range-alignment families launch once per merge level).

    ra_pair0   read every echo row (8N), store its spectrum S0 (8N), one merged energy row per
               pair (4N), the two final profiles of rows 0-1 (8N each) and their column sums (4N)
    ra_upd(s)  read S0 of the M/2 rows that move (8N each), one partial energy row (4N) per
               chunk of R = min(s, chunk) moved rows, final profiles of pair 0's s moved rows (8N
               each) and their per-chunk column sums (4N)
    ra_corr(s) per pair: two group energies read (4N each) plus the chunk partials of each
               group's second half, the first group's energy written back (4N)
    """
    if stage == "ra_pair0":
        return M * 16.0 * N + (M / 2) * 4.0 * N + 2 * 8.0 * N + 4.0 * N
    if stage == "ra_upd":
        per = []
        for s in _ra_levels(M):
            R = min(s, chunk)
            per.append((M / 2) * 8.0 * N + (M / 2 / R) * 4.0 * N + s * 8.0 * N + (s / R) * 4.0 * N)
        return sum(per) / len(per)
    if stage == "ra_corr":
        per = []
        for s in _ra_levels(M):
            pairs = M // (2 * s)
            parts = 0 if s == 2 else (s // 2) // min(s // 2, chunk)
            per.append(pairs * (2 * 4.0 * N + 2 * parts * 4.0 * N + 4.0 * N))
        return sum(per) / len(per)
    return {
        "autofocus": 8.0 * M * U + 8.0 * M * 2,
        "mtrc_czt": 2 * 8.0 * M * W,
        "mtrc_rows": 2 * 8.0 * M * W,
        "mtrc_image": 8.0 * M * W + 4.0 * M * W,
        "pa_colfft": 8.0 * M * W,
        "cut": 2 * 8.0 * M * W,
    }[stage]


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), as lscpu reports it."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(frames: int = 64, runs: int = 3, intra_frames: int = 8) -> dict:
    """FP64 oracle on this host's cores (bounded sample of the same workload), with the two ways
    to spread it over threads: frame-level (one OpenMP thread per frame, `frames` frames, median
    of `runs` timed runs) and intra-frame (every thread inside each stage of one frame,
    `intra_frames` frames in turn). `value` is the faster; the intra-frame runs give the time of
    each of the reference's SEPARATE_TIMING_ stages (isar_main.cu:34-218)."""
    from isar_imaging_cuda_amd import synth
    from oracle import oracle as O

    cfgs = [synth.preset(2, seed=0xC0FFEE + k) for k in range(frames)]
    data = [synth.frame(c) for c in cfgs]
    echo = np.stack([d[0] for d in data])
    vel = np.stack([d[1] for d in data])
    h = synth.radar_header(cfgs[0])
    p = O.params(M, N, W, h["band_width"], h["fc"], h["Tp"], h["Fs"])
    O.isar_batch(echo[:1], vel[:1], p)                      # warm-up (plan caches)
    walls = []
    for _ in range(runs):
        t0 = time.perf_counter()
        _, threads = O.isar_batch(echo, vel, p)
        walls.append(time.perf_counter() - t0)
    dt = float(np.median(walls))
    frame_level = frames / dt
    stages, t_intra = [], []
    for k in range(intra_frames):
        t0 = time.perf_counter()
        _, st = O.isar_stage_times(echo[k], vel[k], p)
        t_intra.append(time.perf_counter() - t0)
        stages.append(st)
    intra = intra_frames / float(np.sum(t_intra))
    stage_ms = {name: round(float(np.median([s[name] for s in stages])), 2) for name in O.STAGE_NAMES}
    best = "frame-level" if frame_level >= intra else "intra-frame"
    return {"value": max(frame_level, intra), "unit": "frames/s", "cores": int(threads), "kind": "port",
            "cpu_model": cpu_model(), "parallelism": best,
            "frame_level_frames_per_s": round(frame_level, 3), "intra_frame_frames_per_s": round(intra, 3),
            "stage_ms_per_frame_intra": stage_ms,
            "sample": f"{frames} distinct 256x4096 frames, FP64 oracle: frame-level one OpenMP thread per frame "
                      f"(median of {runs} runs, {dt:.1f} s wall each) and intra-frame all threads per stage "
                      f"({intra_frames} frames, {float(np.sum(t_intra)):.1f} s); value = the faster ({best})"}


# the reference's SEPARATE_TIMING_ stages (isar_main.cu:34-218) and the device kernel families
# that carry their work; HPC runs inside the first merge level (k_ra_pair0), and the MTRC's last
# row transform is done together with post-processing (k_mtrc_image)
GPU_STAGES = {"hpc": (), "range_alignment": ("ra_pair0", "ra_corr", "ra_upd"), "centre_cut": ("center", "cut"),
              "phase_compensation": ("pa_colfft", "pa_select", "autofocus"), "mtrc": ("mtrc_rows", "mtrc_czt"),
              "post": ("mtrc_image",)}


def gpu_stage_times(im, step, frames: int, steps: int = 3) -> dict:
    """Device time per frame (us) of each SEPARATE_TIMING_ stage: HIP events around every launch
    of the stage's kernel families on the run stream, `steps` steps per family."""
    import torch

    import isar_imaging_cuda_amd as isar
    out = {}
    for stage, fams in GPU_STAGES.items():
        ms = 0.0
        for fam in fams:
            im.probe_enable(isar.STAGES[fam])
            for _ in range(steps):
                step()
            torch.cuda.synchronize()
            t, _ = im.probe_read()
            ms += t / steps
        im.probe_enable(-1)
        out[stage] = round(ms * 1e3 / frames, 3)
    return out


def copy_calibration(mib: int) -> int:
    """A plain device copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4,
                      "launches": 3}))
    return 0


def learned_vs_plain(frames: int = 1000) -> dict:
    """Autofocus iterations to reach the plain 120-iteration image entropy (within 1e-5 relative)
    on held-out 256 x 2048 frames (range-aligned, per-pulse phase errors uniform in [-pi, pi],
    15 dB SNR): plain from the Doppler phase versus the shipped warm-start network."""
    from isar_imaging_cuda_amd import learned as L
    params, front_iters = L.load_weights()
    res = L.evaluate(params, list(range(30_000_001, 30_000_001 + frames)), chunk=min(frames, 256), log=lambda m: None,
                     front_iters=front_iters)
    res["front_iterations"] = front_iters
    res["workload"] = "256 x 2048 complex64 frames, held-out seeds"
    # device time per frame of the autofocus and what follows it, at the mean iteration counts:
    # plain = its iterations + the image; learned = its iterations (the front ones included) +
    # the network + the judgement of the start (three image formations) + the image
    t = L.stage_times(params, list(range(30_000_001, 30_000_257)), front_iters=front_iters)
    plain_ms = res["plain_mean_iterations"] * t["iteration"] + t["image"]
    learned_ms = res["learned_mean_iterations"] * t["iteration"] + t["network"] + t["judge"] + t["image"]
    res["device_ms_per_frame"] = {"iteration": round(t["iteration"], 6), "image": round(t["image"], 6),
                                  "network": round(t["network"], 6), "judge": round(t["judge"], 6),
                                  "plain_chain": round(plain_ms, 6), "learned_chain": round(learned_ms, 6),
                                  "time_ratio": round(plain_ms / learned_ms, 3), "batch_frames": t["batch_frames"]}
    return res


# per pulse count of the sub-aperture plan: chunk rows, autofocus form, chirp-z length and
# threads, as the plan's autotuner chose them for this workload on an MI355X, then the autofocus
# form of the 192- and 384-pulse groups from a sweep of forms and chunk sizes per group, each
# table timed by graph replay with the groups placed on branches by measured times
SA_TUNING = {128: dict(ra_chunk=2, autofocus_mode=3, czt_length=256, czt_threads=256),
             192: dict(ra_chunk=4, autofocus_mode=1, czt_length=512, czt_threads=256),
             256: dict(ra_chunk=4, autofocus_mode=1, czt_length=512, czt_threads=256),
             384: dict(ra_chunk=4, autofocus_mode=1, czt_length=1024, czt_threads=256),
             512: dict(ra_chunk=4, autofocus_mode=3, czt_length=1024, czt_threads=256)}


def subapertures(count: int = 64, replays: int = 10, tuned: bool = True) -> dict:
    """64 sliding windows of {128, 192, 256, 384, 512} pulses of one 8192 x 4096 complex64 dwell
    imaged by one sub-aperture plan, replayed as a graph: with the tuning table SA_TUNING set
    through isar_sa_plan_set_group_tuning, or (tuned=False) the default chunk size and autofocus
    form. The plan's autotuner is not run here."""
    import torch

    import isar_imaging_cuda_amd as isar
    from isar_imaging_cuda_amd import subaperture as SA
    from isar_imaging_cuda_amd import synth

    cfg = synth.preset(3)
    echo, vel = synth.frame(cfg)
    h = synth.radar_header(cfg)
    P, N = echo.shape
    first, pulses = SA.sliding_windows(P, count, seed=7)
    plan = SA.SubaperturePlan(isar.RadarParams(256, N, h["band_width"], h["fc"], h["Tp"], h["Fs"]), P, first, pulses)
    dev = torch.device("cuda:0")
    d_echo = torch.from_numpy(echo.view(np.float32)).to(dev)
    d_vel = torch.from_numpy(vel).to(dev)
    img = torch.zeros(plan.image_floats, dtype=torch.float32, device=dev)
    if tuned:
        plan.set_tuning(SA_TUNING)
        plan.calibrate(d_echo.data_ptr(), d_vel.data_ptr(), img.data_ptr(), repeats=2)   # branches by measured times
    s = torch.cuda.current_stream(dev)
    plan.run(d_echo.data_ptr(), d_vel.data_ptr(), img.data_ptr(), graph=True, stream=s.cuda_stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(replays):
        plan.run(d_echo.data_ptr(), d_vel.data_ptr(), img.data_ptr(), graph=True, stream=s.cuda_stream)
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / replays
    tuning = plan.tuning()
    plan_groups, lanes = plan.group_pulses.tolist(), plan.lanes()
    plan.close()
    # minimum-pass bytes of every window of the launch; U = 64 selected bins per window (the 8MU
    # term is at most 0.13 % of a window's bytes at 512 x 4096, so the assumption barely matters)
    alg = float(sum(algorithmic_bytes_per_frame(int(m), N, W, 64.0) for m in pulses))
    gbs = alg / (ms / 1e3) / 1e9
    return {"workload": f"{count} windows of one {P} x {N} dwell, pulse counts {sorted(set(pulses.tolist()))}",
            "ms_per_graph_launch": round(ms, 4), "windows_per_s": round(count / (ms / 1e3), 1),
            "pulses_per_s": round(float(pulses.sum()) / (ms / 1e3), 1),
            "tuning": {m: f"chunk {t['ra_chunk']}, autofocus form {t['autofocus_mode']}, chirp-z {t['czt_length']} x "
                          f"{t['czt_threads']} threads" for m, t in tuning.items()},
            "tuning_source": ("fixed table (isar_sa_plan_set_group_tuning), branches by measured group times "
                              "(isar_sa_plan_calibrate)" if tuned else "defaults"),
            "branches": {int(m): b for m, b in zip(plan_groups, lanes)},
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                         "algorithmic_bytes_per_launch": alg, "selected_bins_assumed": 64}}


def _two_batches_in_flight(im, run, img, echo, vel, frames: int, steps: int, s) -> dict:
    """Batches of the same size alternating over two contexts on two streams (default
    priority), so that one batch's autofocus iterations, which leave most of the GPU idle at 8
    frames, overlap the other's range alignment. Both contexts image the same frames, and their
    images must agree bit for bit."""
    import torch

    import isar_imaging_cuda_amd as isar
    s2 = torch.cuda.Stream(device=img.device)
    im2 = isar.Imager(im.params, max_frames=frames)
    im2.set_autofocus_mode(im.autofocus_mode)
    img2 = torch.empty_like(img)
    run2 = lambda: im2.run_device(echo.data_ptr(), vel.data_ptr(), img2.data_ptr(), frames, isar.Options(),
                                  stream=s2.cuda_stream)
    run2()
    torch.cuda.synchronize()
    start, end_a, end_b = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    start.record(s)
    s2.wait_event(start)
    for _ in range(steps):
        run()
        run2()
    end_a.record(s)
    end_b.record(s2)
    torch.cuda.synchronize()
    ms = max(start.elapsed_time(end_a), start.elapsed_time(end_b)) / (2 * steps)
    same = bool(torch.equal(img, img2))
    im2.close()
    if not same:
        raise RuntimeError("two-stream batches: the two contexts' images differ")
    return {"ms_per_batch": round(ms, 3), "frames_per_s": round(frames / (ms / 1e3), 1), "batches": 2 * steps,
            "note": "two contexts on two streams, batches alternating; per-batch latency is not lower"}


def large_batch(frames: int = 8, steps: int = 3, learned_frames: int = 16, two_streams: bool = True) -> dict:
    """8 frames of 1024 x 16384 complex64 (128 MiB each; 1000 scatterers, 7 km/s, 100 m/s^2,
    0.08 rad/s) through the full plain pipeline as one device batch (frames generated on the GPU,
    the same as the CPU generator's), and the 1024-pulse learned warm start on `learned_frames`
    held-out frames of that configuration: iterations to the plain 120-iteration entropy."""
    import torch

    import isar_imaging_cuda_amd as isar
    from isar_imaging_cuda_amd import learned as L
    from isar_imaging_cuda_amd import synth

    echo, vel, h = L.synth_frames_device(5, [0x5EED + k for k in range(frames)])
    M5, N5 = echo.shape[1], echo.shape[2]
    dev = torch.device("cuda:0")
    img = torch.empty((frames, M5, W), dtype=torch.float32, device=dev)
    im = isar.Imager(isar.RadarParams(M5, N5, h["band_width"], h["fc"], h["Tp"], h["Fs"]), max_frames=frames)
    im.set_autofocus_mode(isar.AF_RESIDENT)   # all iterations in one launch (a launch per iteration at AUTO)
    s = torch.cuda.current_stream(dev)
    run = lambda: im.run_device(echo.data_ptr(), vel.data_ptr(), img.data_ptr(), frames, isar.Options(),
                                stream=s.cuda_stream)
    run()
    torch.cuda.synchronize()
    U = float(np.mean([im.trace(k).unit2_num for k in range(frames)]))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        run()
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    if any(im.trace(k).status & 2 for k in range(frames)):
        raise RuntimeError("large batch: a resident autofocus wait hit its bound")
    fps = frames / (ms / 1e3)
    fb = algorithmic_bytes_per_frame(M5, N5, W, U)
    out = {"workload": f"{frames} frames of {M5} x {N5} complex64, full plain pipeline",
           "autofocus_form": "resident", "ms_per_batch": round(ms, 3), "frames_per_s": round(fps, 1),
           "mean_selected_bins": U,
           "algorithmic_bytes_per_frame": fb, "achieved_GBs": round(fb * fps / 1e9, 1),
           "frac": round(fb * fps / 1e9 / HBM_PEAK_GBS, 4)}
    if two_streams:
        out["two_streams"] = _two_batches_in_flight(im, run, img, echo, vel, frames, steps, s)
        out["two_streams"]["frac"] = round(fb * out["two_streams"]["frames_per_s"] / 1e9 / HBM_PEAK_GBS, 4)
    im.close()
    del echo, vel, img
    torch.cuda.empty_cache()
    if learned_frames:
        params, front_iters = L.load_weights(M=M5)
        res = L.evaluate(params, list(range(40_000_001, 40_000_001 + learned_frames)), chunk=8, config=5,
                         log=lambda m: None, front_iters=front_iters, gpu_synth=True)
        res["workload"] = f"{learned_frames} held-out {M5} x {N5} frames, network trained on 1024 x 2048 frames of the scene"
        out["learned_vs_plain"] = res
        torch.cuda.empty_cache()
    return out


def single_frame(params, echo, vel, stream, runs: int = 20) -> dict:
    """Latency of one frame alone (a batch of one, device-resident), the drop-in's unit of work:
    with the drop-in session's two-row alignment chunks and resident autofocus, replayed as a HIP
    graph (graph mode, as the drop-in session runs) and launched directly."""
    import torch

    import isar_imaging_cuda_amd as isar
    im = isar.Imager(params, max_frames=1)
    im.set_ra_chunk(2)   # as the drop-in session images its one frame per call
    im.set_autofocus_mode(isar.AF_RESIDENT)
    img1 = torch.empty((1, params.echo_num, W), dtype=torch.float32, device=echo.device)
    run = lambda: im.run_device(echo.data_ptr(), vel.data_ptr(), img1.data_ptr(), 1, isar.Options(),
                                stream=stream.cuda_stream)
    out = {}
    for name, graph in (("ms_per_frame", True), ("ms_per_frame_direct", False)):
        im.set_graph_mode(graph)
        for _ in range(3):
            run()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(runs):
            run()
        e1.record(stream)
        torch.cuda.synchronize()
        out[name] = round(e0.elapsed_time(e1) / runs, 4)
    if im.trace(0).status & 2:
        raise RuntimeError("single frame: a resident autofocus wait hit its bound")
    im.close()
    out["runs"] = runs
    out["autofocus_form"] = "resident"
    return out


def host_transfer(echo, img, ms_per_step: float, stream, max_frames: int = 256) -> dict:
    """The drop-in boundary hands host buffers over: H2D of the echo and D2H of the images of (up
    to max_frames of) one step's batch through pinned host memory, timed apart from the
    device-resident figure; the PCIe-inclusive rate adds the device time of those frames."""
    import torch
    K = echo.shape[0]
    n = min(K, max_frames)
    e, im = echo[:n], img[:n]
    h_echo = torch.empty(e.shape, dtype=e.dtype, pin_memory=True)
    h_img = torch.empty(im.shape, dtype=im.dtype, pin_memory=True)
    h_echo.copy_(e)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        ev[0].record(stream)
        e.copy_(h_echo, non_blocking=True)            # the same data back into the batch buffer
        ev[1].record(stream)
        h_img.copy_(im, non_blocking=True)
        ev[2].record(stream)
    torch.cuda.synchronize()
    h2d, d2h = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
    return {"frames": n, "h2d_echo_ms": round(h2d, 3), "h2d_GBs": round(e.numel() * 4 / (h2d / 1e3) / 1e9, 1),
            "d2h_image_ms": round(d2h, 3), "d2h_GBs": round(im.numel() * 4 / (d2h / 1e3) / 1e9, 1),
            "pcie_inclusive_frames_per_s": round(n / ((h2d + ms_per_step * n / K + d2h) / 1e3), 1)}


def load_traffic(kernel: str, frames: int, chunk: int) -> dict | None:
    """HBM traffic per launch of `kernel` at this workload from a committed PMC summary
    (profiles/traffic_<kernel>_256x4096.json, written from two rocprofv3 --pmc runs and scaled by
    the same counters read on a plain copy). None when no summary exists or when it was measured
    at another range-alignment chunk size."""
    path = ROOT / "profiles" / f"traffic_{kernel}_256x4096.json"
    if not path.exists():
        return None
    d = json.loads(path.read_text())
    if d.get("ra_chunk_rows") != chunk:
        return None
    per_frame = d["hbm_bytes_per_launch"] / d["frames_per_launch"]
    return {"bytes": per_frame * frames, "source": path.name}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=1024, help="frames per batch (one step)")
    ap.add_argument("--distinct", type=int, default=64, help="distinct synthetic frames in the batch")
    ap.add_argument("--probe", default="ra_upd", help="kernel family timed for the roofline figure")
    ap.add_argument("--probe-steps", type=int, default=3, help="extra steps timed per launch for the roofline")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-frames", type=int, default=64)
    ap.add_argument("--ra-chunk", type=int, default=0, help="fix the range-alignment chunk size (no autotuning)")
    ap.add_argument("--ra-split", type=int, default=0, help="range-alignment update launch form (isar_set_ra_split)")
    ap.add_argument("--autotune", action="store_true",
                    help="let the autotuner pick the range-alignment chunk size on this batch (default: 16 rows, "
                         "the size the committed traffic measurement used)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain iteration count")
    ap.add_argument("--no-subapertures", action="store_true", help="skip the sub-aperture plan measurement")
    ap.add_argument("--no-large", action="store_true", help="skip the 8 x 1024 x 16384 batch")
    ap.add_argument("--large-only", action="store_true", help="only the 8 x 1024 x 16384 batch (for profiling)")
    ap.add_argument("--single-only", action="store_true", help="only the one-frame latency (for profiling)")
    ap.add_argument("--subapertures-only", action="store_true", help="only the sub-aperture plan (for profiling)")
    ap.add_argument("--no-transfer", action="store_true", help="skip the host-transfer (PCIe) timing")
    ap.add_argument("--no-single", action="store_true", help="skip the single-frame latency")
    ap.add_argument("--no-stages", action="store_true", help="skip the per-stage device times")
    ap.add_argument("--sa-fixed", action="store_true",
                    help="sub-aperture plan with the default chunk size and autofocus form (no tuning table)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported: the imaging path shards "
              f"by frames with no exchange between GPUs, and this build measures one GPU.", file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.large_only:
        print(json.dumps({"large_batch": large_batch(learned_frames=0 if args.no_learned else 16)}))
        return 0
    if args.subapertures_only:
        print(json.dumps({"subapertures": subapertures(tuned=not args.sa_fixed)}))
        return 0
    if args.single_only:
        import torch

        import isar_imaging_cuda_amd as isar
        from isar_imaging_cuda_amd import synth
        dev = torch.device("cuda:0")
        cfg = synth.preset(2, seed=0xBEEF)
        e1, v1 = synth.frame(cfg)
        h = synth.radar_header(cfg)
        params = isar.RadarParams(cfg.echo_num, cfg.range_num, h["band_width"], h["fc"], h["Tp"], h["Fs"])
        echo1 = torch.from_numpy(e1[None].view(np.float32)).to(dev)
        vel1 = torch.from_numpy(v1[None]).to(dev)
        print(json.dumps({"single_frame": single_frame(params, echo1, vel1, torch.cuda.current_stream(dev))}))
        return 0

    import torch

    import isar_imaging_cuda_amd as isar
    from isar_imaging_cuda_amd import synth

    dev = torch.device("cuda:0")
    K = args.frames
    D = min(args.distinct, K)
    cfgs = [synth.preset(2, seed=0xBEEF + k) for k in range(D)]
    frames = [synth.frame(c) for c in cfgs]
    h = synth.radar_header(cfgs[0])
    # the D distinct frames go to the device once and are repeated there: frame k is frame k mod D
    pick = torch.arange(K, device=dev) % D
    echo = torch.from_numpy(np.stack([f[0] for f in frames]).view(np.float32)).to(dev)[pick].contiguous()
    vel = torch.from_numpy(np.stack([f[1] for f in frames])).to(dev)[pick].contiguous()   # echo: K x M x 2N float32
    img = torch.empty((K, M, W), dtype=torch.float32, device=dev)
    params = isar.RadarParams(M, N, h["band_width"], h["fc"], h["Tp"], h["Fs"])
    opts = isar.Options()
    im = isar.Imager(params, max_frames=K)
    stream = torch.cuda.current_stream(dev)

    def step():
        im.run_device(echo.data_ptr(), vel.data_ptr(), img.data_ptr(), K, opts, stream=stream.cuda_stream)

    # optionally the autotuner picks the range-alignment chunk size on this batch (outside the timed region)
    if args.ra_chunk:
        im.set_ra_chunk(args.ra_chunk)
        ra_chunk = args.ra_chunk
    elif args.autotune:
        ra_chunk = im.autotune(echo.data_ptr(), vel.data_ptr(), img.data_ptr(), K, opts, repeats=2,
                               stream=stream.cuda_stream)
    else:
        ra_chunk = 16
        im.set_ra_chunk(ra_chunk)
    if args.ra_split:
        im.set_ra_split(args.ra_split)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # mean number of selected bins (needed for the algorithmic byte count)
    U = float(np.mean([im.trace(k).unit2_num for k in range(min(K, D))]))
    print(f"[bench] warm-up done: {K} frames/step, mean U = {U:.1f}", file=sys.stderr, flush=True)

    # timed region: exactly `steps` steps, nothing else on the stream
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    for i in range(args.steps):
        step()
        evs[i + 1].record(stream)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3
    total_ms = evs[0].elapsed_time(evs[-1])
    step_ms = [evs[i].elapsed_time(evs[i + 1]) for i in range(args.steps)]
    ms_per_step = total_ms / args.steps
    fps = K * args.steps / (total_ms / 1e3)

    # roofline of the dominant kernel family: HIP events around each of its launches on the run
    # stream, in separate steps after the timed region
    probe = args.probe
    im.probe_enable(isar.STAGES[probe])
    for _ in range(args.probe_steps):
        step()
    torch.cuda.synchronize()
    probe_ms, probe_n = im.probe_read()
    im.probe_enable(-1)
    avg_launch_ms = probe_ms / max(probe_n, 1)
    stage_us = gpu_stage_times(im, step, K) if not args.no_stages else None
    alg = stage_bytes_per_frame(probe, M, N, W, U, ra_chunk) * K
    achieved_gbs = alg / (avg_launch_ms / 1e3) / 1e9 if probe_n else 0.0
    traffic = load_traffic(probe, K, ra_chunk)
    frame_bytes = algorithmic_bytes_per_frame(M, N, W, U)
    out = {
        "metric": "frames/s @256x4096 cf32, % HBM roofline; learned vs plain at equal entropy",
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "ms_per_step_median": round(float(np.median(step_ms)), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 complex; fp64 phase arguments)",
        "data": "synthetic",
        "config": {"workload": "256 pulses x 4096 range cells complex64 STRETCH frames, full plain pipeline "
                               "(HPC, range alignment, centring and cut, 120-iteration fast entropy, MTRC, "
                               "azimuth FFT)",
                   "frames_per_step": K, "distinct_frames": D, "mean_selected_bins": U, "ra_chunk_rows": ra_chunk,
                   "wall_ms_timed_region": round(wall_ms, 3)},
        "roofline": {"bound": "hbm", "kernel": probe, "achieved": round(achieved_gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved_gbs / HBM_PEAK_GBS, 5),
                     "traffic": None if traffic is None else round(traffic["bytes"]),
                     "traffic_source": None if traffic is None else traffic["source"],
                     "avg_launch_ms": round(avg_launch_ms, 5), "launches_timed": probe_n,
                     "algorithmic_bytes_per_launch": alg},
        "pipeline_roofline": {"algorithmic_bytes_per_frame": frame_bytes,
                              "achieved_GBs": round(frame_bytes * fps / 1e9, 2),
                              "frac": round(frame_bytes * fps / 1e9 / HBM_PEAK_GBS, 5)},
    }
    if stage_us is not None:
        out["stage_us_per_frame"] = stage_us
        out["stage_note"] = ("HIP-event time per frame of the kernels that do each SEPARATE_TIMING_ stage; HPC is fused "
                             "into the first range-alignment kernel, the MTRC's last row FFT into post-processing")
    if not args.no_transfer:
        out["host_transfer"] = host_transfer(echo, img, ms_per_step, stream)
    if not args.no_single:
        out["single_frame"] = single_frame(params, echo, vel, stream)
    im.close()
    del echo, vel, img
    torch.cuda.empty_cache()
    if not args.no_large:
        out["large_batch"] = large_batch()
    if not args.no_learned:
        out["learned_vs_plain"] = learned_vs_plain()
    if not args.no_subapertures:
        out["subapertures"] = subapertures()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_frames)
        out["gpu_over_cpu"] = round(fps / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
