"""Synthetic STRETCH and IFDS .wbd dwells for the drop-in ingest tests, and an independent reader.

Layout (common.cu:796-899, 1082-1100): frames of frame_len bytes = 256-byte header + int16 I/Q
pairs. Words (uint32, little endian) of the first header: [4] frame_len/4, [12] fc MHz, [13] B MHz,
[15] Tp us, [17] Fs MHz (low 16 bits). Per frame, words at byte 48: [7] range x0.1 m,
[8] velocity int32 x0.1 m/s, [9] azimuth, [10] pitch (int32 x 360/2^24 deg).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from isar_imaging_cuda_amd import synth


def make_dwell(directory: Path, pulses: int = 256, range_num: int = 1024, seed: int = 77, polar: str = "1101",
               azimuth_rate_deg: float = 0.002, pitch_deg: float | np.ndarray = 5.0):
    """Writes one STRETCH file; returns (path, echo complex64 (P, N), velocity as stored, config).
    pitch_deg: one pitch for every pulse, or one per pulse."""
    cfg = synth.preset(1, echo_num=pulses, range_num=range_num, seed=seed, walk_cells=24.0)
    echo, vel = synth.frame(cfg)
    directory.mkdir(parents=True, exist_ok=True)
    path = directory / f"test_dwell_00_{polar}.wbd"
    az = 30.0 + azimuth_rate_deg * np.arange(pulses)
    pitch = np.broadcast_to(np.asarray(pitch_deg, np.float64), (pulses,)).copy()
    rng_m = 10000.0 + 20.0 * np.arange(pulses) / cfg.prf
    synth.write_wbd(path, echo, cfg, range_m=rng_m, velocity=vel, azimuth=az, pitch=pitch)
    return path, echo, vel, cfg


def make_ifds_dwell(directory: Path, pulses: int = 160, seed: int = 91, polar: str = "1101"):
    """Writes one IFDS file (raw chirp echoes, preset 6); returns (path, iq complex64 (P, len),
    velocity as stored, config)."""
    cfg = synth.preset(6, echo_num=pulses, seed=seed)
    iq, vel = synth.ifds(cfg)
    directory.mkdir(parents=True, exist_ok=True)
    path = directory / f"test_dwell_01_{polar}.wbd"
    az = 30.0 + 0.002 * np.arange(pulses)
    pitch = np.full(pulses, 5.0)
    rng_m = 10000.0 + 20.0 * np.arange(pulses) / cfg.prf
    synth.write_wbd(path, iq, cfg, range_m=rng_m, velocity=vel, azimuth=az, pitch=pitch, ifds=True)
    return path, iq, vel, cfg


def read_wbd(path: Path):
    """Independent reader: returns dict(frame_len, fc, B, Tp, Fs, echo (K, N) complex64, range, vel, az, pitch)."""
    raw = np.fromfile(path, dtype=np.uint8)
    hdr = raw[:144].view("<u4")
    frame_len = int(hdr[4]) * 4
    K = raw.size // frame_len
    frames = raw[: K * frame_len].reshape(K, frame_len)
    words = frames[:, 48:92].copy().view("<u4")
    iq = frames[:, 256:].copy().view("<i2").astype(np.float32)
    echo = (iq[:, 0::2] + 1j * iq[:, 1::2]).astype(np.complex64)

    def s32(w):
        w = w.astype(np.float64)
        return w - np.where(w > 2.0 ** 31, 2.0 ** 32, 0.0)

    az = s32(words[:, 9]) * 360.0 / 2 ** 24
    pt = s32(words[:, 10]) * 360.0 / 2 ** 24
    return dict(frame_len=frame_len, fc=int(hdr[12]) * 1_000_000, B=int(hdr[13]) * 1_000_000, Tp=int(hdr[15]) / 1e6,
                Fs=int(hdr[17] % 65536) * 1_000_000, echo=echo, range=words[:, 7] * 0.1, vel=s32(words[:, 8]) * 0.1,
                az=np.where(az < 0, az + 360, az), pitch=np.where(pt < 0, pt + 360, pt))
