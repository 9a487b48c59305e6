"""Seeded synthetic STRETCH echo frames (binding of libisar_synth.so, see include/isar_synth.h)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "lib" / "libisar_synth.so"


class SceneConfig(C.Structure):
    _fields_ = [("echo_num", C.c_int32), ("range_num", C.c_int32), ("n_scatterers", C.c_int32),
                ("phase_error_model", C.c_int32), ("box_x", C.c_double), ("box_y", C.c_double),
                ("box_z", C.c_double), ("fc", C.c_double), ("band_width", C.c_double), ("prf", C.c_double),
                ("fs", C.c_double), ("tp", C.c_double), ("omega", C.c_double), ("v", C.c_double),
                ("a", C.c_double), ("walk_cells", C.c_double), ("snr_db", C.c_double),
                ("elevation", C.c_double), ("poly_sigma", C.c_double), ("white_sigma", C.c_double),
                ("peak_adc", C.c_double), ("quantize", C.c_int32), ("reserved", C.c_int32),
                ("seed", C.c_uint64)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"{_LIB_PATH} is missing: run __graft_entry__.build()")
        L = C.CDLL(str(_LIB_PATH))
        L.isar_synth_preset.argtypes = [C.c_int, C.POINTER(SceneConfig)]
        L.isar_synth_frame.argtypes = [C.POINTER(SceneConfig), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.isar_synth_radar_header.argtypes = [C.POINTER(SceneConfig), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                              C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.isar_write_wbd.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_double,
                                     C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def preset(config_id: int, **overrides) -> SceneConfig:
    cfg = SceneConfig()
    if lib().isar_synth_preset(int(config_id), C.byref(cfg)) != 0:
        raise ValueError(f"unknown preset {config_id}")
    for k, v in overrides.items():
        if k not in dict(SceneConfig._fields_):
            raise KeyError(k)
        setattr(cfg, k, v)
    if "range_num" in overrides and "tp" not in overrides:
        cfg.tp = cfg.range_num / cfg.fs
    return cfg


def frame(cfg: SceneConfig, with_truth: bool = False):
    """Returns (echo complex64 (M, N), velocity float64 (M,)) [+ (phase, walk) truth]."""
    M, N = cfg.echo_num, cfg.range_num
    echo = np.zeros((M, N), np.complex64)
    vel = np.zeros(M, np.float64)
    ph = np.zeros(M, np.float64)
    wk = np.zeros(M, np.float64)
    rc = lib().isar_synth_frame(C.byref(cfg), echo.ctypes.data_as(C.c_void_p), vel.ctypes.data_as(C.c_void_p),
                                ph.ctypes.data_as(C.c_void_p), wk.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise ValueError("isar_synth_frame failed")
    if with_truth:
        return echo, vel, ph, wk
    return echo, vel


def radar_header(cfg: SceneConfig) -> dict:
    B, fc, fs = C.c_int64(), C.c_int64(), C.c_int64()
    tp = C.c_double()
    lib().isar_synth_radar_header(C.byref(cfg), C.byref(B), C.byref(fc), C.byref(tp), C.byref(fs))
    return {"band_width": B.value, "fc": fc.value, "Tp": tp.value, "Fs": fs.value}


def ifds(cfg: SceneConfig):
    """Raw IFDS echoes: (iq complex64 (M, cfg.range_num), velocity float64 (M,))."""
    M, L = cfg.echo_num, cfg.range_num
    iq = np.zeros((M, L), np.complex64)
    vel = np.zeros(M, np.float64)
    L_ = lib()
    L_.isar_synth_ifds.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    if L_.isar_synth_ifds(C.byref(cfg), L, iq.ctypes.data_as(C.c_void_p), vel.ctypes.data_as(C.c_void_p)) != 0:
        raise ValueError("isar_synth_ifds failed")
    return iq, vel


def write_wbd(path: str | Path, echo: np.ndarray, cfg: SceneConfig, range_m=None, velocity=None,
              azimuth=None, pitch=None, ifds: bool = False) -> None:
    """Frames of 256-byte header + int16 I/Q. IFDS files carry twice the sampling rate in the
    header, which the reader halves (common.cu:823)."""
    e = np.ascontiguousarray(echo, dtype=np.complex64)
    K, N = e.shape
    h = dict(radar_header(cfg))
    if ifds:
        h["Fs"] = 2 * h["Fs"]

    def arr(a):
        return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.c_void_p)

    keep = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
            for a in (range_m, velocity, azimuth, pitch)]
    rc = lib().isar_write_wbd(str(path).encode(), e.ctypes.data_as(C.c_void_p), K, N, h["fc"], h["band_width"],
                              h["Tp"], h["Fs"], *[arr(a) for a in keep])
    if rc != 0:
        raise OSError(f"cannot write {path}")


def frames_device(cfg: SceneConfig, seeds, d_echo: int, d_velocity: int) -> None:
    """The frames `frame` makes for each seed of `seeds`, generated on the GPU into device buffers
    (d_echo: len(seeds) x M x N complex64, d_velocity: len(seeds) x M float64); isar_synth.h's
    isar_synth_frames_device in libisar_imaging.so."""
    from . import _check
    from . import lib as imaging_lib
    L = imaging_lib()
    L.isar_synth_frames_device.argtypes = [C.POINTER(SceneConfig), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    s = np.ascontiguousarray(np.asarray(list(seeds), dtype=np.uint64))
    _check(L.isar_synth_frames_device(C.byref(cfg), int(s.size), s.ctypes.data_as(C.c_void_p), C.c_void_p(d_echo),
                                      C.c_void_p(d_velocity)), "isar_synth_frames_device")
