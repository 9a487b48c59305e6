"""Runs the CPU emulation build of the kernel sources (csrc/emu) on a batch of frames.

Test infrastructure: the emulated pipeline executes the same kernel code as the device
library, with AddressSanitizer, so that indexing and synchronisation can be checked without a
GPU."""
from __future__ import annotations

import subprocess
import tempfile
from dataclasses import dataclass
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
DRIVER = ROOT / "isar_imaging_cuda_amd" / "_build" / "emu" / "isar_emu_driver"

OPT_FIELDS = ["data_type", "option_alignment", "option_phase", "if_hrrp", "if_hpc", "if_mtrc", "quirk_q1",
              "iterations", "warm_start", "reserved"]


def default_options(**kw) -> dict:
    o = dict(data_type=2, option_alignment=0, option_phase=1, if_hrrp=0, if_hpc=1, if_mtrc=1, quirk_q1=1,
             iterations=120, warm_start=0, reserved=0)
    o.update(kw)
    return o


@dataclass
class EmuFrame:
    img: np.ndarray
    n_pairs: int
    center_applied: int
    center_shift: int
    idx1_len: int
    idx2_len: int
    diff_imin: int
    cut_offset: int
    tgt_num: int
    unit2_num: int
    status: int
    center_thr: float
    pa_threshold: float
    ra_lag: np.ndarray
    ra_imax: np.ndarray
    pa_colmax: np.ndarray
    pa_colE: np.ndarray
    pa_colH: np.ndarray
    select_bin: np.ndarray
    phi0: np.ndarray
    phi: np.ndarray


def run(echo: np.ndarray, velocity: np.ndarray, radar: dict, opts: dict | None = None, phi0=None,
        timeout: float = 900.0, autofocus_mode: int | None = None, ra_split: int | None = None,
        static_check: bool = False, czt_length: int | None = None, czt_threads: int | None = None) -> list[EmuFrame]:
    """echo: complex64 (K, M, N); velocity: (K, M). radar: band_width, fc, Tp, Fs.
    autofocus_mode: isar_autofocus_mode (0 auto, 1 one workgroup per frame, 2 split, 4 resident); ra_split:
    isar_set_ra_split (0 by batch size, 1 one grid, 2 two grids). static_check: the driver also
    images the batch and then its first frame through isar_run_device_static and fails unless both
    equal the context's images bit for bit. czt_length, czt_threads: isar_set_czt_length and
    isar_set_czt_threads."""
    if not DRIVER.exists():
        raise RuntimeError(f"{DRIVER} missing: build with isar_imaging_cuda_amd/build.py --emu")
    opts = opts or default_options()
    echo = np.ascontiguousarray(echo, dtype=np.complex64)
    if echo.ndim == 2:
        echo = echo[None]
    K, M, N = echo.shape
    W = 512
    vel = np.ascontiguousarray(np.broadcast_to(velocity, (K, M)), dtype=np.float64)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = Path(td) / "in.bin", Path(td) / "out.bin"
        with open(fin, "wb") as f:
            f.write(np.array([M, N, W, K], np.int32).tobytes())
            f.write(np.array([radar["band_width"], radar["fc"]], np.int64).tobytes())
            f.write(np.array([radar["Tp"]], np.float64).tobytes())
            f.write(np.array([radar["Fs"]], np.int64).tobytes())
            f.write(np.array([opts[k] for k in OPT_FIELDS], np.int32).tobytes())
            f.write(echo.tobytes())
            f.write(vel.tobytes())
            if opts.get("warm_start"):
                f.write(np.ascontiguousarray(phi0, dtype=np.complex64).tobytes())
        # positional driver arguments; every one before the last given takes its default
        args = [autofocus_mode, ra_split, 1 if static_check else None, czt_length, czt_threads]
        defaults = [0, 0, 0, 0, 256]
        last = max((i for i, a in enumerate(args) if a is not None), default=-1)
        extra = [str(defaults[i] if args[i] is None else args[i]) for i in range(last + 1)]
        cmd = [str(DRIVER), str(fin), str(fout)] + extra
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"emu driver failed rc={r.returncode}\n{r.stdout}\n{r.stderr[-4000:]}")
        raw = fout.read_bytes()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a.copy()

    img = take(np.float32, K * M * W).reshape(K, M, W)
    out = []
    for k in range(K):
        sc = take(np.int32, 10)
        fs = take(np.float32, 2)
        lag = take(np.float64, M - 1)
        imax = take(np.int32, M - 1)
        cm, ce, ch = take(np.float32, W), take(np.float32, W), take(np.float32, W)
        sel = take(np.int32, W)
        p0 = take(np.float32, 2 * M)
        p1 = take(np.float32, 2 * M)
        out.append(EmuFrame(img[k], *[int(v) for v in sc], float(fs[0]), float(fs[1]), lag, imax, cm, ce, ch,
                            sel[: int(sc[8])], p0[0::2] + 1j * p0[1::2], p1[0::2] + 1j * p1[1::2]))
    return out
