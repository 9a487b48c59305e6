"""The reference's caller sequence (calling_dll.cpp:48-86) against the drop-in library on a
synthetic .wbd dwell: gpuDevInit -> dataParsing -> imagingMemInit -> {dataExtracting ->
isarMainSingle} with overlapping sliding windows -> imagingMemDest. Each image is compared with
the FP64 oracle run on the same window read back from the file."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from parity import IMG_REL_L2, oracle_run
from wbd import make_dwell, read_wbd

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
APP = ROOT / "isar_imaging_cuda_amd" / "lib" / "isar_calling_app"


def test_calling_sequence_sliding_windows(tmp_path):
    path, echo, vel, cfg = make_dwell(tmp_path / "data", pulses=224, range_num=1024)
    ref = read_wbd(path)
    window, stride, n_img = 128, 32, 4
    print("[case] drop-in app, 4 windows of 128 pulses, stride 32", flush=True)
    r = subprocess.run([str(APP), str(tmp_path / "data"), str(window), str(stride), str(n_img), str(tmp_path / "img"), "1"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    radar = dict(band_width=ref["B"], fc=ref["fc"], Tp=ref["Tp"], Fs=ref["Fs"])
    for i in range(n_img):
        img = np.fromfile(tmp_path / f"img_{i}.bin", dtype=np.float32).reshape(window, 512)
        lo = i * stride
        img_o, tr = oracle_run(ref["echo"][lo:lo + window], ref["vel"][lo:lo + window], radar, window, 1024)
        rel = np.linalg.norm(img - img_o) / np.linalg.norm(img_o)
        assert rel <= IMG_REL_L2, f"window {i}: image rel L2 {rel:.3e}"
    txt = np.loadtxt(tmp_path / "img_0.txt")
    img0 = np.fromfile(tmp_path / "img_0.bin", dtype=np.float32)
    assert np.allclose(txt, img0, atol=6e-6 + 1e-7 * np.abs(img0))   # writeFileFLT: 5 decimals


def test_session_images_again_when_a_resident_wait_gives_up(tmp_path):
    """The session images with the resident autofocus and checks each frame's status; with a wait
    bound that makes every wait give up (status bit 2), each window is imaged again with the split
    form, which gives the resident form's image bit for bit."""
    path, echo, vel, cfg = make_dwell(tmp_path / "data", pulses=192, range_num=1024)
    window, stride, n_img = 128, 32, 3
    out = {}
    for name, polls in (("bounded", None), ("gives_up", "-1")):
        print(f"[case] drop-in app, resident wait bound {polls or 'default'}", flush=True)
        args = [str(APP), str(tmp_path / "data"), str(window), str(stride), str(n_img), str(tmp_path / name),
                "1", "1", "1", "2", "0", ""] + ([polls] if polls else [])
        r = subprocess.run(args, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        out[name] = r.stderr.count("imaging the frame again with the split form")
        for i in range(n_img):
            out[(name, i)] = np.fromfile(tmp_path / f"{name}_{i}.bin", dtype=np.float32)
    assert out["bounded"] == 0 and out["gives_up"] == n_img
    for i in range(n_img):
        assert np.array_equal(out[("bounded", i)], out[("gives_up", i)]), f"window {i}"


def test_extracting_refuses_arguments_that_do_not_match_init(tmp_path):
    """dataExtracting with a frame length or data type other than imagingMemInit's, or an output
    vector smaller than imagingMemInit made it, returns EXIT_FAILURE with a message (the reference
    would write past the caller's vector); a matching call still succeeds afterwards."""
    make_dwell(tmp_path / "data", pulses=160, range_num=1024)
    print("[case] drop-in app, dataExtracting misuse", flush=True)
    r = subprocess.run([str(APP), str(tmp_path / "data"), "128", "32", "-1", str(tmp_path / "img"), "1"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = dict(ln.split(": rc ", 1) for ln in r.stdout.strip().splitlines())
    for what in ("frame_len + 4", "other data_type", "dataW half size"):
        assert lines[what].startswith("1 (") and ("differs" in lines[what] or "smaller" in lines[what]), what
    assert lines["matching"].startswith("0")
