"""High-resolution range profiles (getHRRP, common.cu:541-580, called at isar_main.cu:57-67): the
FP64 oracle pinned by numpy's FFT and the IFDS identity (CPU), the device output against it and
through the drop-in session (MI355X)."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

ROOT = Path(__file__).resolve().parents[1]
APP = ROOT / "isar_imaging_cuda_amd" / "lib" / "isar_calling_app"
HRRP_REL = 1e-5      # FP32 transforms of <= 4096 points against FP64


def _rand(M, N, seed):
    rng = np.random.default_rng(seed)
    e = (rng.standard_normal((M, N)) + 1j * rng.standard_normal((M, N))).astype(np.complex64)
    return e, rng.uniform(-300.0, 300.0, M)


def _ham(n):
    return 0.54 - 0.46 * np.cos(2 * np.pi * np.arange(n) / (n - 1))   # genHammingVec, common.cu:260-266


# ------------------------------------------------------------------------------------------
# CPU: pin the oracle
# ------------------------------------------------------------------------------------------
def test_oracle_hrrp_ifds_returns_the_pulse_compressed_rows():
    """IFDS: the pre-stage ifftshift(IFFT(ifftshift(x))/N) followed by getHRRP's
    ifftshift(FFT(ifftshift(.))) is the identity."""
    e, v = _rand(8, 1024, 1)
    h = O.hrrp(e, v, O.params(8, 1024, 512), O.options(data_type=1))
    assert np.abs(h - e).max() <= 1e-12 * np.abs(e).max()


@pytest.mark.parametrize("if_hpc", [0, 1])
def test_oracle_hrrp_stretch_is_shifted_windowed_fft(if_hpc):
    M, N = 8, 2048
    e, v = _rand(M, N, 2)
    p = O.params(M, N, 512)
    h = O.hrrp(e, v, p, O.options(data_type=2, if_hpc=if_hpc))
    rows = O.hpc(e, v, p) if if_hpc else e.astype(np.complex128)
    ref = np.fft.fftshift(np.fft.fft(rows * _ham(N), axis=1), axes=1)
    assert np.abs(h - ref).max() <= 1e-12 * np.abs(ref).max()


def test_oracle_hrrp_tone_lands_at_its_shifted_bin():
    """A beat tone exp(j 2 pi k0 n / N) puts the profile peak at k0 + N/2 with value sum(hamming)."""
    M, N, k0 = 2, 1024, 100
    n = np.arange(N)
    e = np.tile(np.exp(2j * np.pi * k0 * n / N), (M, 1)).astype(np.complex64)
    h = O.hrrp(e, np.zeros(M), O.params(M, N, 512), O.options(data_type=2, if_hpc=0))
    assert int(np.argmax(np.abs(h[0]))) == k0 + N // 2
    assert abs(abs(h[0, k0 + N // 2]) - _ham(N).sum()) <= 1e-4 * _ham(N).sum()


# ------------------------------------------------------------------------------------------
# GPU
# ------------------------------------------------------------------------------------------
def _frames(K, M, N, seed=5):
    from isar_imaging_cuda_amd import synth
    cfg = synth.preset(1, echo_num=M, range_num=N, seed=seed)
    echoes, vels = [], []
    for k in range(K):
        cfg.seed = seed + k
        e, v = synth.frame(cfg)
        echoes.append(e)
        vels.append(v)
    return np.stack(echoes), np.stack(vels), synth.radar_header(cfg)


@pytest.mark.gpu
@pytest.mark.parametrize("data_type", [2, 1])
def test_device_hrrp_matches_oracle_and_leaves_the_image(data_type):
    import torch
    import isar_imaging_cuda_amd as isar
    K, M, N = 3, 128, 1024
    echo, vel, h = _frames(K, M, N)
    print(f"[case] hrrp data_type {data_type}, {K} frames {M}x{N}", flush=True)
    dev = torch.device("cuda:0")
    d_e = torch.from_numpy(echo.view(np.float32)).to(dev)
    d_v = torch.from_numpy(vel).to(dev)
    img = torch.zeros((K, M, 512), dtype=torch.float32, device=dev)
    img_ref = torch.zeros_like(img)
    prof = torch.zeros((K, M, N, 2), dtype=torch.float32, device=dev)
    radar = isar.RadarParams(M, N, h["band_width"], h["fc"], h["Tp"], h["Fs"])
    opt = isar.Options(data_type=data_type, if_hrrp=True)
    s = torch.cuda.current_stream().cuda_stream
    with isar.Imager(radar, max_frames=K) as im:
        im.run_device(d_e.data_ptr(), d_v.data_ptr(), img_ref.data_ptr(), K, opt, stream=s)
        im.set_hrrp_output(prof.data_ptr())
        im.run_device(d_e.data_ptr(), d_v.data_ptr(), img.data_ptr(), K, opt, stream=s)
        torch.cuda.synchronize()
    assert torch.equal(img, img_ref), "the profile output changed the images"
    got = prof.cpu().numpy().view(np.complex64)[..., 0]
    p = O.params(M, N, 512, h["band_width"], h["fc"], h["Tp"], h["Fs"])
    for k in range(K):
        if data_type == 1:
            assert np.array_equal(got[k], echo[k]), f"frame {k}: IFDS profiles differ from the rows"
            continue
        ref = O.hrrp(echo[k], vel[k], p, O.options(data_type=2, if_hrrp=1))
        rel = np.linalg.norm(got[k] - ref) / np.linalg.norm(ref)
        assert rel <= HRRP_REL, f"frame {k}: profile rel L2 {rel:.2e}"


@pytest.mark.gpu
def test_dropin_session_hrrp(tmp_path):
    from parity import IMG_REL_L2, oracle_run
    from wbd import make_dwell, read_wbd
    path, echo, vel, cfg = make_dwell(tmp_path / "data", pulses=160, range_num=1024)
    ref = read_wbd(path)
    window, stride, n_img = 128, 32, 2
    print("[case] drop-in with if_hrrp, 2 windows of 128 pulses", flush=True)
    r = subprocess.run([str(APP), str(tmp_path / "data"), str(window), str(stride), str(n_img), str(tmp_path / "img"),
                        "1", "1", "1", "2", "1"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stderr
    radar = dict(band_width=ref["B"], fc=ref["fc"], Tp=ref["Tp"], Fs=ref["Fs"])
    p = O.params(window, 1024, 512, ref["B"], ref["fc"], ref["Tp"], ref["Fs"])
    for i in range(n_img):
        lo = i * stride
        e, v = ref["echo"][lo:lo + window], ref["vel"][lo:lo + window]
        got = np.fromfile(tmp_path / f"img_hrrp_{i}.bin", dtype=np.complex64).reshape(window, 1024)
        want = O.hrrp(e, v, p, O.options(if_hrrp=1))
        rel = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert rel <= HRRP_REL, f"window {i}: profile rel L2 {rel:.2e}"
        img = np.fromfile(tmp_path / f"img_{i}.bin", dtype=np.float32).reshape(window, 512)
        img_o, _ = oracle_run(e, v, radar, window, 1024, if_hrrp=1)
        rel = np.linalg.norm(img - img_o) / np.linalg.norm(img_o)
        assert rel <= IMG_REL_L2, f"window {i}: image rel L2 {rel:.3e}"
