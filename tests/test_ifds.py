"""IFDS pulse compression (common.cu:583-604, 635-746): the FP64 oracle pinned by matched-filter
closed forms (CPU), the device kernel against it, and the drop-in caller on an IFDS dwell
(MI355X)."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from parity import rel_l2

ROOT = Path(__file__).resolve().parents[1]
APP = ROOT / "isar_imaging_cuda_amd" / "lib" / "isar_calling_app"

FS, TP, B, FC = 32_000_000, 40e-6, 20_000_000, 10_000_000_000


def _params():
    return O.params(128, 4096, 512, B, FC, TP, FS)


def _chirp(v):
    """The transmitted chirp as the matched filter models it (no window)."""
    S = int(FS * TP)
    c = 3e8
    v2 = 2 * v / c
    c1 = -2 * np.pi * FC * v2 / (1 - v2)
    c2 = np.pi * B / TP
    k = np.arange(S)
    t = (1 - v2) * (-TP / 2 + k * TP / (S - 1))
    return np.exp(1j * (c1 + c2 * t) * t), 0.54 - 0.46 * np.cos(2 * np.pi * k / (S - 1))


# ------------------------------------------------------------------------------------------
# CPU: pin the oracle
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("d", [0, 1500, 2500, 4700])
def test_oracle_matched_filter_peak_and_cut(d):
    """A single return d samples into the echo compresses to sum(hamming) at index d, and the
    cut starts at clamp(d - 2048, 0, NFFT - 4096)."""
    v = 35.0
    u, w = _chirp(v)
    L = 6000
    iq = np.zeros(L, np.complex64)
    n = min(u.size, L - d)
    iq[d:d + n] = (800 * u[:n]).astype(np.complex64)
    out, off = O.pulse_compress(iq, v, _params())
    nfft = O.next_pow2(L)
    assert nfft == 8192
    assert off == min(max(d - 2048, 0), nfft - 4096)
    k = int(np.argmax(np.abs(out)))
    assert off + k == d
    if n == u.size:
        assert abs(out[k] - 800 * w.sum()) <= 1e-3 * 800 * w.sum()


def test_oracle_compression_is_linear_and_velocity_matched():
    u0, _ = _chirp(0.0)
    L = 6000
    a = np.zeros(L, np.complex64)
    a[1000:1000 + u0.size] = 500 * u0
    b = np.zeros(L, np.complex64)
    b[3000:3000 + u0.size] = 300j * u0
    p = _params()
    oa, offa = O.pulse_compress(a, 0.0, p)
    ob, offb = O.pulse_compress(b, 0.0, p)
    oab, offab = O.pulse_compress(a + b, 0.0, p)
    # the stronger return decides the cut; compare where all three cut windows overlap
    assert (offa, offb, offab) == (0, 3000 - 2048, 0)
    lo, hi = max(offa, offb, offab), min(offa, offb, offab) + 4096
    win = lambda o, off: o[lo - off:hi - off]
    assert np.allclose(win(oab, offab), win(oa, offa) + win(ob, offb), atol=1e-3)
    # a target at 3000 m/s: its own velocity compresses it to sum(hamming) at its delay; the
    # zero-velocity filter moves the peak by the range-Doppler coupling fd / (B / Tp) * Fs samples
    v = 3000.0
    u3, w3 = _chirp(v)
    c = np.zeros(L, np.complex64)
    c[2000:2000 + u3.size] = 500 * u3
    om, offm = O.pulse_compress(c, v, p)
    assert offm + int(np.argmax(np.abs(om))) == 2000
    assert abs(abs(om[2000 - offm]) - 500 * w3.sum()) <= 1e-6 * 500 * w3.sum()
    o0, off0 = O.pulse_compress(c, 0.0, p)
    shift = (2 * v * FC / 3e8) / (B / TP) * FS
    assert abs(off0 + int(np.argmax(np.abs(o0))) - 2000 - shift) <= 1.0


def test_oracle_icamax_takes_first_of_equal_peaks():
    u, _ = _chirp(0.0)
    L = 6000
    iq = np.zeros(L, np.complex64)
    iq[2600:2600 + u.size] += 700 * u
    iq[3400:3400 + u.size] += 700 * u
    out, off = O.pulse_compress(iq, 0.0, _params())
    assert off == 2600 - 2048


# ------------------------------------------------------------------------------------------
# GPU
# ------------------------------------------------------------------------------------------
@pytest.mark.gpu
def test_device_pulse_compression_matches_oracle():
    import ctypes as C
    import torch
    import isar_imaging_cuda_amd as isar
    from isar_imaging_cuda_amd import synth
    cfg = synth.preset(6, echo_num=48)
    iq, vel = synth.ifds(cfg)
    M, L = iq.shape
    print(f"[case] pulse compression {M} echoes of {L} samples", flush=True)
    h = synth.radar_header(cfg)
    dev = torch.device("cuda:0")
    raw = np.stack([iq.real, iq.imag], axis=-1).astype(np.int16)     # int16-valued samples
    raw[0] = 0                                                        # an all-zero echo
    raw[1] = np.clip(raw[1].astype(np.int64) * 1000, -32768, 32767)   # clipped at int16 full scale
    iq = raw[..., 0].astype(np.float64) + 1j * raw[..., 1]
    d_iq = torch.from_numpy(raw).to(dev)
    d_v = torch.from_numpy(vel).to(dev)
    out = torch.zeros((M, 4096, 2), dtype=torch.float32, device=dev)
    offs = torch.zeros(M, dtype=torch.int32, device=dev)
    p = isar.RadarParams(M, 4096, h["band_width"], h["fc"], h["Tp"], h["Fs"]).c()
    L_ = isar.lib()
    L_.isar_pulse_compress.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
    rc = L_.isar_pulse_compress(C.byref(p), M, L, C.c_void_p(d_iq.data_ptr()), C.c_void_p(d_v.data_ptr()),
                                C.c_void_p(out.data_ptr()), C.c_void_p(offs.data_ptr()),
                                C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, L_.isar_last_error()
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.complex64)[..., 0]
    go = offs.cpu().numpy()
    op = O.params(M, 4096, 512, h["band_width"], h["fc"], h["Tp"], h["Fs"])
    for m in range(M):
        ref, off = O.pulse_compress(iq[m], vel[m], op)
        assert go[m] == off, f"echo {m}: cut offset {go[m]} vs {off}"
        rel = rel_l2(got[m], ref)
        assert rel < 2e-5, f"echo {m}: rel error {rel:.2e}"
    assert not got[0].any() and go[0] == 0


@pytest.mark.gpu
def test_dropin_ifds_windows_match_oracle(tmp_path):
    from parity import IMG_REL_L2
    from wbd import make_ifds_dwell, read_wbd
    path, iq, vel, cfg = make_ifds_dwell(tmp_path / "data", pulses=160)
    ref = read_wbd(path)
    window, stride, n_img = 128, 16, 3
    print("[case] drop-in IFDS, 3 windows of 128 pulses", flush=True)
    r = subprocess.run([str(APP), str(tmp_path / "data"), str(window), str(stride), str(n_img), str(tmp_path / "img"),
                        "1", "1", "1", "1"], capture_output=True, text=True, timeout=180)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    Fs = ref["Fs"] // 2                         # the IFDS header carries twice the sampling rate
    op = O.params(window, 4096, 512, ref["B"], ref["fc"], ref["Tp"], Fs)
    for i in range(n_img):
        lo = i * stride
        pc = np.stack([O.pulse_compress(ref["echo"][lo + m], ref["vel"][lo + m], op)[0] for m in range(window)])
        img_o = O.isar_main(pc.astype(np.complex64), ref["vel"][lo:lo + window], op, O.options(data_type=1))
        img = np.fromfile(tmp_path / f"img_{i}.bin", dtype=np.float32).reshape(window, 512)
        rel = np.linalg.norm(img - img_o) / np.linalg.norm(img_o)
        assert rel <= IMG_REL_L2, f"window {i}: image rel L2 {rel:.3e}"
