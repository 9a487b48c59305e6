"""CPU checks of the C-ABI library: every declared symbol is exported, the library loads, the
argument validation returns error codes, and the .wbd ingest decodes headers (no GPU needed)."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import isar_imaging_cuda_amd as isar
from isar_imaging_cuda_amd import synth

ROOT = Path(__file__).resolve().parents[1]
APP = ROOT / "isar_imaging_cuda_amd" / "lib" / "isar_calling_app"


def _declared(header: str) -> list[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"(?:ISAR_API|DLL_EXPORT_API)\s+[\w\s\*]+?\b(\w+)\s*\(", text)
    names += re.findall(r"^\s*(?:int|void|size_t)\s+\**\s*(isar_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header", ["isar_imaging_export.h", "isar_imaging_ext.h", "isar_learned.h", "isar_synth.h",
                                    "isar_subaperture.h"])
def test_library_exports_every_declared_entry_point(header):
    names = _declared(header)
    assert names, header
    lib = isar.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{header}: {missing}"


def test_drop_in_names_are_unmangled():
    out = subprocess.run(["nm", "-D", "--defined-only", str(isar.library_path)], capture_output=True, text=True).stdout
    for name in ("gpuDevInit", "dataParsing", "dataExtracting", "imagingMemInit", "isarMainSingle", "writeFileFLT",
                 "imagingMemDest"):
        assert re.search(rf"\bT {name}$", out, flags=re.M), name


def test_invalid_sizes_return_error_codes():
    # imagingMemInit validation (isar_imaging_export.cpp:144-155) without touching a device
    for M, N in [(2048, 4096), (256, 256), (1000, 4096), (256, 3000)]:
        with pytest.raises(isar.ImagingError) as e:
            isar.Imager(isar.RadarParams(M, N))
        assert "rc=-1" in str(e.value)
    msg = isar.lib().isar_last_error()
    assert msg and len(msg) > 0


def test_new_entry_points_reject_bad_arguments_without_a_device():
    """Argument checks that run before any HIP call (error codes of include/isar_imaging_ext.h)."""
    L = isar.lib()
    vp = C.c_void_p
    dummy = vp(16)                                 # never dereferenced: the checks come first
    L.isar_pulse_compress.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]
    p = isar.RadarParams(128, 4096, 20_000_000, 10_000_000_000, 40e-6, 32_000_000).c()
    assert L.isar_pulse_compress(C.byref(p), 1, 100, dummy, dummy, dummy, None, None) == -1
    assert "unsupported transform size 128" in L.isar_last_error().decode()
    p_long = isar.RadarParams(128, 4096, 20_000_000, 10_000_000_000, 1e-3, 32_000_000).c()   # Fs Tp > NFFT
    assert L.isar_pulse_compress(C.byref(p_long), 1, 6000, dummy, dummy, dummy, None, None) == -1
    assert "reference chirp" in L.isar_last_error().decode()
    assert L.isar_pulse_compress(C.byref(p), 0, 6000, dummy, dummy, dummy, None, None) == -1
    L.isar_set_autofocus_mode.argtypes = [vp, C.c_int]
    assert L.isar_set_autofocus_mode(None, 1) == -1
    L.isar_set_hrrp_output.argtypes = [vp, vp]
    assert L.isar_set_hrrp_output(None, dummy) == -1
    L.isar_stage_bytes.argtypes = [vp, C.c_int]
    L.isar_stage_bytes.restype = C.c_size_t
    assert L.isar_stage_bytes(None, 0) == 0
    L.isar_stage_read.argtypes = [vp, C.c_int, C.c_int, vp]
    assert L.isar_stage_read(None, 0, 0, dummy) == -1
    L.isar_select_phase.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, vp]
    assert L.isar_select_phase(0, 256, dummy, dummy, dummy, dummy, dummy, None, None) == -1
    L.isar_half_phase.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp]
    assert L.isar_half_phase(4, 256, dummy, None, dummy, None) == -1
    # frames x M beyond 32-bit indexing is refused, not wrapped
    assert L.isar_half_phase(1 << 22, 1024, dummy, dummy, dummy, None) == -1
    assert L.isar_select_phase(1 << 22, 1024, dummy, dummy, dummy, dummy, dummy, None, None) == -1
    L.isar_phase_change.argtypes = [C.c_int, C.c_int, vp, vp, vp, C.c_int, vp]
    assert L.isar_phase_change(1 << 22, 1024, dummy, dummy, dummy, 2048, None) == -1
    assert L.isar_phase_change(1, 1 << 30, dummy, dummy, dummy, 2048, None) == -1
    L.isar_compose_phase.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp]
    assert L.isar_compose_phase(1 << 22, 1024, dummy, dummy, dummy, None) == -1
    L.isar_mlp_create.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(vp)]
    widths = (C.c_int * 3)(2048, 512, 2048)
    net = vp()
    assert L.isar_mlp_create(2, widths, 1 << 20, C.byref(net)) == -1 and not net.value
    assert "too large" in L.isar_last_error().decode()
    # the drop-in session's extras before imagingMemInit: out of order
    L.isar_session_hrrp.argtypes = [vp]
    assert L.isar_session_hrrp(dummy) == -4
    L.isar_session_dump.argtypes = [C.c_uint, C.c_char_p]
    assert L.isar_session_dump(0x7F, b"/tmp") == -4


def test_last_error_is_per_thread():
    import threading
    L = isar.lib()
    vp = C.c_void_p
    L.isar_half_phase.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp]
    assert L.isar_half_phase(0, 256, None, None, None, None) == -1
    mine = L.isar_last_error().decode()
    seen = []

    def other():
        L.isar_stage_read.argtypes = [vp, C.c_int, C.c_int, vp]
        assert L.isar_stage_read(None, 0, 0, vp(16)) == -1
        seen.append(L.isar_last_error().decode())

    t = threading.Thread(target=other)
    t.start()
    t.join()
    assert seen and seen[0] != mine and "context" in seen[0]
    assert L.isar_last_error().decode() == mine


def test_workspace_bytes_scales_with_frames():
    p = isar.RadarParams(256, 4096)
    one, two, eight = (isar.Imager.workspace_bytes(p, k) for k in (1, 2, 8))
    per_frame = two - one                      # frame-sized buffers; the rest is shared tables
    assert per_frame > 256 * 4096 * 8 and one > per_frame
    assert eight == one + 7 * per_frame


def test_wbd_ingest_decodes_headers(tmp_path):
    from wbd import make_dwell, read_wbd
    path, echo, vel, cfg = make_dwell(tmp_path, pulses=64, range_num=1024, azimuth_rate_deg=0.05)
    ref = read_wbd(path)
    assert np.array_equal(ref["echo"], echo)                      # samples were already int16-valued
    r = subprocess.run([str(APP), str(tmp_path), "64", "16", "0", str(tmp_path / "x"), "1"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    fl, fn = int(lines[0].split()[1]), int(lines[0].split()[3])
    assert fl == 256 + 4 * 1024 and fn == 64
    vals = np.array([[float(v) for v in ln.split()] for ln in lines[1:]])
    assert np.allclose(vals[:, 1], ref["range"], atol=1e-6)
    assert np.allclose(vals[:, 2], ref["vel"], atol=1e-6)
    assert np.allclose(vals[:, 3], ref["az"], atol=1e-8)
    assert np.allclose(vals[:, 4], ref["pitch"], atol=1e-8)
    assert np.allclose(ref["vel"], np.round(vel * 10) / 10, atol=1e-9)   # 0.1 m/s quantisation
    # turn angle (common.cu:902-1020) for a constant pitch p and azimuth rate r: great-circle
    # angles r cos(p) per pulse accumulated at 20-pulse knots 0, 20, 40, linear between knots and
    # held constant after the last knot (no extrapolation)
    i = np.arange(64)
    expect = 0.05 * np.cos(np.deg2rad(5.0)) * np.minimum(i, 40)
    assert np.abs(vals[:, 5] - expect).max() < 2e-3 * expect.max()


def test_wbd_other_polarisation_is_not_picked_up(tmp_path):
    from wbd import make_dwell
    make_dwell(tmp_path, pulses=32, range_num=1024, polar="1100")
    r = subprocess.run([str(APP), str(tmp_path), "32", "8", "0", str(tmp_path / "x"), "1"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 4 and "no valid file" in r.stderr


def _great_circle_deg(d_az, pitch):
    """Angle between two unit vectors at the same pitch whose azimuths differ by d_az degrees."""
    p, d = np.deg2rad(pitch), np.deg2rad(d_az)
    return np.rad2deg(np.arccos(np.sin(p) ** 2 + np.cos(p) ** 2 * np.cos(d)))


def test_turn_angle_curve_across_a_pitch_jump(tmp_path):
    """turnAngleLine (common.cu:902-980) on 100 pulses whose pitch jumps from 5 to 40 deg after
    pulse 59 (a jump > 0.2 deg starts a block): block 0 (pulses 0-59) accumulates great-circle
    angles at the 20-pulse knots 0, 20, 40 and interpolates linearly between them, held after
    the last knot. Block 1 (pulses 60-99, knots 0 and 20) samples the curve at its
    block-relative knot positions taken as absolute indices (common.cu:968-975), i.e. block 0's
    values at pulses 0 and 20, which discards its own accumulated angle; then it adds block 0's
    last value. With A1, A2 the angles of pulses 0-20 and 20-40: i*A1/20 to pulse 20, then
    A1 + (i-20)*A2/20 to 40, A1 + A2 to 59, then A1 + A2 + min(i-60, 20)*A1/20."""
    from wbd import make_dwell
    P, rate = 100, 0.05
    pitch = np.where(np.arange(P) < 60, 5.0, 40.0)
    make_dwell(tmp_path, pulses=P, range_num=1024, azimuth_rate_deg=rate, pitch_deg=pitch)
    r = subprocess.run([str(APP), str(tmp_path), str(P), "16", "0", str(tmp_path / "x"), "1"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    vals = np.array([[float(v) for v in ln.split()] for ln in r.stdout.strip().splitlines()[1:]])
    assert vals.shape[0] == P
    az, pt = vals[:, 3], vals[:, 4]   # the angles as the file holds them (quantised to 360/2^24 deg)
    A1 = _great_circle_deg(az[20] - az[0], pt[0])
    A2 = _great_circle_deg(az[40] - az[20], pt[20])
    i = np.arange(P)
    want = np.where(i <= 20, i * A1 / 20, A1 + np.minimum(i - 20, 20) * A2 / 20)
    want = np.where(i >= 60, A1 + A2 + np.minimum(i - 60, 20) * A1 / 20, want)
    assert np.abs(vals[:, 5] - want).max() < 1e-3 * want.max()
    # block 1's own angle over pulses 60-80 (at 40 deg pitch, about 0.77 A1) is discarded
    own = _great_circle_deg(az[80] - az[60], pt[60])
    assert own < 0.8 * A1 and abs(vals[80, 5] - (A1 + A2 + own)) > 0.2 * A1
