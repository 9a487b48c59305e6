"""Intermediate arrays of a run (isar_stage_read / isar_stage_dump, the reference's
DATA_WRITE_BACK_* dumps, isar_main.cu:10-223) checked stage by stage: every device stage against
the FP64 oracle's stage function applied to the device's own previous stage (MI355X). At
1024 x 16384 this checks the FP64 phase generation, the high-speed quadratic phase and the
chirp-z chirps, stage by stage."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

import isar_imaging_cuda_amd as isar
from isar_imaging_cuda_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
APP = ROOT / "isar_imaging_cuda_amd" / "lib" / "isar_calling_app"

# Bars per stage (relative L2 against the oracle stage fed with the device's previous stage).
HPC_REL = 1e-6       # FP32 samples times an FP64-reduced phase
HRRP_REL = 1e-5      # one FP32 transform of <= 16384 points
RA_REL = 1e-4        # FP32 row transforms, FP64 lags
PC_REL = 1e-4        # cut, Doppler tracking and 120 autofocus iterations
MTRC_REL = 1e-5      # chirp-z transform with FP64 chirps, FP32 transforms


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _run(cid, opts=None, **over):
    cfg = synth.preset(cid, **over)
    echo, vel = synth.frame(cfg)
    h = synth.radar_header(cfg)
    M, N = cfg.echo_num, cfg.range_num
    params = isar.RadarParams(M, N, h["band_width"], h["fc"], h["Tp"], h["Fs"])
    with isar.Imager(params, max_frames=1) as im:
        img = im.run_host(echo, vel, opts or isar.Options())
        st = {n: im.stage(n) for n in isar.DUMP_STAGES}
        t = im.trace(0)
    p = O.params(M, N, 512, h["band_width"], h["fc"], h["Tp"], h["Fs"])
    return echo, vel, p, img[0] if img.ndim == 3 else img, st, t


@pytest.mark.parametrize("name,cid", [("c1", 1), ("c2", 2), ("c5", 5)])
def test_stage_by_stage(name, cid):
    print(f"[case] stages {name}", flush=True)
    echo, vel, p, img, st, t = _run(cid)
    M, N = echo.shape
    r = {}
    assert np.array_equal(st["dataW"], echo)
    r["hpc"] = _rel(st["hpc"], O.hpc(echo, vel, p))
    r["hrrp"] = _rel(st["hrrp"], O.hrrp(echo, vel, p, O.options()))
    tr = O.Trace(M, N, 512)
    ra_o = O.hrrp_center(O.range_alignment(st["hpc"], p, tr), p, tr)
    r["ra"] = _rel(st["ra"], ra_o)
    r["mtrc"] = _rel(st["mtrc"], O.mtrc(st["pc"], p, 2))
    tr2 = O.Trace(M, N, 512)
    r["pc"] = _rel(st["pc"], O.fast_entropy(O.cut_range_profile(st["ra"], p, tr2), 120, tr2))
    print({k: f"{v:.2e}" for k, v in r.items()}, f"shift dev {t.center_shift} oracle {tr.center_shift}")
    assert t.center_shift == tr.center_shift
    assert r["hpc"] <= HPC_REL and r["hrrp"] <= HRRP_REL
    assert r["ra"] <= RA_REL
    assert r["mtrc"] <= MTRC_REL
    assert r["pc"] <= PC_REL
    assert np.array_equal(st["final"], img)


def test_stage_variants_ifds_and_no_mtrc():
    """IFDS: the pre-stage rows and the profiles equal to the input rows; without MTRC the MTRC
    dump is the compensated cut."""
    echo, vel, p, img, st, t = _run(1, isar.Options(data_type=1))
    pre = np.fft.ifftshift(np.fft.ifft(np.fft.ifftshift(echo.astype(np.complex128), axes=1), axis=1), axes=1)
    assert _rel(st["hpc"], pre) <= 1e-5
    assert np.array_equal(st["hrrp"], echo)
    assert _rel(st["mtrc"], O.mtrc(st["pc"], p, 1)) <= MTRC_REL
    echo, vel, p, img, st, t = _run(1, isar.Options(if_mtrc=False, if_hpc=False))
    assert np.array_equal(st["mtrc"], st["pc"])
    assert np.array_equal(st["hpc"], echo)


def test_dropin_session_dump(tmp_path):
    from wbd import make_dwell
    make_dwell(tmp_path / "data", pulses=128, range_num=1024)
    dump = tmp_path / "dump"
    dump.mkdir()
    print("[case] drop-in stage dump, 1 window of 128 pulses", flush=True)
    r = subprocess.run([str(APP), str(tmp_path / "data"), "128", "128", "1", str(tmp_path / "img"), "1", "1", "1", "2",
                        "0", str(dump)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-800:])
    assert r.returncode == 0, r.stderr
    img = np.fromfile(tmp_path / "img_0.bin", dtype=np.float32)
    final = np.loadtxt(dump / "final.dat")
    assert np.allclose(final, img, atol=6e-6 + 1e-7 * np.abs(img))
    for name, n in (("dataW", 2 * 128 * 1024), ("hpc", 2 * 128 * 1024), ("hrrp", 2 * 128 * 1024),
                    ("ra", 2 * 128 * 1024), ("pc", 2 * 128 * 512), ("mtrc", 2 * 128 * 512)):
        v = np.loadtxt(dump / f"{name}.dat")
        assert v.size == n and np.isfinite(v).all(), name
