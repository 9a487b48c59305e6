"""The kernel sources executed on the CPU (csrc/emu/hip_emu.hpp, AddressSanitizer and
UndefinedBehaviorSanitizer on) against the FP64 oracle: indexing, LDS use and barrier structure
checked without a GPU, in both launch forms of the range-alignment update (one grid per merge
level, and the rows of pair 0 apart from the other pairs, isar_set_ra_split). A 192-pulse frame
carries an odd group up a level (range_alignment.cu:34-73 generalised to any even M). Emulated
device memory starts as all-ones bytes (NaN floats), so a kernel reading workspace it did not
write in the same run would show; the static-workspace entry point (isar_run_device_static) must
equal the context run bit for bit, on a first call and on a second, smaller one. The 192-pulse
frame runs the chirp-z transform at the mixed-radix length 384 with 512-thread column tiles. The
first c1 case runs the one-workgroup autofocus, the second the split form and the third the
resident form (every workgroup of the frame run concurrently, exchanging its sums through the
hand-off counter), and the three images must agree bit for bit."""
from __future__ import annotations

import concurrent.futures as cf

import numpy as np
import pytest

import emu_runner
from isar_imaging_cuda_amd import build as B
from isar_imaging_cuda_amd import synth
from parity import compare, oracle_run

pytestmark = pytest.mark.slow

# a few autofocus iterations keep the emulated run to about a minute; the oracle runs as many
ITERATIONS = 4
# (pulses, ra_split, (chirp-z length, threads) or None, autofocus mode or None, name)
CASES = [(128, 1, None, 1, "c1, one grid"), (128, 2, None, 2, "c1, two grids"),
         (128, 2, None, 4, "c1, two grids, resident autofocus"),
         (192, 2, (384, 512), None, "M=192, two grids, chirp-z length 384, 512-thread tiles")]
STATIC_CASE = (64, 2, "M=64 x 2 frames, static workspace")


@pytest.fixture(scope="module")
def emu_driver():
    return B.build_emu(force=False)


def _one(M, split, czt, af):
    cfg = synth.preset(1, echo_num=M, range_num=1024)
    echo, vel = synth.frame(cfg)
    radar = synth.radar_header(cfg)
    length, threads = czt or (None, None)
    (fr,) = emu_runner.run(echo, vel, radar, emu_runner.default_options(iterations=ITERATIONS), autofocus_mode=af,
                           ra_split=split, czt_length=length, czt_threads=threads, timeout=600)
    img_o, tr = oracle_run(echo, vel, radar, M, 1024, iterations=ITERATIONS)
    return fr, compare(fr.img, fr, img_o, tr, M), tr


def _static(M, frames):
    """Two frames through the context, then through isar_run_device_static (both frames, then the
    first alone); the driver exits non-zero unless all three agree bit for bit."""
    cfgs = [synth.preset(1, echo_num=M, range_num=1024, seed=40 + k) for k in range(frames)]
    fr = [synth.frame(c) for c in cfgs]
    radar = synth.radar_header(cfgs[0])
    echo, vel = np.stack([f[0] for f in fr]), np.stack([f[1] for f in fr])
    out = emu_runner.run(echo, vel, radar, emu_runner.default_options(iterations=ITERATIONS), static_check=True,
                         timeout=600)
    reps = []
    for k in range(frames):
        img_o, tr = oracle_run(echo[k], vel[k], radar, M, 1024, iterations=ITERATIONS)
        reps.append(compare(out[k].img, out[k], img_o, tr, M))
    return reps


def test_emulated_kernels_match_oracle(emu_driver):
    with cf.ThreadPoolExecutor(max_workers=len(CASES) + 1) as ex:
        futs = {name: ex.submit(_one, M, split, czt, af) for M, split, czt, af, name in CASES}
        fstatic = ex.submit(_static, STATIC_CASE[0], STATIC_CASE[1])
        res = {name: f.result() for name, f in futs.items()}
        static_reps = fstatic.result()
    for rep in static_reps:
        print(STATIC_CASE[2], rep)
        assert rep.ok(), f"{STATIC_CASE[2]}: {rep}"
    for name, (fr, rep, tr) in res.items():
        print(name, rep)
        assert rep.ok(), f"{name}: {rep}"
        # FP64 parabola on exact direct sums: the lags agree far inside the 1e-3-cell bar
        assert np.abs(fr.ra_lag - tr.ra_lag[: len(fr.ra_lag)]).max() < 1e-5, name
    one, two = res["c1, one grid"][0], res["c1, two grids"][0]
    assert np.array_equal(one.img, two.img) and np.array_equal(one.ra_lag, two.ra_lag)
    resident = res["c1, two grids, resident autofocus"][0]
    assert resident.status & 2 == 0 and np.array_equal(resident.img, two.img)
