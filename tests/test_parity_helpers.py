"""The comparison that every parity test relies on: it must fail on NaN and on a planted wrong
decision, and excuse a decision only where the oracle's own margin shows a rounding tie."""
from __future__ import annotations

import copy

import numpy as np
import pytest

from isar_imaging_cuda_amd import synth
from parity import NEAR_TIE, centring_margin, compare, oracle_run, rel_l2, selection_margin


class _Trace:
    """A device-style trace holding the oracle's own decisions."""

    def __init__(self, tr, M):
        for name in ("center_applied", "center_shift", "idx1_len", "idx2_len", "diff_imin", "cut_offset",
                     "tgt_num", "unit2_num"):
            setattr(self, name, int(getattr(tr.c, name)))
        self.ra_imax = tr.ra_imax[:M - 1].copy()
        self.ra_lag = tr.ra_lag[:M - 1].copy()
        self.select_bin = tr.selected.copy()


@pytest.fixture(scope="module")
def c1():
    cfg = synth.preset(1, seed=321)
    return cfg, synth.radar_header(cfg)


def test_rel_l2_zero_and_nan():
    z = np.zeros((4, 4), np.float32)
    assert rel_l2(z, z) == 0.0
    bad = z.copy()
    bad[1, 2] = np.nan
    assert not rel_l2(bad, z) <= 1e-4
    ref = np.ones((4, 4), np.float32)
    bad = ref.copy()
    bad[0, 0] = np.nan
    assert not rel_l2(bad, ref) <= 1e-4
    assert rel_l2(ref * 2, ref) == pytest.approx(1.0)


def test_compare_passes_the_oracle_and_fails_planted_mistakes(c1):
    cfg, radar = c1
    M, N = cfg.echo_num, cfg.range_num
    echo, vel = synth.frame(cfg)
    img_o, tr = oracle_run(echo, vel, radar, M, N)
    assert centring_margin(tr, N) > 10 * NEAR_TIE and selection_margin(tr) > 10 * NEAR_TIE
    t = _Trace(tr, M)
    assert compare(img_o, t, img_o, tr, M).ok()
    for name, delta in (("center_shift", 1), ("cut_offset", -1), ("unit2_num", 1)):
        w = copy.deepcopy(t)
        setattr(w, name, getattr(w, name) + delta)
        assert not compare(img_o, w, img_o, tr, M).ok(), name
    w = copy.deepcopy(t)
    w.select_bin[0] += 1
    assert not compare(img_o, w, img_o, tr, M).ok()
    w = copy.deepcopy(t)
    w.ra_lag[5] = np.nan
    assert not compare(img_o, w, img_o, tr, M).ok()
    bad = img_o.copy()
    bad[7, 9] = np.nan
    assert not compare(bad, t, img_o, tr, M).ok()
    assert not compare(img_o * 1.001, t, img_o, tr, M).ok()


def test_symmetric_tone_ties_and_zero_frame(c1):
    cfg, radar = c1
    M, N = cfg.echo_num, cfg.range_num
    tone = (1000 * np.tile(np.exp(2j * np.pi * np.arange(N) * 137 / N), (M, 1))).astype(np.complex64)
    img_o, tr = oracle_run(tone, np.zeros(M), radar, M, N)
    assert centring_margin(tr, N) < NEAR_TIE and selection_margin(tr) < NEAR_TIE
    w = _Trace(tr, M)
    w.center_shift += 1
    w.cut_offset += 1
    rep = compare(img_o, w, img_o, tr, M)
    assert rep.ok() and len(rep.near_ties) == 2
    zero = np.zeros((M, N), np.complex64)
    img_z, tz = oracle_run(zero, np.zeros(M), radar, M, N)
    assert not img_z.any() and tz.c.center_applied == 0
    assert compare(img_z, _Trace(tz, M), img_z, tz, M).ok()
    assert not compare(np.ones_like(img_z), _Trace(tz, M), img_z, tz, M).ok()
