"""Comparison of a pipeline run (device library or CPU emulation) against the FP64 oracle.

Bars: integer decisions identical (near-ties logged, not failed), fractional
range-alignment lags within 1e-3 cell, image relative L2 <= 1e-4, image entropy within 1e-5
relative.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from oracle import oracle as O

IMG_REL_L2 = 1e-4
ENTROPY_REL = 1e-5
LAG_ABS = 1e-3
NEAR_TIE = 1e-5


@dataclass
class ParityReport:
    near_ties: list[str] = field(default_factory=list)
    failures: list[str] = field(default_factory=list)
    img_rel_l2: float = float("nan")
    entropy_rel: float = float("nan")

    def ok(self) -> bool:
        return not self.failures

    def __str__(self) -> str:
        return (f"img_rel_l2={self.img_rel_l2:.3e} entropy_rel={self.entropy_rel:.3e} "
                f"near_ties={self.near_ties} failures={self.failures}")


def oracle_run(echo, velocity, radar: dict, M: int, N: int, **opts):
    p = O.params(M, N, 512, radar["band_width"], radar["fc"], radar["Tp"], radar["Fs"])
    tr = O.Trace(M, N, 512)
    img = O.isar_main(echo, velocity, p, O.options(**opts), tr)
    return img, tr


def compare(img, t, img_o, tr, M: int, img_rel_l2: float = IMG_REL_L2) -> ParityReport:
    """t: device/emulation trace (attributes as in isar_imaging_cuda_amd.FrameTrace)."""
    r = ParityReport()
    # range alignment: integer argmax per (level, pair); lags
    n = M - 1
    imax_o = tr.ra_imax[:n]
    diff = np.nonzero(np.asarray(t.ra_imax) != imax_o)[0]
    for i in diff:
        if tr.ra_gap[i] < NEAR_TIE or abs(float(t.ra_lag[i]) - tr.ra_lag[i]) <= LAG_ABS:
            r.near_ties.append(f"ra pair {i}: imax {t.ra_imax[i]} vs {imax_o[i]} (gap {tr.ra_gap[i]:.2e})")
        else:
            r.failures.append(f"ra pair {i}: imax {t.ra_imax[i]} vs {imax_o[i]} (gap {tr.ra_gap[i]:.2e})")
    lag_err = np.abs(np.asarray(t.ra_lag, np.float64) - tr.ra_lag[:n])
    if lag_err.size and lag_err.max() > LAG_ABS:
        bad = int(np.argmax(lag_err))
        r.failures.append(f"ra lag error {lag_err.max():.3e} at pair {bad}")
    # centring and cut
    for name in ("center_applied", "center_shift", "idx1_len", "idx2_len", "diff_imin", "cut_offset"):
        a, b = int(getattr(t, name)), int(getattr(tr.c, name))
        if a != b:
            r.failures.append(f"{name}: {a} vs {b}")
    # bin selection
    for name in ("tgt_num", "unit2_num"):
        a, b = int(getattr(t, name)), int(getattr(tr.c, name))
        if a != b:
            r.failures.append(f"{name}: {a} vs {b}")
    if not r.failures:
        if not np.array_equal(np.asarray(t.select_bin), tr.selected):
            r.failures.append(f"select_bin differs: {sorted(set(t.select_bin) ^ set(tr.selected))}")
    # image and entropy
    img = np.asarray(img, np.float64)
    r.img_rel_l2 = float(np.linalg.norm(img - img_o) / max(np.linalg.norm(img_o), 1e-300))
    e_dev = O.image_entropy(np.asarray(img, np.float32))
    r.entropy_rel = abs(e_dev - tr.c.entropy) / max(abs(tr.c.entropy), 1e-300)
    if not r.failures:
        if r.img_rel_l2 > img_rel_l2:
            r.failures.append(f"image rel L2 {r.img_rel_l2:.3e} > {img_rel_l2}")
        if r.entropy_rel > ENTROPY_REL:
            r.failures.append(f"entropy rel {r.entropy_rel:.3e} > {ENTROPY_REL}")
    return r
