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


def rel_l2(a, ref) -> float:
    """||a - ref|| / ||ref|| in float64 (complex128 for complex data); 0 when both are zero, inf
    when only ref is."""
    dt = np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(ref) else np.float64
    a, ref = np.asarray(a, dt), np.asarray(ref, dt)
    num, den = float(np.linalg.norm(a - ref)), float(np.linalg.norm(ref))
    if den > 0:
        return num / den
    return 0.0 if num == 0 else float("inf")


def centring_margin(tr, N: int, inter_length: int = 30) -> float:
    """Smallest relative margin behind the oracle's centring decisions (range_alignment.cu:402-449):
    the gap between the two best candidates of the sorted-ARP difference search, and the distance
    of every ARP value and every idx_1 window mean from the threshold. Below NEAR_TIE the decision
    rests on rounding (a spectrum symmetric about its peak ties exactly) and FP32 may take either
    side."""
    arp = np.asarray(tr.arp[:N], np.float64)
    s = np.sort(arp)
    d = np.abs(s[inter_length:] - s[:-inter_length] - inter_length * np.abs(s).max() / N)
    d2 = np.sort(d)[:2]
    margins = [(d2[1] - d2[0]) / max(d2[1], 1e-300)]
    thr = float(tr.c.center_thr)
    if thr > 0:
        above = arp[arp != thr]
        if above.size:
            margins.append(float(np.min(np.abs(above - thr))) / thr)
        means = np.asarray(tr.arp_mean[:int(tr.c.idx1_len)], np.float64)
        if means.size:
            margins.append(float(np.min(np.abs(means - thr))) / thr)
    return float(min(margins))


def selection_margin(tr) -> float:
    """Smallest margin behind the oracle's bin selection (phase_adjustment.cu:248-264): the
    relative energy gap at the tgt_num cut of the energy ranking, and the absolute entropy gap
    (entropies are dimensionless; FP32 resolves them to about 1e-6) at the unit2 cut of the
    entropy ranking. A column set whose entropies all round to zero (a lone tone) ties here."""
    W = len(tr.pa_colE)
    E, H = np.asarray(tr.pa_colE, np.float64), np.asarray(tr.pa_colH, np.float64)
    u1, u2 = int(tr.c.tgt_num), int(tr.c.unit2_num)
    margins = [np.inf]
    order = np.argsort(-E, kind="stable")
    if 0 < u1 < W and E[order[u1 - 1]] > 0:
        margins.append((E[order[u1 - 1]] - E[order[u1]]) / E[order[u1 - 1]])
    top = order[:u1]
    by_h = top[np.argsort(H[top], kind="stable")]
    if 0 < u2 < len(by_h):
        margins.append(H[by_h[u2]] - H[by_h[u2 - 1]])
    return float(min(margins))


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
    if lag_err.size and not lag_err.max() <= LAG_ABS:
        bad = int(np.argmax(lag_err))
        r.failures.append(f"ra lag error {lag_err.max():.3e} at pair {bad}")
    # centring and cut (the cut follows the centring shift); the image bars below still apply
    tie = centring_margin(tr, len(tr.arp)) < NEAR_TIE
    for name in ("center_applied", "center_shift", "idx1_len", "idx2_len", "diff_imin", "cut_offset"):
        a, b = int(getattr(t, name)), int(getattr(tr.c, name))
        if a != b:
            (r.near_ties if tie else r.failures).append(f"{name}: {a} vs {b}")
    # bin selection
    for name in ("tgt_num", "unit2_num"):
        a, b = int(getattr(t, name)), int(getattr(tr.c, name))
        if a != b:
            r.failures.append(f"{name}: {a} vs {b}")
    if not r.failures:
        if not np.array_equal(np.asarray(t.select_bin), tr.selected):
            tie = selection_margin(tr) < NEAR_TIE
            (r.near_ties if tie else r.failures).append(
                f"select_bin differs: {sorted(set(t.select_bin) ^ set(tr.selected))}")
    # image and entropy
    r.img_rel_l2 = rel_l2(img, img_o)
    e_dev, e_o = float(O.image_entropy(np.asarray(img, np.float32))), float(tr.c.entropy)
    if not (np.isfinite(e_dev) or np.isfinite(e_o)):
        r.entropy_rel = 0.0 if np.isnan(e_dev) == np.isnan(e_o) else float("inf")   # both undefined (zero image)
    else:
        r.entropy_rel = abs(e_dev - e_o) / abs(e_o) if e_o != 0 else abs(e_dev)
    if not r.failures:
        # written so that a NaN fails
        if not r.img_rel_l2 <= img_rel_l2:
            r.failures.append(f"image rel L2 {r.img_rel_l2:.3e} > {img_rel_l2}")
        if not r.entropy_rel <= ENTROPY_REL:
            r.failures.append(f"entropy rel {r.entropy_rel:.3e} > {ENTROPY_REL}")
    return r
