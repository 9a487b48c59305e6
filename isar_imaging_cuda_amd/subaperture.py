"""Sub-aperture plans (include/isar_subaperture.h): many windows of uneven pulse counts of one
dwell imaged in one HIP graph launch, with a per-pulse-count autotuner."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import Options, RadarParams, _check, lib

PULSE_CHOICES = (128, 192, 256, 384, 512)


def _bind() -> C.CDLL:
    L = lib()
    if getattr(L, "_sa_bound", False):
        return L
    vp, i = C.c_void_p, C.c_int
    L.isar_sa_plan_create.argtypes = [vp, i, i, vp, vp, C.POINTER(vp)]
    L.isar_sa_plan_destroy.argtypes = [vp]
    L.isar_sa_plan_image_floats.argtypes = [vp]
    L.isar_sa_plan_image_floats.restype = C.c_size_t
    L.isar_sa_plan_image_offsets.argtypes = [vp, vp]
    L.isar_sa_plan_run.argtypes = [vp, vp, vp, vp, vp, i, vp]
    L.isar_sa_plan_autotune.argtypes = [vp, vp, vp, vp, vp, i, vp]
    L.isar_sa_plan_groups.argtypes = [vp, vp, vp]
    L.isar_sa_plan_group_tuning.argtypes = [vp, i, vp]
    L.isar_run_frames.argtypes = [vp, i, vp, vp, vp, vp, vp, vp, vp, vp]
    L.isar_set_ra_chunk.argtypes = [vp, i]
    L.isar_get_ra_chunk.argtypes = [vp]
    L._sa_bound = True
    return L


def sliding_windows(dwell_pulses: int, count: int = 64, seed: int = 0,
                    choices: tuple[int, ...] = PULSE_CHOICES) -> tuple[np.ndarray, np.ndarray]:
    """`count` windows with pulse counts drawn uniformly from `choices` and seeded start pulses,
    ordered by start."""
    rng = np.random.default_rng(seed)
    pulses = rng.choice(np.asarray(choices), size=count)
    first = np.array([rng.integers(0, dwell_pulses - m + 1) for m in pulses])
    order = np.argsort(first, kind="stable")
    return first[order].astype(np.int32), pulses[order].astype(np.int32)


class SubaperturePlan:
    def __init__(self, radar: RadarParams, dwell_pulses: int, first_pulse, pulses):
        L = _bind()
        self.first = np.ascontiguousarray(first_pulse, dtype=np.int32)
        self.pulses = np.ascontiguousarray(pulses, dtype=np.int32)
        self.n = self.first.size
        self.W = radar.range_num_cut
        self._h = C.c_void_p()
        p = radar.c()
        _check(L.isar_sa_plan_create(C.byref(p), dwell_pulses, self.n, self.first.ctypes.data_as(C.c_void_p),
                                     self.pulses.ctypes.data_as(C.c_void_p), C.byref(self._h)), "isar_sa_plan_create")
        self.image_floats = int(L.isar_sa_plan_image_floats(self._h))
        self.offsets = np.zeros(self.n, np.int64)
        _check(L.isar_sa_plan_image_offsets(self._h, self.offsets.ctypes.data_as(C.c_void_p)),
               "isar_sa_plan_image_offsets")
        ng = 8
        gp = np.zeros(ng, np.int32)
        gw = np.zeros(ng, np.int32)
        self.n_groups = L.isar_sa_plan_groups(self._h, gp.ctypes.data_as(C.c_void_p), gw.ctypes.data_as(C.c_void_p))
        self.group_pulses = gp[:self.n_groups].copy()
        self.group_windows = gw[:self.n_groups].copy()

    def close(self) -> None:
        if self._h:
            _bind().isar_sa_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, dwell_ptr: int, vel_ptr: int, img_ptr: int, options: Options | None = None,
            graph: bool = True, stream: int | None = None) -> None:
        o = (options or Options()).c()
        _check(_bind().isar_sa_plan_run(self._h, C.c_void_p(dwell_ptr), C.c_void_p(vel_ptr or 0), C.c_void_p(img_ptr),
                                        C.byref(o), int(graph), C.c_void_p(stream or 0)), "isar_sa_plan_run")

    def autotune(self, dwell_ptr: int, vel_ptr: int, img_ptr: int, options: Options | None = None,
                 repeats: int = 3) -> np.ndarray:
        o = (options or Options()).c()
        chosen = np.zeros(self.n_groups, np.int32)
        _check(_bind().isar_sa_plan_autotune(self._h, C.c_void_p(dwell_ptr), C.c_void_p(vel_ptr or 0),
                                             C.c_void_p(img_ptr), C.byref(o), repeats,
                                             chosen.ctypes.data_as(C.c_void_p)), "isar_sa_plan_autotune")
        return chosen

    def tuning(self) -> dict[int, dict[str, int]]:
        """Per pulse count: the group's range-alignment chunk rows, autofocus form, chirp-z length
        and chirp-z workgroup threads."""
        out = {}
        for g, M in enumerate(self.group_pulses):
            t = np.zeros(4, np.int32)
            _check(_bind().isar_sa_plan_group_tuning(self._h, g, t.ctypes.data_as(C.c_void_p)),
                   "isar_sa_plan_group_tuning")
            out[int(M)] = dict(ra_chunk=int(t[0]), autofocus_mode=int(t[1]), czt_length=int(t[2]),
                               czt_threads=int(t[3]))
        return out

    def set_tuning(self, tuning: dict[int, dict[str, int]]) -> None:
        """Applies per pulse count the layout tuning() returns (isar_sa_plan_set_group_tuning)."""
        L = _bind()
        L.isar_sa_plan_set_group_tuning.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        for g, M in enumerate(self.group_pulses):
            if int(M) not in tuning:
                continue
            t = tuning[int(M)]
            a = np.array([t["ra_chunk"], t["autofocus_mode"], t["czt_length"], t["czt_threads"]], np.int32)
            _check(L.isar_sa_plan_set_group_tuning(self._h, g, a.ctypes.data_as(C.c_void_p)),
                   "isar_sa_plan_set_group_tuning")

    def calibrate(self, dwell_ptr: int, vel_ptr: int, img_ptr: int, options: Options | None = None,
                  repeats: int = 2) -> None:
        """Times each group alone with its current tuning and places the groups on the graph's
        branches by those times (isar_sa_plan_calibrate)."""
        o = (options or Options()).c()
        L = _bind()
        L.isar_sa_plan_calibrate.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        _check(L.isar_sa_plan_calibrate(self._h, C.c_void_p(dwell_ptr), C.c_void_p(vel_ptr or 0), C.c_void_p(img_ptr),
                                        C.byref(o), int(repeats)), "isar_sa_plan_calibrate")

    def lanes(self) -> list[int]:
        """Graph branch of each group (groups in increasing pulse count)."""
        L = _bind()
        L.isar_sa_plan_group_lane.argtypes = [C.c_void_p, C.c_int]
        return [int(L.isar_sa_plan_group_lane(self._h, g)) for g in range(self.n_groups)]

    def image(self, flat: np.ndarray, i: int) -> np.ndarray:
        """Window i's image out of the flat output."""
        o = int(self.offsets[i])
        m = int(self.pulses[i])
        return flat[o:o + m * self.W].reshape(m, self.W)
