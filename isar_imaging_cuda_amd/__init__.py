"""MI355X-native ISAR range-Doppler imaging (Python binding of libisar_imaging.so).

The imaging chain is the reference's isarMainSingle path (isar_imaging_export.cpp:208-212 ->
isar_main.cu:3-227): high-speed compensation, range alignment, HRRP centring and cut, Doppler
tracking + fast minimum-entropy autofocus, MTRC (chirp-z keystone) and the azimuth FFT. All
compute runs in the HIP kernels of libisar_imaging.so; this module only moves pointers. There is
no CPU fallback: if the library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

__all__ = ["RadarParams", "Options", "Imager", "lib", "fft", "ImagingError", "library_path"]

_LIB_DIR = Path(__file__).resolve().parent / "lib"
library_path = _LIB_DIR / "libisar_imaging.so"


class ImagingError(RuntimeError):
    pass


class _Params(C.Structure):
    _fields_ = [("echo_num", C.c_int32), ("range_num", C.c_int32), ("range_num_cut", C.c_int32),
                ("reserved", C.c_int32), ("band_width", C.c_int64), ("fc", C.c_int64), ("Tp", C.c_double),
                ("Fs", C.c_int64)]


class _Options(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("data_type", "option_alignment", "option_phase", "if_hrrp", "if_hpc",
                                          "if_mtrc", "quirk_q1", "iterations", "warm_start", "stage_mode")]


_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)


class _Trace(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("n_pairs", "center_applied", "center_shift", "idx1_len", "idx2_len",
                                          "diff_imin", "cut_offset", "tgt_num", "unit2_num", "status")] + \
               [("center_thr", C.c_float), ("pa_threshold", C.c_float), ("ra_lag", _dp), ("ra_imax", _ip),
                ("pa_colmax", _fp), ("pa_colE", _fp), ("pa_colH", _fp), ("select_bin", _ip), ("phi0", _fp),
                ("phi", _fp)]


# kernel families (enum isar_stage in include/isar_imaging_ext.h)
# isar_autofocus_mode
AF_AUTO, AF_FUSED, AF_SPLIT, AF_SPLIT_PHASE, AF_RESIDENT = 0, 1, 2, 3, 4

# arrays of isar_stage_read (the reference's DATA_WRITE_BACK_* dumps): name -> (id, complex?, columns)
DUMP_STAGES = {"dataW": 0, "hpc": 1, "hrrp": 2, "ra": 3, "pc": 4, "mtrc": 5, "final": 6}

STAGES = {"ifds_pre": 0, "ra_pair0": 1, "ra_corr": 2, "ra_upd": 3, "center": 4, "cut": 5, "pa_colfft": 6,
          "pa_select": 7, "autofocus": 8, "mtrc_rows": 9, "mtrc_czt": 10, "col_window": 11, "mtrc_image": 12,
          "post_nomtrc": 13, "hrrp": 14}

_lib = None


def _torch_runtime_first() -> None:
    """PyTorch's ROCm wheel carries its own HIP and HSA runtimes next to the system ones this
    library links. Two runtimes coexist in one process when PyTorch's initialises first; the
    other order leaves PyTorch without a GPU. When PyTorch is importable, initialise it first."""
    try:
        import torch
    except Exception:
        return
    try:
        if torch.cuda.is_available():
            torch.cuda.init()
    except Exception:
        pass


def lib() -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not library_path.exists():
            raise ImagingError(f"{library_path} is missing: run __graft_entry__.build() first")
        _torch_runtime_first()
        L = C.CDLL(str(library_path))
        vp = C.c_void_p
        L.isar_default_options.argtypes = [C.POINTER(_Options)]
        L.isar_create.argtypes = [C.POINTER(_Params), C.c_int, C.POINTER(vp)]
        L.isar_destroy.argtypes = [vp]
        L.isar_workspace_bytes.argtypes = [C.POINTER(_Params), C.c_int]
        L.isar_workspace_bytes.restype = C.c_size_t
        L.isar_run_device.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.POINTER(_Options), vp]
        L.isar_run_host.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.POINTER(_Options)]
        L.isar_get_trace.argtypes = [vp, C.c_int, C.POINTER(_Trace)]
        L.isar_get_phases.argtypes = [vp, C.c_int, vp, vp, vp]
        L.isar_fft_host.argtypes = [C.c_int, C.c_int, vp, vp, C.c_int]
        L.isar_probe_enable.argtypes = [vp, C.c_int]
        L.isar_probe_read.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.isar_last_error.restype = C.c_char_p
        L.isar_device_count.restype = C.c_int
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().isar_last_error()
        raise ImagingError(f"{what} failed (rc={rc}): {msg.decode() if msg else ''}")


@dataclass
class RadarParams:
    """RadarParameters of the reference (common.cuh:96-117)."""
    echo_num: int
    range_num: int
    band_width: int = 1_000_000_000
    fc: int = 10_000_000_000
    Tp: float | None = None
    Fs: int = 64_000_000
    range_num_cut: int = 512

    def c(self) -> _Params:
        tp = self.Tp if self.Tp is not None else self.range_num / self.Fs
        return _Params(self.echo_num, self.range_num, self.range_num_cut, 0, int(self.band_width), int(self.fc),
                       float(tp), int(self.Fs))


@dataclass
class Options:
    """Switches of ISAR_RD_Imaging_Main_Ku (isar_main.cu:3-4); STRETCH with HPC and MTRC by default."""
    data_type: int = 2
    option_alignment: int = 0
    option_phase: int = 1
    if_hrrp: bool = False
    if_hpc: bool = True
    if_mtrc: bool = True
    quirk_q1: bool = True
    iterations: int = 120
    warm_start: bool = False
    stage_mode: int = 0            # 0 full chain, 1 front (stop after the autofocus), 2 back

    def c(self) -> _Options:
        return _Options(self.data_type, self.option_alignment, self.option_phase, int(self.if_hrrp),
                        int(self.if_hpc), int(self.if_mtrc), int(self.quirk_q1), self.iterations,
                        int(self.warm_start), int(self.stage_mode))


@dataclass
class FrameTrace:
    n_pairs: int
    center_applied: int
    center_shift: int
    idx1_len: int
    idx2_len: int
    diff_imin: int
    cut_offset: int
    tgt_num: int
    unit2_num: int
    status: int
    center_thr: float
    pa_threshold: float
    ra_lag: np.ndarray = field(repr=False)
    ra_imax: np.ndarray = field(repr=False)
    pa_colmax: np.ndarray = field(repr=False)
    pa_colE: np.ndarray = field(repr=False)
    pa_colH: np.ndarray = field(repr=False)
    select_bin: np.ndarray = field(repr=False)
    phi0: np.ndarray = field(repr=False)
    phi: np.ndarray = field(repr=False)


class Imager:
    """One imaging context: device workspaces for up to `max_frames` frames of one shape."""

    def __init__(self, params: RadarParams, max_frames: int = 1):
        self.params = params
        self.max_frames = max_frames
        self._h = C.c_void_p()
        _check(lib().isar_create(C.byref(params.c()), max_frames, C.byref(self._h)), "isar_create")

    def close(self) -> None:
        if self._h:
            lib().isar_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def workspace_bytes(params: RadarParams, frames: int) -> int:
        return int(lib().isar_workspace_bytes(C.byref(params.c()), frames))

    def run_host(self, echo: np.ndarray, velocity: np.ndarray | None, options: Options | None = None,
                 phi0: np.ndarray | None = None) -> np.ndarray:
        """echo: complex64 (K, M, N) or (M, N). Returns float32 images (K, M, 512)."""
        options = options or Options()
        p = self.params
        e = np.ascontiguousarray(echo, dtype=np.complex64)
        squeeze = e.ndim == 2
        if squeeze:
            e = e[None]
        K = e.shape[0]
        if e.shape[1:] != (p.echo_num, p.range_num):
            raise ValueError(f"echo shape {e.shape} does not match ({p.echo_num}, {p.range_num})")
        v = None if velocity is None else np.ascontiguousarray(
            np.broadcast_to(velocity, (K, p.echo_num)), dtype=np.float64)
        ph = None if phi0 is None else np.ascontiguousarray(np.broadcast_to(phi0, (K, p.echo_num)),
                                                            dtype=np.complex64)
        img = np.zeros((K, p.echo_num, p.range_num_cut), np.float32)
        ptr = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        _check(lib().isar_run_host(self._h, K, ptr(e), ptr(v), ptr(ph), ptr(img), C.byref(options.c())),
               "isar_run_host")
        return img[0] if squeeze else img

    def run_device(self, echo_ptr: int, velocity_ptr: int | None, img_ptr: int, frames: int,
                   options: Options | None = None, phi0_ptr: int | None = None, stream: int | None = None) -> None:
        """Device-resident batch (raw device pointers, e.g. torch tensor data_ptr()); asynchronous."""
        options = options or Options()
        _check(lib().isar_run_device(self._h, frames, C.c_void_p(echo_ptr), C.c_void_p(velocity_ptr or 0),
                                     C.c_void_p(phi0_ptr or 0), C.c_void_p(img_ptr), C.byref(options.c()),
                                     C.c_void_p(stream or 0)), "isar_run_device")

    def phases(self, frames: int, start_ptr: int | None, final_ptr: int | None, stream: int | None = None) -> None:
        """Copy the autofocus start and final phases of the last run (K x M complex64 each) into
        device buffers; asynchronous on `stream`."""
        _check(lib().isar_get_phases(self._h, frames, C.c_void_p(start_ptr or 0), C.c_void_p(final_ptr or 0),
                                     C.c_void_p(stream or 0)), "isar_get_phases")

    def stage(self, name: str, frame: int = 0) -> np.ndarray:
        """Frame `frame`'s array of a stage of the last full run (DUMP_STAGES): complex64 (M, N)
        for dataW/hpc/hrrp/ra, complex64 (M, W) for pc/mtrc, float32 (M, W) for final."""
        L = lib()
        L.isar_stage_bytes.argtypes = [C.c_void_p, C.c_int]
        L.isar_stage_bytes.restype = C.c_size_t
        L.isar_stage_read.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        sid = DUMP_STAGES[name]
        p = self.params
        cols = p.range_num if sid <= 3 else p.range_num_cut
        out = np.zeros((p.echo_num, cols), np.float32 if name == "final" else np.complex64)
        assert out.nbytes == L.isar_stage_bytes(self._h, sid)
        _check(L.isar_stage_read(self._h, sid, int(frame), out.ctypes.data_as(C.c_void_p)), "isar_stage_read")
        return out

    def stage_dump(self, directory: str, frame: int = 0, names=tuple(DUMP_STAGES)) -> None:
        """The reference's text dumps (dataW.dat ... final.dat) of the named stages."""
        L = lib()
        L.isar_stage_dump.argtypes = [C.c_void_p, C.c_int, C.c_uint, C.c_char_p]
        mask = sum(1 << DUMP_STAGES[n] for n in names)
        _check(L.isar_stage_dump(self._h, int(frame), mask, str(directory).encode()), "isar_stage_dump")

    def set_autofocus_mode(self, mode: int) -> None:
        """AF_AUTO (split batches of at most 32 frames), AF_FUSED (one workgroup per frame),
        AF_SPLIT (several workgroups per frame, one launch per iteration), AF_SPLIT_PHASE (the
        same with a phase launch per iteration) or AF_RESIDENT (several workgroups per frame, all
        iterations in one launch); the phases are the same bit for bit."""
        L = lib()
        L.isar_set_autofocus_mode.argtypes = [C.c_void_p, C.c_int]
        _check(L.isar_set_autofocus_mode(self._h, int(mode)), "isar_set_autofocus_mode")

    def set_af_wait_polls(self, polls: int) -> None:
        """Bound of each resident-autofocus wait in polls (default 2^20); negative makes every wait
        give up at once and sets status bit 2 of every frame (a test of that path)."""
        L = lib()
        L.isar_set_af_wait_polls.argtypes = [C.c_void_p, C.c_int]
        _check(L.isar_set_af_wait_polls(self._h, int(polls)), "isar_set_af_wait_polls")

    @property
    def autofocus_mode(self) -> int:
        L = lib()
        L.isar_get_autofocus_mode.argtypes = [C.c_void_p]
        return int(L.isar_get_autofocus_mode(self._h))

    def set_graph_mode(self, on: bool) -> None:
        """Replay a captured HIP graph for full runs whose arguments repeat."""
        L = lib()
        L.isar_set_graph_mode.argtypes = [C.c_void_p, C.c_int]
        _check(L.isar_set_graph_mode(self._h, int(bool(on))), "isar_set_graph_mode")

    def set_hrrp_output(self, hrrp_ptr: int | None) -> None:
        """Device buffer (max_frames x M x N complex64) for the range profiles that runs with
        if_hrrp write; None stops the output."""
        L = lib()
        L.isar_set_hrrp_output.argtypes = [C.c_void_p, C.c_void_p]
        _check(L.isar_set_hrrp_output(self._h, C.c_void_p(hrrp_ptr or 0)), "isar_set_hrrp_output")

    def set_ra_chunk(self, rows: int) -> None:
        """Rows per workgroup in the range-alignment update (2, 4, 8 or 16)."""
        L = lib()
        L.isar_set_ra_chunk.argtypes = [C.c_void_p, C.c_int]
        _check(L.isar_set_ra_chunk(self._h, int(rows)), "isar_set_ra_chunk")

    def set_ra_split(self, mode: int) -> None:
        """Alignment update grids per merge level: 0 by batch size, 1 one grid, 2 two grids (the
        rows of pair 0 apart); bit-identical results."""
        L = lib()
        L.isar_set_ra_split.argtypes = [C.c_void_p, C.c_int]
        _check(L.isar_set_ra_split(self._h, int(mode)), "isar_set_ra_split")

    def czt_lengths(self) -> list[int]:
        """Chirp-z transform lengths this context accepts, the default (nextPow2(2M - 1)) first."""
        L = lib()
        L.isar_czt_lengths.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int]
        out = (C.c_int32 * 4)()
        n = L.isar_czt_lengths(self._h, out, 4)
        if n < 1:
            _check(n, "isar_czt_lengths")
        return [int(out[i]) for i in range(n)]

    def set_czt_length(self, length: int) -> None:
        """Chirp-z transform length (one of czt_lengths(); 0 = the default). Waits for the device
        and re-forms the chirp-z tables; the transform is the same up to FP32 rounding."""
        L = lib()
        L.isar_set_czt_length.argtypes = [C.c_void_p, C.c_int]
        _check(L.isar_set_czt_length(self._h, int(length)), "isar_set_czt_length")

    @property
    def czt_length(self) -> int:
        L = lib()
        L.isar_get_czt_length.argtypes = [C.c_void_p]
        return int(L.isar_get_czt_length(self._h))

    def set_czt_threads(self, threads: int) -> None:
        """Chirp-z workgroup width, 256 or 512 threads (bit-identical results)."""
        L = lib()
        L.isar_set_czt_threads.argtypes = [C.c_void_p, C.c_int]
        _check(L.isar_set_czt_threads(self._h, int(threads)), "isar_set_czt_threads")

    @property
    def czt_threads(self) -> int:
        L = lib()
        L.isar_get_czt_threads.argtypes = [C.c_void_p]
        return int(L.isar_get_czt_threads(self._h))

    def autotune(self, echo_ptr: int, velocity_ptr: int | None, img_ptr: int, frames: int,
                 options: Options | None = None, repeats: int = 3, stream: int | None = None) -> int:
        """Time every range-alignment chunk size on these frames, keep the fastest and return it;
        then the autofocus forms, and with MTRC the chirp-z lengths and column tiles (czt_length,
        czt_threads)."""
        L = lib()
        L.isar_autotune.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                    C.c_void_p, C.c_void_p]
        chosen = C.c_int32(0)
        o = (options or Options()).c()
        _check(L.isar_autotune(self._h, frames, C.c_void_p(echo_ptr), C.c_void_p(velocity_ptr or 0),
                               C.c_void_p(img_ptr), C.byref(o), repeats, C.byref(chosen), C.c_void_p(stream or 0)),
               "isar_autotune")
        return int(chosen.value)

    def probe_enable(self, stage: int) -> None:
        """Time every launch of kernel family `stage` (STAGES) with HIP events on the run stream."""
        _check(lib().isar_probe_enable(self._h, int(stage)), "isar_probe_enable")

    def probe_read(self) -> tuple[float, int]:
        """(summed milliseconds, launches) since the last enable/read."""
        ms, n = C.c_double(), C.c_int()
        _check(lib().isar_probe_read(self._h, C.byref(ms), C.byref(n)), "isar_probe_read")
        return ms.value, n.value

    def trace(self, frame: int = 0) -> FrameTrace:
        M, W = self.params.echo_num, self.params.range_num_cut
        lag = np.zeros(M - 1, np.float64)
        imax = np.zeros(M - 1, np.int32)
        cm, ce, ch = (np.zeros(W, np.float32) for _ in range(3))
        sel = np.full(W, -1, np.int32)
        p0 = np.zeros(2 * M, np.float32)
        p1 = np.zeros(2 * M, np.float32)
        t = _Trace()
        t.ra_lag = lag.ctypes.data_as(_dp)
        t.ra_imax = imax.ctypes.data_as(_ip)
        t.pa_colmax = cm.ctypes.data_as(_fp)
        t.pa_colE = ce.ctypes.data_as(_fp)
        t.pa_colH = ch.ctypes.data_as(_fp)
        t.select_bin = sel.ctypes.data_as(_ip)
        t.phi0 = p0.ctypes.data_as(_fp)
        t.phi = p1.ctypes.data_as(_fp)
        _check(lib().isar_get_trace(self._h, frame, C.byref(t)), "isar_get_trace")
        sc = {n: getattr(t, n) for n, _ in _Trace._fields_[:12]}
        return FrameTrace(**sc, ra_lag=lag, ra_imax=imax, pa_colmax=cm, pa_colE=ce, pa_colH=ch,
                          select_bin=sel[: t.unit2_num].copy(), phi0=p0[0::2] + 1j * p0[1::2],
                          phi=p1[0::2] + 1j * p1[1::2])


def static_workspace_bytes() -> int:
    """Size of the per-GPU device array isar_run_device_static images in."""
    L = lib()
    L.isar_static_workspace_bytes.restype = C.c_size_t
    return int(L.isar_static_workspace_bytes())


def run_device_static(params: RadarParams, echo_ptr: int, velocity_ptr: int | None, img_ptr: int, frames: int,
                      options: Options | None = None, phi0_ptr: int | None = None, stream: int | None = None) -> None:
    """isar_run_device without a context or any device allocation: the workspace is the library's
    per-GPU static array (static_workspace_bytes()); calls overlapping on one GPU race."""
    options = options or Options()
    L = lib()
    vp = C.c_void_p
    L.isar_run_device_static.argtypes = [C.POINTER(_Params), C.c_int, vp, vp, vp, vp, C.POINTER(_Options), vp]
    _check(L.isar_run_device_static(C.byref(params.c()), frames, vp(echo_ptr), vp(velocity_ptr or 0), vp(phi0_ptr or 0),
                                    vp(img_ptr), C.byref(options.c()), vp(stream or 0)), "isar_run_device_static")


def fft(x: np.ndarray, inverse: bool = False, grouped: bool = False) -> np.ndarray:
    """Unnormalised device FFT of each row of a 2-D complex array (tests and diagnostics). With
    grouped=True the rows are transformed several per workgroup, each within one wave
    (isar_fft_host_grouped, lengths 64 ... 1024)."""
    a = np.ascontiguousarray(x, dtype=np.complex64)
    if a.ndim == 1:
        a = a[None]
    out = np.zeros_like(a)
    name = "isar_fft_host_grouped" if grouped else "isar_fft_host"
    fn = getattr(lib(), name)
    fn.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
    _check(fn(a.shape[1], 1 if inverse else -1, a.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
              a.shape[0]), name)
    return out
