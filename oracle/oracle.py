"""ctypes binding of the FP64 oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, to
compare against or to time it. The product package never imports it.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "build" / "libisar_oracle.so"


class OrcParams(C.Structure):
    _fields_ = [("echo_num", C.c_int32), ("range_num", C.c_int32), ("range_num_cut", C.c_int32),
                ("reserved", C.c_int32), ("band_width", C.c_int64), ("fc", C.c_int64),
                ("Tp", C.c_double), ("Fs", C.c_int64)]


class OrcOptions(C.Structure):
    _fields_ = [("data_type", C.c_int32), ("if_hpc", C.c_int32), ("if_hrrp", C.c_int32),
                ("if_mtrc", C.c_int32), ("quirk_q1", C.c_int32), ("iterations", C.c_int32),
                ("phi0_given", C.c_int32), ("range_fft32", C.c_int32)]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class OrcTrace(C.Structure):
    _fields_ = [("n_pairs", C.c_int32), ("reserved0", C.c_int32), ("ra_imax", _ip), ("ra_lag", _dp),
                ("ra_gap", _dp), ("center_applied", C.c_int32), ("center_shift", C.c_int32),
                ("idx1_len", C.c_int32), ("idx2_len", C.c_int32), ("diff_imin", C.c_int32),
                ("cut_offset", C.c_int32), ("center_thr", C.c_double), ("arp", _dp), ("arp_mean", _dp),
                ("cut_key", _dp), ("tgt_num", C.c_int32), ("unit2_num", C.c_int32),
                ("pa_threshold", C.c_double), ("pa_colmax", _dp), ("pa_colE", _dp), ("pa_colH", _dp),
                ("select_bin", _ip), ("phi0", _dp), ("phi", _dp), ("entropy", C.c_double),
                ("reserved1", C.c_double)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(str(_LIB_PATH))
        L.orc_isar_main.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(OrcParams), C.POINTER(OrcOptions),
                                    C.c_void_p, C.POINTER(OrcTrace)]
        L.orc_isar_batch.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.POINTER(OrcParams),
                                     C.POINTER(OrcOptions), C.c_void_p]
        L.orc_fft.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_czt.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                              C.c_void_p]
        L.orc_czt_alpha.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_void_p]
        L.orc_hpc.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(OrcParams)]
        L.orc_range_alignment.argtypes = [C.c_void_p, C.POINTER(OrcParams), C.POINTER(OrcTrace)]
        L.orc_hrrp_center.argtypes = [C.c_void_p, C.POINTER(OrcParams), C.POINTER(OrcTrace)]
        L.orc_cut_range_profile.argtypes = [C.c_void_p, C.POINTER(OrcParams), C.c_void_p, C.POINTER(OrcTrace)]
        L.orc_doppler_tracking.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.orc_fast_entropy.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(OrcTrace)]
        L.orc_mtrc.argtypes = [C.c_void_p, C.POINTER(OrcParams), C.c_int]
        L.orc_next_pow2.argtypes = [C.c_int]
        L.orc_next_pow2.restype = C.c_int
        L.orc_image_entropy.argtypes = [C.c_void_p, C.c_int]
        L.orc_image_entropy.restype = C.c_double
        _lib = L
    return _lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def params(M, N, W=512, band_width=1_000_000_000, fc=10_000_000_000, Tp=None, Fs=64_000_000) -> OrcParams:
    if Tp is None:
        Tp = N / Fs
    return OrcParams(M, N, W, 0, int(band_width), int(fc), float(Tp), int(Fs))


def options(data_type=2, if_hpc=1, if_hrrp=0, if_mtrc=1, quirk_q1=1, iterations=120, phi0_given=0,
            range_fft32=0) -> OrcOptions:
    return OrcOptions(data_type, if_hpc, if_hrrp, if_mtrc, quirk_q1, iterations, phi0_given, range_fft32)


@dataclass
class Trace:
    """Numpy-backed trace of every integer decision and key vector of one oracle run."""
    M: int
    N: int
    W: int

    def __post_init__(self):
        M, N, W = self.M, self.N, self.W
        self.ra_imax = np.zeros(M, np.int32)
        self.ra_lag = np.zeros(M, np.float64)
        self.ra_gap = np.zeros(M, np.float64)
        self.arp = np.zeros(N, np.float64)
        self.arp_mean = np.zeros(N, np.float64)
        self.cut_key = np.zeros(N, np.float64)
        self.pa_colmax = np.zeros(W, np.float64)
        self.pa_colE = np.zeros(W, np.float64)
        self.pa_colH = np.zeros(W, np.float64)
        self.select_bin = np.zeros(max(W, 2), np.int32)
        self.phi0 = np.zeros(2 * M, np.float64)
        self.phi = np.zeros(2 * M, np.float64)
        self.c = OrcTrace()
        t = self.c
        t.ra_imax = self.ra_imax.ctypes.data_as(_ip)
        t.ra_lag = self.ra_lag.ctypes.data_as(_dp)
        t.ra_gap = self.ra_gap.ctypes.data_as(_dp)
        t.arp = self.arp.ctypes.data_as(_dp)
        t.arp_mean = self.arp_mean.ctypes.data_as(_dp)
        t.cut_key = self.cut_key.ctypes.data_as(_dp)
        t.pa_colmax = self.pa_colmax.ctypes.data_as(_dp)
        t.pa_colE = self.pa_colE.ctypes.data_as(_dp)
        t.pa_colH = self.pa_colH.ctypes.data_as(_dp)
        t.select_bin = self.select_bin.ctypes.data_as(_ip)
        t.phi0 = self.phi0.ctypes.data_as(_dp)
        t.phi = self.phi.ctypes.data_as(_dp)

    def __getattr__(self, name):
        # scalar fields live in the ctypes struct
        c = self.__dict__.get("c")
        if c is not None and name in dict(OrcTrace._fields_):
            return getattr(c, name)
        raise AttributeError(name)

    @property
    def n_pairs(self):
        return self.c.n_pairs

    @property
    def phi_c(self) -> np.ndarray:
        return self.phi[0::2] + 1j * self.phi[1::2]

    @property
    def phi0_c(self) -> np.ndarray:
        return self.phi0[0::2] + 1j * self.phi0[1::2]

    @property
    def selected(self) -> np.ndarray:
        return self.select_bin[: self.c.unit2_num].copy()


def isar_main(echo: np.ndarray, velocity: np.ndarray, p: OrcParams, opt: OrcOptions | None = None,
              trace: Trace | None = None) -> np.ndarray:
    """Run the oracle chain on one frame. echo: complex64 (M, N). Returns the (M, W) image."""
    opt = opt or options()
    M, N, W = p.echo_num, p.range_num, p.range_num_cut
    e = np.ascontiguousarray(echo, dtype=np.complex64)
    v = np.ascontiguousarray(velocity, dtype=np.float64)
    img = np.zeros((M, W), np.float32)
    rc = lib().orc_isar_main(_ptr(e), _ptr(v), C.byref(p), C.byref(opt), _ptr(img),
                             C.byref(trace.c) if trace is not None else None)
    if rc != 0:
        raise ValueError(f"oracle rejected the frame shape (rc={rc})")
    return img


STAGE_NAMES = ("hpc", "range_alignment", "centre_cut", "phase_compensation", "mtrc", "post")


def isar_stage_times(echo: np.ndarray, velocity: np.ndarray, p: OrcParams, opt: OrcOptions | None = None
                     ) -> tuple[np.ndarray, dict]:
    """One frame through the oracle with the wall time (ms) of each SEPARATE_TIMING_ stage
    (isar_main.cu:34-218); the stages use every OpenMP thread (intra-frame parallelism)."""
    opt = opt or options()
    M, W = p.echo_num, p.range_num_cut
    e = np.ascontiguousarray(echo, dtype=np.complex64)
    v = np.ascontiguousarray(velocity, dtype=np.float64)
    img = np.zeros((M, W), np.float32)
    ms = np.zeros(6, np.float64)
    L = lib()
    L.orc_isar_stage_times.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(OrcParams), C.POINTER(OrcOptions),
                                       C.c_void_p, C.c_void_p, C.c_void_p]
    if L.orc_isar_stage_times(_ptr(e), _ptr(v), C.byref(p), C.byref(opt), _ptr(img), None, _ptr(ms)) != 0:
        raise ValueError("oracle rejected the frame shape")
    return img, dict(zip(STAGE_NAMES, ms.tolist()))


def isar_batch(echo: np.ndarray, velocity: np.ndarray, p: OrcParams, opt: OrcOptions | None = None) -> tuple[np.ndarray, int]:
    opt = opt or options()
    K = echo.shape[0]
    e = np.ascontiguousarray(echo, dtype=np.complex64)
    v = np.ascontiguousarray(velocity, dtype=np.float64)
    img = np.zeros((K, p.echo_num, p.range_num_cut), np.float32)
    threads = lib().orc_isar_batch(K, _ptr(e), _ptr(v), C.byref(p), C.byref(opt), _ptr(img))
    return img, threads


def fft(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft(_ptr(a), a.size, 1 if inverse else -1)
    return a


def czt(x: np.ndarray, k: int, w: complex, a: complex) -> np.ndarray:
    xa = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros(k, np.complex128)
    lib().orc_czt(_ptr(xa), xa.size, k, w.real, w.imag, a.real, a.imag, _ptr(out))
    return out


def czt_alpha(x: np.ndarray, alpha: float) -> np.ndarray:
    xa = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros(xa.size, np.complex128)
    lib().orc_czt_alpha(_ptr(xa), xa.size, float(alpha), _ptr(out))
    return out


def hpc(s: np.ndarray, velocity: np.ndarray, p: OrcParams) -> np.ndarray:
    a = np.ascontiguousarray(s, dtype=np.complex128).copy()
    v = np.ascontiguousarray(velocity, dtype=np.float64)
    lib().orc_hpc(_ptr(a), _ptr(v), C.byref(p))
    return a


def range_alignment(s: np.ndarray, p: OrcParams, trace: Trace | None = None) -> np.ndarray:
    a = np.ascontiguousarray(s, dtype=np.complex128).copy()
    lib().orc_range_alignment(_ptr(a), C.byref(p), C.byref(trace.c) if trace is not None else None)
    return a


def hrrp_center(s: np.ndarray, p: OrcParams, trace: Trace | None = None) -> np.ndarray:
    a = np.ascontiguousarray(s, dtype=np.complex128).copy()
    lib().orc_hrrp_center(_ptr(a), C.byref(p), C.byref(trace.c) if trace is not None else None)
    return a


def cut_range_profile(s: np.ndarray, p: OrcParams, trace: Trace | None = None) -> np.ndarray:
    a = np.ascontiguousarray(s, dtype=np.complex128)
    out = np.zeros((p.echo_num, p.range_num_cut), np.complex128)
    lib().orc_cut_range_profile(_ptr(a), C.byref(p), _ptr(out), C.byref(trace.c) if trace is not None else None)
    return out


def doppler_tracking(x: np.ndarray, compensate: bool = True) -> tuple[np.ndarray, np.ndarray]:
    a = np.ascontiguousarray(x, dtype=np.complex128)
    rows, cols = a.shape
    out = np.zeros_like(a)
    ph = np.zeros(rows, np.complex128)
    lib().orc_doppler_tracking(_ptr(out), _ptr(ph), _ptr(a), rows, cols, 1 if compensate else 0)
    return out, ph


def fast_entropy(x: np.ndarray, iterations: int = 120, trace: Trace | None = None,
                 phi0: np.ndarray | None = None) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    rows, cols = a.shape
    given = 0
    if phi0 is not None:
        if trace is None:
            trace = Trace(rows, 2, cols)
        trace.phi0[0::2] = np.real(phi0)
        trace.phi0[1::2] = np.imag(phi0)
        given = 1
    lib().orc_fast_entropy(_ptr(a), rows, cols, iterations, given, C.byref(trace.c) if trace is not None else None)
    return a


def mtrc(x: np.ndarray, p: OrcParams, data_type: int = 2) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_mtrc(_ptr(a), C.byref(p), data_type)
    return a


def post_process(x: np.ndarray, quirk_q1: bool = True, data_type: int = 2) -> np.ndarray:
    """isar_main.cu:191-211 on an (M, W) complex array: window, slow-time FFT, shift, |.|."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    M, W = a.shape
    img = np.zeros((M, W), np.float32)
    L = lib()
    L.orc_post_process.restype = None
    L.orc_post_process.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.orc_post_process(_ptr(a), M, W, int(quirk_q1), int(data_type), _ptr(img))
    return img


def next_pow2(n: int) -> int:
    return lib().orc_next_pow2(n)


def pulse_compress(iq: np.ndarray, velocity: float, p: OrcParams, cut: int = 4096) -> tuple[np.ndarray, int]:
    """IFDS pulse compression of one echo (complex64 samples) -> (cut complex128 samples, offset)."""
    a = np.ascontiguousarray(iq, dtype=np.complex64)
    out = np.zeros(cut, np.complex128)
    L = lib()
    L.orc_pulse_compress.restype = C.c_int
    L.orc_pulse_compress.argtypes = [C.c_void_p, C.c_int, C.c_double, C.POINTER(OrcParams), C.c_int, C.c_void_p]
    off = L.orc_pulse_compress(_ptr(a), a.size, float(velocity), C.byref(p), cut, _ptr(out))
    if off < 0:
        raise ValueError("pulse compression: unsupported sizes")
    return out, off


def hrrp(echo: np.ndarray, velocity: np.ndarray, p: OrcParams, opt: OrcOptions | None = None) -> np.ndarray:
    """getHRRP of one frame: complex128 (M, N)."""
    opt = opt or options()
    e = np.ascontiguousarray(echo, dtype=np.complex64)
    v = np.ascontiguousarray(velocity, dtype=np.float64)
    out = np.zeros((p.echo_num, p.range_num), np.complex128)
    L = lib()
    L.orc_hrrp.restype = C.c_int
    L.orc_hrrp.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(OrcParams), C.POINTER(OrcOptions), C.c_void_p]
    if L.orc_hrrp(_ptr(e), _ptr(v), C.byref(p), C.byref(opt), _ptr(out)) != 0:
        raise ValueError("hrrp: unsupported sizes or data type")
    return out


def image_entropy(img: np.ndarray) -> float:
    a = np.ascontiguousarray(img, dtype=np.float32)
    return lib().orc_image_entropy(_ptr(a), a.size)
