"""ctypes wrapper of the FP64 C oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, by
``__graft_entry__.smoke()`` and by the CPU-baseline leg of ``bench.py``,
always as the checker.  The product library (``rawspec_amd``) never imports
it.  Reference lines restated by each C function are cited in
``oracle/rawspec_oracle.c``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
# the same code built in single precision: the FP32 CPU port the CPU
# baseline times beside the FP64 oracle; never the checker
_LIB32_PATH = os.path.join(_HERE, "liboracle32.so")
MAX_OUTPUTS = 4

# Rule codes returned by ro_derive (the reference returns 1 for all).
ERRORS = {
    0: "ok", 1: "No", 2: "Np", 3: "Ntpb", 4: "Nbps", 5: "Nt0", 6: "NtDiv",
    7: "Nb", 8: "NtpbDiv", 9: "NtmaxDiv", 10: "NbHost", 11: "Na0",
    12: "NaDiv", 13: "Nbc", 14: "Naws",
}


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not (os.path.exists(_LIB_PATH) and os.path.exists(_LIB32_PATH)):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


class Cfg(C.Structure):
    _fields_ = [
        ("No", C.c_uint), ("Np", C.c_uint), ("Nant", C.c_uint),
        ("Nc", C.c_uint), ("Ntpb", C.c_uint), ("Nbps", C.c_uint),
        ("Nbc", C.c_uint), ("Nb", C.c_uint), ("Nb_host", C.c_uint),
        ("Npolout", C.c_int * MAX_OUTPUTS),
        ("Nts", C.c_uint * MAX_OUTPUTS),
        ("Nas", C.c_uint * MAX_OUTPUTS),
        ("input_conjugated", C.c_int),
        ("incoherently_sum", C.c_int),
        ("Naws", C.c_int),
        ("have_blkbufs", C.c_int),
        ("expand4", C.c_int),
        ("Ntmax", C.c_uint),
        ("Nss", C.c_uint * MAX_OUTPUTS),
        ("Nds", C.c_uint * MAX_OUTPUTS),
        ("Nis", C.c_uint * MAX_OUTPUTS),
        ("pwr_len", C.c_size_t * MAX_OUTPUTS),
        ("ics_len", C.c_size_t * MAX_OUTPUTS),
    ]


_libs = {}


def lib(fp32: bool = False):
    """The FP64 oracle library, or with fp32 the single-precision port."""
    if fp32 not in _libs:
        build()
        L = C.CDLL(_LIB32_PATH if fp32 else _LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.ro_derive.argtypes = [C.POINTER(Cfg)]
        L.ro_derive.restype = C.c_int
        L.ro_snorm8.argtypes = [C.c_int8]
        L.ro_snorm8.restype = C.c_double
        L.ro_snorm16.argtypes = [C.c_int16]
        L.ro_snorm16.restype = C.c_double
        L.ro_expand4.argtypes = [C.c_uint8, C.POINTER(C.c_int8), C.POINTER(C.c_int8)]
        L.ro_fft.argtypes = [dp, C.c_size_t, C.c_int]
        L.ro_dft_naive.argtypes = [dp, dp, C.c_size_t, C.c_int]
        L.ro_create.argtypes = [C.POINTER(Cfg), C.POINTER(C.c_float)]
        L.ro_create.restype = C.c_void_p
        L.ro_destroy.argtypes = [C.c_void_p]
        L.ro_reset.argtypes = [C.c_void_p]
        L.ro_set_channels.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.c_int]
        L.ro_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.ro_process.argtypes = [C.c_void_p, C.c_void_p, C.c_int,
                                 C.POINTER(dp), C.POINTER(dp)]
        L.ro_process.restype = C.c_int
        L.ro_dump_events.argtypes = [C.POINTER(Cfg), C.c_int, C.c_ulong]
        L.ro_dump_events.restype = C.c_ulong
        L.ro_splitmix64.argtypes = [C.c_uint64]
        L.ro_splitmix64.restype = C.c_uint64
        L.ro_gen_block.argtypes = [C.c_void_p, C.c_uint, C.c_uint, C.c_uint,
                                   C.c_uint, C.c_uint64, C.c_double, C.c_int,
                                   C.c_ulong, C.c_int, C.c_double, C.c_double,
                                   C.c_double]
        _libs[fp32] = L
    return _libs[fp32]


def make_cfg(No=1, Np=2, Nant=1, Nc=1, Ntpb=1024, Nbps=8, Nbc=0, Nb=0,
             Nb_host=0, Npolout=(1,), Nts=(1024,), Nas=(1,),
             input_conjugated=0, incoherently_sum=0, Naws=0,
             have_blkbufs=0) -> Cfg:
    c = Cfg()
    c.No, c.Np, c.Nant, c.Nc, c.Ntpb, c.Nbps = No, Np, Nant, Nc, Ntpb, Nbps
    c.Nbc, c.Nb, c.Nb_host = Nbc, Nb, Nb_host
    for i in range(MAX_OUTPUTS):
        c.Npolout[i] = Npolout[i] if i < len(Npolout) else 0
        c.Nts[i] = Nts[i] if i < len(Nts) else 0
        c.Nas[i] = Nas[i] if i < len(Nas) else 0
    c.input_conjugated = input_conjugated
    c.incoherently_sum = incoherently_sum
    c.Naws = Naws
    c.have_blkbufs = have_blkbufs
    return c


def derive(cfg: Cfg) -> int:
    return lib().ro_derive(C.byref(cfg))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def fft(x: np.ndarray, sign: int = -1) -> np.ndarray:
    """Unnormalised DFT with kernel exp(sign*2*pi*i*k*m/n), FP64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).copy()
    lib().ro_fft(_dptr(a.view(np.float64)), a.size, sign)
    return a


def dft_naive(x: np.ndarray, sign: int = -1) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    out = np.empty_like(a)
    lib().ro_dft_naive(_dptr(a.view(np.float64)), _dptr(out.view(np.float64)),
                       a.size, sign)
    return out


def snorm8(v: int) -> float:
    return lib().ro_snorm8(v)


def snorm16(v: int) -> float:
    return lib().ro_snorm16(v)


def expand4(b: int):
    re, im = C.c_int8(), C.c_int8()
    lib().ro_expand4(b, C.byref(re), C.byref(im))
    return re.value, im.value


def splitmix64(x: int) -> int:
    return lib().ro_splitmix64(x)


def gen_block(Nc, Ntpb, Np, Nbps, seed, sigma, clip, blk, tone_chan=-1,
              tone_amp=0.0, tone_bin=0.0, tone_nfft=1.0) -> np.ndarray:
    """One host block [Nc][Ntpb][Np][re,im] of int8 or int16 (flat array)."""
    dt = np.int16 if Nbps == 16 else np.int8
    out = np.empty(Nc * Ntpb * Np * 2, dtype=dt)
    lib().ro_gen_block(out.ctypes.data, Nc, Ntpb, Np, Nbps, seed, sigma, clip,
                       blk, tone_chan, tone_amp, tone_bin, tone_nfft)
    return out


def device_layout(blocks, Nc, Ntpb, Np) -> np.ndarray:
    """Concatenates host blocks into the device buffer layout
    [Nc][Nb*Ntpb][Np][re,im] that rawspec_copy_blocks_to_gpu builds."""
    per = [b.reshape(Nc, Ntpb * Np * 2) for b in blocks]
    return np.ascontiguousarray(np.concatenate(per, axis=1)).reshape(-1)


class State:
    """Streaming FP64 oracle for one configuration (fp32: the single-
    precision port of the same code, for CPU-baseline timing only)."""

    def __init__(self, cfg: Cfg, weights=None, threads: int = 0, fp32: bool = False):
        self.cfg = cfg
        self._L = lib(fp32)
        self._w = None
        wp = None
        if weights is not None:
            self._w = np.ascontiguousarray(np.asarray(weights, dtype=np.float32))
            wp = self._w.ctypes.data_as(C.POINTER(C.c_float))
        self._h = self._L.ro_create(C.byref(cfg), wp)
        if threads:
            self._L.ro_set_threads(self._h, threads)
        self.pwr = [np.zeros(cfg.pwr_len[i]) for i in range(cfg.No)]
        self.ics = [np.zeros(max(cfg.ics_len[i], 1)) for i in range(cfg.No)]

    def set_channels(self, chans):
        arr = (C.c_int * len(chans))(*chans)
        self._L.ro_set_channels(self._h, arr, len(chans))

    def reset(self):
        self._L.ro_reset(self._h)

    def process(self, buf: np.ndarray, fft_dir: int = 1):
        """Returns (mask, {i: pwr}, {i: ics}) for products dumped this buffer."""
        dp = C.POINTER(C.c_double)
        P = (dp * MAX_OUTPUTS)()
        Q = (dp * MAX_OUTPUTS)()
        for i in range(self.cfg.No):
            P[i] = _dptr(self.pwr[i])
            Q[i] = _dptr(self.ics[i])
        buf = np.ascontiguousarray(buf)
        mask = self._L.ro_process(self._h, buf.ctypes.data, fft_dir, P, Q)
        if mask < 0:
            raise RuntimeError("ro_process failed")
        outs = {i: self.pwr[i].copy() for i in range(self.cfg.No) if mask >> i & 1}
        ics = {i: self.ics[i][: self.cfg.ics_len[i]].copy()
               for i in range(self.cfg.No) if mask >> i & 1}
        return mask, outs, ics

    def close(self):
        if self._h:
            self._L.ro_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
