"""Python mirror of librawspec's C API (include/rawspec.h, include/rawspec_amd.h).

The product is the C-ABI shared library ``rawspec_amd/librawspec.so`` (C host
layer + gfx950 HIP kernels).  This module binds it with ctypes so tests,
benchmarks and Python clients drive exactly the entry points a C client
would: ``RawspecContext`` is field-for-field ``struct rawspec_context_s``
(reference rawspec.h:32-148) and every ``rawspec_*`` function keeps the
reference's name, arguments, return codes and error behaviour.

Loading fails loudly when the library has not been built; there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = [
    "MAX_OUTPUTS", "RAWSPEC_FORWARD_FFT", "RAWSPEC_INVERSE_FFT",
    "RAWSPEC_CALLBACK_PRE_DUMP", "RAWSPEC_CALLBACK_POST_DUMP",
    "RawspecContext", "DumpCallback", "lib", "lib_path", "Spectrometer",
    "RawHdr", "FbHdr", "host_format_prototypes",
]

MAX_OUTPUTS = 4
RAWSPEC_FORWARD_FFT = +1
RAWSPEC_INVERSE_FFT = -1
RAWSPEC_CALLBACK_PRE_DUMP = 0
RAWSPEC_CALLBACK_POST_DUMP = 1

_HERE = os.path.dirname(os.path.abspath(__file__))
# RAWSPEC_AMD_LIB names another build of the library (e.g. a sanitizer build)
_LIB_PATH = os.environ.get("RAWSPEC_AMD_LIB") or os.path.join(_HERE, "librawspec.so")


class RawspecContext(C.Structure):
    pass


# void (*)(rawspec_context *ctx, int output_product, int callback_type)
DumpCallback = C.CFUNCTYPE(None, C.POINTER(RawspecContext), C.c_int, C.c_int)

RawspecContext._fields_ = [
    ("No", C.c_uint),
    ("Np", C.c_uint),
    ("Nant", C.c_uint),
    ("Nc", C.c_uint),
    ("Ntpb", C.c_uint),
    ("exit_soon", C.c_uint),
    ("Nbc", C.c_uint),
    ("Nbps", C.c_uint),
    ("Npolout", C.c_int * MAX_OUTPUTS),
    ("Nts", C.c_uint * MAX_OUTPUTS),
    ("Nas", C.c_uint * MAX_OUTPUTS),
    ("dump_callback", DumpCallback),
    ("user_data", C.c_void_p),
    ("Nb", C.c_uint),
    ("Nb_host", C.c_uint),
    ("h_blkbufs", C.POINTER(C.c_void_p)),   # char **
    ("gpu_index", C.c_int),
    ("input_conjugated", C.c_int),
    ("incoherently_sum", C.c_int),
    ("Naws", C.c_int),
    ("Aws", C.POINTER(C.c_float)),
    ("h_pwrbuf", C.POINTER(C.c_float) * MAX_OUTPUTS),
    ("h_pwrbuf_size", C.c_size_t * MAX_OUTPUTS),
    ("h_icsbuf", C.POINTER(C.c_float) * MAX_OUTPUTS),
    ("Nds", C.c_uint * MAX_OUTPUTS),
    ("Ntmax", C.c_uint),
    ("gpu_ctx", C.c_void_p),
]


class RawHdr(C.Structure):
    """rawspec_raw_hdr_t (include/rawspec_rawutils.h)."""
    _fields_ = [
        ("directio", C.c_int), ("blocsize", C.c_size_t), ("npol", C.c_uint),
        ("obsnchan", C.c_uint), ("nbits", C.c_uint), ("pktidx", C.c_int64),
        ("obsfreq", C.c_double), ("obsbw", C.c_double), ("tbin", C.c_double),
        ("ra", C.c_double), ("dec", C.c_double), ("mjd", C.c_double),
        ("beam_id", C.c_int), ("nbeam", C.c_int), ("refbeam", C.c_int), ("nants", C.c_uint),
        ("src_name", C.c_char * 81), ("telescop", C.c_char * 81),
        ("hdr_pos", C.c_long), ("hdr_size", C.c_size_t),
    ]


class FbHdr(C.Structure):
    """fb_hdr_t (include/rawspec_fbutils.h)."""
    _fields_ = [
        ("machine_id", C.c_int), ("telescope_id", C.c_int), ("data_type", C.c_int),
        ("barycentric", C.c_int), ("pulsarcentric", C.c_int),
        ("src_raj", C.c_double), ("src_dej", C.c_double), ("az_start", C.c_double),
        ("za_start", C.c_double), ("fch1", C.c_double), ("foff", C.c_double),
        ("nchans", C.c_int), ("nbeams", C.c_int), ("ibeam", C.c_int), ("nbits", C.c_int),
        ("tstart", C.c_double), ("tsamp", C.c_double), ("nifs", C.c_int),
        ("source_name", C.c_char * 81), ("rawdatafile", C.c_char * 81),
        ("nfpc", C.c_int), ("refbeam", C.c_int),
    ]


def host_format_prototypes(L, only=None):
    """Sets ctypes prototypes of the GUPPI RAW (rawspec_raw_*) and SIGPROC
    filterbank (fb_*) header functions on a loaded library; `only`, a name
    prefix, restricts them to one family (for a library exporting just it)."""
    cp, vp, i32p, dp = C.c_char_p, C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double)
    RH, FH = C.POINTER(RawHdr), C.POINTER(FbHdr)
    for name, args, res in [
        ("rawspec_raw_get_s32", [cp, cp, C.c_int32], C.c_int32),
        ("rawspec_raw_get_u32", [cp, cp, C.c_uint32], C.c_uint32),
        ("rawspec_raw_get_s64", [cp, cp, C.c_int64], C.c_int64),
        ("rawspec_raw_get_u64", [cp, cp, C.c_uint64], C.c_uint64),
        ("rawspec_raw_get_dbl", [cp, cp, C.c_double], C.c_double),
        ("rawspec_raw_get_str", [cp, cp, cp, vp, C.c_size_t], None),
        ("rawspec_raw_dmsstr_to_d", [vp], C.c_double),
        ("rawspec_raw_header_size", [vp, C.c_size_t, C.c_int], C.c_int),
        ("rawspec_raw_parse_header", [cp, RH], None),
        ("rawspec_raw_read_header", [C.c_int, RH], C.c_long),
        ("fb_ddd_to_dms", [C.c_double], C.c_double),
        ("fb_dms_to_ddd", [C.c_double], C.c_double),
        ("fb_fd_write_int", [C.c_int, C.c_int32], C.c_ssize_t),
        ("fb_buf_write_int", [vp, C.c_int32], vp),
        ("fb_fd_write_double", [C.c_int, C.c_double], C.c_ssize_t),
        ("fb_buf_write_double", [vp, C.c_double], vp),
        ("fb_fd_write_angle", [C.c_int, C.c_double], C.c_ssize_t),
        ("fb_buf_write_angle", [vp, C.c_double], vp),
        ("fb_fd_write_string", [C.c_int, cp], C.c_ssize_t),
        ("fb_buf_write_string", [vp, cp], vp),
        ("fb_fd_read_int", [C.c_int, i32p], C.c_int32),
        ("fb_buf_read_int", [vp, i32p], vp),
        ("fb_fd_read_double", [C.c_int, dp], C.c_double),
        ("fb_buf_read_double", [vp, dp], vp),
        ("fb_fd_read_angle", [C.c_int, dp], C.c_double),
        ("fb_buf_read_angle", [vp, dp], vp),
        ("fb_fd_read_string", [C.c_int, vp, i32p], C.c_ssize_t),
        ("fb_buf_peek_string", [vp, C.POINTER(C.c_void_p), i32p], vp),
        ("fb_buf_read_string", [vp, vp, i32p], vp),
        ("fb_fd_write_padded_header", [C.c_int, FH, C.c_int32], C.c_ssize_t),
        ("fb_fd_write_header", [C.c_int, FH], C.c_ssize_t),
        ("fb_buf_write_padded_header", [vp, FH, C.c_int32], vp),
        ("fb_buf_write_header", [vp, FH], vp),
        ("fb_fd_read_header", [C.c_int, FH, C.POINTER(C.c_size_t)], C.c_ssize_t),
        ("fb_buf_read_header", [vp, FH, C.POINTER(C.c_size_t)], vp),
        ("fb_telescope_id", [cp], C.c_int),
    ]:
        if only is not None and not name.startswith(only):
            continue
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    return L


_lib = None


def lib_path() -> str:
    return _LIB_PATH


def lib():
    """Loads librawspec.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                f"{_LIB_PATH} is missing: build it with `python -m rawspec_amd.build` "
                "(there is no fallback implementation)")
        # One HIP runtime per process.  torch's libraries resolve the
        # unversioned libamdhip64.so from torch's own directory, while this
        # library binds libamdhip64.so.7 by soname.  Loaded after torch, this
        # library shares torch's runtime; loaded first, it would make torch
        # load a second runtime, which then finds no GPU.  So torch, when it
        # is installed, is loaded first.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER(RawspecContext)
        L.get_librawspec_version.restype = C.c_char_p
        L.get_librawspec_version.argtypes = []
        L.get_cufft_version.restype = C.c_char_p
        L.get_cufft_version.argtypes = []
        for name, args, res in [
            ("rawspec_initialize", [P], C.c_int),
            ("rawspec_cleanup", [P], None),
            ("rawspec_copy_blocks_to_gpu", [P, C.c_long, C.c_long, C.c_size_t], C.c_int),
            ("rawspec_copy_blocks_to_gpu_expanding_complex4",
             [P, C.c_long, C.c_long, C.c_size_t], C.c_int),
            ("rawspec_zero_blocks_to_gpu", [P, C.c_long, C.c_size_t], C.c_int),
            ("rawspec_start_processing", [P, C.c_int], C.c_int),
            ("rawspec_copy_blocks_to_gpu_and_start_processing",
             [P, C.c_size_t, C.c_char, C.c_int], C.c_int),
            ("rawspec_reset_integration", [P], C.c_int),
            ("rawspec_check_for_completion", [P], C.c_uint),
            ("rawspec_wait_for_completion", [P], C.c_int),
            ("rawspec_amd_plan", [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], C.c_int),
            ("rawspec_amd_device_input", [P], C.c_void_p),
            ("rawspec_amd_device_input_size", [P], C.c_size_t),
            ("rawspec_amd_start_processing_device", [P, C.c_void_p, C.c_int], C.c_int),
            ("rawspec_amd_start_processing_device_noalloc", [P, C.c_void_p, C.c_int], C.c_int),
            ("rawspec_amd_start_processing_device_batch",
             [P, C.POINTER(C.c_void_p), C.c_uint, C.c_int], C.c_int),
            ("rawspec_amd_set_device_output", [P, C.c_int], C.c_int),
            ("rawspec_amd_device_pwrbuf", [P, C.c_int], C.c_void_p),
            ("rawspec_amd_device_icsbuf", [P, C.c_int], C.c_void_p),
            ("rawspec_amd_stream", [P], C.c_void_p),
            ("rawspec_amd_set_stream", [P, C.c_void_p], C.c_int),
            ("rawspec_amd_profile", [P, C.c_int], C.c_int),
            ("rawspec_amd_set_mfma", [P, C.c_int], C.c_int),
            ("rawspec_amd_get_mfma", [P], C.c_int),
            ("rawspec_amd_set_held_buffers", [P, C.c_uint], C.c_int),
            ("rawspec_amd_autotune", [P, C.c_void_p, C.c_uint, C.c_uint, C.c_char_p, C.c_size_t],
             C.c_int),
            ("rawspec_amd_get_held_buffers", [P], C.c_int),
            ("rawspec_amd_set_lanes", [P, C.c_int], C.c_int),
            ("rawspec_amd_get_lanes", [P], C.c_int),
            ("rawspec_amd_profile_report", [P, C.c_char_p, C.c_size_t], C.c_int),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        host_format_prototypes(L)
        _lib = L
    return _lib


class Spectrometer:
    """Thin convenience owner of one ``rawspec_context``.

    Fields are set exactly as a C client would set them; methods call the
    C entry points one to one.  Host block buffers, power buffers and the
    ICS buffers are exposed as numpy views of the library's pinned memory.
    """

    def __init__(self, *, Nc, Ntpb, Nts, Nas, Np=2, Nbps=8, Npolout=None,
                 Nant=1, Nbc=0, Nb=0, Nb_host=0, gpu_index=0,
                 input_conjugated=0, incoherently_sum=0, Aws=None,
                 callback=None, record_events=True, init=True):
        self.ctx = RawspecContext()
        c = self.ctx
        No = len(Nts)
        c.No, c.Np, c.Nant, c.Nc, c.Ntpb = No, Np, Nant, Nc, Ntpb
        c.Nbc, c.Nbps, c.Nb, c.Nb_host = Nbc, Nbps, Nb, Nb_host
        Npolout = list(Npolout) if Npolout is not None else [1] * No
        for i in range(No):
            c.Nts[i] = Nts[i]
            c.Nas[i] = Nas[i]
            c.Npolout[i] = Npolout[i]
        c.gpu_index = gpu_index
        c.input_conjugated = input_conjugated
        c.incoherently_sum = incoherently_sum
        self._aws = None
        if Aws is not None:
            self._aws = (C.c_float * len(Aws))(*Aws)
            c.Aws = C.cast(self._aws, C.POINTER(C.c_float))
            c.Naws = len(Aws)
        self.events = []   # (product, type) in callback order
        self._user_cb = callback
        self._cb = DumpCallback(self._on_dump)
        if record_events or callback is not None:
            c.dump_callback = self._cb
        self.initialized = False
        if init:
            rc = self.initialize()
            if rc != 0:
                raise RuntimeError(f"rawspec_initialize failed ({rc})")

    def _on_dump(self, ctxp, product, cbtype):
        self.events.append((product, cbtype))
        if self._user_cb is not None:
            self._user_cb(self, product, cbtype)

    # -- C entry points --------------------------------------------------
    def initialize(self) -> int:
        rc = lib().rawspec_initialize(C.byref(self.ctx))
        self.initialized = rc == 0
        return rc

    def cleanup(self):
        if self.initialized:
            lib().rawspec_cleanup(C.byref(self.ctx))
            self.initialized = False

    def copy_blocks(self, src_idx=0, dst_idx=0, num_blocks=None) -> int:
        n = self.ctx.Nb if num_blocks is None else num_blocks
        return lib().rawspec_copy_blocks_to_gpu(C.byref(self.ctx), src_idx, dst_idx, n)

    def copy_blocks_expanding_complex4(self, src_idx=0, dst_idx=0, num_blocks=None) -> int:
        n = self.ctx.Nb if num_blocks is None else num_blocks
        return lib().rawspec_copy_blocks_to_gpu_expanding_complex4(
            C.byref(self.ctx), src_idx, dst_idx, n)

    def zero_blocks(self, dst_idx=0, num_blocks=1) -> int:
        return lib().rawspec_zero_blocks_to_gpu(C.byref(self.ctx), dst_idx, num_blocks)

    def start_processing(self, fft_dir=RAWSPEC_FORWARD_FFT) -> int:
        return lib().rawspec_start_processing(C.byref(self.ctx), fft_dir)

    def copy_and_start(self, num_blocks=None, expand4=False, fft_dir=RAWSPEC_FORWARD_FFT) -> int:
        n = self.ctx.Nb if num_blocks is None else num_blocks
        return lib().rawspec_copy_blocks_to_gpu_and_start_processing(
            C.byref(self.ctx), n, b"\x01" if expand4 else b"\x00", fft_dir)

    def start_processing_device(self, dev_ptr: int, fft_dir=RAWSPEC_FORWARD_FFT) -> int:
        return lib().rawspec_amd_start_processing_device(C.byref(self.ctx), C.c_void_p(dev_ptr),
                                                         fft_dir)

    def start_processing_device_noalloc(self, dev_ptr: int, fft_dir=RAWSPEC_FORWARD_FFT) -> int:
        return lib().rawspec_amd_start_processing_device_noalloc(C.byref(self.ctx),
                                                                 C.c_void_p(dev_ptr), fft_dir)

    def start_processing_device_batch(self, dev_ptrs, fft_dir=RAWSPEC_FORWARD_FFT) -> int:
        arr = (C.c_void_p * len(dev_ptrs))(*dev_ptrs)
        return lib().rawspec_amd_start_processing_device_batch(C.byref(self.ctx), arr,
                                                               len(dev_ptrs), fft_dir)

    def set_device_output(self, enable: bool) -> int:
        return lib().rawspec_amd_set_device_output(C.byref(self.ctx), 1 if enable else 0)

    def device_input(self) -> int:
        return lib().rawspec_amd_device_input(C.byref(self.ctx)) or 0

    def device_input_size(self) -> int:
        return lib().rawspec_amd_device_input_size(C.byref(self.ctx))

    def device_pwrbuf(self, i) -> int:
        return lib().rawspec_amd_device_pwrbuf(C.byref(self.ctx), i) or 0

    def stream(self) -> int:
        return lib().rawspec_amd_stream(C.byref(self.ctx)) or 0

    def set_stream(self, stream: int) -> int:
        """Queue this context's work on the hipStream_t `stream` (0: its own)."""
        return lib().rawspec_amd_set_stream(C.byref(self.ctx), C.c_void_p(stream or None))

    def set_mfma(self, level) -> int:
        """0 / False: vector units; 1 / True: matrix cores at the 1e-3 bar
        (unsplit operands for long int8 integrations, the vector units
        where they are faster); 2: matrix cores with
        split operands everywhere."""
        return lib().rawspec_amd_set_mfma(C.byref(self.ctx), int(level))

    def get_mfma(self) -> int:
        return lib().rawspec_amd_get_mfma(C.byref(self.ctx))

    def autotune(self, dev_ptr: int | None = None, per_call: int = 1, rounds: int = 3) -> str:
        """Runs rawspec_amd_autotune; returns its report (raises on failure)."""
        buf = C.create_string_buffer(4096)
        if lib().rawspec_amd_autotune(C.byref(self.ctx), dev_ptr, per_call, rounds, buf, len(buf)):
            raise RuntimeError("rawspec_amd_autotune failed")
        return buf.value.decode()

    def set_held_buffers(self, max_buffers: int) -> int:
        return lib().rawspec_amd_set_held_buffers(C.byref(self.ctx), max_buffers)

    def get_held_buffers(self) -> int:
        return lib().rawspec_amd_get_held_buffers(C.byref(self.ctx))

    def set_lanes(self, enable: bool) -> int:
        return lib().rawspec_amd_set_lanes(C.byref(self.ctx), 1 if enable else 0)

    def get_lanes(self) -> int:
        return lib().rawspec_amd_get_lanes(C.byref(self.ctx))

    def profile(self, enable: bool) -> int:
        return lib().rawspec_amd_profile(C.byref(self.ctx), 1 if enable else 0)

    def profile_report(self) -> dict:
        """{kernel: (launches, total_ms)} since the last report."""
        buf = C.create_string_buffer(1 << 16)
        n = lib().rawspec_amd_profile_report(C.byref(self.ctx), buf, len(buf))
        if n < 0:
            raise RuntimeError("rawspec_amd_profile_report failed")
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            out[name] = (int(cnt), float(ms))
        return out

    def reset_integration(self) -> int:
        return lib().rawspec_reset_integration(C.byref(self.ctx))

    def check_for_completion(self) -> int:
        return lib().rawspec_check_for_completion(C.byref(self.ctx))

    def wait_for_completion(self) -> int:
        return lib().rawspec_wait_for_completion(C.byref(self.ctx))

    # -- numpy views -----------------------------------------------------
    @property
    def blocsize(self) -> int:
        c = self.ctx
        return c.Nc * c.Ntpb * c.Np * 2 * c.Nbps // 8

    def block(self, i) -> np.ndarray:
        """Writable uint8 view of host block buffer i."""
        ptr = self.ctx.h_blkbufs[i]
        if not ptr:
            raise RuntimeError(f"host block {i} is not allocated")
        return np.ctypeslib.as_array((C.c_uint8 * self.blocsize).from_address(ptr))

    def pwrbuf(self, i) -> np.ndarray:
        n = self.ctx.h_pwrbuf_size[i] // 4
        ptr = C.cast(self.ctx.h_pwrbuf[i], C.c_void_p).value
        return np.ctypeslib.as_array((C.c_float * n).from_address(ptr))

    def icsbuf(self, i) -> np.ndarray:
        n = self.ctx.h_pwrbuf_size[i] // 4 // self.ctx.Nant
        ptr = C.cast(self.ctx.h_icsbuf[i], C.c_void_p).value
        return np.ctypeslib.as_array((C.c_float * n).from_address(ptr))

    def close(self):
        self.cleanup()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
