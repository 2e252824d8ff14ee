"""ctypes binding of libmarvl_align.so (declarations: include/marvl_align.h)."""

from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmarvl_align.so")


class LibraryMissing(RuntimeError):
    pass


class MiAlignParams(C.Structure):
    _fields_ = [("ave_corr", C.c_double), ("trace_space", C.c_int), ("freq", C.c_float * 4)]


class MiLaTask(C.Structure):
    _fields_ = [("aword", C.c_longlong), ("bword", C.c_longlong), ("alen", C.c_int), ("blen", C.c_int),
                ("comp", C.c_int), ("low", C.c_int), ("hgh", C.c_int), ("anti", C.c_int),
                ("lbord", C.c_int), ("hbord", C.c_int), ("selfie", C.c_int)]


class MiLaResult(C.Structure):
    _fields_ = [("status", C.c_int), ("abpos", C.c_int), ("bbpos", C.c_int), ("aepos", C.c_int),
                ("bepos", C.c_int), ("diffs", C.c_int), ("b_abpos", C.c_int), ("b_bbpos", C.c_int),
                ("b_aepos", C.c_int), ("b_bepos", C.c_int), ("atlen", C.c_int), ("btlen", C.c_int),
                ("atrace", C.c_longlong), ("btrace", C.c_longlong), ("cells", C.c_longlong),
                ("snake", C.c_longlong)]


class MiGenTaskParams(C.Structure):
    _fields_ = [("seed", C.c_ulonglong), ("lmin", C.c_int), ("lmax", C.c_int), ("emin_ppm", C.c_int),
                ("emax_ppm", C.c_int), ("offmax", C.c_int)]


class MiTuning(C.Structure):
    _fields_ = [("report_waves_per_cu", C.c_int), ("report_pool_log2", C.c_int), ("report_l1_waves_per_cu", C.c_int),
                ("batch_waves_per_cu", C.c_int), ("batch_wide_waves_per_cu", C.c_int)]


class MiTraceTask(C.Structure):
    _fields_ = [("aword", C.c_longlong), ("bword", C.c_longlong), ("alen", C.c_int), ("blen", C.c_int),
                ("abpos", C.c_int), ("bbpos", C.c_int), ("aepos", C.c_int), ("bepos", C.c_int),
                ("tlen", C.c_int), ("points", C.c_longlong), ("selfie", C.c_int)]


class MiTraceResult(C.Structure):
    _fields_ = [("status", C.c_int), ("diffs", C.c_int), ("ntrace", C.c_int), ("trace", C.c_longlong)]


# every symbol declared in include/marvl_align.h (checked by the CPU tests)
EXPORTS = [
    "New_Align_Spec", "Free_Align_Spec", "Trace_Spacing", "Average_Correlation", "Base_Frequencies",
    "New_Work_Data", "Free_Work_Data", "Local_Alignment",
    "Read_Overlap", "Read_Trace", "Write_Overlap", "Compress_TraceTo8", "Decompress_TraceTo16",
    "Check_Trace_Points",
    "mi_la_batch_host", "mi_la_batch_device", "mi_la_batch_device_noalloc", "mi_gen_la_tasks", "mi_abort_device_work", "mi_device_alloc", "mi_device_free",
    "mi_device_copy", "mi_device_memset", "mi_set_tuning", "mi_get_tuning", "mi_device_info", "mi_align_params", "mi_last_error", "mi_device_count", "mi_selftest",
]


class _Lib:
    def __init__(self):
        self._h = None

    def handle(self) -> C.CDLL:
        if self._h is None:
            if not os.path.exists(LIB_PATH):
                raise LibraryMissing(f"{LIB_PATH} not built: run `python -m marvel_amd.build`")
            h = C.CDLL(LIB_PATH)
            h.mi_la_batch_host.argtypes = [C.POINTER(MiAlignParams), C.c_void_p, C.c_longlong,
                                           C.POINTER(MiLaTask), C.c_int, C.POINTER(MiLaResult),
                                           C.c_void_p, C.c_longlong]
            h.mi_la_batch_host.restype = C.c_int
            h.mi_la_batch_device.argtypes = [C.POINTER(MiAlignParams), C.c_void_p, C.c_longlong, C.c_void_p, C.c_int,
                                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p,
                                             C.POINTER(C.c_longlong)]
            h.mi_la_batch_device.restype = C.c_int
            h.mi_la_batch_device_noalloc.argtypes = [C.POINTER(MiAlignParams), C.c_void_p, C.c_longlong, C.c_void_p,
                                                     C.c_int, C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p]
            h.mi_la_batch_device_noalloc.restype = C.c_int
            h.mi_gen_la_tasks.argtypes = [C.POINTER(MiGenTaskParams), C.c_int, C.c_void_p, C.c_longlong,
                                          C.POINTER(C.c_longlong), C.c_void_p]
            h.mi_gen_la_tasks.restype = C.c_int
            h.mi_device_alloc.argtypes = [C.c_longlong]
            h.mi_device_alloc.restype = C.c_void_p
            h.mi_device_free.argtypes = [C.c_void_p]
            h.mi_device_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong, C.c_int]
            h.mi_device_memset.argtypes = [C.c_void_p, C.c_int, C.c_longlong]
            h.mi_set_tuning.argtypes = [C.c_void_p]
            h.mi_get_tuning.argtypes = [C.c_void_p]
            h.mi_device_info.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int)]
            h.mi_align_params.argtypes = [C.POINTER(MiAlignParams), C.POINTER(C.c_int),
                                          C.POINTER(C.c_int), C.POINTER(C.c_int)]
            h.mi_last_error.restype = C.c_char_p
            h.mi_device_count.restype = C.c_int
            h.mi_trace_pts_host.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_longlong, C.c_void_p, C.c_longlong,
                                            C.POINTER(MiTraceTask), C.c_int, C.POINTER(MiTraceResult), C.c_void_p,
                                            C.c_longlong]
            h.mi_trace_pts_host.restype = C.c_int
            h.mi_selftest.argtypes = [C.c_int, C.c_void_p, C.c_longlong, C.c_void_p, C.c_longlong, C.c_void_p,
                                      C.c_int, C.c_int, C.c_void_p, C.c_longlong]
            h.mi_selftest.restype = C.c_int
            self._h = h
        return self._h

    def __getattr__(self, name):
        return getattr(self.handle(), name)

    def last_error(self) -> str:
        return self.handle().mi_last_error().decode()


lib = _Lib()


def params(ave_corr: float, trace_space: int, freq) -> MiAlignParams:
    p = MiAlignParams()
    p.ave_corr = ave_corr
    p.trace_space = trace_space
    for i in range(4):
        p.freq[i] = float(freq[i])
    return p


class DeviceBuffer:
    """A device allocation of the library (mi_device_alloc); freed with the object."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lib.mi_device_alloc(max(self.nbytes, 1))
        if not self.ptr:
            raise MemoryError(f"device allocation of {self.nbytes} bytes: {lib.last_error()}")

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p:
            lib.mi_device_free(p)
            self.ptr = None

    def offset(self, nbytes: int) -> int:
        return self.ptr + int(nbytes)

    def upload(self, arr, at: int = 0):
        import numpy as np
        a = np.ascontiguousarray(arr)
        if at + a.nbytes > self.nbytes:
            raise ValueError("upload past the end of the buffer")
        if lib.mi_device_copy(self.ptr + at, a.ctypes.data, a.nbytes, 1) != 0:
            raise RuntimeError(lib.last_error())

    def download(self, dtype, count: int, at: int = 0):
        import numpy as np
        out = np.empty(int(count), dtype=dtype)
        if at + out.nbytes > self.nbytes:
            raise ValueError("download past the end of the buffer")
        if out.nbytes and lib.mi_device_copy(out.ctypes.data, self.ptr + at, out.nbytes, 0) != 0:
            raise RuntimeError(lib.last_error())
        return out

    def zero(self):
        if lib.mi_device_memset(self.ptr, 0, self.nbytes) != 0:
            raise RuntimeError(lib.last_error())
