"""ctypes binding of libusbs_hip.so (the C ABI declared in include/usbs_hip.h).

The library is built in-tree by ``python -m usbs_amd.build``.  There is no
CPU fallback: if the shared library is missing or fails to load, importing
anything that needs it raises ``RuntimeError``.
"""
import atexit
import ctypes
import os
import threading
import weakref

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libusbs_hip.so")

USBS_OK = 0
ERRORS = {-1: "invalid argument", -2: "HIP error", -3: "out of device memory",
          -4: "non-finite input", -5: "bad call order", -6: "device wait expired"}

PRECISION = {"fp64": 0, "mixed": 1}
EXEC_MODE = {"perop": 0, "persistent": 1, "auto": 2, "tune": 3}

_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


class usbs_stats(ctypes.Structure):
    """Timing counters of the persistent Lanczos step kernel (usbs_get_stats)."""
    _fields_ = [("lz_kernel_ms", ctypes.c_double), ("lz_launches", ctypes.c_int64),
                ("lz_steps", ctypes.c_int64), ("lz_sum_cols", ctypes.c_int64),
                ("n", ctypes.c_int64), ("nnz", ctypes.c_int64), ("value_bytes", ctypes.c_int32),
                ("lz_form", ctypes.c_int32)]


class USBSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("usbs-hip error %d (%s): %s" % (code, ERRORS.get(code, "?"), msg))
        self.code = code


def _declare(lib):
    lib.usbs_version.restype = ctypes.c_char_p
    lib.usbs_last_error.restype = ctypes.c_char_p
    lib.usbs_last_error.argtypes = [_vp]
    lib.usbs_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int]
    lib.usbs_destroy.argtypes = [_vp]
    lib.usbs_destroy.restype = None
    lib.usbs_set_problem.argtypes = [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _ip, _ip,
                                     _dp, ctypes.c_int64, _ip, _dp, _dp, _dp, _dp]
    lib.usbs_lanczos_smallest.argtypes = [_vp, ctypes.c_double, _dp, _dp, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_int, _dp, _dp, _ip]
    lib.usbs_set_timing.argtypes = [_vp, ctypes.c_int]
    lib.usbs_device_memory.argtypes = [_vp, _i64p, _i64p]
    lib.usbs_get_stats.argtypes = [_vp, ctypes.POINTER(usbs_stats)]
    for name, args in EXTRA_SIGNATURES.items():
        if hasattr(lib, name):
            getattr(lib, name).argtypes = args
    _declared[0] = len(EXTRA_SIGNATURES)


# Filled by modules that bind further entry points (kept here so one loader
# declares every symbol before first use).
EXTRA_SIGNATURES = {}


def lib():
    """Load (once) and return the shared library; raise if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libusbs_hip.so not built (run `python -m usbs_amd.build`); "
                               "there is no CPU fallback for the HIP backend")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    elif len(EXTRA_SIGNATURES) != _declared[0]:
        _declare(_lib)
    return _lib


_declared = [0]


def dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def as_f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def as_i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


# Every live context, so that they are destroyed (streams, graphs, device
# memory) at interpreter exit while the HIP runtime is still up, rather than
# from __del__ during module teardown.
_live = weakref.WeakSet()
_live_lock = threading.Lock()


def close_all():
    """Destroy every context that is still open (registered with atexit)."""
    with _live_lock:
        ctxs = list(_live)
    for c in ctxs:
        try:
            c.close()
        except Exception:
            pass


atexit.register(close_all)


class Context:
    """An owned device context (one HIP stream + resident problem data).
    Usable as a context manager; close() is idempotent."""

    def __init__(self, device=0):
        self._lib = lib()
        h = _vp()
        rc = self._lib.usbs_create(ctypes.byref(h), int(device))
        if rc != USBS_OK:
            raise USBSError(rc, "usbs_create(device=%d) failed" % device)
        self.h = h
        self.n = None
        self.m = None
        with _live_lock:
            _live.add(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def check(self, rc):
        if rc != USBS_OK:
            raise USBSError(rc, self._lib.usbs_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self._lib.usbs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_memory(self):
        """(free, total) bytes of the context's device."""
        free, total = ctypes.c_int64(0), ctypes.c_int64(0)
        self.check(self._lib.usbs_device_memory(self.h, ctypes.byref(free), ctypes.byref(total)))
        return free.value, total.value

    def set_timing(self, enable):
        """Bracket persistent Lanczos launches with HIP events (resets counters)."""
        self.check(self._lib.usbs_set_timing(self.h, 1 if enable else 0))

    def stats(self):
        st = usbs_stats()
        self.check(self._lib.usbs_get_stats(self.h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in usbs_stats._fields_}

    def set_problem(self, n, m, C_rows, C_cols, C_vals, A_indices, A_data, b, b_ineq_mask, SCALE_A):
        C_rows, C_cols, C_vals = as_i32(C_rows), as_i32(C_cols), as_f64(C_vals)
        A_indices = as_i32(np.asarray(A_indices).reshape(-1, 3))
        A_data, b, mask, sa = as_f64(A_data), as_f64(b), as_f64(b_ineq_mask), as_f64(SCALE_A)
        self.check(self._lib.usbs_set_problem(self.h, n, m, C_vals.size, iptr(C_rows), iptr(C_cols),
                                              dptr(C_vals), A_data.size, iptr(A_indices),
                                              dptr(A_data), dptr(b), dptr(mask), dptr(sa)))
        self.n, self.m = int(n), int(m)
