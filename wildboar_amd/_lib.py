"""ctypes binding of libwbx.so (the C ABI declared in include/wbx.h).

The library is the only compute path: when it is missing, or when a call needs
a GPU and none is present, the call raises -- there is no CPU fallback.
"""

import ctypes
import os
import subprocess
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libwbx.so")
CSRC = os.path.join(_HERE, "csrc")

WBX_OK, WBX_EINVAL, WBX_ENOMEM, WBX_EDEVICE, WBX_EUNSUPPORTED = 0, -1, -2, -3, -4
EUCLIDEAN, DTW, WDTW, DDTW, WDDTW, ADTW = 0, 1, 2, 3, 4, 5
FP64, FP32, FP32_ACC64 = 0, 1, 2
F64, F32 = 0, 1

KINDS = {"euclidean": EUCLIDEAN, "dtw": DTW, "wdtw": WDTW, "ddtw": DDTW, "wddtw": WDDTW,
         "adtw": ADTW}
PRECISIONS = {"fp64": FP64, "fp32": FP32, "fp32_acc64": FP32_ACC64}

EXPORTS = (
    "wbx_init", "wbx_last_error", "wbx_version", "wbx_device_count",
    "wbx_pairwise", "wbx_paired", "wbx_argmin", "wbx_subsequence_pairwise",
    "wbx_dev_pairwise", "wbx_dev_pairwise_noalloc", "wbx_set_acc64_length", "wbx_dev_paired", "wbx_dev_argmin",
    "wbx_dev_subsequence_pairwise", "wbx_set_policy", "wbx_prof_enable", "wbx_prof_take", "wbx_prof_pruned",
    "wbx_prof_valu_peak", "wbx_dev_eos_lengths", "wbx_batch_create", "wbx_batch_free",
    "wbx_batch_set_launch", "wbx_batch_autotune", "wbx_batch_bucket_info", "wbx_batch_info", "wbx_batch_run", "wbx_dtw_train_sizes", "wbx_dev_dtw_train_forward",
    "wbx_dev_dtw_train_backward", "wbx_argmin_subsequence", "wbx_subsequence_profile",
    "wbx_dev_argmin_subsequence", "wbx_dev_subsequence_profile", "wbx_dev_dtw_path",
    "wbx_dev_dtw_mm_update", "wbx_dev_dtw_envelope", "wbx_dev_lb_keogh", "wbx_dev_lb_kim",
    "wbx_scaled_subsequence_pairwise", "wbx_dev_scaled_subsequence_pairwise",
    "wbx_scaled_subsequence_profile", "wbx_dev_scaled_subsequence_profile",
    "wbx_scaled_argmin_subsequence", "wbx_dev_scaled_argmin_subsequence",
    "wbx_dev_dtw_ssg_update",
)


class WbxMetric(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("r", ctypes.c_double), ("g", ctypes.c_double), ("p", ctypes.c_double)]


class WbxError(RuntimeError):
    pass


_lib = None
_lock = threading.Lock()


def build(jobs=8):
    """Compile every HIP source under csrc/ for gfx950 into lib/libwbx.so."""
    subprocess.run(["make", "-s", "-C", CSRC, f"-j{jobs}"], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise WbxError(
                f"{LIB_PATH} is not built; build it with wildboar_amd._lib.build() "
                "(hipcc --offload-arch=gfx950)")
        # libwbx.so and torch both need libamdhip64.so.7; whichever loads
        # first serves both.  Importing torch first keeps one HIP runtime, the
        # one torch was built with, in the process.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(LIB_PATH)
        p, i64, i32, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
        m = ctypes.POINTER(WbxMetric)
        sig = {
            "wbx_init": (i32, [i32]),
            "wbx_last_error": (ctypes.c_char_p, []),
            "wbx_version": (ctypes.c_char_p, []),
            "wbx_device_count": (i32, []),
            "wbx_set_policy": (i32, [i32]),
            "wbx_prof_enable": (i32, [i32]),
            "wbx_prof_take": (i32, [p, p, p]),
            "wbx_prof_pruned": (i32, [p]),
            "wbx_prof_valu_peak": (i32, [p]),
            "wbx_dev_eos_lengths": (i32, [i32, p, i64, i64, i64, p, p]),
            "wbx_batch_create": (i32, [m, i64, i64, p, p, p, i32, p]),
            "wbx_dev_dtw_path": (i32, [p, i64, i64, p, i64, i64, i64, i64, p, p, p]),
            "wbx_dev_dtw_mm_update": (i32, [p, i64, p, i64, i64, i64, i64, p, p, p, p]),
            "wbx_dev_dtw_envelope": (i32, [p, i64, i64, i64, i64, p, p, p]),
            "wbx_dev_lb_keogh": (i32, [p, i64, i64, p, p, i64, i64, p, p]),
            "wbx_dev_lb_kim": (i32, [p, i64, i64, p, i64, i64, i64, p, p]),
            "wbx_batch_free": (i32, [p]),
            "wbx_batch_set_launch": (i32, [p, i32, i32]),
            "wbx_batch_autotune": (i32, [p, i32, p, i64, p, i64, p, i32, p]),
            "wbx_batch_bucket_info": (i32, [p, i32, p, p, p, p, p, p]),
            "wbx_batch_info": (i32, [p, p, p, p]),
            "wbx_batch_run": (i32, [p, i32, p, i64, p, i64, p, p, i32]),
            "wbx_dtw_train_sizes": (i32, [m, i32, i64, i64, i64, i64, p, p]),
            "wbx_dev_dtw_train_forward": (i32, [m, i32, p, i64, i64, i64, p, i64, i64, i64, p, p, p]),
            "wbx_dev_dtw_train_backward": (i32, [m, i32, p, i64, i64, i64, p, i64, i64, i64, p, p,
                                                 p, i64, p, i64, p, p]),
            "wbx_pairwise": (i32, [m, p, i64, i64, i64, p, i64, i64, i64, p]),
            "wbx_paired": (i32, [m, p, p, i64, i64, i64, i64, p]),
            "wbx_argmin": (i32, [m, p, i64, i64, i64, p, i64, i64, i64, i64, p, p, p]),
            "wbx_subsequence_pairwise": (i32, [m, p, p, p, i64, p, i64, i64, i64, i64, p, p]),
            "wbx_dev_pairwise": (i32, [m, i32, p, i64, i64, i64, p, i64, i64, i64, p, p]),
            "wbx_dev_pairwise_noalloc": (i32, [m, i32, p, i64, i64, i64, p, i64, i64, i64, p, p]),
            "wbx_set_acc64_length": (i32, [i64]),
            "wbx_dev_paired": (i32, [m, i32, p, i64, p, i64, i64, i64, p, p]),
            "wbx_dev_argmin": (i32, [m, i32, p, i64, i64, i64, p, i64, i64, i64, i64, p, p, p, p]),
            "wbx_dev_subsequence_pairwise": (i32, [m, i32, p, p, p, i64, p, i64, i64, i64, p, p, p]),
            "wbx_argmin_subsequence": (i32, [m, p, p, p, i64, p, i64, i64, i64, i64, p, p]),
            "wbx_subsequence_profile": (i32, [m, p, p, p, i64, i32, p, i64, i64, i64, i64, i64, p]),
            "wbx_dev_argmin_subsequence": (i32, [m, i32, p, p, p, p, i64, i64, i64, i64, p, p, p]),
            "wbx_dev_subsequence_profile": (i32, [m, i32, p, p, p, i64, i32, p, i64, i64, i64, i64,
                                                  p, p]),
            "wbx_scaled_subsequence_pairwise": (i32, [m, p, p, p, p, p, i64, p, i64, i64, i64, i64,
                                                      p, p]),
            "wbx_dev_scaled_subsequence_pairwise": (i32, [m, i32, p, p, p, p, p, i64, p, i64, i64,
                                                          i64, p, p, p]),
            "wbx_scaled_subsequence_profile": (i32, [m, p, p, p, p, p, i64, i32, p, i64, i64, i64,
                                                     i64, i64, d, p]),
            "wbx_dev_scaled_subsequence_profile": (i32, [m, i32, p, p, p, p, p, i64, i32, p, i64,
                                                         i64, i64, i64, d, p, p]),
            "wbx_scaled_argmin_subsequence": (i32, [m, p, p, p, i64, p, i64, i64, i64, i64, p, p]),
            "wbx_dev_dtw_ssg_update": (i32, [p, i64, p, p, i64, d, d, p]),
            "wbx_dev_scaled_argmin_subsequence": (i32, [m, i32, p, p, p, p, i64, i64, i64, i64, p,
                                                        p, p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc):
    """Raise the Python exception the reference raises for a status code."""
    if rc == WBX_OK:
        return
    msg = load().wbx_last_error().decode("utf-8", "replace")
    if rc == WBX_EINVAL:
        raise ValueError(msg)
    if rc == WBX_ENOMEM:
        raise MemoryError(msg)
    if rc == WBX_EUNSUPPORTED:
        raise NotImplementedError(msg)
    raise WbxError(msg or f"wbx call failed with status {rc}")


def metric_struct(kind, precision, r=1.0, g=0.05, p=1.0):
    return WbxMetric(int(kind), int(precision), float(r), float(g), float(p))


def set_policy(policy):
    """0: fastest kernels (default); 1: the plain one-pair-per-block kernel only;
    2: the group kernel (n lanes per pair) wherever it applies."""
    check(load().wbx_set_policy(int(policy)))


def set_acc64_length(n):
    """FP32 requests on series longer than ``n`` run with float64 accumulation
    (default 2048; 0 never promotes)."""
    check(load().wbx_set_acc64_length(int(n)))


def prof_pruned():
    """Pairs whose distance argmin with a lower bound computed since the last
    call (current device)."""
    n = ctypes.c_int64()
    check(load().wbx_prof_pruned(ctypes.byref(n)))
    return n.value


def prof_enable(on=True):
    check(load().wbx_prof_enable(1 if on else 0))


def prof_take():
    """(summed kernel milliseconds, number of kernels, kind of the last one)."""
    ms = ctypes.c_double(0.0)
    n = ctypes.c_int64(0)
    kind = ctypes.c_int(0)
    check(load().wbx_prof_take(ctypes.byref(ms), ctypes.byref(n), ctypes.byref(kind)))
    return ms.value, n.value, kind.value


def valu_peak():
    """Measured VALU lane-instructions per second of the DTW cell mix."""
    v = ctypes.c_double(0.0)
    check(load().wbx_prof_valu_peak(ctypes.byref(v)))
    return v.value
