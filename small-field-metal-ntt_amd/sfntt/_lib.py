"""ctypes declarations of libsfntt.so (include/sfntt.h).

The library is the only compute path: if it is missing, importing the
product API raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsfntt.so")

# enums (include/sfntt.h)
OK = 0
ERR_INVALID_SIZE, ERR_DEVICE_NOT_FOUND, ERR_BUFFER_ALLOC_FAILED = 1, 2, 3
ERR_KERNEL_LOAD, ERR_GPU_EXECUTION, ERR_CORRECTNESS, ERR_UNSUPPORTED = 4, 5, 6, 7
ERR_INVALID_ARGUMENT, ERR_BUSY, ERR_NOT_READY = 8, 9, 10
FIELD_M31, FIELD_BABYBEAR, FIELD_KOALABEAR = 0, 1, 2
TUNE_LDE = 16  # sfn_autotune's field value for the coset LDE
CT_DIT_R2, CT_GS_R2, STOCKHAM_R2, CT_GS_R4, VARIANT_AUTO = 0, 1, 2, 3, 4
FORWARD, INVERSE = 0, 1
LAYOUT_SOA, LAYOUT_ROW_MAJOR = 0, 1
JOB_NTT, JOB_LDE = 0, 1

u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
vp = ctypes.c_void_p


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("gcn_arch", ctypes.c_char * 64),
                ("compute_units", ctypes.c_int), ("lds_bytes_per_cu", ctypes.c_int),
                ("hbm_bytes", ctypes.c_size_t), ("clock_khz", ctypes.c_int),
                ("mem_clock_khz", ctypes.c_int), ("mem_bus_width", ctypes.c_int)]


class JobDesc(ctypes.Structure):
    _fields_ = [("field", ctypes.c_int), ("dir", ctypes.c_int), ("log_n", ctypes.c_uint32),
                ("batch", ctypes.c_uint32), ("d_in", vp), ("d_out", vp)]


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("field", ctypes.c_int), ("dir", ctypes.c_int),
                ("log_n", ctypes.c_uint32), ("batch", ctypes.c_uint32),
                ("added_bits", ctypes.c_uint32), ("shift_raw", ctypes.c_uint32),
                ("device_ptrs", ctypes.c_int), ("inp", vp), ("out", vp)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("mem_budget_bytes", ctypes.c_size_t), ("max_inflight", ctypes.c_uint32),
                ("max_queue", ctypes.c_uint32), ("max_group", ctypes.c_uint32),
                ("max_group_cols", ctypes.c_uint32)]


class JobTiming(ctypes.Structure):
    _fields_ = [("submit_ns", ctypes.c_uint64), ("start_ns", ctypes.c_uint64),
                ("end_ns", ctypes.c_uint64), ("group_size", ctypes.c_uint32), ("status", ctypes.c_int)]


# name -> (restype, argtypes); every symbol include/sfntt.h declares
SIGNATURES = {
    "sfn_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
    "sfn_ctx_destroy": (None, [vp]),
    "sfn_last_error": (ctypes.c_char_p, [vp]),
    "sfn_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "sfn_device_info_get": (ctypes.c_int, [vp, ctypes.POINTER(DeviceInfo)]),
    "sfn_ntt": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                               ctypes.c_uint32, ctypes.c_int, u32p, u32p, u64p]),
    "sfn_coset_lde_batch": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                           ctypes.c_uint32, ctypes.c_int, u32p, u32p, u64p]),
    "sfn_coset_lde_natural": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                             ctypes.c_uint32, ctypes.c_int, u32p, u32p, u64p]),
    "sfn_pointwise_mul": (ctypes.c_int, [vp, ctypes.c_int, u32p, u32p, u32p, ctypes.c_size_t]),
    "sfn_buf_alloc": (ctypes.c_int, [vp, ctypes.c_size_t, ctypes.POINTER(vp)]),
    "sfn_buf_free": (ctypes.c_int, [vp, vp]),
    "sfn_upload": (ctypes.c_int, [vp, vp, vp, ctypes.c_size_t, vp]),
    "sfn_download": (ctypes.c_int, [vp, vp, vp, ctypes.c_size_t, vp]),
    "sfn_synchronize": (ctypes.c_int, [vp, vp]),
    "sfn_twiddles_prepare": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_uint32]),
    "sfn_twiddles_bytes": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_size_t)]),
    "sfn_twiddles_limit": (ctypes.c_int, [vp, ctypes.c_size_t]),
    "sfn_ntt_device": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_int, vp, vp, vp]),
    "sfn_coset_lde_batch_device": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_uint32,
                                                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                                  vp, vp, vp]),
    "sfn_coset_lde_natural_device": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_uint32,
                                                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                                  vp, vp, vp]),
    "sfn_coset_lde_natural_device_noalloc": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_uint32,
                                                         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                                         vp, vp, vp]),
    "sfn_last_device_ns": (ctypes.c_int, [vp, u64p]),
    "sfn_call_timing": (ctypes.c_int, [vp, ctypes.c_int]),
    "sfn_pass_timing": (ctypes.c_int, [vp, ctypes.c_int]),
    "sfn_pass_times": (ctypes.c_int, [vp, u64p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
    "sfn_pass_label": (ctypes.c_char_p, [vp, ctypes.c_uint32]),
    "sfn_plan_create": (ctypes.c_int, [vp, ctypes.POINTER(JobDesc), ctypes.c_uint32, ctypes.c_int,
                                       ctypes.POINTER(vp)]),
    "sfn_plan_launch": (ctypes.c_int, [vp, vp]),
    "sfn_plan_info": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_uint32),
                                     ctypes.POINTER(ctypes.c_uint32)]),
    "sfn_plan_destroy": (None, [vp]),
    "sfn_autotune": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_uint32)]),
    "sfn_tuning_save": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_size_t)]),
    "sfn_tuning_load": (ctypes.c_int, [vp, ctypes.c_char_p]),
    "sfn_host_alloc": (ctypes.c_int, [vp, ctypes.c_size_t, ctypes.POINTER(vp)]),
    "sfn_host_free": (ctypes.c_int, [vp, vp]),
    "sfn_now": (ctypes.c_uint64, []),
    "sfn_server_create": (ctypes.c_int, [vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]),
    "sfn_server_submit": (ctypes.c_int, [vp, ctypes.POINTER(Request), u64p]),
    "sfn_server_poll": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.POINTER(JobTiming)]),
    "sfn_server_wait": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.POINTER(JobTiming)]),
    "sfn_server_drain": (ctypes.c_int, [vp]),
    "sfn_server_destroy": (None, [vp]),
}

_lib = None


def load():
    """Loads libsfntt.so; raises ImportError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run small-field-metal-ntt_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib
