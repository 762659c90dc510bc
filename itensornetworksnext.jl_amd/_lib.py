"""ctypes binding of the C ABI in include/bp_mi355x.h (libbpmi355x.so).

Loading fails loudly: there is no CPU fallback for any compute entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "csrc", "libbpmi355x.so")

BP_OK, BP_ERR_ARG, BP_ERR_HIP, BP_ERR_STATE, BP_ERR_UNSUPPORTED, BP_ERR_NOMEM = range(6)
BP_F64, BP_F32, BP_BF16, BP_C128 = 0, 1, 2, 3
BP_INIT_IDENTITY, BP_INIT_ONES, BP_INIT_UPLOAD = 0, 1, 2
BP_CRIT_FIDELITY, BP_CRIT_MAXABS = 0, 1
BP_PATH_AUTO, BP_PATH_GENERIC, BP_PATH_VERTEX, BP_PATH_MIXED = 0, 1, 2, 3

DTYPES = {"f64": BP_F64, "float64": BP_F64, "f32": BP_F32, "float32": BP_F32,
          "bf16": BP_BF16, "bfloat16": BP_BF16, "c128": BP_C128, "complex128": BP_C128}
PATHS = {"auto": BP_PATH_AUTO, "generic": BP_PATH_GENERIC, "vertex": BP_PATH_VERTEX}

# Every exported symbol and its signature: (restype, argtypes).
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p
SIGNATURES = {
    "bp_last_error": (C.c_char_p, []),
    "bp_version": (C.c_int, []),
    "bp_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int]),
    "bp_destroy": (C.c_int, [_vp]),
    "bp_set_stream": (C.c_int, [_vp, _vp]),
    "bp_set_topology": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _ip,
                                  C.c_int, C.c_int]),
    "bp_upload_sites": (C.c_int, [_vp, _dp]),
    "bp_init_messages": (C.c_int, [_vp, C.c_int, _dp]),
    "bp_download_messages": (C.c_int, [_vp, _dp]),
    "bp_set_slot_messages": (C.c_int, [_vp, C.c_int, _ip, _dp]),
    "bp_get_slot_messages": (C.c_int, [_vp, C.c_int, _ip, _dp]),
    "bp_set_path": (C.c_int, [_vp, C.c_int]),
    "bp_get_path": (C.c_int, [_vp, _ip]),
    "bp_sweep": (C.c_int, [_vp]),
    "bp_run": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int, C.c_int, _ip, _dp, _ip]),
    "bp_prepare_run": (C.c_int, [_vp, C.c_int]),
    "bp_update_slot": (C.c_int, [_vp, C.c_int]),
    "bp_run_sequential": (C.c_int, [_vp, _ip, C.c_int, C.c_int, C.c_double, C.c_int, _ip, _dp,
                                    _ip]),
    "bp_residuals": (C.c_int, [_vp, _dp, _dp]),
    "bp_iterate_diff": (C.c_int, [_vp, _dp, _dp]),
    "bp_iterate_diff_noalloc": (C.c_int, [_vp, _dp, _dp]),
    "bp_vertex_scalars": (C.c_int, [_vp, _dp]),
    "bp_edge_scalars": (C.c_int, [_vp, _dp]),
    "bp_lognorm": (C.c_int, [_vp, _dp, _dp]),
    "bp_polish_fp64": (C.c_int, [_vp, C.c_int, _dp]),
    "bp_grad_lognorm": (C.c_int, [_vp, _dp]),
    "bp_tfim_energy": (C.c_int, [_vp, C.c_double, C.c_double, _dp, _dp, C.c_double, C.c_int,
                                 _ip, _dp]),
    "bp_grad_lognorm_device": (C.c_int, [_vp, _vp]),
    "bp_tfim_energy_device": (C.c_int, [_vp, C.c_double, C.c_double, _dp, _vp, C.c_double,
                                        C.c_int, _ip, _dp]),
    "bp_expect_vertex": (C.c_int, [_vp, _dp, C.c_int, _dp]),
    "bp_expect_edge": (C.c_int, [_vp, _dp, _dp, C.c_int, _dp]),
    "bp_local_energy": (C.c_int, [_vp, _dp, _dp, _dp, C.c_int, _dp, _dp, C.c_double, C.c_int,
                                  _ip, _dp]),
    "bp_local_energy_device": (C.c_int, [_vp, _dp, _dp, _dp, C.c_int, _dp, _vp, C.c_double,
                                         C.c_int, _ip, _dp]),
    "bp_autotune": (C.c_int, [_vp, C.c_char_p, C.c_int, _ip]),
    "bp_set_option": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "bp_get_option": (C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_int)]),
    "bp_last_run_ms": (C.c_int, [_vp, _dp]),
    "bp_time_sweeps": (C.c_int, [_vp, C.c_int, _dp]),
    "bp_cn_plan_create": (C.c_int, [_vp, C.c_int, _ip, _ip, C.c_int, _ip, _ip, C.c_int,
                                    C.c_longlong, C.POINTER(C.c_void_p)]),
    "bp_cn_plan_info": (C.c_int, [_vp, _ip, _ip, C.POINTER(C.c_longlong),
                                  C.POINTER(C.c_longlong), _dp, _dp, C.POINTER(C.c_longlong)]),
    "bp_cn_execute": (C.c_int, [_vp, _vp, _dp, _dp]),
    "bp_cn_execute_device": (C.c_int, [_vp, _vp, _vp, _vp]),
    "bp_cn_plan_destroy": (C.c_int, [_vp]),
    "bp_contract_network": (C.c_int, [_vp, C.c_int, _ip, _ip, C.c_int, _ip, _dp, _ip, C.c_int,
                                      C.c_longlong, _ip, _ip, _dp, C.c_longlong]),
    "bp_gate_sizes": (C.c_int, [C.c_int, _ip, C.POINTER(C.c_longlong),
                                C.POINTER(C.c_longlong)]),
    "bp_apply_gates": (C.c_int, [_vp, C.c_int, _ip, _dp, _dp, _ip, C.c_double]),
    "bp_apply_gates_resident": (C.c_int, [_vp, C.c_int, _ip, _ip, _dp, C.c_int, C.c_int,
                                          C.c_double, _ip, _dp, C.c_int]),
    "bp_download_sites": (C.c_int, [_vp, _dp]),
}
BP_CN_ORDER_LEFT, BP_CN_ORDER_GREEDY, BP_CN_ORDER_CHEAPEST = 0, 1, 2
BP_GATE_DESC = 25

_lib = None


class BPError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: status {code}: {msg}")
        self.code = code


def load(path: str = LIB_PATH):
    """Load the engine library; raises if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP engine library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, where: str) -> None:
    if rc != BP_OK:
        msg = load().bp_last_error().decode(errors="replace")
        raise BPError(rc, where, msg)


# ---------------------------------------------------------------- device buffers
# Device memory for the device-resident entry points comes from the HIP runtime that
# libbpmi355x.so itself links (ROCm's libamdhip64.so.7).  PyTorch wheels bundle their own
# HIP runtime; pointers are not exchanged between the two.
_hip = None


def hip_runtime():
    global _hip
    if _hip is None:
        load()  # the engine first, so the same runtime instance is returned below
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipMalloc.restype = C.c_int
        h.hipFree.argtypes = [C.c_void_p]
        h.hipFree.restype = C.c_int
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpy.restype = C.c_int
        h.hipSetDevice.argtypes = [C.c_int]
        h.hipSetDevice.restype = C.c_int
        _hip = h
    return _hip


class DeviceArray:
    """A 1-D device buffer of `n` float32 or float64 entries on HIP device `device`."""

    def __init__(self, n: int, dtype, device: int = 0):
        import numpy as np

        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.device = int(device)
        self.nbytes = self.n * self.dtype.itemsize
        h = hip_runtime()
        if h.hipSetDevice(self.device) != 0:
            raise BPError(BP_ERR_HIP, "hipSetDevice", f"device {self.device}")
        p = C.c_void_p()
        rc = h.hipMalloc(C.byref(p), max(self.nbytes, 1))
        if rc != 0:
            raise BPError(BP_ERR_NOMEM, "hipMalloc", f"{self.nbytes} bytes (hip error {rc})")
        self.ptr = p.value

    def numpy(self):
        import numpy as np

        out = np.empty(self.n, dtype=self.dtype)
        rc = hip_runtime().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2)  # to host
        if rc != 0:
            raise BPError(BP_ERR_HIP, "hipMemcpy", f"hip error {rc}")
        return out

    def upload(self, arr):
        import numpy as np

        a = np.ascontiguousarray(arr, dtype=self.dtype).reshape(-1)
        if a.size != self.n:
            raise ValueError(f"upload of {a.size} entries into a buffer of {self.n}")
        rc = hip_runtime().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1)  # to device
        if rc != 0:
            raise BPError(BP_ERR_HIP, "hipMemcpy", f"hip error {rc}")

    def free(self):
        if getattr(self, "ptr", None):
            hip_runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()
