"""ctypes binding of libhora_mi355x.so (the C ABI in include/hora_mi355x.h).

The product path has no CPU fallback: if the library is missing or cannot
be loaded, every entry raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhora_mi355x.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "hora_mi355x.h")


class HoraError(RuntimeError):
    """A non-OK status from the C ABI; .status is the hm_status value."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status
        self.message = message


STATUS = {
    0: "HM_OK", 1: "HM_ERR_INVALID_ARG", 2: "HM_ERR_DIM_MISMATCH",
    3: "HM_ERR_NONFINITE_INPUT", 4: "HM_ERR_UNKNOWN_METRIC", 5: "HM_ERR_NOT_BUILT",
    6: "HM_ERR_CAPACITY", 7: "HM_ERR_HIP", 8: "HM_ERR_NAN_DISTANCE", 9: "HM_ERR_EMPTY",
    10: "HM_ERR_NOT_IMPLEMENTED", 11: "HM_ERR_OUT_OF_MEMORY",
}


class HmParams(C.Structure):
    _fields_ = [
        ("n_sub", C.c_uint32),
        ("sub_bits", C.c_uint32),
        ("n_kmeans_center", C.c_uint32),
        ("search_n_center", C.c_uint32),
        ("train_epoch", C.c_uint32),
        ("kmeans_accumulate", C.c_uint32),
        ("seed", C.c_uint64),
    ]


def build(jobs: int = 8) -> str:
    """Compile every HIP source for gfx950 into LIB_PATH (hipcc, no GPU needed)."""
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", os.path.join(_HERE, "csrc")], check=True)
    return LIB_PATH


def declared_symbols() -> list[str]:
    """Entry points declared in include/hora_mi355x.h."""
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"HM_API\s+[\w\s\*]+?\b(hm_\w+)\s*\(", text)))


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HoraError(7, f"{LIB_PATH} is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, u64, u32, i32 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int32
    L.hm_params_default.argtypes = [i32, C.POINTER(HmParams)]
    L.hm_index_create.argtypes = [i32, u64, C.POINTER(HmParams), C.POINTER(vp)]
    L.hm_index_destroy.argtypes = [vp]
    L.hm_index_add.argtypes = [vp, vp, u64]
    L.hm_index_add_device.argtypes = [vp, vp, u64]
    L.hm_index_build.argtypes = [vp, i32]
    L.hm_index_built.argtypes = [vp, C.POINTER(i32)]
    L.hm_index_search.argtypes = [vp, vp, u64, u32, vp, vp, vp]
    L.hm_index_search_device.argtypes = [vp, vp, u64, u32, vp, vp, vp, vp]
    L.hm_index_search_device_noalloc.argtypes = [vp, vp, u64, u32, vp, vp, vp, vp]
    L.hm_index_set_search_mode.argtypes = [vp, i32, u32]
    L.hm_index_search_ragged_device.argtypes = [vp, vp, u32, vp, vp, vp, vp, vp, vp]
    L.hm_index_autotune.argtypes = [vp, vp, u64, u32, C.POINTER(u32)]
    L.hm_index_tuning.argtypes = [vp, C.POINTER(u32), C.POINTER(u32)]
    L.hm_index_search_stats.argtypes = [vp, C.POINTER(u64), C.POINTER(u64)]
    L.hm_index_nodes_size.argtypes = [vp, C.POINTER(u64)]
    L.hm_index_dimension.argtypes = [vp, C.POINTER(u64)]
    L.hm_index_name.argtypes = [vp]
    L.hm_index_name.restype = C.c_char_p
    L.hm_index_clear.argtypes = [vp]
    L.hm_index_rebuild.argtypes = [vp, i32]
    L.hm_index_idx_info.argtypes = [vp, C.c_char_p, u64, C.POINTER(u64)]
    L.hm_index_get_rows.argtypes = [vp, vp, u64, vp]
    L.hm_index_set_nprobe.argtypes = [vp, u32]
    L.hm_ivfpq_info.argtypes = [vp, vp]
    L.hm_ivfpq_export.argtypes = [vp, vp, vp, vp, vp, vp]
    L.hm_ivfpq_import.argtypes = [vp, i32, vp, vp, vp, vp, vp]
    L.hm_index_dump.argtypes = [vp, C.c_char_p, vp, vp]
    L.hm_index_load.argtypes = [i32, C.c_char_p, i32, C.POINTER(vp)]
    L.hm_index_loaded_ids.argtypes = [vp, vp, vp, C.POINTER(u64)]
    L.hm_index_knn_graph.argtypes = [vp, u32, vp, vp, vp]
    L.hm_index_distances.argtypes = [vp, vp, u64, vp, vp, vp]
    L.hm_index_distances_device.argtypes = [vp, vp, u64, vp, vp, u64, vp, vp]
    L.hm_pq_info.argtypes = [vp, vp]
    L.hm_pq_export.argtypes = [vp, vp, vp]
    L.hm_pq_import.argtypes = [vp, i32, vp, vp, u64]
    L.hm_ivfpq_train_probe.argtypes = [vp, vp, u64, u32, u32, u32, u32, C.c_float, u64,
                                       C.POINTER(C.c_double)]
    L.hm_ivfpq_set_probe_policy.argtypes = [vp, i32, C.c_float, u32, u32]
    L.hm_ivfpq_set_scan_form.argtypes = [vp, i32]
    L.hm_ivfpq_set_rerank.argtypes = [vp, u32]
    L.hm_ivfpq_mean_nprobe.argtypes = [vp, C.POINTER(C.c_double)]
    L.hm_ivfpq_last_probes.argtypes = [vp, u64, u32, vp, vp]
    L.hm_mlp_loss_grad.argtypes = [u32, u32, u32, u32] + [vp] * 12
    L.hm_profile_enable.argtypes = [i32]
    L.hm_profile_reset.argtypes = []
    L.hm_profile_only.argtypes = [C.c_char_p]
    L.hm_profile_query.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.POINTER(u64)]
    L.hm_last_error.argtypes = []
    L.hm_last_error.restype = C.c_char_p
    for name in declared_symbols():
        fn = getattr(L, name)
        if fn.restype is C.c_int:  # default restype: every status entry returns int32
            fn.restype = C.c_int32
    _lib = L
    return L


def check(status: int) -> None:
    if status != 0:
        msg = lib().hm_last_error().decode(errors="replace")
        raise HoraError(status, msg)


def profile_enable(on: bool = True) -> None:
    check(lib().hm_profile_enable(1 if on else 0))


def profile_only(kernel=None) -> None:
    """Time only this kernel family (None: every family)."""
    check(lib().hm_profile_only(kernel.encode() if kernel else None))


def profile_reset() -> None:
    check(lib().hm_profile_reset())


def profile_query(kernel: str):
    """(total_ms, launches) of a kernel family timed with HIP events."""
    ms, n = C.c_double(), C.c_uint64()
    check(lib().hm_profile_query(kernel.encode(), C.byref(ms), C.byref(n)))
    return ms.value, int(n.value)
