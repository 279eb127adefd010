"""ctypes binding of libbmi.so (include/bmi.h).

The library is built in-tree (beastgpu_amd/libbmi.so) by ``__graft_entry__.build()``
or ``make -C beastgpu_amd/csrc``.  There is no fallback: if the library or a
GPU is missing, every call raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbmi.so")

OK, EINVAL, ENOMEM, EDEVICE, EBUSY, ECANCELED, EUNSUPPORTED, EINTERNAL = 0, -1, -2, -3, -4, -5, -6, -7
SPACE_LAGRANGE_C0D1, SPACE_RWG = 1, 2
OP_HH3D_SINGLELAYER, OP_MW_SINGLELAYER, OP_MW_DOUBLELAYER = 1, 2, 3
PREC_FP64, PREC_MIXED = 0, 1
C64, C32 = 0, 1
CLASS_NAMES = ("far", "common_vertex", "common_edge", "common_face", "near", "far_fp64")

# every entry point declared in include/bmi.h
EXPORTS = (
    "bmi_version", "bmi_last_error", "bmi_options_default",
    "bmi_context_create", "bmi_context_destroy",
    "bmi_mesh_create", "bmi_mesh_free",
    "bmi_space_create", "bmi_space_create_shaped", "bmi_space_free",
    "bmi_plan_create", "bmi_plan_create_rules", "bmi_plan_free",
    "bmi_matrix_create", "bmi_matrix_free", "bmi_matrix_device_ptr",
    "bmi_assemble", "bmi_plan_stats", "bmi_matrix_to_host", "bmi_assemble_to_host",
    "bmi_matrix_apply", "bmi_matrix_apply_noalloc", "bmi_matrix_compare", "bmi_device_memory",
    "bmi_server_create", "bmi_server_destroy", "bmi_server_submit", "bmi_server_wait",
    "bmi_server_release", "bmi_server_memory", "bmi_request_bytes",
    "bmi_batch_create", "bmi_batch_assemble", "bmi_batch_stats_get", "bmi_batch_set_schedule", "bmi_batch_free",
)
BATCH_MAX_BUCKETS = 8
REQ_ASSEMBLE, REQ_ASSEMBLE_APPLY, REQ_APPLY = 0, 1, 2


class BMIError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"bmi status {status}: {message}")
        self.status = status


class Operator(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("alpha", ctypes.c_double * 2),
        ("beta", ctypes.c_double * 2),
        ("gamma", ctypes.c_double * 2),
    ]


class Options(ctypes.Structure):
    _fields_ = [
        ("far_test_order", ctypes.c_int32),
        ("far_trial_order", ctypes.c_int32),
        ("ss_order", ctypes.c_int32 * 3),
        ("near_order", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("near_factor", ctypes.c_double),
    ]


class Rules(ctypes.Structure):
    _fields_ = [
        ("test_points", ctypes.c_int32),
        ("trial_points", ctypes.c_int32),
        ("test_uv", (ctypes.c_double * 2) * 4),
        ("test_weight", ctypes.c_double * 4),
        ("trial_uv", (ctypes.c_double * 2) * 4),
        ("trial_weight", ctypes.c_double * 4),
        ("gl_points", ctypes.c_int32 * 3),
        ("gl_node", (ctypes.c_double * 8) * 3),
        ("gl_weight", (ctypes.c_double * 8) * 3),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("classify_ms", ctypes.c_double),
        ("special_ms", ctypes.c_double),
        ("far_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("pairs", ctypes.c_int64 * 6),
        ("far_launches", ctypes.c_int32),
        ("n_colours_test", ctypes.c_int32),
        ("n_colours_trial", ctypes.c_int32),
        ("plan_bytes", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        return {
            "plan_bytes": int(self.plan_bytes),
            "classify_ms": self.classify_ms,
            "special_ms": self.special_ms,
            "far_ms": self.far_ms,
            "total_ms": self.total_ms,
            "pairs": {CLASS_NAMES[i]: int(self.pairs[i]) for i in range(6)},
            "far_launches": int(self.far_launches),
            "n_colours_test": int(self.n_colours_test),
            "n_colours_trial": int(self.n_colours_trial),
        }


class ServerOptions(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int32), ("memory_budget", ctypes.c_int64)]


class Request(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("xyz", ctypes.c_void_p),
        ("nv", ctypes.c_int64),
        ("tri", ctypes.c_void_p),
        ("nt", ctypes.c_int64),
        ("space_kind", ctypes.c_int32),
        ("gidx", ctypes.c_void_p),
        ("coeff", ctypes.c_void_p),
        ("nfun", ctypes.c_int64),
        ("op", Operator),
        ("opts", Options),
        ("dtype", ctypes.c_int32),
        ("matrix", ctypes.c_uint64),
        ("nrhs", ctypes.c_int32),
        ("x", ctypes.c_void_p),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("message", ctypes.c_char * 160),
        ("queue_ms", ctypes.c_double),
        ("prep_ms", ctypes.c_double),
        ("device_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("entries", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        return {"status": self.status, "message": self.message.decode(errors="replace"),
                "queue_ms": self.queue_ms, "prep_ms": self.prep_ms, "device_ms": self.device_ms,
                "total_ms": self.total_ms, "entries": int(self.entries)}


BATCH_MODES = ("per-colour", "persistent", "per-problem")


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("last_ms", ctypes.c_double),
        ("problems", ctypes.c_int32),
        ("buckets", ctypes.c_int32),
        ("singles", ctypes.c_int32),
        ("bucket_problems", ctypes.c_int32 * 8),
        ("bucket_mode", ctypes.c_int32 * 8),
        ("bucket_wgs_per_cu", ctypes.c_int32 * 8),
        ("bucket_tuned_ms", ctypes.c_double * 8),
        ("bucket_candidate_ms", (ctypes.c_double * 5) * 8),
    ]

    def as_dict(self) -> dict:
        nb = self.buckets
        return {"last_ms": self.last_ms, "problems": self.problems, "buckets": nb, "singles": self.singles,
                "bucket_problems": list(self.bucket_problems[:nb]),
                "bucket_mode": [BATCH_MODES[m] for m in self.bucket_mode[:nb]],
                "bucket_wgs_per_cu": list(self.bucket_wgs_per_cu[:nb]),
                "bucket_tuned_ms": list(self.bucket_tuned_ms[:nb]),
                # [per-colour, persistent 1, 2, 3 work-groups per CU, per-problem]; 0: not timed
                "bucket_candidate_ms": [list(self.bucket_candidate_ms[b]) for b in range(nb)]}


_lib = None
vp = ctypes.c_void_p
pp = ctypes.POINTER(ctypes.c_void_p)


def lib():
    """Load libbmi.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BMIError(EDEVICE, f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    i32, i64, st = ctypes.c_int32, ctypes.c_int64, ctypes.c_int32
    L.bmi_version.restype = ctypes.c_char_p
    L.bmi_last_error.restype = ctypes.c_char_p
    L.bmi_options_default.argtypes = [ctypes.POINTER(Options)]
    L.bmi_options_default.restype = None
    L.bmi_context_create.argtypes = [i32, pp]
    L.bmi_context_destroy.argtypes = [vp]
    L.bmi_mesh_create.argtypes = [vp, vp, i64, vp, i64, pp]
    L.bmi_mesh_free.argtypes = [vp]
    L.bmi_space_create.argtypes = [vp, i32, vp, vp, i64, pp]
    L.bmi_space_create_shaped.argtypes = [vp, i32, vp, vp, i64, i32, i32, pp]
    L.bmi_space_free.argtypes = [vp]
    L.bmi_plan_create.argtypes = [vp, ctypes.POINTER(Operator), vp, vp, ctypes.POINTER(Options), pp]
    L.bmi_plan_create_rules.argtypes = [vp, ctypes.POINTER(Operator), vp, vp, ctypes.POINTER(Options),
                                        ctypes.POINTER(Rules), pp]
    L.bmi_plan_free.argtypes = [vp]
    L.bmi_matrix_create.argtypes = [vp, i64, i64, i32, pp]
    L.bmi_matrix_free.argtypes = [vp]
    L.bmi_matrix_device_ptr.argtypes = [vp, pp, ctypes.POINTER(i64)]
    L.bmi_assemble.argtypes = [vp, vp]
    L.bmi_plan_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.bmi_matrix_to_host.argtypes = [vp, vp, i32]
    L.bmi_assemble_to_host.argtypes = [vp, vp, i32, i64]
    L.bmi_matrix_apply.argtypes = [vp, vp, vp, i32]
    L.bmi_matrix_apply_noalloc.argtypes = [vp, vp, vp, i32]
    L.bmi_matrix_compare.argtypes = [vp, vp, vp]
    L.bmi_device_memory.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
    L.bmi_server_create.argtypes = [i32, ctypes.POINTER(ServerOptions), pp]
    L.bmi_server_destroy.argtypes = [vp]
    L.bmi_server_submit.argtypes = [vp, ctypes.POINTER(Request), vp, ctypes.POINTER(ctypes.c_uint64)]
    L.bmi_server_wait.argtypes = [vp, ctypes.c_uint64, i32, ctypes.POINTER(Result)]
    L.bmi_server_release.argtypes = [vp, ctypes.c_uint64]
    L.bmi_server_memory.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
    L.bmi_request_bytes.argtypes = [ctypes.POINTER(Request), ctypes.POINTER(i64), ctypes.POINTER(i64)]
    L.bmi_batch_create.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), i32, pp]
    L.bmi_batch_assemble.argtypes = [vp]
    L.bmi_batch_stats_get.argtypes = [vp, ctypes.POINTER(BatchStats)]
    L.bmi_batch_set_schedule.argtypes = [vp, i32, i32, i32]
    L.bmi_batch_free.argtypes = [vp]
    for name in EXPORTS:
        f = getattr(L, name)
        if name not in ("bmi_version", "bmi_last_error", "bmi_options_default"):
            f.restype = st
    _lib = L
    return L


def check(status: int) -> None:
    if status != OK:
        msg = lib().bmi_last_error()
        raise BMIError(status, msg.decode() if msg else "")


def default_options() -> Options:
    o = Options()
    lib().bmi_options_default(ctypes.byref(o))
    return o
