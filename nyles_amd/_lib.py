"""ctypes binding of libnyles_hip.so (the C ABI of include/nyles_hip.h).

The library is the only compute path: importing this module fails loudly
when the shared object is missing, and every entry point raises
``NylesHipError`` with the library's message on a non-zero status.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnyles_hip.so")

# status codes, precision modes, field ids, reductions and ops (mirror the header)
ABI_VERSION = 3        # NYL_ABI_VERSION of include/nyles_hip.h
NYL_OK = 0
NYL_E_ARG, NYL_E_STATE, NYL_E_HIP, NYL_E_SIZE = -1, -2, -3, -4
NYL_PREC_FP64 = 0
NYL_PREC_MIXED = 1

(NYL_F_B, NYL_F_U, NYL_F_V, NYL_F_W, NYL_F_KE, NYL_F_VORX, NYL_F_VORY, NYL_F_VORZ,
 NYL_F_UCX, NYL_F_UCY, NYL_F_UCZ, NYL_F_P, NYL_F_DIV,
 NYL_F_DS0_B, NYL_F_DS0_U, NYL_F_DS0_V, NYL_F_DS0_W,
 NYL_F_DS1_B, NYL_F_DS1_U, NYL_F_DS1_V, NYL_F_DS1_W,
 NYL_F_DS2_B, NYL_F_DS2_U, NYL_F_DS2_V, NYL_F_DS2_W,
 NYL_F_MGX, NYL_F_MGB, NYL_F_MGY, NYL_NFIELDS) = range(29)

NYL_DT_F64 = 0
NYL_DT_F32 = 1

NYL_MAX_TRACERS = 8
NYL_TILE_CANDIDATES = 15
NYL_MODEL_LES, NYL_MODEL_LINEAR, NYL_MODEL_EULER = 0, 1, 2


def NYL_F_TRACER(q):
    return 64 + q


def NYL_F_DS_TRACER(slot, q):
    return 80 + NYL_MAX_TRACERS * slot + q


NYL_RED_MAXU2 = 0
NYL_RED_KE = 1
NYL_RED_BSUM = 2
NYL_RED_NONFINITE = 3

NYL_OP_U_FROM_U = 0
NYL_OP_DIV = 1
NYL_OP_CORRECT = 2
NYL_OP_VORTICITY = 3
NYL_OP_KINENERGY = 4
NYL_OP_PRESSURE = 5

# every symbol include/nyles_hip.h declares
EXPORTED = [
    "nyl_abi_version", "nyl_last_error", "nyl_ctx_create", "nyl_ctx_destroy",
    "nyl_workspace_bytes", "nyl_bind_workspace", "nyl_set_stream", "nyl_set_graphs",
    "nyl_field_offset",
    "nyl_mg_levels", "nyl_mg_level_shape", "nyl_rhs", "nyl_rk3_stage", "nyl_stage_update",
    "nyl_diagnose",
    "nyl_step_rk3", "nyl_step_ef", "nyl_step_lfam3", "nyl_lfam3_reset", "nyl_mg_solve", "nyl_apply", "nyl_reduce", "nyl_prof_start",
    "nyl_prof_stop", "nyl_debug_mg", "nyl_reduce_last", "nyl_sync",
    "nyl_batch_create", "nyl_batch_destroy", "nyl_batch_workspace_bytes",
    "nyl_batch_bind_workspace", "nyl_batch_set_stream", "nyl_batch_set_graphs",
    "nyl_batch_tile_candidate", "nyl_batch_autotune", "nyl_batch_set_tile",
    "nyl_batch_step_rk3",
    "nyl_reduce_noalloc",
]

# kernel classes of nyl_prof_start
KERNEL_CLASSES = {"stage": 1, "mg_sweep": 2, "mg_restrict": 3, "mg_prolong": 4, "diag": 5,
                  "div": 6, "correct": 7, "resnorm": 8, "mg_update": 9, "mg_sweep_fine": 10, "mg_sweep2": 11,
                  "mg_coarse": 12, "mg_sweep2_fine": 13}


class NylesHipError(RuntimeError):
    pass


class NylDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("periodic_x", ctypes.c_int32), ("periodic_y", ctypes.c_int32),
        ("periodic_z", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("idx2", ctypes.c_double), ("idy2", ctypes.c_double), ("idz2", ctypes.c_double),
        ("dx2", ctypes.c_double), ("fparameter", ctypes.c_double),
        ("precision", ctypes.c_int32), ("mg_maxite", ctypes.c_int32),
        ("mg_tol", ctypes.c_double), ("mg_omega", ctypes.c_double),
        ("viscosity", ctypes.c_double), ("n_tracers", ctypes.c_int32),
        ("model", ctypes.c_int32), ("upwind_a", ctypes.c_int32), ("upwind_vf", ctypes.c_int32),
    ]


class NylMGStats(ctypes.Structure):
    _fields_ = [("nite", ctypes.c_int32), ("total_cycles", ctypes.c_int32),
                ("res", ctypes.c_double), ("normb", ctypes.c_double)]

    def as_dict(self):
        return {"nite": self.nite, "total_cycles": self.total_cycles,
                "res": self.res, "normb": self.normb}


_lib = None


def load():
    """Load libnyles_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NylesHipError(
            "libnyles_hip.so not found at %s: build it with `python -m nyles_amd.build` "
            "(there is no CPU fallback)" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, ci, u64, i64, dbl = (ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int64,
                             ctypes.c_double)
    P = ctypes.POINTER
    sig = {
        "nyl_abi_version": (ci, []),
        "nyl_last_error": (ctypes.c_char_p, []),
        "nyl_ctx_create": (ci, [P(NylDesc), P(vp)]),
        "nyl_ctx_destroy": (ci, [vp]),
        "nyl_workspace_bytes": (ci, [vp, P(u64)]),
        "nyl_bind_workspace": (ci, [vp, vp, u64]),
        "nyl_set_stream": (ci, [vp, vp]),
        "nyl_set_graphs": (ci, [vp, ci]),
        "nyl_field_offset": (ci, [vp, ci, P(u64), P(i64), P(ctypes.c_int32)]),
        "nyl_mg_levels": (ci, [vp, P(ctypes.c_int32)]),
        "nyl_mg_level_shape": (ci, [vp, ci, P(ctypes.c_int32), P(ctypes.c_int32),
                                    P(ctypes.c_int32)]),
        "nyl_rhs": (ci, [vp, ci, ci]),
        "nyl_rk3_stage": (ci, [vp, ci, dbl]),
        "nyl_stage_update": (ci, [vp, ci, dbl, dbl]),
        "nyl_diagnose": (ci, [vp, P(NylMGStats)]),
        "nyl_step_rk3": (ci, [vp, dbl, P(NylMGStats)]),
        "nyl_step_ef": (ci, [vp, dbl, P(NylMGStats)]),
        "nyl_step_lfam3": (ci, [vp, dbl, P(NylMGStats)]),
        "nyl_lfam3_reset": (ci, [vp]),
        "nyl_mg_solve": (ci, [vp, P(NylMGStats)]),
        "nyl_apply": (ci, [vp, ci]),
        "nyl_reduce": (ci, [vp, ci, P(dbl)]),
        "nyl_prof_start": (ci, [vp, ci]),
        "nyl_prof_stop": (ci, [vp, P(dbl), P(i64), P(dbl)]),
        "nyl_debug_mg": (ci, [vp, ci]),
        "nyl_reduce_last": (ci, [vp, P(dbl)]),
        "nyl_sync": (ci, [vp]),
        "nyl_batch_create": (ci, [P(vp), ctypes.c_int32, P(vp)]),
        "nyl_batch_destroy": (ci, [vp]),
        "nyl_batch_workspace_bytes": (ci, [vp, P(u64)]),
        "nyl_batch_bind_workspace": (ci, [vp, vp, u64]),
        "nyl_batch_set_stream": (ci, [vp, vp]),
        "nyl_batch_set_graphs": (ci, [vp, ci]),
        "nyl_batch_tile_candidate": (ci, [ctypes.c_int32, P(ctypes.c_int32)]),
        "nyl_batch_autotune": (ci, [vp, P(ctypes.c_int32), P(dbl)]),
        "nyl_batch_set_tile": (ci, [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
        "nyl_batch_step_rk3": (ci, [vp, P(dbl), P(NylMGStats)]),
        "nyl_reduce_noalloc": (ci, [P(NylDesc), ctypes.c_int32, vp, vp, vp, vp, ctypes.c_int32,
                                    vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.nyl_abi_version() != ABI_VERSION:
        raise NylesHipError("libnyles_hip.so has ABI %d, this binding expects %d: rebuild it"
                            % (L.nyl_abi_version(), ABI_VERSION))
    _lib = L
    return L


def check(rc):
    if rc != NYL_OK:
        msg = load().nyl_last_error()
        raise NylesHipError("libnyles_hip error %d: %s" % (rc, msg.decode() if msg else ""))
    return rc
