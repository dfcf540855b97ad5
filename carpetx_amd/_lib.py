"""ctypes binding of libcxp.so (the C ABI declared in include/cxp.h and
include/cxp_tag.h).

There is no fallback: importing the binding without the built library, or
calling a device entry point without a GPU, raises.
"""
import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcxp.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "cxp.h")
# every header whose functions libcxp.so exports
HEADERS = [HEADER, os.path.join(os.path.dirname(_HERE), "include", "cxp_tag.h")]

c_int = ctypes.c_int
c_int64 = ctypes.c_int64
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p
P_INT = ctypes.POINTER(ctypes.c_int)
P_INT64 = ctypes.POINTER(ctypes.c_int64)
P_DOUBLE = ctypes.POINTER(ctypes.c_double)

CXP_OK = 0
STATUS = {
    0: "CXP_OK", -1: "CXP_ERR_TYPE", -2: "CXP_ERR_ORDER", -3: "CXP_ERR_ARG",
    -4: "CXP_ERR_REGION", -5: "CXP_ERR_COARSE", -6: "CXP_ERR_NCOMP",
    -7: "CXP_ERR_RATIO", -8: "CXP_ERR_HIP", -9: "CXP_ERR_COMP",
    -10: "CXP_ERR_STATE", -11: "CXP_ERR_UNSUPPORTED", -12: "CXP_ERR_SCRATCH",
}


class CxpFab(ctypes.Structure):
    _fields_ = [("data", c_void_p), ("lo", c_int * 3), ("hi", c_int * 3),
                ("comp_stride", c_int64), ("ncomp", c_int)]


class CxpLevel(ctypes.Structure):
    _fields_ = [("nfabs", c_int), ("fabs", ctypes.POINTER(CxpFab)),
                ("valid", P_INT), ("domain_lo", c_int * 3), ("domain_hi", c_int * 3)]


class CxpBc(ctypes.Structure):
    _fields_ = [("type", (c_int * 3) * 2), ("dirichlet_values", P_DOUBLE),
                ("robin_values", P_DOUBLE), ("parities", P_INT),
                ("xmin", ctypes.c_double * 3), ("dx", ctypes.c_double * 3)]


class CxpBibRegion(ctypes.Structure):
    _fields_ = [("active", c_int), ("shape", c_int), ("num_levels", c_int),
                ("position", ctypes.c_double * 3), ("radius", ctypes.c_double * 3)]


class CxpError(RuntimeError):
    def __init__(self, func, status, msg):
        super().__init__("%s failed: %s (%d) %s" % (func, STATUS.get(status, "?"),
                                                    status, msg))
        self.status = status


_lib = None

_SIGS = {
    "cxp_version": (c_int, []),
    "cxp_status_string": (c_char_p, [c_int]),
    "cxp_op_lookup": (c_int, [c_char_p, c_int, P_INT, ctypes.POINTER(c_void_p)]),
    "cxp_op_lookup_fallback": (c_int, [c_char_p, c_int, P_INT, c_int,
                                       ctypes.POINTER(c_void_p)]),
    "cxp_group_op_lookup": (c_int, [c_char_p, c_char_p, c_int, P_INT,
                                    ctypes.POINTER(c_void_p)]),
    "cxp_op_describe": (c_int, [c_void_p, P_INT, P_INT, P_INT, P_INT, P_INT]),
    "cxp_coarse_box": (c_int, [c_void_p, P_INT, P_INT, P_INT, P_INT]),
    "cxp_prolongate": (c_int, [c_void_p, ctypes.POINTER(CxpFab), c_int,
                               ctypes.POINTER(CxpFab), c_int, c_int, P_INT, P_INT,
                               c_int, c_void_p]),
    "cxp_prolongate_regions": (c_int, [c_void_p, ctypes.POINTER(CxpFab), c_int,
                                       ctypes.POINTER(CxpFab), c_int, c_int, c_int,
                                       P_INT, P_INT, c_int, c_void_p]),
    "cxp_prolongate_adjoint_regions": (c_int, [c_void_p, ctypes.POINTER(CxpFab), c_int,
                                               ctypes.POINTER(CxpFab), c_int, c_int,
                                               c_int, P_INT, P_INT, c_int, c_void_p]),
    "cxp_restrict": (c_int, [P_INT, ctypes.POINTER(CxpFab), c_int,
                             ctypes.POINTER(CxpFab), c_int, c_int, P_INT, P_INT,
                             c_void_p]),
    "cxp_prolongate_adjoint": (c_int, [c_void_p, ctypes.POINTER(CxpFab),
                                       ctypes.POINTER(CxpFab), c_int,
                                       ctypes.POINTER(CxpFab), c_int, c_int,
                                       P_INT, P_INT, c_int, c_void_p]),
    "cxp_prolongate_adjoint_noalloc": (c_int, [c_void_p, ctypes.POINTER(CxpFab),
                                               ctypes.POINTER(CxpFab), c_int,
                                               ctypes.POINTER(CxpFab), c_int, c_int,
                                               P_INT, P_INT, c_int, c_void_p]),
    "cxp_branch_bytes": (c_int, [c_void_p, P_INT, P_INT, c_int, c_int, P_INT64]),
    "cxp_prolongate_record": (c_int, [c_void_p, ctypes.POINTER(CxpFab), c_int,
                                      ctypes.POINTER(CxpFab), c_int, c_int, P_INT,
                                      P_INT, c_int, c_void_p, c_void_p]),
    "cxp_prolongate_adjoint_recorded": (c_int, [c_void_p, c_void_p,
                                                ctypes.POINTER(CxpFab), c_int,
                                                ctypes.POINTER(CxpFab), c_int, c_int,
                                                P_INT, P_INT, c_int, c_void_p]),
    "cxp_restrict_adjoint": (c_int, [P_INT, ctypes.POINTER(CxpFab), c_int,
                                     ctypes.POINTER(CxpFab), c_int, c_int, P_INT,
                                     P_INT, c_void_p]),
    "cxp_worklist_create": (c_int, [ctypes.POINTER(c_void_p)]),
    "cxp_worklist_destroy": (c_int, [c_void_p]),
    "cxp_worklist_clear": (c_int, [c_void_p]),
    "cxp_worklist_append": (c_int, [c_void_p, c_void_p, ctypes.POINTER(CxpFab),
                                    c_int, ctypes.POINTER(CxpFab), c_int, c_int,
                                    P_INT, P_INT]),
    "cxp_worklist_append_shell": (c_int, [c_void_p, c_void_p,
                                          ctypes.POINTER(CxpFab), c_int,
                                          ctypes.POINTER(CxpFab), c_int, c_int,
                                          P_INT, P_INT]),
    "cxp_worklist_append_level": (c_int, [c_void_p, c_void_p,
                                          ctypes.POINTER(CxpLevel), c_int,
                                          ctypes.POINTER(CxpLevel), c_int, c_int]),
    "cxp_worklist_append_level_bc": (c_int, [c_void_p, c_void_p,
                                             ctypes.POINTER(CxpLevel), c_int,
                                             ctypes.POINTER(CxpLevel), c_int, c_int,
                                             ctypes.POINTER(CxpBc)]),
    "cxp_worklist_append_new_level": (c_int, [c_void_p, c_void_p,
                                              ctypes.POINTER(CxpLevel), c_int,
                                              ctypes.POINTER(CxpLevel), c_int, c_int,
                                              ctypes.POINTER(CxpBc)]),
    "cxp_worklist_append_remake_level": (c_int, [c_void_p, c_void_p,
                                                 ctypes.POINTER(CxpLevel), c_int,
                                                 ctypes.POINTER(CxpLevel),
                                                 ctypes.POINTER(CxpLevel), c_int, c_int,
                                                 ctypes.POINTER(CxpBc)]),
    "cxp_worklist_append_boundaries": (c_int, [c_void_p, ctypes.POINTER(CxpLevel), P_INT,
                                               c_int, c_int, ctypes.POINTER(CxpBc)]),
    "cxp_worklist_set_per_group": (c_int, [c_void_p, c_int]),
    "cxp_worklist_upload": (c_int, [c_void_p, c_void_p]),
    "cxp_worklist_stats": (c_int, [c_void_p, P_INT64, P_INT64, P_INT64, P_INT64]),
    "cxp_prolongate_batched": (c_int, [c_void_p, c_void_p]),
    "cxp_prolongate_per_patch": (c_int, [c_void_p, c_void_p]),
    "cxp_worklist_poison": (c_int, [c_void_p, ctypes.c_double, c_void_p]),
    "cxp_graph_capture_begin": (c_int, [c_void_p]),
    "cxp_graph_capture_end": (c_int, [c_void_p, ctypes.POINTER(c_void_p)]),
    "cxp_graph_launch": (c_int, [c_void_p, c_void_p]),
    "cxp_graph_destroy": (c_int, [c_void_p]),
    "cxp_autotune": (c_int, [c_void_p, c_int, c_void_p]),
    "cxp_tune_save": (c_int, [c_char_p]),
    "cxp_tune_load": (c_int, [c_char_p]),
    "cxp_tune_set": (c_int, [c_void_p, P_INT, P_INT, c_int, c_int]),
    "cxp_tune_clear": (c_int, []),
    "cxp_tile_config_count": (c_int, []),
    "cxp_tile_config": (c_int, [c_int, P_INT, P_INT, P_INT]),
    # include/cxp_tag.h
    "cxp_boxinbox_tag": (c_int, [ctypes.POINTER(CxpBibRegion), c_int, c_int, P_DOUBLE,
                                 P_DOUBLE, c_void_p, P_INT, P_INT, P_INT, P_INT, c_void_p]),
    "cxp_error_est": (c_int, [c_void_p, P_INT, P_INT, P_INT, P_INT, ctypes.c_double,
                              c_void_p, c_void_p, c_void_p]),
    "cxp_reflux_plan_create": (c_int, [ctypes.POINTER(CxpLevel),
                                       ctypes.POINTER(ctypes.POINTER(CxpFab)),
                                       ctypes.POINTER(CxpLevel),
                                       ctypes.POINTER(ctypes.POINTER(CxpFab)), c_int, c_int,
                                       P_DOUBLE, P_INT, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.POINTER(c_void_p)]),
    "cxp_reflux_run": (c_int, [c_void_p, c_void_p]),
    "cxp_reflux_plan_stats": (c_int, [c_void_p, P_INT64, P_INT64]),
    "cxp_reflux_plan_destroy": (c_int, [c_void_p]),
}


def header_symbols():
    """Function names declared in include/cxp.h and include/cxp_tag.h."""
    names = set()
    for h in HEADERS:
        with open(h) as f:
            names |= set(re.findall(r"\b(cxp_[a-z0-9_]+)\s*\(", f.read()))
    return sorted(names)


def lib():
    """Load libcxp.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcxp.so is not built (%s); run "
                               "`python -c 'import __graft_entry__ as g; g.build()'`"
                               % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(func, status):
    if status != CXP_OK:
        msg = lib().cxp_status_string(status)
        raise CxpError(func, status, msg.decode() if msg else "")
    return status


def i3(v):
    return (c_int * 3)(*[int(x) for x in v])
