"""ctypes binding of libwaterlily_hip.so (the C ABI of include/wl.h).

The library is required: importing this module raises when it is missing or cannot be loaded,
so no code path silently falls back to anything else.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwaterlily_hip.so")

WL_F32, WL_F64 = 0, 1
WL_QUICK, WL_VANLEER, WL_CDS = 0, 1, 2
FIELDS = {"u": 0, "u0": 1, "f": 2, "p": 3, "sigma": 4, "V": 5, "mu0": 6, "mu1": 7}
PFIELDS = {"x": 0, "L": 1, "D": 2, "iD": 3, "eps": 4, "r": 5, "z": 6}
WL_SOLVER_PLAIN, WL_SOLVER_MIXED = 0, 1
WL_BODY_SPHERE, WL_BODY_ELLIPSOID = 1, 2
WL_ENAN = -3
MFIELDS = {"P": 0, "U": 1, "UU": 2, "tau": 3}
WL_FN_UBC, WL_FN_ACC0, WL_FN_ACC1 = 0, 1, 2
WL_WARM_GUESS, WL_WARM_RECORD, WL_WARM_STATS, WL_WARM_PREDICTOR = 1, 2, 4, 8


class WlBody(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("center", C.c_double * 3),
                ("radius", C.c_double), ("axes", C.c_double * 3)]


class WlFlowDesc(C.Structure):
    _fields_ = [("D", C.c_int32), ("n", C.c_int32 * 3), ("dtype", C.c_int32), ("scheme", C.c_int32),
                ("perdir_mask", C.c_int32), ("exit_bc", C.c_int32), ("has_g", C.c_int32),
                ("has_uBC_rate", C.c_int32), ("uBC", C.c_double * 3), ("uBC_rate", C.c_double * 3),
                ("g0", C.c_double * 3), ("g1", C.c_double * 3), ("nu", C.c_double), ("dt0", C.c_double)]


class WlKernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 40), ("level", C.c_int32), ("reserved", C.c_int32), ("launches", C.c_int64),
                ("total_ms", C.c_double), ("bytes_per_launch", C.c_double)]


class WaterLilyError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_P = C.POINTER
_vp, _i, _d, _sz = C.c_void_p, C.c_int, C.c_double, C.c_size_t
# wl_udf_fn: int (*)(void* user, wl_flow_t h, int phase, double t)
UDF_FN = C.CFUNCTYPE(_i, _vp, _vp, _i, _d)
SIGNATURES = {
    "wl_last_error": (C.c_char_p, []),
    "wl_device_count": (_i, [_P(_i)]),
    "wl_flow_create": (_i, [_P(_vp), _P(WlFlowDesc)]),
    "wl_flow_destroy": (_i, [_vp]),
    "wl_flow_field_size": (_i, [_vp, _i, _P(_sz)]),
    "wl_flow_upload": (_i, [_vp, _i, _vp, _sz]),
    "wl_flow_download": (_i, [_vp, _i, _vp, _sz]),
    "wl_flow_measure": (_i, [_vp, _P(WlBody), _i, _d, _d]),
    "wl_flow_pois_update": (_i, [_vp]),
    "wl_flow_set_solver": (_i, [_vp, _i, _d, _i]),
    "wl_mom_step": (_i, [_vp, _d, _d, _d, _P(_d), _P(C.c_int16)]),
    "wl_pressure_force": (_i, [_vp, _P(WlBody), _i, _d, _P(_d)]),
    "wl_flow_metrics": (_i, [_vp, _P(WlBody), _i, _d, _P(_d), _P(_d), _P(_d), _P(_d), _P(_d)]),
    "wl_flow_metrics_noalloc": (_i, [_vp, _P(WlBody), _i, _d, _P(_d), _P(_d), _P(_d), _P(_d), _P(_d)]),
    "wl_meanflow_create": (_i, [_P(_vp), _vp, _i]),
    "wl_meanflow_destroy": (_i, [_vp]),
    "wl_meanflow_update": (_i, [_vp, _vp, _d]),
    "wl_meanflow_reset": (_i, [_vp]),
    "wl_meanflow_download": (_i, [_vp, _i, _vp, _sz]),
    "wl_meanflow_copy_to_flow": (_i, [_vp, _vp]),
    "wl_meanflow_upload": (_i, [_vp, _i, _vp, _sz]),
    "wl_flow_set_field_fn": (_i, [_vp, _i, _vp, _sz]),
    "wl_flow_apply_ic": (_i, [_vp]),
    "wl_flow_cfl": (_i, [_vp, _P(_d)]),
    "wl_flow_set_udf": (_i, [_vp, UDF_FN, _vp]),
    "wl_flow_autotune": (_i, [_vp, _i, _P(C.c_int32)]),
    "wl_flow_set_tiles": (_i, [_vp, C.c_int32, C.c_int32]),
    "wl_batch_create": (_i, [_P(_vp), _P(_vp), _i]),
    "wl_batch_destroy": (_i, [_vp]),
    "wl_batch_step": (_i, [_vp, _P(_d), _P(_d), _P(_d), _P(_d), _P(C.c_int16)]),
    "wl_flow_last_residuals": (_i, [_vp, _P(_d), _i, _P(_i)]),
    "wl_flow_profile": (_i, [_vp, _i]),
    "wl_flow_profile_read": (_i, [_vp, _P(WlKernelStat), _i, _P(_i)]),
    "wl_self_test": (_i, [_i, _P(C.c_int64)]),
    "wl_pois_create": (_i, [_P(_vp), _i, _P(C.c_int32), _i, _vp, _i, _i]),
    "wl_pois_destroy": (_i, [_vp]),
    "wl_pois_nlevels": (_i, [_vp, _P(_i)]),
    "wl_pois_level_dims": (_i, [_vp, _i, _P(C.c_int32)]),
    "wl_pois_upload": (_i, [_vp, _i, _i, _vp, _sz]),
    "wl_pois_download": (_i, [_vp, _i, _i, _vp, _sz]),
    "wl_pois_mult": (_i, [_vp, _vp, _vp, _sz]),
    "wl_pois_update": (_i, [_vp]),
    "wl_pois_solve": (_i, [_vp, _d, _i, _P(C.c_int32)]),
    "wl_net_create": (_i, [_P(_vp), _P(C.c_int32), _i]),
    "wl_net_destroy": (_i, [_vp]),
    "wl_net_num_weights": (_i, [_vp, _P(C.c_int64)]),
    "wl_net_set_weights": (_i, [_vp, _vp, C.c_int64]),
    "wl_net_get_weights": (_i, [_vp, _vp, C.c_int64]),
    "wl_net_forward": (_i, [_vp, _vp, _vp]),
    "wl_net_gradient": (_i, [_vp, _vp, _vp, _P(_d), _vp]),
    "wl_net_adam": (_i, [_vp, _d]),
    "wl_net_train": (_i, [_vp, _i, _d, C.c_uint64, _P(_d)]),
    "wl_net_pairs": (_i, [_vp, _P(_i)]),
    "wl_net_get_pair": (_i, [_vp, _i, _vp, _vp]),
    "wl_flow_set_warmstart": (_i, [_vp, _vp, _i]),
    "wl_flow_warm_stats": (_i, [_vp, _P(_d), _i, _P(_i)]),
}

_lib = None


def load():
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = load().wl_last_error().decode(errors="replace")
        if rc == -1:
            raise AssertionError(msg)
        if rc == WL_ENAN:
            raise FloatingPointError(msg)
        raise WaterLilyError(rc, msg)
    return rc
