"""ctypes binding of libtbfgpu.so (the C ABI declared in include/tbfgpu.h).

The library is loaded from this package directory; a missing library raises
instead of falling back to anything else.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtbfgpu.so")

TBFG_OK = 0
TBFG_EINVAL = 1
TBFG_ENOMEM = 2
TBFG_ENOTCONVERGED = 3
TBFG_ENUMERIC = 4
TBFG_EDEVICE = 5
TBFG_EIO = 6
PREC_FP64 = 0
PREC_MIXED = 1
PREC_MIXED_BF16 = 2

STATUS_NAMES = {0: "TBFG_OK", 1: "TBFG_EINVAL", 2: "TBFG_ENOMEM", 3: "TBFG_ENOTCONVERGED",
                4: "TBFG_ENUMERIC", 5: "TBFG_EDEVICE", 6: "TBFG_EIO"}


class TbfgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (STATUS_NAMES.get(code, code), msg))
        self.code = code


class GridDesc(C.Structure):
    _fields_ = [("n", C.c_int32 * 3), ("L", C.c_double * 3),
                ("stretch", C.c_int32 * 3), ("gamma", C.c_double)]

    @property
    def shape(self):
        return tuple(int(v) for v in self.n)

    @property
    def cells(self):
        return int(self.n[0]) * int(self.n[1]) * int(self.n[2])

    def __repr__(self):
        return "GridDesc(n=%s, L=%s, stretch=%s, gamma=%g)" % (
            list(self.n), list(self.L), list(self.stretch), self.gamma)


class FpsOpts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("refine_sweeps", C.c_int32)]


class PcgParams(C.Structure):
    _fields_ = [("n_levels", C.c_int32), ("n_pre", C.c_int32), ("n_post", C.c_int32),
                ("max_iter", C.c_int32), ("tol", C.c_double), ("mg_cycles", C.c_int32),
                ("precision", C.c_int32), ("coarse_max_iter", C.c_int32), ("flexible", C.c_int32),
                ("coarse_tol", C.c_double), ("preconditioner", C.c_int32), ("mg_max_iter", C.c_int32),
                ("mg_tol", C.c_double)]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I64x3 = C.POINTER(C.c_int64)
_P3 = C.POINTER(C.c_void_p)  # three device pointers (x, y, z components of a staggered field), or NULL

# name -> (restype, argtypes); mirrors include/tbfgpu.h
SIGNATURES = {
    "tbfg_last_error": (C.c_char_p, []),
    "tbfg_version": (C.c_char_p, []),
    "tbfg_device_count": (C.c_int, []),
    "tbfg_fps_create": (C.c_int, [C.POINTER(GridDesc), C.POINTER(FpsOpts), C.POINTER(_P)]),
    "tbfg_fps_solve": (C.c_int, [_P, _D, _I64x3, _I64x3, _D, _I64x3, _I64x3]),
    "tbfg_fps_solve_dev": (C.c_int, [_P, _P, _P, _P]),
    "tbfg_fps_apply_operator_dev": (C.c_int, [_P, _P, _P, _P]),
    "tbfg_fps_footprint": (C.c_int64, [_P]),
    "tbfg_fps_destroy": (C.c_int, [_P]),
    "tbfg_pcg_default_params": (None, [C.POINTER(PcgParams)]),
    "tbfg_pcg_create": (C.c_int, [C.POINTER(GridDesc), C.POINTER(PcgParams), C.POINTER(_P)]),
    "tbfg_pcg_solve": (C.c_int, [_P, _D, _I64x3, _I64x3, _D, _I64x3, _I64x3, _D, _I64x3, _I64x3,
                                 C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "tbfg_pcg_solve_dev": (C.c_int, [_P, _P, _P, _P, C.POINTER(C.c_int32), C.POINTER(C.c_double), _P]),
    "tbfg_pcg_solve_queued_dev": (C.c_int, [_P, _P, _P, _P, C.c_int32, _P]),
    "tbfg_pcg_result": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "tbfg_pcg_precondition_dev": (C.c_int, [_P, _P, _P, _P, _P]),
    "tbfg_pcg_footprint": (C.c_int64, [_P]),
    "tbfg_pcg_destroy": (C.c_int, [_P]),
    "tbfg_pcg_profile_begin": (C.c_int, [_P, C.c_int32]),
    "tbfg_pcg_coarse_stats": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_double)]),
    "tbfg_pcg_mg_stats": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "tbfg_pcg_profile_end": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "tbfg_apply_operator_dev": (C.c_int, [C.POINTER(GridDesc), _P, _P, _P, _P]),
    "tbfg_divergence_dev": (C.c_int, [C.POINTER(GridDesc), _P, _P, _P, _P, _P]),
    "tbfg_beta_dev": (C.c_int, [C.POINTER(GridDesc), C.c_double, _P, _P, _P]),
    "tbfg_project_mg_dev": (C.c_int, [C.POINTER(GridDesc), C.c_double, _P, _P, _P, _P, _P, _P]),
    "tbfg_max_divergence_dev": (C.c_int, [C.POINTER(GridDesc), _P, _P, _P, _D, _P]),
    "tbfg_max_divergence_noalloc_dev": (C.c_int, [C.POINTER(GridDesc), _P, _P, _P, _D, _P]),
    "tbfg_source_fast_dev": (C.c_int, [C.POINTER(GridDesc), C.c_double, C.c_double, _P, _P, _P, _P, _P, _P, _P3,
                                       _P3, _P, _P]),
    "tbfg_project_fast_dev": (C.c_int, [C.POINTER(GridDesc), C.c_double, C.c_double, _P, _P, _P, _P, _P3, _P3, _P,
                                        _P, _P, _P]),
    "tbfg_update_pressure_dev": (C.c_int, [C.POINTER(GridDesc), _P, _P, _P, _P, _P3, _P3, _P3, _P]),
    "tbfg_material_work_doubles": (C.c_int64, [C.POINTER(GridDesc)]),
    "tbfg_material_props_dev": (C.c_int, [C.POINTER(GridDesc), C.c_double, C.c_double, C.c_double, C.c_double, _P,
                                          _P, _P, _P, _P, _P]),
    "tbfg_tune_set": (C.c_int, [C.c_int32, C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32)]),
    "tbfg_tune_get": (C.c_int, [C.c_int32, C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32)]),
    "tbfg_tune_clear": (C.c_int, []),
    "tbfg_field_count": (C.c_int, [C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "tbfg_field_write": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int32), _D]),
    "tbfg_field_read": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int32), _D]),
    "tbfg_time_folder_name": (C.c_int, [C.c_int32, C.c_char_p, C.c_int32]),
    "tbfg_make_velocity": (C.c_int, [C.POINTER(GridDesc), C.c_uint64, _D, _D, _D]),
    "tbfg_make_bubble_density": (C.c_int, [C.POINTER(GridDesc), C.c_int32, C.c_double, C.c_uint64,
                                           C.c_int32, C.c_double, C.c_double, _D, _D]),
}

_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libtbfgpu.so is not built (run `python -m tbfsolver_amd.build`)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().tbfg_last_error().decode()


def check(rc, allow=()):
    if rc != TBFG_OK and rc not in allow:
        raise TbfgError(rc, last_error())
    return rc


def ptr(x):
    """Device pointer of a torch tensor (or an int address)."""
    if x is None:
        return None
    if isinstance(x, int):
        return C.c_void_p(x)
    return C.c_void_p(x.data_ptr())


def ptr3(v):
    """A staggered vector field (three tensors, or None) as a C array of three
    device pointers; None gives NULL."""
    if v is None:
        return None
    return (C.c_void_p * 3)(*(None if t is None else t.data_ptr() for t in v))


def stream_ptr(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(stream.cuda_stream)
