"""ctypes binding of libtwoldg.so, the HIP library behind include/twoldg.h.

The library must be built (``__graft_entry__.build()`` or ``make -C
treewaveoldg.jl_amd/csrc``); importing this module never falls back to anything
else: a missing library raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libtwoldg.so")

ABI_VERSION = 4
F64, F32C, SPLIT = 0, 1, 2
FACE_SAME, FACE_COARSER, FACE_FINER, FACE_MIRROR, FACE_GLUE, FACE_DIRICHLET = 1, 2, 3, 4, 5, 6
FACE_RADIATIVE = 7
SLOT_STATE, SLOT_DU, SLOT_AUX, SLOT_ARG = 0, 1, 2, 3
FLAG_GENERIC = 1
XFER_COPY, XFER_REFINE, XFER_COARSEN = 0, 1, 2


class Face(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("nface", ctypes.c_int32),
                ("map", ctypes.c_int32 * 2), ("off", ctypes.c_int32 * 3),
                ("pad", ctypes.c_int32), ("nb", ctypes.c_int64 * 4)]


class Desc(ctypes.Structure):
    _fields_ = [("abi_version", ctypes.c_int32), ("D", ctypes.c_int32), ("P", ctypes.c_int32),
                ("E", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("truncate", ctypes.c_int32), ("nblocks", ctypes.c_int64),
                ("spacing", ctypes.c_void_p), ("faces", ctypes.c_void_p),
                ("coef", ctypes.c_void_p), ("flat", ctypes.c_void_p),
                ("tslot", ctypes.c_void_p), ("ntables", ctypes.c_int64),
                ("tmodes", ctypes.c_void_p), ("tweights", ctypes.c_void_p),
                ("tchol", ctypes.c_void_p), ("dirichlet_x", ctypes.c_void_p),
                ("wave_k", ctypes.c_double * 3), ("wave_phi", ctypes.c_double),
                ("hspeed", ctypes.c_double), ("flags", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("dcoef", ctypes.c_void_p),
                ("lmode", ctypes.c_void_p), ("ldual", ctypes.c_void_p),
                ("fslot", ctypes.c_void_p), ("nfactored", ctypes.c_int64),
                ("fradial", ctypes.c_void_p), ("frmodes", ctypes.c_void_p),
                ("frduals", ctypes.c_void_p), ("faslot", ctypes.c_void_p),
                ("nangular", ctypes.c_int64), ("famodes", ctypes.c_void_p),
                ("faweights", ctypes.c_void_p), ("fachol", ctypes.c_void_p),
                ("rad_kappa", ctypes.c_double), ("rad_invR", ctypes.c_double)]


class TransferBlock(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("child", ctypes.c_int32),
                ("src", ctypes.c_int64 * 8)]


class TwoldgError(RuntimeError):
    pass


_LIB = None
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_dbl = ctypes.c_double


def lib():
    """The loaded library; raises if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise TwoldgError(
                f"{LIBPATH} is not built: run __graft_entry__.build() (hipcc for gfx950)")
        L = ctypes.CDLL(LIBPATH)
        sig = {
            "twoldg_abi_version": ([], ctypes.c_int),
            "twoldg_last_error": ([], ctypes.c_char_p),
            "twoldg_device_count": ([_vp], ctypes.c_int),
            "twoldg_problem_create": ([_vp, _vp], ctypes.c_int),
            "twoldg_problem_destroy": ([_vp], ctypes.c_int),
            "twoldg_problem_info": ([_vp, _vp, _vp, _vp], ctypes.c_int),
            "twoldg_state_upload": ([_vp, ctypes.c_int, _vp], ctypes.c_int),
            "twoldg_state_download": ([_vp, ctypes.c_int, _vp], ctypes.c_int),
            "twoldg_state_copy": ([_vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
            "twoldg_rhs": ([_vp, ctypes.c_int, ctypes.c_int, _dbl], ctypes.c_int),
            "twoldg_truncate": ([_vp, ctypes.c_int], ctypes.c_int),
            "twoldg_rk4": ([_vp, _dbl, _dbl, _i64], ctypes.c_int),
            "twoldg_energy": ([_vp, ctypes.c_int, _dbl, _vp], ctypes.c_int),
            "twoldg_norms": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp,
                              _vp], ctypes.c_int),
            "twoldg_norms_noalloc": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, _vp, _vp], ctypes.c_int),
            "twoldg_timestep": ([_vp, _dbl, _vp], ctypes.c_int),
            "twoldg_synchronize": ([_vp], ctypes.c_int),
            "twoldg_problem_poison": ([_vp], ctypes.c_int),
            "twoldg_modes": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp,
                              _dbl, _vp], ctypes.c_int),
            "twoldg_transfer": ([_vp, ctypes.c_int, _vp, ctypes.c_int, _vp], ctypes.c_int),
            "twoldg_exchange_plan": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
            "twoldg_node_coordinates": ([_vp, _vp, _vp, _vp], ctypes.c_int),
            "twoldg_rk4_timed": ([_vp, _dbl, _dbl, _i64, ctypes.c_int, _vp, _vp, _vp],
                                 ctypes.c_int),
            "twoldg_batch_create": ([_vp, _i64, _vp, _vp], ctypes.c_int),
            "twoldg_batch_destroy": ([_vp], ctypes.c_int),
            "twoldg_batch_upload": ([_vp, _i64, _vp], ctypes.c_int),
            "twoldg_batch_download": ([_vp, _i64, _vp], ctypes.c_int),
            "twoldg_batch_rk4": ([_vp, _i64, _vp], ctypes.c_int),
            "twoldg_batch_info": ([_vp, _vp, _vp, _vp, _vp], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        if L.twoldg_abi_version() != ABI_VERSION:
            raise TwoldgError("libtwoldg ABI version mismatch")
        _LIB = L
    return _LIB


def check(rc):
    """Raise with the library's reason on a negative status (the reference's
    ArgumentError for refusals)."""
    if rc != 0:
        msg = lib().twoldg_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(msg)
        raise TwoldgError(f"twoldg error {rc}: {msg}")
    return rc


def exported_symbols():
    """Names declared in include/twoldg.h that the library must export."""
    return ["twoldg_abi_version", "twoldg_last_error", "twoldg_device_count",
            "twoldg_problem_create", "twoldg_problem_destroy", "twoldg_problem_info",
            "twoldg_state_upload", "twoldg_state_download", "twoldg_state_copy", "twoldg_rhs",
            "twoldg_truncate", "twoldg_rk4", "twoldg_energy", "twoldg_norms",
            "twoldg_norms_noalloc", "twoldg_timestep",
            "twoldg_synchronize", "twoldg_problem_poison", "twoldg_modes", "twoldg_transfer", "twoldg_exchange_plan", "twoldg_node_coordinates",
            "twoldg_rk4_timed",
            "twoldg_batch_create", "twoldg_batch_destroy", "twoldg_batch_upload",
            "twoldg_batch_download", "twoldg_batch_rk4", "twoldg_batch_info"]
