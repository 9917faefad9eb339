"""ctypes declarations of the C ABI in include/rsfm_amd.h."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librsfm_amd.so")


class RsfmError(RuntimeError):
    pass


if not os.path.exists(LIB_PATH):
    raise ImportError(
        "librsfm_amd.so is not built (%s); run `python rapidsfm_amd/build.py` "
        "(or __graft_entry__.build()) first" % LIB_PATH
    )

lib = ctypes.CDLL(LIB_PATH)

c_void_p = ctypes.c_void_p
c_u32 = ctypes.c_uint32
c_size = ctypes.c_size_t
c_int = ctypes.c_int
c_float = ctypes.c_float


class BestMatch(ctypes.Structure):
    _fields_ = [("index", c_u32), ("distance", c_u32)]


class ValidMatch(ctypes.Structure):
    _fields_ = [("idxQuery", c_u32), ("idxTrain", c_u32), ("distance", c_float)]


class BfTask(ctypes.Structure):
    _fields_ = [
        ("nbDescA", c_u32),
        ("nbDescB", c_u32),
        ("descA", c_void_p),
        ("descB", c_void_p),
        ("bestMatchA", c_void_p),
        ("bestMatchB", c_void_p),
        ("sqrNormDescA", c_void_p),
        ("sqrNormDescB", c_void_p),
    ]


class BfTop2Task(ctypes.Structure):
    _fields_ = [("base", BfTask), ("secondMatchA", c_void_p)]


class RansacTask(ctypes.Structure):
    _fields_ = [
        ("pts", c_void_p),
        ("nbPts", c_u32),
        ("nbPtsDev", c_void_p),
        ("sampleIdx", c_void_p),
        ("nbHyp", c_u32),
        ("counts", c_void_p),
        ("models", c_void_p),
        ("best", c_void_p),
        ("inlierMask", c_void_p),
    ]


class FinishTask(ctypes.Structure):
    _fields_ = [
        ("pts", c_void_p),
        ("nbPts", c_u32),
        ("nbPtsDev", c_void_p),
        ("sampleIdx", c_void_p),
        ("counts", c_void_p),
        ("models", c_void_p),
        ("nbHyp", c_u32),
        ("F", c_void_p),
        ("result", c_void_p),
        ("inlierMask", c_void_p),
    ]


class DegensacTask(ctypes.Structure):
    _fields_ = [
        ("pts", c_void_p),
        ("nbPts", c_u32),
        ("nbPtsDev", c_void_p),
        ("sampleIdx", c_void_p),
        ("counts", c_void_p),
        ("nbHyp", c_u32),
        ("F", c_void_p),
        ("H", c_void_p),
        ("result", c_void_p),
        ("inlierMaskF", c_void_p),
        ("inlierMaskH", c_void_p),
    ]


class Camera(ctypes.Structure):
    _fields_ = [("fx", c_float), ("fy", c_float), ("cx", c_float), ("cy", c_float)]


class PoseHTask(ctypes.Structure):
    _fields_ = [
        ("H", c_void_p),
        ("cam0", Camera),
        ("cam1", Camera),
        ("pts", c_void_p),
        ("nbPts", c_u32),
        ("nbPtsDev", c_void_p),
        ("inlierMask", c_void_p),
        ("poses", c_void_p),
        ("normals", c_void_p),
        ("result", c_void_p),
    ]


class RealCamera(ctypes.Structure):
    _fields_ = [("pinHole", Camera), ("distortion", c_float * 5)]


class UndistortTask(ctypes.Structure):
    _fields_ = [("pts", c_void_p), ("nbPts", c_u32), ("nbPtsDev", c_void_p), ("cam0", RealCamera),
                ("cam1", RealCamera)]


class PoseTask(ctypes.Structure):
    _fields_ = [
        ("F", c_void_p),
        ("cam0", Camera),
        ("cam1", Camera),
        ("pts", c_void_p),
        ("nbPts", c_u32),
        ("nbPtsDev", c_void_p),
        ("inlierMask", c_void_p),
        ("poses", c_void_p),
        ("result", c_void_p),
    ]


class FilterTask(ctypes.Structure):
    _fields_ = [
        ("fwd", c_void_p),
        ("second", c_void_p),
        ("nbQueryDesc", c_u32),
        ("bwd", c_void_p),
        ("nbTrainDesc", c_u32),
        ("out", c_void_p),
        ("count", c_void_p),
        ("kpQuery", c_void_p),
        ("kpTrain", c_void_p),
        ("pts", c_void_p),
    ]


class SampleTask(ctypes.Structure):
    _fields_ = [
        ("sampleIdx", c_void_p),
        ("nbHyp", c_u32),
        ("nbPts", c_u32),
        ("nbPtsDev", c_void_p),
        ("seed", ctypes.c_uint64),
    ]


class WSampleTask(ctypes.Structure):
    _fields_ = [
        ("sampleIdx", c_void_p),
        ("nbHyp", c_u32),
        ("nbPts", c_u32),
        ("nbPtsDev", c_void_p),
        ("weights", c_void_p),
        ("cdf", c_void_p),
        ("seed", ctypes.c_uint64),
    ]


assert ctypes.sizeof(BfTask) == 56
assert ctypes.sizeof(BestMatch) == 8
assert ctypes.sizeof(ValidMatch) == 12

_SIGS = {
    "rsfm_last_error": (ctypes.c_char_p, []),
    "rsfm_version": (ctypes.c_char_p, []),
    "rsfm_sqrnorm_u8": (c_int, [c_void_p, c_void_p, c_u32, c_void_p]),
    "rsfm_sqrnorm_u4": (c_int, [c_void_p, c_void_p, c_u32, c_void_p]),
    "rsfm_desc8b_to_4b": (c_int, [c_void_p, c_void_p, c_u32, c_void_p]),
    "rsfm_bf_match_u8": (c_int, [c_void_p, c_size, c_void_p]),
    "rsfm_bf_match_u4": (c_int, [c_void_p, c_size, c_void_p]),
    "rsfm_bf_match_idp_u8": (c_int, [c_void_p, c_size, c_void_p]),
    "rsfm_bf_match_top2_workspace_size": (c_int, [c_void_p, c_size, ctypes.POINTER(c_size)]),
    "rsfm_bf_match_top2": (c_int, [c_void_p, c_size, c_void_p, c_size, c_void_p]),
    "rsfm_cross_check_host": (c_size, [c_void_p, c_u32, c_void_p, c_u32, c_void_p]),
    "rsfm_remove_conflicts_host": (c_size, [c_void_p, c_u32, c_u32, c_void_p]),
    "rsfm_filter_matches_workspace_size": (c_int, [c_u32, ctypes.POINTER(c_size)]),
    "rsfm_filter_matches": (
        c_int,
        [c_void_p, c_void_p, c_u32, c_void_p, c_u32, c_u32, c_u32, c_void_p, c_void_p, c_void_p, c_size, c_void_p],
    ),
    "rsfm_ransac_f8": (c_int, [c_void_p, c_size, c_float, c_int, c_float, c_float, c_void_p]),
    "rsfm_ransac_f8_finish": (c_int, [c_void_p, c_size, c_float, c_int, c_float, c_u32, c_int, c_void_p]),
    "rsfm_ransac_f7": (c_int, [c_void_p, c_size, c_float, c_int, c_float, c_float, c_void_p]),
    "rsfm_ransac_h4": (c_int, [c_void_p, c_size, c_float, c_float, c_float, c_void_p]),
    "rsfm_pose_from_f": (c_int, [c_void_p, c_size, c_float, c_void_p]),
    "rsfm_pose_from_h": (c_int, [c_void_p, c_size, c_float, c_void_p]),
    "rsfm_undistort_in_place": (c_int, [c_void_p, c_void_p, c_size, c_void_p, c_u32, c_void_p]),
    "rsfm_undistort_pt_pairs": (c_int, [c_void_p, c_size, c_u32, c_void_p]),
    "rsfm_degensac_refine": (
        c_int, [c_void_p, c_size, c_float, c_float, c_int, c_u32, c_u32, ctypes.c_uint64, c_void_p]
    ),
    "rsfm_ransac_sample_indices": (c_int, [c_void_p, c_u32, c_u32, c_void_p, ctypes.c_uint64, c_void_p]),
    "rsfm_filter_matches_batch": (c_int, [c_void_p, c_size, c_u32, c_u32, c_void_p]),
    "rsfm_ransac_sample_indices_batch": (c_int, [c_void_p, c_size, c_void_p]),
    "rsfm_ransac_sample_weighted_batch": (c_int, [c_void_p, c_size, c_void_p]),
    "rsfm_mlp_param_count": (c_int, []),
    "rsfm_mlp_forward": (c_int, [c_void_p, c_void_p, c_u32, c_void_p, c_void_p]),
    "rsfm_mlp_grad_workspace_size": (c_int, [c_u32, ctypes.POINTER(c_size)]),
    "rsfm_mlp_grad": (c_int, [c_void_p, c_void_p, c_void_p, c_u32, c_void_p, c_void_p, c_void_p, c_size, c_void_p]),
    "rsfm_mlp_adam_step": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_u32, c_float, c_float, c_float, c_float,
                                   c_void_p]),
}


def _bind(name, restype, argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


for _name, (_res, _args) in _SIGS.items():
    _bind(_name, _res, _args)


def bind_extra(name, restype, argtypes):
    """Declare a symbol not listed above (used by the other op modules)."""
    return _bind(name, restype, argtypes)


def check(rc, what=""):
    if rc != 0:
        msg = lib.rsfm_last_error().decode(errors="replace")
        raise RsfmError("%s failed (%d): %s" % (what, rc, msg))
