"""ctypes binding of libvolprim_io.so (C ABI in include/vp_io.h): PLY scenes,
Radiance RGBE environments, OpenEXR images.  CPU only."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvolprim_io.so")
_LIB = None
EXPORTED = ["vp_io_load_ply", "vp_io_save_ply", "vp_io_load_hdr", "vp_io_save_hdr", "vp_io_save_exr",
            "vp_io_load_exr", "vp_io_free", "vp_io_last_error"]


class IoError(RuntimeError):
    pass


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise IoError(f"{LIB_PATH} is not built: run `make -C {HERE}`")
        lib = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        lib.vp_io_last_error.restype = ctypes.c_char_p
        lib.vp_io_load_ply.argtypes = [ctypes.c_char_p, ctypes.c_float, vp, ctypes.POINTER(vp),
                                       ctypes.POINTER(ctypes.c_uint32)]
        lib.vp_io_save_ply.argtypes = [ctypes.c_char_p, vp, ctypes.c_uint32, ctypes.c_float]
        lib.vp_io_load_hdr.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_uint32),
                                       ctypes.POINTER(ctypes.c_uint32)]
        lib.vp_io_save_hdr.argtypes = [ctypes.c_char_p, vp, ctypes.c_uint32, ctypes.c_uint32]
        lib.vp_io_save_exr.argtypes = [ctypes.c_char_p, vp, ctypes.c_uint32, ctypes.c_uint32]
        lib.vp_io_load_exr.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_uint32),
                                       ctypes.POINTER(ctypes.c_uint32)]
        lib.vp_io_free.argtypes = [vp]
        _LIB = lib
    return _LIB


def _check(rc, what):
    if rc != 0:
        raise IoError(f"{what}: {load().vp_io_last_error().decode(errors='replace')}")


def load_ply(path, sigma_multiplier=7.5, albedo_override=None):
    """-> N x 14 rows (centre, quat wxyz, linear scale, albedo, mass)."""
    out = ctypes.c_void_p()
    n = ctypes.c_uint32()
    ov = None if albedo_override is None else np.ascontiguousarray(albedo_override, np.float32)
    _check(load().vp_io_load_ply(path.encode(), sigma_multiplier, ov.ctypes.data if ov is not None else None,
                                 ctypes.byref(out), ctypes.byref(n)), "vp_io_load_ply")
    rows = np.ctypeslib.as_array(ctypes.cast(out, ctypes.POINTER(ctypes.c_float)), shape=(n.value, 14)).copy()
    load().vp_io_free(out)
    return rows


def save_ply(path, rows14, sigma_multiplier=1.0):
    rows14 = np.ascontiguousarray(rows14, np.float32).reshape(-1, 14)
    _check(load().vp_io_save_ply(path.encode(), rows14.ctypes.data, rows14.shape[0], sigma_multiplier),
           "vp_io_save_ply")


def load_hdr(path):
    out = ctypes.c_void_p()
    w, h = ctypes.c_uint32(), ctypes.c_uint32()
    _check(load().vp_io_load_hdr(path.encode(), ctypes.byref(out), ctypes.byref(w), ctypes.byref(h)),
           "vp_io_load_hdr")
    img = np.ctypeslib.as_array(ctypes.cast(out, ctypes.POINTER(ctypes.c_float)), shape=(h.value, w.value, 4)).copy()
    load().vp_io_free(out)
    return img


def save_hdr(path, rgba):
    rgba = np.ascontiguousarray(rgba, np.float32)
    _check(load().vp_io_save_hdr(path.encode(), rgba.ctypes.data, rgba.shape[1], rgba.shape[0]), "vp_io_save_hdr")


def save_exr(path, rgba):
    rgba = np.ascontiguousarray(rgba, np.float32)
    _check(load().vp_io_save_exr(path.encode(), rgba.ctypes.data, rgba.shape[1], rgba.shape[0]), "vp_io_save_exr")


def load_exr(path):
    out = ctypes.c_void_p()
    w, h = ctypes.c_uint32(), ctypes.c_uint32()
    _check(load().vp_io_load_exr(path.encode(), ctypes.byref(out), ctypes.byref(w), ctypes.byref(h)),
           "vp_io_load_exr")
    img = np.ctypeslib.as_array(ctypes.cast(out, ctypes.POINTER(ctypes.c_float)), shape=(h.value, w.value, 4)).copy()
    load().vp_io_free(out)
    return img
