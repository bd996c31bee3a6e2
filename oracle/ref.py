"""TEST INFRASTRUCTURE ONLY: the reference's own random-number and orientation
code (SCORE_Random.cpp, SCORE_Math.cpp), built from the reference tree into
oracle/_ref/libscore_ref.so by `make -C oracle ref` with the C entry points of
oracle/ref_harness.cpp.  Absent where the reference tree is absent; the tests
that use it skip then."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_PATH = os.path.join(HERE, "_ref", "libscore_ref.so")
_lib = None


def available() -> bool:
    return os.path.exists(REF_PATH)


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        lib = C.CDLL(REF_PATH)
        f64p = C.POINTER(C.c_double)
        lib.ref_mt_stream.argtypes = [C.c_long, C.c_uint64, f64p]
        lib.ref_shoemake.argtypes = [C.c_long, C.c_uint64, f64p]
        lib.ref_disori_angle.argtypes = [C.c_uint64, f64p, f64p, f64p]
        lib.ref_disori_quat.argtypes = [C.c_uint64, f64p, f64p, f64p]
        lib.ref_euler2quat.argtypes = [C.c_uint64, f64p, f64p]
        for f in ("ref_mt_stream", "ref_shoemake", "ref_disori_angle", "ref_disori_quat", "ref_euler2quat"):
            getattr(lib, f).restype = C.c_int
        _lib = lib
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def mt_stream(seed: int, n: int) -> np.ndarray:
    """randomClass::init(seed), then n MersenneTwister() values"""
    out = np.empty(n)
    load().ref_mt_stream(seed, n, _p(out))
    return out


def shoemake(seed: int, n: int) -> np.ndarray:
    out = np.empty((n, 4))
    load().ref_shoemake(seed, n, _p(out))
    return out


def disori_angle(p: np.ndarray, q: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(p, np.float64).reshape(-1, 4)
    q = np.ascontiguousarray(q, np.float64).reshape(-1, 4)
    out = np.empty(len(p))
    load().ref_disori_angle(len(p), _p(p), _p(q), _p(out))
    return out


def disori_quat(p: np.ndarray, q: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(p, np.float64).reshape(-1, 4)
    q = np.ascontiguousarray(q, np.float64).reshape(-1, 4)
    out = np.empty((len(p), 4))
    load().ref_disori_quat(len(p), _p(p), _p(q), _p(out))
    return out


def euler2quat(bunge: np.ndarray) -> np.ndarray:
    b = np.ascontiguousarray(bunge, np.float64).reshape(-1, 3)
    out = np.empty((len(b), 4))
    load().ref_euler2quat(len(b), _p(b), _p(out))
    return out
