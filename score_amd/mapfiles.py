"""Grain-map files in SCORE's formats, through libscore_io.so (include/score_io.h).

The map is what GrowthMachine.grainmap(abi.MAP_RAW2GID) returns: uint32, 1 for
an infected cell, else 2 + grain id, x fastest.  SCORE writes it as a headerless
uint32 stack (ompcrit_write_voxeldata_coloring_grainids,
SCORE_Io_Characterization.cpp:583-676), as an HDF5 dataset
"/GROWTH3DGrainID/grainid_<step>" in "SCORE.<simid>.VoxelData.h5"
(SCORE_Io_HDF5.cpp:126-262, :862-935) and as an XDMF description beside it
(SCORE_Io_HDF5.cpp:715-797).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscore_io.so")
GROUP = "/GROWTH3DGrainID"
_lib = None


class ScoreIoError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"score_io error {code}: {what}")
        self.code = code


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is missing (built by score_amd.build when HDF5 is present)")
        lib = C.CDLL(LIB_PATH)
        u32p = C.POINTER(C.c_uint32)
        lib.score_io_raw_name.argtypes = [C.c_char_p, C.c_size_t, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double,
                                          C.c_char_p]
        lib.score_io_write_raw.argtypes = [C.c_char_p, u32p, C.c_uint32, C.c_uint32, C.c_uint32]
        lib.score_io_h5_name.argtypes = [C.c_char_p, C.c_size_t, C.c_uint32]
        lib.score_io_xdmf_name.argtypes = [C.c_char_p, C.c_size_t, C.c_uint32]
        lib.score_io_write_h5_grainid.argtypes = [C.c_char_p, C.c_char_p, C.c_uint32, u32p, C.c_uint32, C.c_uint32,
                                                  C.c_uint32, u32p]
        lib.score_io_read_h5_dims.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_uint64)]
        lib.score_io_read_h5_u32.argtypes = [C.c_char_p, C.c_char_p, u32p, C.c_size_t]
        lib.score_io_write_xdmf.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, u32p, C.c_uint32]
        for f in ("score_io_raw_name", "score_io_write_raw", "score_io_h5_name", "score_io_xdmf_name",
                  "score_io_write_h5_grainid", "score_io_read_h5_dims", "score_io_read_h5_u32", "score_io_write_xdmf"):
            getattr(lib, f).restype = C.c_int
        _lib = lib
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise ScoreIoError(rc, what)


def _u32(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _name(fn, *args) -> str:
    buf = C.create_string_buffer(512)
    _check(fn(buf, 512, *args), "file name")
    return buf.value.decode()


def raw_name(simid: int, jobid: int, step: int, X: float, postfix: str = "GROWTH.3DGrainID") -> str:
    return _name(load().score_io_raw_name, simid, jobid, step, X, postfix.encode())


def h5_name(simid: int) -> str:
    return _name(load().score_io_h5_name, simid)


def xdmf_name(simid: int) -> str:
    return _name(load().score_io_xdmf_name, simid)


def _map3(gmap: np.ndarray, shape) -> np.ndarray:
    nx, ny, nz = shape
    a = np.ascontiguousarray(gmap, dtype=np.uint32).reshape(-1)
    if a.size != nx * ny * nz:
        raise ValueError("map size does not match (nx, ny, nz)")
    return a


def write_raw(path: str, gmap: np.ndarray, shape) -> None:
    a = _map3(gmap, shape)
    _check(load().score_io_write_raw(path.encode(), _u32(a), *shape), f"write {path}")


def write_h5_grainid(h5path: str, step: int, gmap: np.ndarray, shape, window=None, group: str = GROUP) -> str:
    """Writes the map (or its inclusive window (xmin, xmax, ymin, ymax, zmin, zmax));
    returns the dataset path."""
    a = _map3(gmap, shape)
    w = None if window is None else np.ascontiguousarray(window, np.uint32)
    _check(load().score_io_write_h5_grainid(h5path.encode(), group.encode(), step, _u32(a), *shape,
                                            None if w is None else _u32(w)), f"write {h5path}")
    return f"{group}/grainid_{step}"


def read_h5(h5path: str, dataset: str) -> np.ndarray:
    """The dataset's values (x fastest) and its HDF5 dimensions, as (values, dims)."""
    lib = load()
    dims = (C.c_uint64 * 3)()
    _check(lib.score_io_read_h5_dims(h5path.encode(), dataset.encode(), dims), f"read {h5path}:{dataset}")
    d = tuple(int(x) for x in dims)
    out = np.empty(d[0] * d[1] * d[2], np.uint32)
    _check(lib.score_io_read_h5_u32(h5path.encode(), dataset.encode(), _u32(out), out.size), f"read {dataset}")
    return out, d


def write_xdmf(path: str, h5name: str, snapshots, group: str = GROUP) -> None:
    """snapshots: (nx, ny, nz, step) per grain-ID snapshot in the HDF5 file"""
    s = np.ascontiguousarray(np.asarray(snapshots, np.uint32).reshape(-1, 4))
    _check(load().score_io_write_xdmf(path.encode(), h5name.encode(), group.encode(), _u32(s), len(s)),
           f"write {path}")
