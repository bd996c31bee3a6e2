"""ctypes mirror of include/score_gm.h and loader of the in-tree libscore_gm.so.

The structs below must match the C header byte for byte; tests/test_abi.py
checks every offset against the C compiler's view.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscore_gm.so")

ABI_VERSION = 11

OK = 0
ERR_ARG = -1
ERR_OOM = -2
ERR_HIP = -3
ERR_CAPACITY = -4
ERR_UNSUPPORTED = -5
ERR_STATE = -6
ERR_NODEVICE = -7

MOB_SEBALDGOTTSTEIN = 1
MOB_ROLLETTHOLM = 2

MAP_CELLGRID = 0
MAP_RAW2GID = 1

LIST_INSIDE = 0
LIST_BORDER = 1

CURRENTLY_INFECTED = 4096002000

STATUS_NOALLOC_MAX = 1 << 22

# every symbol include/score_gm.h declares
EXPORTED = (
    "score_gm_abi_version",
    "score_gm_strerror",
    "score_gm_device_count",
    "score_gm_create",
    "score_gm_step",
    "score_gm_run",
    "score_gm_get_result",
    "score_gm_get_grainmap",
    "score_gm_get_counts",
    "score_gm_get_trajectory",
    "score_gm_get_grainlog",
    "score_gm_get_front",
    "score_gm_get_status",
    "score_gm_get_status_noalloc",
    "score_gm_batch_run",
    "score_gm_footprint",
    "score_gm_server_create",
    "score_gm_server_submit",
    "score_gm_server_wait",
    "score_gm_server_destroy",
    "score_gm_set_profiling",
    "score_gm_get_profile",
    "score_gm_destroy",
)

_u32p = C.POINTER(C.c_uint32)
_f64p = C.POINTER(C.c_double)


class Domain(C.Structure):
    _fields_ = [
        ("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
        ("cellsize", C.c_double),
        ("cellgrid", _u32p),
        ("n_ori", C.c_uint32),
        ("ori_quat", _f64p),
        ("n_defg", C.c_uint32),
        ("defg_ori", _u32p),
        ("defg_rho", _f64p),
        ("n_rxg", C.c_uint32),
        ("rxg_ori", _u32p),
        ("rxg_site", _u32p),
        ("rxg_step", _u32p),
    ]


class Params(C.Structure):
    _fields_ = [
        ("mobility_model", C.c_int32), ("zener_mode", C.c_int32),
        ("recovery_mode", C.c_int32), ("n_regions", C.c_int32),
        ("rh_hagb_m0", C.c_double), ("rh_hagb_hact", C.c_double),
        ("rh_cut", C.c_double), ("rh_trans", C.c_double), ("rh_exponent", C.c_double),
        ("lagb_m0", C.c_double), ("lagb_hact", C.c_double),
        ("hagb_m0", C.c_double), ("hagb_hact", C.c_double),
        ("gs_m0", C.c_double), ("gs_hact", C.c_double),
        ("G0", C.c_double), ("b0", C.c_double),
        ("n_sched", C.c_uint32), ("reserved1", C.c_uint32),
        ("sched_time", _f64p), ("sched_temp", _f64p),
        ("maxfill", C.c_double), ("xmax", C.c_double), ("tmax", C.c_double),
        ("nmax", C.c_int64),
        ("rel_cell_caching", C.c_double), ("rel_cell_recaching", C.c_double),
        ("n_defrag", C.c_uint32), ("reserved2", C.c_uint32),
        ("defrag_at", _f64p),
        ("zener_alpha", C.c_double), ("zener_gamma", C.c_double),
        ("n_disp", C.c_uint32), ("reserved3", C.c_uint32),
        ("disp_time", _f64p), ("disp_fr", _f64p),
        ("rec_vacancy_d0", C.c_double), ("rec_vacancy_hact", C.c_double),
        ("rec_solute_d0", C.c_double), ("rec_solute_hact", C.c_double),
        ("rec_solute_ls", C.c_double), ("rec_solute_conc", C.c_double),
        ("rec_alpha3", C.c_double), ("rec_kappa2", C.c_double), ("rec_c3", C.c_double), ("rec_c5", C.c_double),
        ("rec_tau", C.c_double),
        ("n_grainlog", C.c_uint32), ("reserved4", C.c_uint32),
        ("grainlog_at", _f64p),
    ]


class TrajRow(C.Structure):
    _fields_ = [
        ("step", C.c_uint32), ("sv", C.c_uint32),
        ("nfull_inside", C.c_uint32), ("nfull_border", C.c_uint32),
        ("nslots_inside", C.c_uint32), ("nslots_border", C.c_uint32),
        ("ninfected", C.c_uint32), ("npending", C.c_uint32),
        ("t", C.c_double), ("dt", C.c_double), ("X", C.c_double),
        ("xcells", C.c_double), ("temperature", C.c_double), ("rhomax", C.c_double),
        ("pmax", C.c_double), ("nplaced", C.c_uint32), ("reserved", C.c_uint32),
    ]


class GrainLogRow(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("X", C.c_double),
        ("logcnt", C.c_uint32), ("step", C.c_uint32), ("sv", C.c_uint32), ("ngrains", C.c_uint32),
    ]


class Result(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("X", C.c_double), ("xcells", C.c_double),
        ("step", C.c_uint32), ("completed", C.c_int32),
        ("sum_sv", C.c_uint64),
        ("nuclei_placed", C.c_uint32), ("windows", C.c_uint32),
        ("pmax", C.c_double), ("loop_seconds", C.c_double),
    ]


class BatchItem(C.Structure):
    _fields_ = [
        ("domain", C.POINTER(Domain)), ("params", C.POINTER(Params)),
        ("defg_counts", C.POINTER(C.c_uint32)), ("rxg_counts", C.POINTER(C.c_uint32)),
        ("grainmap", C.POINTER(C.c_uint32)),
        ("map_encoding", C.c_int32), ("status", C.c_int32),
        ("result", Result), ("seconds", C.c_double), ("queued_seconds", C.c_double),
    ]


# kernels of one growth step, in launch order (SCORE_GM_K_* in the header)
KERNELS = ("k_reanchor", "k_sweep", "k_sweep_scan", "k_compact", "k_propose", "k_decide",
           "k_decide_scan", "k_recycle", "k_commit")
NKERNELS = len(KERNELS)


class Profile(C.Structure):
    _fields_ = [
        ("steps", C.c_uint64),
        ("ms_kernel", C.c_double * NKERNELS), ("ms_step", C.c_double), ("ms_host", C.c_double),
        ("sum_sv", C.c_uint64), ("slots_scanned", C.c_uint64), ("n_full", C.c_uint64), ("n_recycled", C.c_uint64),
        ("n_attempts", C.c_uint64), ("n_new", C.c_uint64), ("n_predicted", C.c_uint64),
        ("n_reanchor_steps", C.c_uint64), ("n_reanchored", C.c_uint64), ("pending_resolved", C.c_uint64),
    ]


class ScoreGmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"score_gm error {code}: {msg}")
        self.code = code


_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libscore_gm.so (built in-tree by score_amd.build); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(path)
    lib.score_gm_abi_version.restype = C.c_int
    lib.score_gm_strerror.restype = C.c_char_p
    lib.score_gm_strerror.argtypes = [C.c_int]
    lib.score_gm_device_count.restype = C.c_int
    lib.score_gm_create.restype = C.c_int
    lib.score_gm_create.argtypes = [C.POINTER(Domain), C.POINTER(Params), C.c_int, C.POINTER(C.c_void_p)]
    lib.score_gm_step.restype = C.c_int
    lib.score_gm_step.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_int32)]
    lib.score_gm_run.restype = C.c_int
    lib.score_gm_run.argtypes = [C.c_void_p, C.POINTER(Result)]
    lib.score_gm_get_result.restype = C.c_int
    lib.score_gm_get_result.argtypes = [C.c_void_p, C.POINTER(Result)]
    lib.score_gm_get_grainmap.restype = C.c_int
    lib.score_gm_get_grainmap.argtypes = [C.c_void_p, _u32p, C.c_int32]
    lib.score_gm_get_counts.restype = C.c_int
    lib.score_gm_get_counts.argtypes = [C.c_void_p, _u32p, _u32p]
    lib.score_gm_get_trajectory.restype = C.c_int
    lib.score_gm_get_trajectory.argtypes = [C.c_void_p, C.POINTER(TrajRow), C.c_size_t, C.POINTER(C.c_size_t)]
    lib.score_gm_get_grainlog.restype = C.c_int
    lib.score_gm_get_grainlog.argtypes = [C.c_void_p, C.POINTER(GrainLogRow), _u32p, C.c_size_t,
                                          C.POINTER(C.c_size_t)]
    lib.score_gm_get_front.restype = C.c_int
    lib.score_gm_get_front.argtypes = [C.c_void_p, C.c_int32, C.c_uint32, C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint8), C.POINTER(C.c_uint8), _u32p, _f64p, _f64p, _u32p, _u32p]
    lib.score_gm_get_status.restype = C.c_int
    lib.score_gm_get_status.argtypes = [C.c_void_p, C.c_double, C.c_uint32, C.POINTER(C.c_uint32), _u32p, _u32p]
    lib.score_gm_get_status_noalloc.restype = C.c_int
    lib.score_gm_get_status_noalloc.argtypes = [C.c_void_p, C.c_double, C.c_uint32, C.POINTER(C.c_uint32), _u32p,
                                                _u32p]
    lib.score_gm_batch_run.restype = C.c_int
    lib.score_gm_batch_run.argtypes = [C.POINTER(BatchItem), C.c_uint32, C.c_int32, C.c_uint32]
    lib.score_gm_footprint.restype = C.c_int
    lib.score_gm_footprint.argtypes = [C.POINTER(Domain), C.POINTER(Params), C.POINTER(C.c_uint64)]
    lib.score_gm_server_create.restype = C.c_int
    lib.score_gm_server_create.argtypes = [C.c_int32, C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p)]
    lib.score_gm_server_submit.restype = C.c_int
    lib.score_gm_server_submit.argtypes = [C.c_void_p, C.POINTER(BatchItem), C.POINTER(C.c_uint64)]
    lib.score_gm_server_wait.restype = C.c_int
    lib.score_gm_server_wait.argtypes = [C.c_void_p, C.c_uint64, C.c_int32]
    lib.score_gm_server_destroy.restype = C.c_int
    lib.score_gm_server_destroy.argtypes = [C.c_void_p]
    lib.score_gm_set_profiling.restype = C.c_int
    lib.score_gm_set_profiling.argtypes = [C.c_void_p, C.c_int32]
    lib.score_gm_get_profile.restype = C.c_int
    lib.score_gm_get_profile.argtypes = [C.c_void_p, C.POINTER(Profile)]
    lib.score_gm_destroy.restype = None
    lib.score_gm_destroy.argtypes = [C.c_void_p]
    if lib.score_gm_abi_version() != ABI_VERSION:
        raise ScoreGmError(ERR_STATE, "ABI version mismatch between score_amd/abi.py and libscore_gm.so")
    _lib = lib
    return lib


def check(lib: C.CDLL, rc: int) -> int:
    if rc < 0:
        raise ScoreGmError(rc, lib.score_gm_strerror(rc).decode())
    return rc
