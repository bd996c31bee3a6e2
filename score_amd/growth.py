"""Host-side mirror of SCORE's growth-machine call on one domain.

`GrowthMachine` is the Python face of caHdl::solve_RXGROWTH
(SCORE_Kernel.cpp:6497-6792): construct it from the domain state after
solve_REPLACE_CA_STRUCTURE, call `solve_rxgrowth()` (or `step()`), then read
the grain map, grain counts and trajectory.  Every call goes through the C ABI
of libscore_gm.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .synth import DomainArrays, ParamsSpec

TRAJ_DTYPE = np.dtype([
    ("step", np.uint32), ("sv", np.uint32), ("nfull_inside", np.uint32), ("nfull_border", np.uint32),
    ("nslots_inside", np.uint32), ("nslots_border", np.uint32), ("ninfected", np.uint32), ("npending", np.uint32),
    ("t", np.float64), ("dt", np.float64), ("X", np.float64), ("xcells", np.float64),
    ("temperature", np.float64), ("rhomax", np.float64), ("pmax", np.float64), ("nplaced", np.uint32),
    ("reserved", np.uint32),
], align=True)
assert TRAJ_DTYPE.itemsize == C.sizeof(abi.TrajRow)

GRAINLOG_DTYPE = np.dtype([("t", np.float64), ("X", np.float64), ("logcnt", np.uint32), ("step", np.uint32),
                           ("sv", np.uint32), ("ngrains", np.uint32)], align=True)
assert GRAINLOG_DTYPE.itemsize == C.sizeof(abi.GrainLogRow)


def result_dict(r: abi.Result) -> dict:
    return {k: getattr(r, k) for k, _ in abi.Result._fields_ if not k.startswith("reserved")}


class GrowthMachine:
    """One domain on one GPU; see include/score_gm.h for the contract."""

    def __init__(self, dom: DomainArrays, prm: ParamsSpec | None = None, device: int = 0):
        self._lib = abi.load()
        prm = prm or ParamsSpec()
        self.dom = dom
        self.prm = prm
        cd, self._dkeep = dom.to_c()
        cp, self._pkeep = prm.to_c()
        h = C.c_void_p()
        abi.check(self._lib, self._lib.score_gm_create(C.byref(cd), C.byref(cp), device, C.byref(h)))
        self._h = h

    def _handle(self):
        if self._h is None:
            raise abi.ScoreGmError(abi.ERR_STATE, "handle closed")
        return self._h

    def step(self, n: int = 1) -> bool:
        """Advance up to n growth steps; True once the termination test fired."""
        done = C.c_int32(0)
        abi.check(self._lib, self._lib.score_gm_step(self._handle(), n, C.byref(done)))
        return bool(done.value)

    def solve_rxgrowth(self) -> dict:
        """Run to completion (SCORE's solve_RXGROWTH)."""
        r = abi.Result()
        abi.check(self._lib, self._lib.score_gm_run(self._handle(), C.byref(r)))
        return result_dict(r)

    def result(self) -> dict:
        r = abi.Result()
        abi.check(self._lib, self._lib.score_gm_get_result(self._handle(), C.byref(r)))
        return result_dict(r)

    def grainmap(self, encoding: int = abi.MAP_CELLGRID) -> np.ndarray:
        out = np.empty(self.dom.N, np.uint32)
        abi.check(self._lib, self._lib.score_gm_get_grainmap(
            self._handle(), out.ctypes.data_as(C.POINTER(C.c_uint32)), encoding))
        return out

    def counts(self) -> tuple[np.ndarray, np.ndarray]:
        dg = np.empty(self.dom.n_defg, np.uint32)
        rg = np.empty(max(1, self.dom.n_rxg), np.uint32)
        abi.check(self._lib, self._lib.score_gm_get_counts(
            self._handle(), dg.ctypes.data_as(C.POINTER(C.c_uint32)), rg.ctypes.data_as(C.POINTER(C.c_uint32))))
        return dg, rg[: self.dom.n_rxg]

    def trajectory(self) -> np.ndarray:
        n = C.c_size_t(0)
        abi.check(self._lib, self._lib.score_gm_get_trajectory(self._handle(), None, 0, C.byref(n)))
        out = np.zeros(n.value, TRAJ_DTYPE)
        if n.value:
            abi.check(self._lib, self._lib.score_gm_get_trajectory(
                self._handle(), out.ctypes.data_as(C.POINTER(abi.TrajRow)), n.value, C.byref(n)))
        return out

    def grainlog(self) -> tuple[np.ndarray, np.ndarray]:
        """The grain-evolution log (mygrainevolution): one row per log point and
        the cell counts at it, deformed grains first (rows x (n_defg + n_rxg))."""
        n = C.c_size_t(0)
        abi.check(self._lib, self._lib.score_gm_get_grainlog(self._handle(), None, None, 0, C.byref(n)))
        rows = np.zeros(n.value, GRAINLOG_DTYPE)
        counts = np.zeros((n.value, self.dom.n_defg + self.dom.n_rxg), np.uint32)
        if n.value:
            abi.check(self._lib, self._lib.score_gm_get_grainlog(
                self._handle(), rows.ctypes.data_as(C.POINTER(abi.GrainLogRow)),
                counts.ctypes.data_as(C.POINTER(C.c_uint32)), n.value, C.byref(n)))
        return rows, counts

    def front(self, which: int) -> dict:
        n = C.c_uint32(0)
        nul = lambda t: C.cast(None, C.POINTER(t))  # noqa: E731
        abi.check(self._lib, self._lib.score_gm_get_front(self._handle(), which, 0, C.byref(n), nul(C.c_uint8),
                                                          nul(C.c_uint8), nul(C.c_uint32), nul(C.c_double),
                                                          nul(C.c_double), nul(C.c_uint32), nul(C.c_uint32)))
        m = n.value
        out = dict(activity=np.zeros(m, np.uint8), infector=np.zeros(m, np.uint8), cell=np.zeros(m, np.uint32),
                   rxfrac=np.zeros(m, np.float64), P=np.zeros(m, np.float64), defg=np.zeros(m, np.uint32),
                   rxg=np.zeros(m, np.uint32))
        if m:
            p = lambda k, t: out[k].ctypes.data_as(C.POINTER(t))  # noqa: E731
            abi.check(self._lib, self._lib.score_gm_get_front(
                self._handle(), which, m, C.byref(n), p("activity", C.c_uint8), p("infector", C.c_uint8),
                p("cell", C.c_uint32), p("rxfrac", C.c_double), p("P", C.c_double), p("defg", C.c_uint32),
                p("rxg", C.c_uint32)))
        return out

    def status(self, rxthreshold: float = 0.5):
        """(linear cell index, grain id shown) of every ACTIVE front record, Inside
        list then Border list in slot order (ompcrit_clarify_status_infectedcells,
        SCORE_Io_Characterization.cpp:925-997)."""
        n = C.c_uint32(0)
        nul = C.cast(None, C.POINTER(C.c_uint32))
        abi.check(self._lib, self._lib.score_gm_get_status(self._handle(), rxthreshold, 0, C.byref(n), nul, nul))
        ixyz, gid = np.zeros(n.value, np.uint32), np.zeros(n.value, np.uint32)
        if n.value:
            p = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))  # noqa: E731
            abi.check(self._lib, self._lib.score_gm_get_status(self._handle(), rxthreshold, n.value, C.byref(n),
                                                               p(ixyz), p(gid)))
        return ixyz, gid

    def set_profiling(self, on: bool = True) -> None:
        abi.check(self._lib, self._lib.score_gm_set_profiling(self._handle(), 1 if on else 0))

    def profile(self) -> dict:
        p = abi.Profile()
        abi.check(self._lib, self._lib.score_gm_get_profile(self._handle(), C.byref(p)))
        out = {k: getattr(p, k) for k, _ in abi.Profile._fields_ if k != "ms_kernel"}
        out["ms_kernel"] = {name: p.ms_kernel[i] for i, name in enumerate(abi.KERNELS)}
        return out

    def close(self) -> None:
        if self._h is not None:
            self._lib.score_gm_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_count() -> int:
    return abi.load().score_gm_device_count()


def same_trajectory(a: np.ndarray, b: np.ndarray) -> bool:
    """Bitwise equality of two trajectories, ignoring the host-resolution counter
    (the serial oracle resolves weights inline and always reports 0 there)."""
    if len(a) != len(b):
        return False
    a, b = a.copy(), b.copy()
    a["npending"] = 0
    b["npending"] = 0
    return a.tobytes() == b.tobytes()


def footprint(dom: DomainArrays, prm: ParamsSpec | None = None) -> int:
    """Device bytes one domain needs at most (score_gm_footprint)."""
    lib = abi.load()
    cd, kd = dom.to_c()
    cp, kp = (prm or ParamsSpec()).to_c()
    b = C.c_uint64(0)
    abi.check(lib, lib.score_gm_footprint(C.byref(cd), C.byref(cp), C.byref(b)))
    return b.value


class Request:
    """One submitted domain; its arrays live here until the request is complete."""

    def __init__(self, dom: DomainArrays, prm: ParamsSpec, grainmap: bool, encoding: int):
        self.dom = dom
        self._cd, self._kd = dom.to_c()
        self._cp, self._kp = prm.to_c()
        self.defg_counts = np.zeros(dom.n_defg, np.uint32)
        self._rxg = np.zeros(max(dom.n_rxg, 1), np.uint32)
        self.grainmap = np.zeros(dom.N, np.uint32) if grainmap else None
        it = abi.BatchItem()
        it.domain = C.pointer(self._cd)
        it.params = C.pointer(self._cp)
        it.defg_counts = self.defg_counts.ctypes.data_as(C.POINTER(C.c_uint32))
        it.rxg_counts = self._rxg.ctypes.data_as(C.POINTER(C.c_uint32))
        it.grainmap = self.grainmap.ctypes.data_as(C.POINTER(C.c_uint32)) if grainmap else None
        it.map_encoding = encoding
        self.item = it
        self.ticket = None

    @property
    def rxg_counts(self):
        return self._rxg[: self.dom.n_rxg]

    def outcome(self) -> dict:
        it = self.item
        return dict(status=it.status, result=result_dict(it.result), seconds=it.seconds,
                    queued_seconds=it.queued_seconds, defg_counts=self.defg_counts, rxg_counts=self.rxg_counts,
                    grainmap=self.grainmap)


class Server:
    """Request loop over arriving domains (score_gm_server_*): each starts when a
    worker is free and its footprint fits the memory budget next to the running ones."""

    def __init__(self, max_inflight: int = 8, hbm_budget: int = 0, device: int = 0):
        self._lib = abi.load()
        s = C.c_void_p()
        abi.check(self._lib, self._lib.score_gm_server_create(device, max_inflight, hbm_budget, C.byref(s)))
        self._s = s
        self._reqs = []

    def submit(self, dom: DomainArrays, prm: ParamsSpec | None = None, grainmap: bool = False,
               encoding: int = abi.MAP_CELLGRID) -> Request:
        r = Request(dom, prm or ParamsSpec(), grainmap, encoding)
        t = C.c_uint64(0)
        abi.check(self._lib, self._lib.score_gm_server_submit(self._s, C.byref(r.item), C.byref(t)))
        r.ticket = t.value
        self._reqs.append(r)
        return r

    def wait(self, req: Request, timeout: float | None = None) -> bool:
        ms = -1 if timeout is None else int(timeout * 1000)
        return bool(abi.check(self._lib, self._lib.score_gm_server_wait(self._s, req.ticket, ms)))

    def close(self) -> None:
        if self._s is not None:
            self._lib.score_gm_server_destroy(self._s)
            self._s = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_run(domains, params=None, max_inflight: int = 8, device: int = 0, grainmaps: bool = False,
              encoding: int = abi.MAP_CELLGRID) -> list[dict]:
    """Every domain to completion through score_gm_batch_run (the rank's loop
    over its domains in ensembleHdl::SIMULATE_myCAs, SCORE_Kernel.cpp:1693-1757),
    up to `max_inflight` at a time.  `params` is one ParamsSpec for all or one per
    domain.  Returns per domain: status, result, seconds, counts and optionally the
    grain map.  Raises on the first failed domain, after all have run."""
    lib = abi.load()
    n = len(domains)
    if params is None or isinstance(params, ParamsSpec):
        params = [params or ParamsSpec()] * n
    if len(params) != n:
        raise ValueError("one ParamsSpec per domain, or one for all")
    keep, items = [], (abi.BatchItem * max(n, 1))()
    outs = []
    for i, (dom, prm) in enumerate(zip(domains, params)):
        cd, kd = dom.to_c()
        cp, kp = prm.to_c()
        dc = np.zeros(dom.n_defg, np.uint32)
        rc = np.zeros(max(dom.n_rxg, 1), np.uint32)
        gm = np.zeros(dom.N, np.uint32) if grainmaps else None
        keep.append((cd, kd, cp, kp))
        it = items[i]
        it.domain = C.pointer(cd)
        it.params = C.pointer(cp)
        it.defg_counts = dc.ctypes.data_as(C.POINTER(C.c_uint32))
        it.rxg_counts = rc.ctypes.data_as(C.POINTER(C.c_uint32))
        it.grainmap = gm.ctypes.data_as(C.POINTER(C.c_uint32)) if gm is not None else None
        it.map_encoding = encoding
        outs.append((dc, rc, gm))
    status = lib.score_gm_batch_run(items, n, device, max_inflight)
    res = []
    for i, (dom, (dc, rc, gm)) in enumerate(zip(domains, outs)):
        it = items[i]
        res.append(dict(status=it.status, result=result_dict(it.result), seconds=it.seconds,
                        queued_seconds=it.queued_seconds, defg_counts=dc, rxg_counts=rc[: dom.n_rxg], grainmap=gm))
    if status != abi.OK:
        abi.check(lib, status)
    return res
