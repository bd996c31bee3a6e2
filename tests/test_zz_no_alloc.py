"""score_gm_get_status_noalloc, the status entry point that allocates no device
memory, called through the C ABI: its output equals score_gm_get_status's on
two calls in a row (the second on a smaller front, so that the fixed staging
array still holds the first call's longer contents), and a front with more
occupied slots than the array holds is refused."""
import ctypes as C

import numpy as np
import pytest

from score_amd import abi, synth
from score_amd.synth import ParamsSpec

pytestmark = pytest.mark.gpu


def _status(lib, h, fn, thr):
    n = C.c_uint32(0)
    nul = C.cast(None, C.POINTER(C.c_uint32))
    abi.check(lib, fn(h, thr, 0, C.byref(n), nul, nul))
    ixyz, gid = np.full(n.value, 0xDEADBEEF, np.uint32), np.full(n.value, 0xDEADBEEF, np.uint32)
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))  # noqa: E731
    m = C.c_uint32(0)
    abi.check(lib, fn(h, thr, n.value, C.byref(m), p(ixyz), p(gid)))
    assert m.value == n.value
    return ixyz, gid


def _compare(gm, thr=0.5):
    lib = abi.load()
    h = gm._handle()
    a = _status(lib, h, lib.score_gm_get_status, thr)
    b = _status(lib, h, lib.score_gm_get_status_noalloc, thr)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    return len(a[0])


def test_noalloc_status_equals_status_twice_then_smaller(require_gpu):
    from score_amd.growth import GrowthMachine
    big = synth.make_voronoi_domain(72, 64, 56, n_grains=20, n_nuclei=90, seed=31)
    small = synth.make_voronoi_domain(30, 28, 26, n_grains=6, n_nuclei=12, seed=2)
    gb, gs = GrowthMachine(big, ParamsSpec()), GrowthMachine(small, ParamsSpec())
    gb.step(40)
    gs.step(25)
    n_big = _compare(gb)
    n_small = _compare(gs)
    print(f"status entries: {n_big} then {n_small}", flush=True)
    assert n_big > n_small > 0
    # again on the first handle, a few steps later, and at another threshold
    gb.step(3)
    _compare(gb, 0.25)
    gs.step(2)
    _compare(gs, 0.75)
    gb.close()
    gs.close()


def test_noalloc_status_refuses_more_slots_than_its_array(require_gpu):
    from score_amd.growth import GrowthMachine
    n = abi.STATUS_NOALLOC_MAX + 1000
    e = 170
    assert e ** 3 > n
    dom = synth.make_voronoi_domain(e, e, e, n_grains=4, n_nuclei=n, seed=7)
    g = GrowthMachine(dom, ParamsSpec())
    lib = abi.load()
    cnt = C.c_uint32(0)
    nul = C.cast(None, C.POINTER(C.c_uint32))
    rc = lib.score_gm_get_status_noalloc(g._handle(), 0.5, 0, C.byref(cnt), nul, nul)
    assert rc == abi.ERR_CAPACITY
    # the allocating entry point serves the same handle
    ixyz, gid = g.status()
    assert len(ixyz) == g.result()["nuclei_placed"] > abi.STATUS_NOALLOC_MAX
    g.close()
