"""The C ABI library: it loads, exports every function include/score_gm.h declares,
its ctypes mirror matches the C struct layouts, and it refuses to run without a
gfx950 device instead of falling back to anything."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from score_amd import abi, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "score_gm.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(score_gm_\w+)\s*\(", txt, flags=re.M)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(abi.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    if shutil.which("nm"):
        out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True, text=True).stdout
        exported = set(re.findall(r"\bT (score_gm_\w+)", out))
        assert set(declared_functions()) <= exported


def test_library_has_gfx950_code_object():
    data = open(abi.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_pure_host_entry_points():
    lib = abi.load()
    assert lib.score_gm_abi_version() == abi.ABI_VERSION
    assert lib.score_gm_strerror(abi.ERR_CAPACITY)
    assert lib.score_gm_strerror(0) == b"ok"


LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "score_gm.h"
#define F(S, f) printf("%s.%s %zu\n", #S, #f, offsetof(S, f));
#define Z(S) printf("%s.sizeof %zu\n", #S, sizeof(S));
int main(void) {
@BODY@
  return 0;
}
"""


def test_struct_layouts_match_c():
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    structs = {"score_gm_domain": abi.Domain, "score_gm_params": abi.Params, "score_gm_traj_row": abi.TrajRow,
               "score_gm_result": abi.Result, "score_gm_profile": abi.Profile,
               "score_gm_batch_item": abi.BatchItem, "score_gm_grainlog_row": abi.GrainLogRow}
    body = []
    for cname, py in structs.items():
        for f, _ in py._fields_:
            body.append(f"F({cname}, {f})")
        body.append(f"Z({cname})")
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write(LAYOUT_C.replace("@BODY@", "\n".join(body)))
        subprocess.run([cc, "-std=c99", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.strip().splitlines())
    for cname, py in structs.items():
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)
        assert int(got[f"{cname}.sizeof"]) == C.sizeof(py), cname


def _has_gpu():
    return os.path.exists("/dev/kfd")


@pytest.mark.skipif(_has_gpu(), reason="checks the behaviour without a GPU")
def test_create_fails_loudly_without_device():
    from score_amd.growth import GrowthMachine
    dom = synth.make_voronoi_domain(8, 8, 8, n_grains=2, n_nuclei=2, seed=1)
    with pytest.raises(abi.ScoreGmError) as e:
        GrowthMachine(dom)
    assert e.value.code in (abi.ERR_NODEVICE, abi.ERR_HIP)


def test_invalid_arguments_rejected_before_device_use():
    from score_amd.growth import GrowthMachine
    dom = synth.make_voronoi_domain(8, 8, 8, n_grains=2, n_nuclei=2, seed=1)
    dom.cellgrid = dom.cellgrid.copy()
    dom.cellgrid[5] = 7            # not a deformed-grain id
    with pytest.raises(abi.ScoreGmError) as e:
        GrowthMachine(dom)
    assert e.value.code == abi.ERR_ARG
    # thresholds announced but not given; drag without its table or with a table
    # that does not start at t = 0; a recovery model SCORE does not have
    lib = abi.load()
    dom = synth.make_voronoi_domain(8, 8, 8, n_grains=2, n_nuclei=2, seed=1)
    cd, keep_d = dom.to_c()
    cp, keep_p = synth.ParamsSpec(defrag_at=(0.5,)).to_c()
    cp.defrag_at = C.cast(None, C.POINTER(C.c_double))
    h = C.c_void_p()
    assert lib.score_gm_create(C.byref(cd), C.byref(cp), 0, C.byref(h)) == abi.ERR_ARG
    cp, keep_p = synth.ParamsSpec().to_c()
    cp.zener_mode = 1
    assert lib.score_gm_create(C.byref(cd), C.byref(cp), 0, C.byref(h)) == abi.ERR_ARG
    cp, keep_p = synth.ParamsSpec(zener_mode=2, disp_time=(10.0, 20.0), disp_fr=(1e5, 2e5)).to_c()
    assert lib.score_gm_create(C.byref(cd), C.byref(cp), 0, C.byref(h)) == abi.ERR_ARG
    cp, keep_p = synth.ParamsSpec(zener_mode=2, disp_time=(0.0, 0.0), disp_fr=(1e5, 2e5)).to_c()
    assert lib.score_gm_create(C.byref(cd), C.byref(cp), 0, C.byref(h)) == abi.ERR_ARG
    cp, keep_p = synth.ParamsSpec().to_c()
    cp.zener_mode = 3
    assert lib.score_gm_create(C.byref(cd), C.byref(cp), 0, C.byref(h)) == abi.ERR_ARG
    cp, keep_p = synth.ParamsSpec().to_c()
    cp.recovery_mode = 5
    assert lib.score_gm_create(C.byref(cd), C.byref(cp), 0, C.byref(h)) == abi.ERR_ARG
    # exponential recovery without a positive time constant; log points announced but not given
    cp, keep_p = synth.ParamsSpec(recovery_mode=4, rec_tau=0.0).to_c()
    assert lib.score_gm_create(C.byref(cd), C.byref(cp), 0, C.byref(h)) == abi.ERR_ARG
    cp, keep_p = synth.ParamsSpec(grainlog_at=(10.0,)).to_c()
    cp.grainlog_at = C.cast(None, C.POINTER(C.c_double))
    assert lib.score_gm_create(C.byref(cd), C.byref(cp), 0, C.byref(h)) == abi.ERR_ARG
    # NMAX outside the range SCORE accepts, [1, DEFAULT_NMAX] (SCORE_Kernel.cpp:274)
    for nmax in (0, -5, 1000001):
        cp, keep_p = synth.ParamsSpec(nmax=nmax).to_c()
        assert lib.score_gm_create(C.byref(cd), C.byref(cp), 0, C.byref(h)) == abi.ERR_ARG
    # region counts outside [0, 1024], or regions 10 cells thick or less (8^3 at T = 2)
    for T in (-1, 1025, 2):
        cp, keep_p = synth.ParamsSpec(n_regions=T).to_c()
        assert lib.score_gm_create(C.byref(cd), C.byref(cp), 0, C.byref(h)) == abi.ERR_ARG
    # nuclei arriving after step 0 have no counterpart in SCORE's region path
    dom2 = synth.make_voronoi_domain(12, 12, 24, n_grains=2, n_nuclei=3, seed=1)
    dom2.rxg_step = np.array([0, 2, 0], np.uint32)
    cd2, keep_d2 = dom2.to_c()
    cp, keep_p = synth.ParamsSpec(n_regions=2).to_c()
    assert lib.score_gm_create(C.byref(cd2), C.byref(cp), 0, C.byref(h)) == abi.ERR_UNSUPPORTED
    assert not h.value


def test_oracle_nmax_range_and_loop_bound():
    """the oracle rejects NMAX outside [1, DEFAULT_NMAX] as SCORE's parser does;
    NMAX = 1 runs steps 0 and 1 (the termination test is step > NMAX)"""
    from oracle.oracle import Oracle, OracleError
    dom = synth.make_voronoi_domain(8, 8, 8, n_grains=2, n_nuclei=2, seed=1)
    for nmax in (0, 1000001):
        with pytest.raises(OracleError) as e:
            Oracle(dom, synth.ParamsSpec(nmax=nmax))
        assert e.value.code == abi.ERR_ARG
    o = Oracle(dom, synth.ParamsSpec(nmax=1, xmax=1.0))
    r = o.solve_rxgrowth()
    assert r["step"] == 2 and len(o.trajectory()) == 2
    o.close()


@pytest.mark.skipif(_has_gpu(), reason="checks the behaviour without a GPU")
def test_batch_reports_every_domain_without_device():
    lib = abi.load()
    doms = [synth.make_voronoi_domain(8, 8, 8, n_grains=2, n_nuclei=2, seed=s) for s in (1, 2, 3)]
    keep, items = [], (abi.BatchItem * 3)()
    for i, d in enumerate(doms):
        cd, kd = d.to_c()
        cp, kp = synth.ParamsSpec().to_c()
        keep.append((cd, kd, cp, kp))
        items[i].domain, items[i].params = C.pointer(cd), C.pointer(cp)
    rc = lib.score_gm_batch_run(items, 3, 0, 2)
    assert rc in (abi.ERR_NODEVICE, abi.ERR_HIP)
    assert all(items[i].status == rc for i in range(3))
    assert lib.score_gm_batch_run(None, 0, 0, 0) == abi.OK


def test_footprint_covers_the_domain_arrays():
    from score_amd.growth import footprint
    small = synth.make_voronoi_domain(32, 32, 32, n_grains=4, n_nuclei=8, seed=1)
    big = synth.make_voronoi_domain(64, 48, 40, n_grains=20, n_nuclei=30, seed=1)
    fs, fb = footprint(small), footprint(big)
    assert fb > fs
    # grid and claims (8 B), completion steps and 64-byte records (68 B) per cell at least
    assert fb >= 76 * big.N and fs >= 76 * small.N
    dom = synth.make_voronoi_domain(8, 8, 8, n_grains=2, n_nuclei=2, seed=1)
    dom.cellgrid = dom.cellgrid.copy()
    dom.cellgrid[0] = 9
    with pytest.raises(abi.ScoreGmError) as e:
        footprint(dom)
    assert e.value.code == abi.ERR_ARG


@pytest.mark.skipif(_has_gpu(), reason="checks the behaviour without a GPU")
def test_server_needs_a_device():
    from score_amd.growth import Server
    with pytest.raises(abi.ScoreGmError) as e:
        Server(max_inflight=2)
    assert e.value.code in (abi.ERR_NODEVICE, abi.ERR_HIP)
