"""A C++ caller of the C ABI (tests/abi_caller/caller.cpp, built by
score_amd.build with g++ -std=c++11 against include/score_gm.h and
include/score_io.h and linked to libscore_gm.so and libscore_io.so) runs a
domain the way SCORE's adapter would; what it hands back -- scalars, grain
counts, the map, the trajectory, the grain-evolution log, the status vector and
SCORE's map files -- equals the oracle's."""
import os
import struct
import subprocess

import numpy as np
import pytest

from score_amd import abi, synth
from score_amd.synth import ParamsSpec

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "abi_caller", "score_gm_caller")


def test_caller_is_built_and_linked():
    assert os.path.exists(CALLER), "tests/abi_caller/score_gm_caller missing: run score_amd.build"
    r = subprocess.run([CALLER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def _write_input(path, dom, T, m0, hact, xmax, log_at):
    with open(path, "wb") as f:
        f.write(struct.pack("<3Id4I", dom.nx, dom.ny, dom.nz, dom.cellsize, len(dom.ori_quat), dom.n_defg,
                            dom.n_rxg, len(log_at)))
        for a, t in ((dom.ori_quat, np.float64), (dom.defg_ori, np.uint32), (dom.defg_rho, np.float64),
                     (dom.rxg_ori, np.uint32), (dom.rxg_site, np.uint32), (dom.cellgrid, np.uint32),
                     (np.asarray(log_at, np.float64), np.float64)):
            f.write(np.ascontiguousarray(a, t).tobytes())
        f.write(struct.pack("<4d", T, m0, hact, xmax))


@pytest.mark.gpu
def test_cpp_caller_against_oracle(require_gpu, tmp_path):
    from oracle.oracle import Oracle
    from score_amd import mapfiles
    from score_amd.growth import GRAINLOG_DTYPE, TRAJ_DTYPE

    dom = synth.make_voronoi_domain(56, 48, 40, n_grains=14, n_nuclei=60, seed=17)
    T, m0, hact_ev = 1000.0, 0.1, 2.8
    log_at = tuple(f * dom.N for f in (0.0, 0.2, 0.4, 0.6, 0.8, 0.95))
    _write_input(tmp_path / "in.bin", dom, T, m0, hact_ev * synth.ECHARGE, 0.99, log_at)
    out = tmp_path / "out"
    out.mkdir()
    print(f"caller case: {dom.nx}x{dom.ny}x{dom.nz}, {dom.n_defg} grains, {dom.n_rxg} nuclei", flush=True)
    r = subprocess.run([CALLER, str(tmp_path / "in.bin"), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    prm = ParamsSpec(sched_temp=(T, T), rh_hagb_m0=m0, rh_hagb_hact_ev=hact_ev, grainlog_at=log_at)
    o = Oracle(dom, prm)
    ro = o.solve_rxgrowth()
    rb = (out / "result.bin").read_bytes()
    t, X, xcells, pmax = struct.unpack_from("<4d", rb, 0)
    step, completed, nuclei, _ = struct.unpack_from("<4I", rb, 32)
    (sum_sv,) = struct.unpack_from("<Q", rb, 48)
    assert (t, X, xcells, pmax, step, completed, nuclei, sum_sv) == \
        (ro["t"], ro["X"], ro["xcells"], ro["pmax"], ro["step"], ro["completed"], ro["nuclei_placed"], ro["sum_sv"])
    dc, rc = o.counts()
    assert np.array_equal(np.fromfile(out / "counts.bin", np.uint32), np.concatenate([dc, rc]))
    omap = o.grainmap(abi.MAP_RAW2GID)
    assert np.array_equal(np.fromfile(out / "map.bin", np.uint32), omap)
    traj = np.fromfile(out / "traj.bin", TRAJ_DTYPE)
    tro = o.trajectory()
    traj["npending"] = 0
    tro["npending"] = 0
    assert traj.tobytes() == tro.tobytes()
    glog = np.fromfile(out / "glog.bin", GRAINLOG_DTYPE)
    lo, lc = o.grainlog()
    assert glog.tobytes() == lo.tobytes() and len(glog) == 6
    assert np.array_equal(np.fromfile(out / "glogcounts.bin", np.uint32).reshape(lc.shape), lc)
    st = np.fromfile(out / "status.bin", np.uint32)
    ix, gid = o.status(0.5)
    assert np.array_equal(st[:len(ix)], ix) and np.array_equal(st[len(ix):], gid)
    # SCORE's files of the final map, as the C++ program wrote them
    raw = out / mapfiles.raw_name(1, 0, ro["step"], ro["X"])
    assert raw.read_bytes() == omap.astype("<u4").tobytes()
    h5 = out / mapfiles.h5_name(1)
    back, dims = mapfiles.read_h5(str(h5), f"/GROWTH3DGrainID/grainid_{ro['step']}")
    assert dims == (dom.nx, dom.ny, dom.nz) and np.array_equal(back, omap)
    xmf = (out / mapfiles.xdmf_name(1)).read_text()
    assert f"SCORE.1.VoxelData.h5:/GROWTH3DGrainID/{ro['step']}</DataItem>" in xmf
    assert f'Dimensions="{dom.nx} {dom.ny} {dom.nz}"' in xmf
    o.close()


@pytest.mark.gpu
def test_gpu_map_through_the_map_writer(require_gpu, tmp_path):
    """The GPU's final map written in-process through libscore_io (raw stack,
    HDF5 dataset with a render window, XDMF) gives the same bytes and values as
    the oracle's map written the same way."""
    from oracle.oracle import Oracle
    from score_amd import mapfiles
    from score_amd.growth import GrowthMachine

    dom = synth.make_voronoi_domain(48, 40, 36, n_grains=10, n_nuclei=40, seed=23)
    g, o = GrowthMachine(dom, ParamsSpec()), Oracle(dom, ParamsSpec())
    rg, ro = g.solve_rxgrowth(), o.solve_rxgrowth()
    assert rg["step"] == ro["step"]
    shape = (dom.nx, dom.ny, dom.nz)
    win = (3, 40, 0, 39, 5, 30)
    for who, m in (("gpu", g.grainmap(abi.MAP_RAW2GID)), ("oracle", o.grainmap(abi.MAP_RAW2GID))):
        d = tmp_path / who
        d.mkdir()
        mapfiles.write_raw(str(d / mapfiles.raw_name(1, 0, rg["step"], rg["X"])), m, shape)
        mapfiles.write_h5_grainid(str(d / mapfiles.h5_name(1)), rg["step"], m, shape)
        mapfiles.write_h5_grainid(str(d / mapfiles.h5_name(1)), rg["step"] + 1, m, shape, window=win)
        mapfiles.write_xdmf(str(d / mapfiles.xdmf_name(1)), mapfiles.h5_name(1),
                            [(*shape, rg["step"]), (38, 40, 26, rg["step"] + 1)])
    name = mapfiles.raw_name(1, 0, rg["step"], rg["X"])
    assert (tmp_path / "gpu" / name).read_bytes() == (tmp_path / "oracle" / name).read_bytes()
    for step in (rg["step"], rg["step"] + 1):
        a = mapfiles.read_h5(str(tmp_path / "gpu" / mapfiles.h5_name(1)), f"/GROWTH3DGrainID/grainid_{step}")
        b = mapfiles.read_h5(str(tmp_path / "oracle" / mapfiles.h5_name(1)), f"/GROWTH3DGrainID/grainid_{step}")
        assert a[1] == b[1] and np.array_equal(a[0], b[0])
    assert (tmp_path / "gpu" / mapfiles.xdmf_name(1)).read_text() == \
        (tmp_path / "oracle" / mapfiles.xdmf_name(1)).read_text()
    g.close()
    o.close()
