"""Grain-map files in SCORE's formats (score_amd/mapfiles.py over libscore_io.so).

The reference ships four grain-ID maps of its example/jmak run as raw files
(summarised in tests/golden/jmak_raw_maps.json by make_raw_maps_fixture.py).
They fix the file names, the layout (200^3 uint32, nothing else in the file)
and the encoding: 1 marks the cells of the front (as many as the next step's
Sv), 2 + gid every other cell, nuclei after the deformed grains (as many
recrystallized cells as Xcells).  The oracle's maps must read the same way,
and the writers must put exactly the map's bytes, in x-fastest order, into the
raw file and the HDF5 dataset."""
import json
import os
import re

import numpy as np
import pytest

from score_amd import abi, synth
from oracle.oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = json.load(open(os.path.join(ROOT, "tests", "golden", "jmak_raw_maps.json")))
JMAK = json.load(open(os.path.join(ROOT, "tests", "golden", "jmak_stats.json")))

mapfiles = pytest.importorskip("score_amd.mapfiles")
if not os.path.exists(mapfiles.LIB_PATH):
    pytest.skip("libscore_io.so not built (no HDF5)", allow_module_level=True)


def test_raw_file_names_follow_reference():
    for m in FIX["maps"]:
        assert mapfiles.raw_name(m["simid"], m["jobid"], m["step"], m["X"], m["postfix"]) == m["name"]
    assert mapfiles.raw_name(7, 3, 9, 0.0042, "GROWTH.3DGrainID") == "SCORE.7.JobID.3.Iter.9.X.0004.GROWTH.3DGrainID.raw"
    assert mapfiles.h5_name(1) == "SCORE.1.VoxelData.h5"
    assert mapfiles.xdmf_name(12) == "SCORE.12.VoxelData.xdmf"


def test_reference_maps_layout_and_encoding():
    N = FIX["nx"] * FIX["ny"] * FIX["nz"]
    ndefg, nrxg = JMAK["setup"]["n_defg"], JMAK["setup"]["n_nuclei"]
    sv = dict(zip(JMAK["steps"], JMAK["Sv"]))
    for m in FIX["maps"]:
        assert m["bytes"] == 4 * N
        assert m["n_infected"] + m["n_deformed"] + m["n_recrystallized"] == N
        assert m["max"] <= 2 + ndefg + nrxg - 1 and m["min"] >= 1
        # the front's cells are the ACTIVE records the next step sweeps
        assert m["n_infected"] == sv.get(m["step"] + 1, 0)
        # recrystallized cells are Xcells (X printed with six digits)
        assert abs(m["n_recrystallized"] - m["X"] * N) <= 5e-6 * N
    final = [m for m in FIX["maps"] if m["grain_volumes_sorted"]][0]
    assert final["grain_volumes_sorted"] == sorted(JMAK["final_grain_volumes"])


def test_oracle_maps_read_like_the_reference():
    dom = synth.make_voronoi_domain(30, 28, 26, n_grains=6, n_nuclei=12, seed=2)
    o = Oracle(dom, synth.ParamsSpec())
    o.step(40)
    g = o.grainmap(abi.MAP_RAW2GID)
    front = sum(int(o.front(lst)["activity"].sum()) for lst in (abi.LIST_INSIDE, abi.LIST_BORDER))
    assert int((g == 1).sum()) == front
    assert int((g >= 2 + dom.n_defg).sum()) == int(o.result()["xcells"])
    assert g.min() >= 1 and g.max() < 2 + dom.n_defg + dom.n_rxg
    o.close()


def test_raw_writer_bytes(tmp_path):
    rng = np.random.default_rng(3)
    shape = (13, 7, 5)
    gmap = rng.integers(1, 1 << 31, size=np.prod(shape), dtype=np.uint32)
    p = str(tmp_path / mapfiles.raw_name(1, 0, 5, 0.25, "GROWTH.3DGrainID"))
    mapfiles.write_raw(p, gmap, shape)
    assert open(p, "rb").read() == gmap.astype("<u4").tobytes()


def test_h5_dataset_like_reference(tmp_path):
    rng = np.random.default_rng(4)
    nx, ny, nz = 11, 9, 6
    gmap = rng.integers(1, 5000, size=nx * ny * nz, dtype=np.uint32)
    h5 = str(tmp_path / mapfiles.h5_name(1))
    ds = mapfiles.write_h5_grainid(h5, 308, gmap, (nx, ny, nz))
    assert ds == "/GROWTH3DGrainID/grainid_308"
    vals, dims = mapfiles.read_h5(h5, ds)
    assert dims == (nx, ny, nz)
    assert np.array_equal(vals, gmap)
    # a render window: cells of the inclusive box, x fastest; a second dataset in the same group
    win = (2, 8, 1, 4, 3, 5)
    ds2 = mapfiles.write_h5_grainid(h5, 309, gmap, (nx, ny, nz), window=win)
    vals2, dims2 = mapfiles.read_h5(h5, ds2)
    assert dims2 == (7, 4, 3)
    cube = gmap.reshape(nz, ny, nx)[3:6, 1:5, 2:9]
    assert np.array_equal(vals2, cube.reshape(-1))
    with pytest.raises(mapfiles.ScoreIoError):
        mapfiles.write_h5_grainid(h5, 308, gmap, (nx, ny, nz))
    with pytest.raises(mapfiles.ScoreIoError):
        mapfiles.write_h5_grainid(h5, 1, gmap, (nx, ny, nz), window=(0, nx, 0, 1, 0, 1))


def test_xdmf_text(tmp_path):
    p = str(tmp_path / mapfiles.xdmf_name(1))
    mapfiles.write_xdmf(p, "SCORE.1.VoxelData.h5", [(200, 200, 200, 115), (200, 200, 200, 308)])
    txt = open(p).read()
    head = ('<?xml version="1.0" ?>\n<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>\n'
            '<Xdmf xmlns:xi="http://www.w3.org/2003/XInclude" Version="2.2">\n\t<Domain>\n'
            '\t\t<Grid Name="SCORE_v1.2.0_TimeSeries" GridType="Collection" CollectionType="Temporal">\n')
    assert txt.startswith(head)
    assert txt.endswith("\t\t</Grid>\n\t</Domain>\n</Xdmf>")
    assert txt.count('<Grid Name="SCORE_v1.2.0_Result" GridType="Uniform">') == 2
    assert '<Time Type="Single" Value="   1" />' in txt and '<Time Type="Single" Value="   2" />' in txt
    assert txt.count('<Topology TopologyType="3DCoRectMesh" Dimensions="200 200 200"></Topology>') == 2
    assert "\t\t\t\t\tSCORE.1.VoxelData.h5:/GROWTH3DGrainID/308</DataItem>\n" in txt
    assert re.search(r'<Attribute Name="GrainID\+2" AttributeType="Scalar" Center="Cell">\n\t+<DataItem '
                     r'Dimensions="200 200 200" Format="HDF" NumberType="UInt" Precision="4">', txt)


def test_library_exports_every_declared_function():
    txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "score_io.h")).read(), flags=re.S)
    names = set(re.findall(r"^\s*int\s+(score_io_\w+)\s*\(", txt, flags=re.M))
    lib = mapfiles.load()
    assert names and all(hasattr(lib, n) for n in names)
