"""The reference's input format: XML (test/minimc.xsd) with continuous-energy
and thermal-scattering tables in pandas-HDF5 files (HDF5DataSet<D>,
src/HDF5DataSet.hpp:18-167; the <tsl> element of minimc_basictypes.xsd:
96-133).  The synthetic benchmark problems are written in that format and
read back; the World read from the files must give the oracle the same
results, bit for bit, as the World built in memory.  (The reference's own
.hdf5 fixtures are git-lfs pointers, so the layout is pinned by the items
HDF5DataSet reads, not by a file pandas wrote.)"""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from minimc_amd import configs, load_xml, pandas_h5, xml_writer
from oracle.oracle import Oracle


def test_pandas_h5_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    for D in (1, 2, 3):
        axes = [np.sort(rng.random(n)) for n in (7, 3, 5)[:D]]
        v = rng.random(tuple(len(a) for a in axes))
        p = tmp_path / ("t%d.hdf5" % D)
        pandas_h5.write(p, axes, v)
        a, w = pandas_h5.read(p, D)
        assert all(np.array_equal(x, y) for x, y in zip(a, axes))
        assert np.array_equal(w, v)
    # HDF5DataSet's checks and messages (ReadPandasAxis, :89-101)
    with pytest.raises(RuntimeError, match="Expected 2 dimensions, but got 3"):
        pandas_h5.read(tmp_path / "t3.hdf5", 2)
    with pytest.raises(RuntimeError, match="File not found: "):
        pandas_h5.read(tmp_path / "missing.hdf5", 1)
    # a one-level file is a ContinuousMap table (ToContinuousMap, :49-63)
    x, y = pandas_h5.read_table(tmp_path / "t1.hdf5")
    assert x.shape == y.shape == (7,)


def _same_cycles(pr, pr2, cycles=2):
    out = []
    for x in (pr, pr2):
        o = Oracle(x)
        o.set_entropy_mesh(*pr.entropy_mesh)
        o.keig_init(x.histories)
        st = [o.run_cycle(active=True, threads=1) for _ in range(cycles)]
        out.append([{k: v for k, v in s.items() if k != "seconds"} for s in st])
    assert out[0] == out[1]


def test_multizone_problem_through_xml_and_hdf5(tmp_path):
    pr = configs.c5_multizone(histories=1500)
    path = xml_writer.write_xml(pr, str(tmp_path / "multizone.xml"))
    # the <tsl> element carries the reference's attributes and partitions
    tsl = ET.parse(path).getroot().find(".//tsl")
    assert set(tsl.keys()) == {"majorant", "total_E", "total_T", "total_S", "beta_cutoff",
                               "alpha_cutoff"}
    for tag, x in (("beta_partitions", "E_T"), ("alpha_partitions", "beta_T")):
        parts = tsl.find(tag)
        assert len(parts) == 2
        assert all(set(p.keys()) == {"CDF", "S", x} for p in parts)
    pr2 = load_xml(path)
    assert pr2.problem == "keigenvalue" and (pr2.inactive, pr2.active) == (20, 100)
    t, t2 = pr.tsl[0], pr2.tsl[0]
    for k in ("total_E_axis", "total_E", "total_T_axis", "total_T", "total_S"):
        assert np.array_equal(t[k], t2[k]), k
    for k in ("beta_partitions", "alpha_partitions"):
        for p, q in zip(t[k], t2[k]):
            assert all(np.array_equal(p[f], q[f]) for f in p), k
    _same_cycles(pr, pr2)


def test_ce_slab_through_xml_and_hdf5(tmp_path):
    pr = configs.c2_slab(histories=2000, n_points=6000, n_res=20)
    pr2 = load_xml(xml_writer.write_xml(pr, str(tmp_path / "slab.xml")))
    res = []
    for x in (pr, pr2):
        o = Oracle(x)
        o.run_fixed_source(2000, threads=1)
        res.append([o.tallies(e) for e in range(2)])
    for (a, b), (c, d) in zip(*res):
        assert np.array_equal(a, c) and np.array_equal(b, d)


def test_broomstick_through_xml_and_hdf5(tmp_path):
    pr = configs.c3_broomstick(histories=3000, tables=configs.tsl_tables(reduced=True))
    pr2 = load_xml(xml_writer.write_xml(pr, str(tmp_path / "broomstick.xml")))
    res = []
    for x in (pr, pr2):
        o = Oracle(x)
        o.run_fixed_source(3000, threads=1)
        res.append([o.tallies(e) for e in range(len(pr.estimators))])
    for (a, b), (c, d) in zip(*res):
        assert np.array_equal(a, c) and np.array_equal(b, d)


def test_missing_tsl_table_names_the_file(tmp_path):
    pr = configs.tsl_box(configs.tsl_tables(reduced=True))
    pr.source = dict(position=(0.0, 0.0, 0.0), E=1e-6)
    path = xml_writer.write_xml(pr, str(tmp_path / "box.xml"))
    gone = os.path.join(str(tmp_path), "box_data", "nuclide-0-tsl0-total_S.hdf5")
    os.remove(gone)
    with pytest.raises(RuntimeError, match="File not found: .*total_S.hdf5"):
        load_xml(path)
