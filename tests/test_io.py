"""Particle CSV input and legacy-VTK output (include/pairs_hip_io.h), on the
host: the file formats of runtime/read_from_file.hpp and runtime/vtk.hpp."""
import os

import numpy as np
import pytest

from pairs_amd import io

HERE = os.path.dirname(os.path.abspath(__file__))
DEM_DOMAIN = [(0.0, 0.8), (0.0, 0.015), (0.0, 0.2)]


def test_reads_the_reference_planes_file():
    """data/planes.input with examples/dem.py's property list: both rows are
    kept -- the second lies outside the domain but is Infinite|Fixed|Global."""
    d = io.read_particle_data(os.path.join(HERE, "golden", "planes.input"),
                              ["uid", "type", "mass", "position", "normal", "flags"], DEM_DOMAIN)
    assert d["uid"].tolist() == [100000, 100001]
    assert d["type"].tolist() == [0, 0] and d["flags"].tolist() == [13, 13]
    assert np.array_equal(d["position"], [[0.0, 0.0, 0.0], [0.8, 0.015, 0.2]])
    assert np.array_equal(d["normal"], [[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])
    assert np.array_equal(d["mass"], [1.0, 1.0])


def test_domain_filter_and_column_order(tmp_path):
    f = tmp_path / "spheres.input"
    rows = [
        "1,0,2.5e-05,0.0013,0.0025,0.0025,0.0075,0.1,0.2,-1,0",     # inside
        "2,0,2.5e-05,0.0013,0.7999999,0.0025,0.0075,0,0,-1,0",      # x >= hi - 1e-5: outside
        "3,0,2.5e-05,0.0013,-0.001,0.0025,0.0075,0,0,-1,4",         # outside but Fixed: kept
        "4,1,2.5e-05,0.0013,0.79998,0.0149,0.19998,0,0,0,0",        # just inside
    ]
    f.write_text("\n".join(rows) + "\n")
    props = ["uid", "type", "mass", "radius", "position", "linear_velocity", "flags"]
    d = io.read_particle_data(f, props, DEM_DOMAIN)
    assert d["uid"].tolist() == [1, 3, 4]
    assert d["type"].tolist() == [0, 0, 1] and d["flags"].tolist() == [0, 4, 0]
    assert np.allclose(d["position"][0], [0.0025, 0.0025, 0.0075])
    assert np.allclose(d["linear_velocity"][0], [0.1, 0.2, -1.0])
    assert np.allclose(d["radius"], 0.0013) and d["position"].shape == (3, 3)
    # more rows than the first buffer: the reader asks for the count and retries
    d2 = io.read_particle_data(f, props, DEM_DOMAIN, max_rows=1)
    assert d2["uid"].tolist() == [1, 3, 4]


def test_malformed_rows_raise(tmp_path):
    f = tmp_path / "bad.input"
    f.write_text("1,0,1.0\n")
    with pytest.raises(Exception):
        io.read_particle_data(f, ["uid", "type", "mass", "position"], DEM_DOMAIN)
    with pytest.raises(Exception):
        io.read_particle_data(tmp_path / "missing.input", ["uid"], DEM_DOMAIN)


def test_vtk_file_layout(tmp_path):
    x = np.array([[0.1, 0.2, 0.3], [1.0, 2.0, 3.0], [5.5, 6.5, 7.5]])
    m = np.array([1.0, 2.5, 3.0])
    fl = np.array([0, 1, 0])  # the second is Infinite: skipped
    prefix = str(tmp_path / "out")
    assert not io.vtk_write_arrays(prefix, x, m, fl, timestep=7, frequency=5)
    assert io.vtk_write_arrays(prefix, x, m, fl, timestep=10, frequency=5)
    text = open(prefix + "_10.vtk").read()
    expect = ("# vtk DataFile Version 2.0\nParticle data\nASCII\nDATASET UNSTRUCTURED_GRID\n"
              "POINTS 2 double\n"
              "0.10000000 0.20000000 0.30000000\n5.50000000 6.50000000 7.50000000\n"
              "\n\nCELLS 2 4\n1 0\n1 2\n"
              "\n\nCELL_TYPES 2\n1\n1\n"
              "\n\nPOINT_DATA 2\nSCALARS mass double\nLOOKUP_TABLE default\n1.00000000\n3.00000000\n\n\n")
    assert text == expect
