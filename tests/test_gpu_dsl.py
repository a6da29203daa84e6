"""DSL scripts (tests/dsl/) run through the front end on the GPU: the same
trajectories, bit for bit, as the hand-built systems of the same set-up."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_lj_script_equals_direct_system():
    from pairs_amd import dsl, md
    sims = dsl.run_script(os.path.join(HERE, "dsl", "lj_fcc.py"), ["gpu", "4", "100"], quiet=True)
    s = sims[0].system
    ref = md.fcc_system(4)
    rows = ref.run(0, 101, 100)
    assert np.array_equal(s.positions(True), ref.positions(True))
    assert np.array_equal(s.velocities(), ref.velocities())
    assert [(t, T, P) for t, T, P in sims[0].thermo] == [(0, rows[0][0], rows[0][1]), (99, rows[1][0], rows[1][1])]


def test_dem_script_equals_direct_system():
    from pairs_amd import dsl, dem
    sims = dsl.run_script(os.path.join(HERE, "dsl", "dem_settle.py"), ["gpu", "200"], quiet=True)
    s = sims[0].system
    ref = dem.reference_system((0.04, 0.015, 0.04))
    ref.run(0, 201)
    for name in ("position", "linear_velocity", "angular_velocity", "uid", "num_contacts"):
        assert np.array_equal(s.get(name), ref.get(name)), name


def test_older_api_script_equals_oracle():
    """tests/dsl/older_api.py (the older API: from_file, bare rsq / delta, its
    own pair expression, Euler) with the 4x4x4 particle file of the
    reference's data/ and a box of 4 lattice constants: 101 iterations,
    rebuilt every step, bit-identical to the oracle's restatement."""
    from pairs_amd import dsl
    from oracle import pyoracle as po
    data = os.path.join(HERE, "golden", "minimd_setup_4x4x4_onetype.input")
    L = 4 * 1.6796
    sims = dsl.run_script(os.path.join(HERE, "dsl", "older_api.py"), [data, repr(L)], quiet=True)
    s = sims[0].system
    rows = np.loadtxt(data, delimiter=",")
    o = po.OracleMD([(0.0, L)] * 3, 4096, ncap=100, reneighbor_every=1, thermo_every=0)
    o.set_particles(rows[:, 1:4], rows[:, 4:7], m=rows[:, 0])
    o.setup_cells()
    for _ in range(101):
        o.rebuild()
        o.lj_older()
        o.euler()
    assert s.nlocal == o.nlocal == 256
    assert np.array_equal(s.positions(), o.positions())
    assert np.array_equal(s.velocities(), o.velocities())
    assert np.array_equal(s.forces(), o.forces())
