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
