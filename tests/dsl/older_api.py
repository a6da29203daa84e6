"""A program in the older pairs API (the calls of examples/lj_onetype.py and
lj_embedded.py that the reference's current Simulation no longer defines):
properties declared one kind at a time, particles from a file, a pair module
reading the bare names rsq and delta, and an Euler update.  argv[1]: the
particle file (mass, position, velocity columns); argv[2], if given: the edge
of the cubic periodic box (the older API names none)."""
import sys

import pairs

SIGMA6, EPSILON, DT = 1.0, 1.0, 0.005


def lj(i, j):
    sr2 = 1.0 / rsq
    sr6 = sr2 * sr2 * sr2 * sigma6
    force[i] += delta * 48.0 * sr6 * (sr6 - 0.5) * sr2 * epsilon


def euler(i):
    velocity[i] += dt * force[i] / mass[i]
    position[i] += dt * velocity[i]


sim = pairs.simulation("older_api")
sim.add_real_property("mass", 1.0)
sim.add_position("position")
sim.add_vector_property("velocity")
sim.add_vector_property("force", vol=True)
sim.from_file(sys.argv[1], ["mass", "position", "velocity"])
if len(sys.argv) > 2:
    edge = float(sys.argv[2])
    sim.set_domain([0.0, 0.0, 0.0, edge, edge, edge])
sim.create_cell_lists(2.8, 2.5)
sim.create_neighbor_lists()
sim.periodic(2.8)
sim.compute(lj, 2.5, {"sigma6": SIGMA6, "epsilon": EPSILON})
sim.compute(euler, symbols={"dt": DT})
sim.target(pairs.target_gpu())
sim.generate()
