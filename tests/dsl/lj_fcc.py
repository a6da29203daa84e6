"""A Lennard-Jones simulation written in the pairs DSL (the calls of
examples/md.py), run by `python -m pairs_amd.dsl tests/dsl/lj_fcc.py gpu [n steps]`."""
import sys

import pairs


def lennard_jones(i, j):
    inv = 1.0 / squared_distance(i, j)
    s6 = inv * inv * inv * sigma6[i, j]
    apply(force, delta(i, j) * (48.0 * s6 * (s6 - 0.5) * inv * epsilon[i, j]))


def initial_integrate(i):
    linear_velocity[i] += (dt * 0.5) * force[i] / mass[i]
    position[i] += dt * linear_velocity[i]


def final_integrate(i):
    linear_velocity[i] += (dt * 0.5) * force[i] / mass[i]


target = sys.argv[1] if len(sys.argv) > 1 else "gpu"
n = int(sys.argv[2]) if len(sys.argv) > 2 else 4
steps = int(sys.argv[3]) if len(sys.argv) > 3 else 100
ntypes = 1

psim = pairs.simulation("lj_fcc", [pairs.point_mass()], timesteps=steps, double_prec=True)
psim.target(pairs.target_gpu() if target == "gpu" else pairs.target_cpu())
psim.add_position('position')
psim.add_property('mass', pairs.real(), 1.0)
psim.add_property('linear_velocity', pairs.vector())
psim.add_property('force', pairs.vector(), volatile=True)
psim.add_feature('type', ntypes)
psim.add_feature_property('type', 'epsilon', pairs.real(), [1.0] * ntypes * ntypes)
psim.add_feature_property('type', 'sigma6', pairs.real(), [1.0] * ntypes * ntypes)
psim.copper_fcc_lattice(n, n, n, 0.8442, 1.44, ntypes)
psim.set_domain_partitioner(pairs.regular_domain_partitioner())
psim.compute_thermo(100)
psim.reneighbor_every(20)
psim.build_neighbor_lists(2.5 + 0.3)
psim.compute(initial_integrate, symbols={'dt': 0.005}, pre_step=True, skip_first=True)
psim.compute(lennard_jones, 2.5)
psim.compute(final_integrate, symbols={'dt': 0.005}, skip_first=True)
psim.generate()
