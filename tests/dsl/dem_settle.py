"""A DEM simulation written in the pairs DSL (the calls of examples/dem.py on a
small domain), run by `python -m pairs_amd.dsl tests/dsl/dem_settle.py gpu [steps]`."""
import math
import os
import sys

import pairs


def update_mass_and_inertia(i):
    rotation_matrix[i] = diagonal_matrix(1.0)
    rotation_quat[i] = default_quaternion()
    if is_sphere(i):
        inv_inertia[i] = inversed(diagonal_matrix(0.4 * mass[i] * radius[i] * radius[i]))
    else:
        mass[i] = infinity
        inv_inertia[i] = 0.0


def linear_spring_dashpot(i, j):
    overlap = -penetration_depth(i, j)
    skip_when(overlap < 0.0)
    mr = 1.0 / ((1.0 / mass[i]) + (1.0 / mass[j]))
    kn = mr * (pi * pi + lnDryResCoeff * lnDryResCoeff) / (collisionTime_SI * collisionTime_SI)
    kt = kappa * kn
    gn = -2.0 * mr * lnDryResCoeff / collisionTime_SI
    gt = sqrt(kappa) * gn
    vi = linear_velocity[i] + cross(angular_velocity[i], contact_point(i, j) - position[i])
    vj = linear_velocity[j] + cross(angular_velocity[j], contact_point(i, j) - position[j])
    vr = -(vi - vj)
    vrn = dot(vr, contact_normal(i, j)) * contact_normal(i, j)
    vrt = vr - vrn
    fn = kn * overlap * contact_normal(i, j) + gn * vrn
    spring = tangential_spring_displacement[i, j]
    v_imp = impact_velocity_magnitude[i, j]
    v_imp_new = select(v_imp > 0.0, v_imp, length(vr))
    stick = is_sticking[i, j]
    spring_rot = spring - contact_normal(i, j) * dot(spring, contact_normal(i, j))
    l2 = squared_length(spring_rot)
    spring_new = dt * vrt + select(l2 <= 0.0, zero_vector(), spring_rot * sqrt(squared_length(spring) / l2))
    ft = kt * spring_new + gt * vrt
    ft_len = length(ft)
    tdir = normalized(ft)
    fs_max = friction_static[i, j] * length(fn)
    fd_max = friction_dynamic[i, j] * length(fn)
    eps_v = 1e-8
    keep1 = stick == 1 and length(vrt) < eps_v and ft_len < fs_max
    keep2 = stick == 1 and ft_len < fd_max
    f_cap = select(keep1, fs_max, fd_max)
    stick_new = select(keep1 or keep2 or ft_len < fd_max, 1, 0)
    tangential_spring_displacement[i, j] = \
        select(not keep1 and not keep2 and kt > 0.0, (f_cap * tdir - gt * vrt) / kt, spring_new)
    impact_velocity_magnitude[i, j] = v_imp_new
    is_sticking[i, j] = stick_new
    f = fn + min(ft_len, f_cap) * tdir
    apply(force, f)
    apply(torque, cross(contact_point(i, j) - position, f))


def euler(i):
    im = 1.0 / mass[i]
    position[i] += 0.5 * im * force[i] * dt * dt + linear_velocity[i] * dt
    linear_velocity[i] += im * force[i] * dt
    wd = rotation_matrix[i] * (inv_inertia[i] * torque[i]) * transposed(rotation_matrix[i])
    rot = angular_velocity[i] * dt + 0.5 * wd * dt * dt
    rotation_quat[i] = quaternion(rot, length(rot)) * rotation_quat[i]
    rotation_matrix[i] = quaternion_to_rotation_matrix(rotation_quat[i])
    angular_velocity[i] += wd * dt


def gravity(i):
    v = (4.0 / 3.0) * pi * radius[i] * radius[i] * radius[i]
    force[i][2] = force[i][2] - (densityParticle_SI - densityFluid_SI) * v * gravity_SI


target = sys.argv[1] if len(sys.argv) > 1 else "gpu"
steps = int(sys.argv[2]) if len(sys.argv) > 2 else 200
dom = [0.04, 0.015, 0.04]
d, rho_p, rho_f, g, dt = 0.0029, 2550, 1000, 9.81, 5e-5
e, tc, nu, mu = 0.1, 5e-4, 0.22, 0.5
kap = 2.0 * (1.0 - nu) / (2.0 - nu)
width = 1.01 * d * 1.1
ntypes = 1

psim = pairs.simulation("dem_settle", [pairs.sphere(), pairs.halfspace()], timesteps=steps, double_prec=True,
                        use_contact_history=True, particle_capacity=8192, neighbor_capacity=20)
psim.target(pairs.target_gpu() if target == "gpu" else pairs.target_cpu())
psim.add_position('position')
for name, t in (('mass', pairs.real()), ('linear_velocity', pairs.vector()), ('angular_velocity', pairs.vector()),
                ('radius', pairs.real()), ('normal', pairs.vector()), ('inv_inertia', pairs.matrix()),
                ('rotation_matrix', pairs.matrix()), ('rotation_quat', pairs.quaternion())):
    psim.add_property(name, t)
psim.add_property('force', pairs.vector(), volatile=True)
psim.add_property('torque', pairs.vector(), volatile=True)
psim.add_feature('type', ntypes)
psim.add_feature_property('type', 'friction_static', pairs.real(), [0.0])
psim.add_feature_property('type', 'friction_dynamic', pairs.real(), [mu])
psim.add_contact_property('is_sticking', pairs.int32(), 0)
psim.add_contact_property('tangential_spring_displacement', pairs.vector(), [0.0, 0.0, 0.0])
psim.add_contact_property('impact_velocity_magnitude', pairs.real(), 0.0)
psim.set_domain([0.0, 0.0, 0.0] + dom)
psim.set_domain_partitioner(pairs.regular_domain_partitioner_xy())
psim.pbc([True, True, False])
psim.dem_sc_grid(dom[0], dom[1], dom[2], 0.005, d, d * 0.9, d * 1.1, 1, rho_p, ntypes)
psim.read_particle_data(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "golden", "planes.input"),
                        ['uid', 'type', 'mass', 'position', 'normal', 'flags'], pairs.halfspace())
psim.setup(update_mass_and_inertia, {'densityParticle_SI': rho_p, 'pi': math.pi, 'infinity': math.inf})
psim.build_cell_lists(width)
psim.compute(gravity, symbols={'densityParticle_SI': rho_p, 'densityFluid_SI': rho_f, 'gravity_SI': g,
                               'pi': math.pi})
psim.compute(linear_spring_dashpot, width, symbols={'dt': dt, 'pi': math.pi, 'kappa': kap,
                                                     'lnDryResCoeff': math.log(e), 'collisionTime_SI': tc})
psim.compute(euler, symbols={'dt': dt})
psim.generate()
