"""The reference's Python DSL as a front end of the HIP library pairs_amd (a GPU target).

``import pairs_amd.dsl as pairs`` (or run a script through ``python -m
pairs_amd.dsl script.py gpu``, which installs this module as ``pairs``) gives
the calls that examples/md.py and examples/dem.py make
(src/pairs/__init__.py:9-67, src/pairs/sim/simulation.py:104-452): the
simulation object, its properties, features, contact properties, set-up
(copper_fcc_lattice, dem_sc_grid, read_particle_data), cell / neighbour lists,
``compute`` / ``setup`` of user modules, thermo, VTK output, ``target`` and
``generate``.

The reference lowers each user module (a Python function over particle
properties) to IR with ``ast`` and generates C++/CUDA.  This front end does
not generate kernels: it recognises the modules of the hot path and maps
them to the hand-written HIP kernels behind the C ABI.  A module is traced
symbolically -- its source is parsed, ``and`` / ``or`` / ``not`` / ``if``
rewritten into recording calls, and the body run over proxy objects -- into
an expression graph with hash-consed nodes, which is compared with the graph
of the library's own definition of each supported module (Lennard-Jones with
feature tables, the velocity-Verlet pair, the DEM linear spring-dashpot with
tangential history, gravity, the Euler integrator with rotation, and the DEM
mass / inertia set-up).  Operation order is part of the graph, so a module
is accepted only if it computes exactly what the kernels compute; any other
module fails loudly.

``generate()`` then runs the simulation on the GPU (``timesteps`` + 1
iterations of the reference's schedule, thermo rows printed like the
generated program's, VTK files at the requested frequency) and, if asked,
writes the host C++ driver that calls the same C ABI (``emit_cpp``).
"""
import ast
import builtins
import inspect
import math
import sys
import textwrap

import numpy as np

# ---------------------------------------------------------------- tags (src/pairs/__init__.py)


class Target:
    def __init__(self, kind, parallel=False):
        self.kind = kind          # "cpu" or "gpu"
        self.parallel = parallel

    def __repr__(self):
        return f"Target({self.kind})"


def target_cpu(parallel=False):
    return Target("cpu", parallel)


def target_gpu():
    """The GPU target: the MI355X HIP kernels of pairs_amd."""
    return Target("gpu")


def target_hip():
    return Target("gpu")


class _Tag:
    def __init__(self, name):
        self.name = name

    def __repr__(self):
        return self.name


_INT32, _FLOAT, _DOUBLE, _REAL, _VECTOR, _MATRIX, _QUAT = (_Tag(n) for n in (
    "int32", "float", "double", "real", "vector", "matrix", "quaternion"))
_POINT_MASS, _SPHERE, _HALFSPACE = _Tag("point_mass"), _Tag("sphere"), _Tag("halfspace")
_REGULAR, _REGULAR_XY = _Tag("regular"), _Tag("regular_xy")


def int32(): return _INT32
def float(): return _FLOAT            # noqa: A001 (the DSL's own name)
def double(): return _DOUBLE
def real(): return _REAL
def vector(): return _VECTOR
def matrix(): return _MATRIX
def quaternion(): return _QUAT
def point_mass(): return _POINT_MASS
def sphere(): return _SPHERE
def halfspace(): return _HALFSPACE
def regular_domain_partitioner(): return _REGULAR
def regular_domain_partitioner_xy(): return _REGULAR_XY


class UnsupportedModule(RuntimeError):
    """A user module that is not one of the hot-path modules the HIP kernels implement."""


# ---------------------------------------------------------------- symbolic tracing of modules

class _Node:
    """An expression node; arithmetic and comparisons build new nodes."""
    __slots__ = ("op", "args")

    def __init__(self, op, *args):
        self.op = op
        self.args = args

    def _b(self, op, other, rev=False):
        o = _wrap(other)
        return _Node(op, o, self) if rev else _Node(op, self, o)

    def __add__(self, o): return self._b("add", o)
    def __radd__(self, o): return self._b("add", o, True)
    def __sub__(self, o): return self._b("sub", o)
    def __rsub__(self, o): return self._b("sub", o, True)
    def __mul__(self, o): return self._b("mul", o)
    def __rmul__(self, o): return self._b("mul", o, True)
    def __truediv__(self, o): return self._b("div", o)
    def __rtruediv__(self, o): return self._b("div", o, True)
    def __pow__(self, o): return self._b("pow", o)
    def __neg__(self): return _Node("neg", self)
    def __lt__(self, o): return self._b("lt", o)
    def __le__(self, o): return self._b("le", o)
    def __gt__(self, o): return self._b("gt", o)
    def __ge__(self, o): return self._b("ge", o)
    def __eq__(self, o): return self._b("eq", o)      # noqa: a node, not a truth value
    def __ne__(self, o): return self._b("ne", o)
    __hash__ = object.__hash__

    def __bool__(self):
        raise UnsupportedModule("a traced value was used as a Python truth value")

    def __getitem__(self, k):
        return _Node("elem", self, _wrap(k))

    def __setitem__(self, k, v):
        _T.record("set_elem", self, _wrap(k), _wrap(v))


def _wrap(x):
    if isinstance(x, _Node):
        return x
    if isinstance(x, _Prop):
        return x._self()
    if isinstance(x, (bool, builtins.int, np.integer)):
        return _Node("const", builtins.int(x))
    if isinstance(x, (builtins.float, np.floating)):
        return _Node("const", repr(builtins.float(x)))
    if isinstance(x, (list, tuple)):
        return _Node("tuple", *[_wrap(y) for y in x])
    raise UnsupportedModule(f"cannot trace a value of type {type(x).__name__}")


class _Prop:
    """A particle, feature or contact property of the traced module."""

    def __init__(self, name):
        self.name = name

    def __getitem__(self, idx):
        return _Node("prop", _Node("name", self.name), _wrap(idx))

    def __setitem__(self, idx, v):
        _T.record("set", _Node("name", self.name), _wrap(idx), _wrap(v))

    # a property used without an index stands for its value at i
    # (src/pairs/mapping/funcs.py: properties in expressions of i)
    def _self(self):
        return _Node("prop", _Node("name", self.name), _Node("i"))

    def __add__(self, o): return self._self() + o
    def __radd__(self, o): return o + self._self()
    def __sub__(self, o): return self._self() - o
    def __rsub__(self, o): return _wrap(o) - self._self()
    def __mul__(self, o): return self._self() * o
    def __rmul__(self, o): return _wrap(o) * self._self()


class _Tracer:
    def __init__(self):
        self.stmts = []
        self.conds = []

    def record(self, kind, *args):
        self.stmts.append((kind, tuple(self.conds), args))


_T = _Tracer()


class _When:
    def __init__(self, cond, negate=False):
        self.cond = cond
        self.negate = negate

    def __enter__(self):
        _T.conds.append(_Node("not", self.cond) if self.negate else self.cond)

    def __exit__(self, *a):
        _T.conds.pop()
        return False


class _Rewrite(ast.NodeTransformer):
    """and/or/not -> recording calls; if/else -> condition blocks."""

    def visit_BoolOp(self, node):
        self.generic_visit(node)
        f = "_and" if isinstance(node.op, ast.And) else "_or"
        return ast.copy_location(ast.Call(ast.Name(f, ast.Load()), node.values, []), node)

    def visit_UnaryOp(self, node):
        self.generic_visit(node)
        if isinstance(node.op, ast.Not):
            return ast.copy_location(ast.Call(ast.Name("_not", ast.Load()), [node.operand], []), node)
        return node

    def visit_If(self, node):
        self.generic_visit(node)
        out = [ast.With([ast.withitem(ast.Call(ast.Name("_when", ast.Load()), [node.test], []))], node.body)]
        if node.orelse:
            neg = ast.Call(ast.Name("_when", ast.Load()), [node.test, ast.Constant(True)], [])
            out.append(ast.With([ast.withitem(neg)], node.orelse))
        return [ast.copy_location(o, node) for o in out]


def _k(op):
    return lambda *a: _Node(op, *[_wrap(x) for x in a])


def _keywords():
    """The DSL's keywords (src/pairs/mapping/keywords.py) as node builders."""
    ns = {}
    for name in ("squared_distance", "delta", "sqrt", "cross", "dot", "length", "squared_length", "normalized",
                 "select", "min", "max", "penetration_depth", "contact_point", "contact_normal", "zero_vector",
                 "diagonal_matrix", "inversed", "transposed", "default_quaternion", "quaternion",
                 "quaternion_to_rotation_matrix", "is_sphere", "is_halfspace", "abs", "exp"):
        ns[name] = _k(name)
    ns["apply"] = lambda prop, v: _T.record("apply", _Node("name", prop.name), _wrap(v))
    ns["skip_when"] = lambda c: _T.record("skip", _wrap(c))
    ns["_and"] = lambda *v: _Node("and", *[_wrap(x) for x in v])
    ns["_or"] = lambda *v: _Node("or", *[_wrap(x) for x in v])
    ns["_not"] = lambda v: _Node("not", _wrap(v))
    ns["_when"] = lambda c, neg=False: _When(_wrap(c), neg)
    return ns


def trace(fn_or_src, properties, symbols, name=None):
    """Trace a module into its canonical form: a list of statements whose
    expressions are hash-consed node ids in first-use order."""
    if callable(fn_or_src):
        src = textwrap.dedent(inspect.getsource(fn_or_src))
        name = fn_or_src.__name__
    else:
        src = fn_or_src
    tree = ast.parse(src)
    fdef = next(n for n in tree.body if isinstance(n, ast.FunctionDef) and (name is None or n.name == name))
    fdef.decorator_list = []
    mod = ast.fix_missing_locations(ast.Module(body=[_Rewrite().visit(fdef)], type_ignores=[]))
    ns = _keywords()
    for p in properties:
        ns[p] = _Prop(p)
    for s in symbols:
        ns[s] = _Node("sym", s)
    # older-API pair modules read the pair's squared distance and separation
    # as bare names (examples/lj_onetype.py:5-8)
    if len(fdef.args.args) == 2:
        called = {n.func.id for n in ast.walk(fdef) if isinstance(n, ast.Call) and isinstance(n.func, ast.Name)}
        used = {n.id for n in ast.walk(fdef) if isinstance(n, ast.Name) and isinstance(n.ctx, ast.Load)}
        if "rsq" in used and "rsq" not in ns:
            ns["rsq"] = _Node("squared_distance", _Node("i"), _Node("j"))
        if "delta" in used and "delta" not in called:
            ns["delta"] = _Node("delta", _Node("i"), _Node("j"))
    ns["__builtins__"] = {}
    exec(compile(mod, f"<module {fdef.name}>", "exec"), ns)
    f = ns[fdef.name]
    global _T
    _T = _Tracer()
    args = [_Node("i"), _Node("j")][:len(fdef.args.args)]
    try:
        f(*args)
    except NameError as e:
        raise UnsupportedModule(f"module {fdef.name}: {e}") from None
    return _canonical(_T.stmts)


def _canonical(stmts):
    ids, lines = {}, []

    def nid(n):
        if n.op == "const" or n.op in ("name", "sym", "i", "j"):
            return f"{n.op}:{n.args[0]}" if n.args else n.op
        key = (n.op,) + tuple(nid(a) for a in n.args)
        if key not in ids:
            ids[key] = f"n{len(ids)}"
            lines.append(f"{ids[key]}={key[0]}({','.join(key[1:])})")
        return ids[key]

    out = []
    for kind, conds, args in stmts:
        c = [nid(x) for x in conds]
        a = [nid(x) for x in args]
        out.append(f"{kind}[{';'.join(c)}]({','.join(a)})")
    return "\n".join(lines + out)


# ---------------------------------------------------------------- the library's modules (the kernels' math)
#
# Each definition is what a hand-written kernel computes, written in the
# DSL with its operation order (the traced graph fixes the order):
#   lennard_jones       -> k_lj_full (full and half lists) (md_kernels.hip)
#   initial/final VV    -> fused into k_lj_full, k_vv_* (md_kernels.hip)
#   linear_spring_dashpot, gravity -> k_contact_* + k_contacts (dem_kernels.hip)
#   euler               -> k_euler                    (dem_kernels.hip)
#   update_mass_and_inertia -> k_mass_inertia         (dem_kernels.hip)

# the older API's scalars (examples/lj_onetype.py passes them as symbols)
_BUILTIN_SYMS = {"lj_older": ("sigma6", "epsilon")}

_BUILTIN_SRC = {
    # the older API's modules (examples/lj_onetype.py:5-13): its pair force
    # evaluated left to right, and an Euler update
    "lj_older": '''
def lj_older(i, j):
    inv = 1.0 / rsq
    p6 = inv * inv * inv * sigma6
    force[i] = force[i] + delta * 48.0 * p6 * (p6 - 0.5) * inv * epsilon
''',
    "euler_md": '''
def euler_md(i):
    velocity[i] = velocity[i] + dt * force[i] / mass[i]
    position[i] = position[i] + dt * velocity[i]
''',
    "lennard_jones": '''
def lennard_jones(i, j):
    inv_rsq = 1.0 / squared_distance(i, j)
    s6 = inv_rsq * inv_rsq
    s6 = s6 * inv_rsq
    s6 = s6 * sigma6[i, j]
    mag = 48.0 * s6
    mag = mag * (s6 - 0.5)
    mag = mag * inv_rsq
    apply(force, delta(i, j) * (mag * epsilon[i, j]))
''',
    "initial_integrate": '''
def initial_integrate(i):
    half_kick = (dt * 0.5) * force[i] / mass[i]
    linear_velocity[i] = linear_velocity[i] + half_kick
    position[i] = position[i] + dt * linear_velocity[i]
''',
    "final_integrate": '''
def final_integrate(i):
    linear_velocity[i] = linear_velocity[i] + (dt * 0.5) * force[i] / mass[i]
''',
    "update_mass_and_inertia": '''
def update_mass_and_inertia(i):
    rotation_matrix[i] = diagonal_matrix(1.0)
    rotation_quat[i] = default_quaternion()
    if is_sphere(i):
        inv_inertia[i] = inversed(diagonal_matrix(0.4 * mass[i] * radius[i] * radius[i]))
    else:
        mass[i] = infinity
        inv_inertia[i] = 0.0
''',
    "gravity": '''
def gravity(i):
    r = radius[i]
    vol = (4.0 / 3.0) * pi * r
    vol = vol * r
    vol = vol * r
    force[i][2] = force[i][2] - (densityParticle_SI - densityFluid_SI) * vol * gravity_SI
''',
    "euler": '''
def euler(i):
    im = 1.0 / mass[i]
    position[i] = position[i] + (0.5 * im * force[i] * dt * dt + linear_velocity[i] * dt)
    linear_velocity[i] = linear_velocity[i] + im * force[i] * dt
    R = rotation_matrix[i]
    wd = R * (inv_inertia[i] * torque[i]) * transposed(R)
    rot = angular_velocity[i] * dt + 0.5 * wd * dt * dt
    rotation_quat[i] = quaternion(rot, length(rot)) * rotation_quat[i]
    rotation_matrix[i] = quaternion_to_rotation_matrix(rotation_quat[i])
    angular_velocity[i] = angular_velocity[i] + wd * dt
''',
    "linear_spring_dashpot": '''
def linear_spring_dashpot(i, j):
    pen = -penetration_depth(i, j)
    skip_when(pen < 0.0)
    m_eff = 1.0 / ((1.0 / mass[i]) + (1.0 / mass[j]))
    k_n = m_eff * (pi * pi + lnDryResCoeff * lnDryResCoeff) / (collisionTime_SI * collisionTime_SI)
    k_t = kappa * k_n
    c_n = -2.0 * m_eff * lnDryResCoeff / collisionTime_SI
    c_t = sqrt(kappa) * c_n
    n = contact_normal(i, j)
    cp = contact_point(i, j)
    v_i = linear_velocity[i] + cross(angular_velocity[i], cp - position[i])
    v_j = linear_velocity[j] + cross(angular_velocity[j], cp - position[j])
    v_rel = -(v_i - v_j)
    v_n = dot(v_rel, n) * n
    v_t = v_rel - v_n
    f_n = k_n * pen * n + c_n * v_n
    old_disp = tangential_spring_displacement[i, j]
    old_impact = impact_velocity_magnitude[i, j]
    impact = select(old_impact > 0.0, old_impact, length(v_rel))
    was_sticking = is_sticking[i, j]
    rot_disp = old_disp - n * dot(old_disp, n)
    rot_len2 = squared_length(rot_disp)
    disp = dt * v_t + select(rot_len2 <= 0.0, zero_vector(), rot_disp * sqrt(squared_length(old_disp) / rot_len2))
    f_t_trial = k_t * disp + c_t * v_t
    f_t_len = length(f_t_trial)
    t_dir = normalized(f_t_trial)
    f_static = friction_static[i, j] * length(f_n)
    f_dynamic = friction_dynamic[i, j] * length(f_n)
    threshold = 1e-8
    c1 = was_sticking == 1 and length(v_t) < threshold and f_t_len < f_static
    c2 = was_sticking == 1 and f_t_len < f_dynamic
    f_limit = select(c1, f_static, f_dynamic)
    sticking_now = select(c1 or c2 or f_t_len < f_dynamic, 1, 0)
    tangential_spring_displacement[i, j] = select(not c1 and not c2 and k_t > 0.0,
                                                  (f_limit * t_dir - c_t * v_t) / k_t, disp)
    impact_velocity_magnitude[i, j] = impact
    is_sticking[i, j] = sticking_now
    f_tot = f_n + min(f_t_len, f_limit) * t_dir
    apply(force, f_tot)
    apply(torque, cross(cp - position, f_tot))
''',
}

_MD_PROPS = ("position", "mass", "linear_velocity", "velocity", "force", "epsilon", "sigma6", "type")
_DEM_PROPS = ("position", "mass", "linear_velocity", "angular_velocity", "force", "torque", "radius", "normal",
              "inv_inertia", "rotation_matrix", "rotation_quat", "friction_static", "friction_dynamic",
              "is_sticking", "tangential_spring_displacement", "impact_velocity_magnitude", "type")
_SYMS = ("dt", "pi", "kappa", "lnDryResCoeff", "collisionTime_SI", "densityParticle_SI", "densityFluid_SI",
         "gravity_SI", "infinity")


def _builtin_forms():
    props = tuple(sorted(set(_MD_PROPS) | set(_DEM_PROPS)))
    return {k: trace(src, props, _SYMS + _BUILTIN_SYMS.get(k, ()), k) for k, src in _BUILTIN_SRC.items()}


_FORMS = None


def recognise(fn, symbols=()):
    """Name of the built-in module that fn computes, or UnsupportedModule."""
    global _FORMS
    if _FORMS is None:
        _FORMS = _builtin_forms()
    props = tuple(sorted(set(_MD_PROPS) | set(_DEM_PROPS)))
    form = trace(fn, props, tuple(set(_SYMS) | set(symbols)))
    for k, f in _FORMS.items():
        if f == form:
            return k
    raise UnsupportedModule(f"module {getattr(fn, '__name__', fn)} is not one of the hot-path modules the HIP "
                            f"kernels implement ({', '.join(_BUILTIN_SRC)}) with their operation order")


# ---------------------------------------------------------------- the simulation object (sim/simulation.py)

class Simulation:
    """What examples/md.py and examples/dem.py configure; ``generate()`` runs
    it on the GPU through the C ABI."""

    def __init__(self, ref, shapes, dims=3, timesteps=100, double_prec=False, use_contact_history=False,
                 particle_capacity=800000, neighbor_capacity=100, debug=False):
        if dims != 3:
            raise UnsupportedModule("pairs_amd runs three-dimensional simulations")
        if not double_prec:
            raise UnsupportedModule("pairs_amd runs the double-precision program (double_prec=True)")
        self.ref, self.shapes, self.timesteps = ref, list(shapes), timesteps
        self.use_contact_history = use_contact_history
        self.particle_capacity, self.neighbor_capacity = particle_capacity, neighbor_capacity
        self.props, self.features, self.feature_props, self.contact_props = {}, {}, {}, {}
        self.position_name = None
        self.domain = None
        self.partitioner = None
        self.pbc_flags = [True, True, True]
        self.setups = []            # ("fcc", ...), ("sc_grid", ...), ("read", ...)
        self.setup_modules = []     # (fn, symbols)
        self.modules = []           # dict(fn, cutoff, symbols, pre_step, skip_first)
        self.cell_spacing = None
        self.neighbor_radius = None
        self.half = False
        self.reneighbor = 1
        self.thermo_every = 0
        self.vtk_file, self.vtk_frequency = None, 0
        self._target = None
        self.system = None          # the GPU system after generate()
        self.thermo = []            # (ts, T, P) rows printed by the run

    # ---- configuration (sim/simulation.py:104-381)
    def target(self, t):
        self._target = t

    def set_domain_partitioner(self, p):
        self.partitioner = p

    def compute_half(self):
        self.half = True

    def pbc(self, config):
        self.pbc_flags = [bool(x) for x in config]

    def add_position(self, name, value=(0.0, 0.0, 0.0), volatile=False, layout=None):
        self.position_name = name
        self.props[name] = (_VECTOR, value, volatile)

    def add_property(self, name, type_, value=0.0, volatile=False):
        self.props[name] = (type_, value, volatile)

    def add_feature(self, name, nkinds):
        self.features[name] = nkinds

    def add_feature_property(self, feature, name, type_, data):
        self.feature_props[name] = (feature, type_, list(data))

    def add_contact_property(self, name, type_, default, layout=None):
        self.contact_props[name] = (type_, default)

    def set_domain(self, grid):
        self.domain = [(grid[0], grid[3]), (grid[1], grid[4]), (grid[2], grid[5])]

    def reneighbor_every(self, n):
        self.reneighbor = n

    def compute_thermo(self, every=0):
        self.thermo_every = every

    def vtk_output(self, filename, frequency=0):
        self.vtk_file, self.vtk_frequency = filename, frequency

    def copper_fcc_lattice(self, nx, ny, nz, rho, temperature, ntypes):
        self.setups.append(("fcc", nx, ny, nz, rho, temperature, ntypes))

    def dem_sc_grid(self, xmax, ymax, zmax, spacing, diameter, min_diameter, max_diameter, initial_velocity,
                    particle_density, ntypes):
        self.setups.append(("sc_grid", xmax, ymax, zmax, spacing, diameter, min_diameter, max_diameter,
                            initial_velocity, particle_density, ntypes))

    def read_particle_data(self, filename, prop_names, shape):
        self.setups.append(("read", filename, list(prop_names), shape))

    def build_cell_lists(self, spacing, store_neighbors_per_cell=False):
        self.cell_spacing = spacing

    def build_neighbor_lists(self, spacing):
        self.neighbor_radius = spacing
        self.cell_spacing = spacing

    def compute(self, func, cutoff_radius=None, symbols={}, pre_step=False, skip_first=False):
        self.modules.append(dict(fn=func, cutoff=cutoff_radius, symbols=dict(symbols), pre_step=pre_step,
                                 skip_first=skip_first))

    def setup(self, func, symbols={}):
        self.setup_modules.append((func, dict(symbols)))

    # ---- the older API (examples/lj_onetype.py:29-45, lj_embedded.py:13-30).
    # The reference's current Simulation defines none of these calls; they
    # are thin aliases onto the calls above, so such scripts configure.
    # Their modules are traced and recognised like any other: modules
    # outside the library's kernels fail at generate() with the reason.
    def add_real_property(self, name, value=0.0, vol=False):
        self.add_property(name, _REAL, value, vol)
        return _Prop(name)

    def add_int_property(self, name, value=0, vol=False):
        self.add_property(name, _INT32, value, vol)
        return _Prop(name)

    def add_vector_property(self, name, value=(0.0, 0.0, 0.0), vol=False):
        self.add_property(name, _VECTOR, value, vol)
        return _Prop(name)

    def add_matrix_property(self, name, value=(0.0,) * 9, vol=False):
        self.add_property(name, _MATRIX, value, vol)
        return _Prop(name)

    def from_file(self, filename, prop_names, shape=None):
        self.read_particle_data(filename, prop_names, shape or _POINT_MASS)

    def create_cell_lists(self, spacing, cutoff_radius=None):
        self.build_cell_lists(spacing)

    def create_neighbor_lists(self):
        if self.cell_spacing is None:
            raise UnsupportedModule("create_neighbor_lists() needs create_cell_lists() first")
        self.neighbor_radius = self.cell_spacing

    def periodic(self, cutoff_radius=None, flags=(True, True, True)):
        self.pbc(flags)

    def particle_pairs(self, *args, **kw):
        raise UnsupportedModule("loop-style kernels (particle_pairs / particles) are not traced: write the "
                                "module as a function of (i, j) or (i) and pass it to compute()")

    def particles(self, *args, **kw):
        return self.particle_pairs()

    # ---- lowering onto the C ABI
    def _plan(self):
        if self._target is not None and self._target.kind != "gpu":
            raise UnsupportedModule("pairs_amd has the GPU target only (pairs.target_gpu())")
        names = [recognise(m["fn"], m["symbols"]) for m in self.modules]
        setups = [recognise(f, s) for f, s in self.setup_modules]
        md = ["initial_integrate", "lennard_jones", "final_integrate"]
        dem = ["gravity", "linear_spring_dashpot", "euler"]
        older = ["lj_older", "euler_md"]
        if names == older and not setups:
            if self.modules[0]["cutoff"] is None or self.neighbor_radius is None:
                raise UnsupportedModule("the pair module needs a cutoff and build_neighbor_lists")
            if self.domain is None:
                raise UnsupportedModule("a program that reads its particles from a file needs set_domain(): the "
                                        "older API leaves the box implicit and no rule for it survives in the "
                                        "reference")
            return "md_euler", setups
        if names == md:
            pre, lj, fin = self.modules
            if not (pre["pre_step"] and pre["skip_first"] and fin["skip_first"] and not fin["pre_step"]):
                raise UnsupportedModule("the velocity-Verlet pair runs as pre_step + skip_first (initial) and "
                                        "skip_first (final), as in examples/md.py")
            if pre["symbols"].get("dt") != fin["symbols"].get("dt"):
                raise UnsupportedModule("both integrators must use the same dt")
            if lj["cutoff"] is None or self.neighbor_radius is None:
                raise UnsupportedModule("lennard_jones needs a cutoff and build_neighbor_lists")
            return "md", setups
        if names == dem and setups == ["update_mass_and_inertia"]:
            return "dem", setups
        raise UnsupportedModule(f"unsupported module sequence {names} (setup {setups}); pairs_amd runs "
                                f"{md} or setup update_mass_and_inertia + {dem}")

    def _md_system(self):
        from . import md
        fcc = [s for s in self.setups if s[0] == "fcc"]
        if len(fcc) != 1 or len(self.setups) != 1:
            raise UnsupportedModule("the LJ path is set up by one copper_fcc_lattice call")
        _, nx, ny, nz, rho, temp, ntypes = fcc[0]
        a = (4.0 / rho) ** (1.0 / 3.0)
        box = [(0.0, nx * a), (0.0, ny * a), (0.0, nz * a)]
        lj = self.modules[1]
        nt = self.features.get("type", 1)
        eps = self.feature_props.get("epsilon", (None, None, [1.0] * nt * nt))[2]
        sg6 = self.feature_props.get("sigma6", (None, None, [1.0] * nt * nt))[2]
        natoms = 4 * nx * ny * nz
        sim = md.LJSystem(box, particle_capacity=max(self.particle_capacity, 4 * natoms),
                          neighbor_capacity=self.neighbor_capacity, send_capacity=max(200000, 2 * natoms),
                          cutoff=lj["cutoff"], skin=self.neighbor_radius - lj["cutoff"],
                          dt=self.modules[0]["symbols"]["dt"], half=self.half, ntypes=nt, epsilon=eps,
                          sigma6=sg6, reneighbor_every=self.reneighbor, thermo_every=self.thermo_every,
                          periodic=self.pbc_flags)
        sim.copper_fcc_lattice(nx, ny, nz, rho, nt)
        sim.adjust_thermo(temp)
        sim._call("pairs_hip_setup_cells")
        return sim

    def _md_euler_system(self):
        """The older API's LJ program: particles from one from_file call,
        full lists, the module lj with its own pair expression, Euler."""
        from . import io, md
        reads = [s for s in self.setups if s[0] == "read"]
        if len(reads) != 1 or len(self.setups) != 1:
            raise UnsupportedModule("the older API's LJ program is set up by one from_file call")
        _, path, names, _ = reads[0]
        kinds = {"mass": io.PROP_REAL, "position": io.PROP_VECTOR, "velocity": io.PROP_VECTOR}
        data = io.read_particle_data(path, names, self.domain, kinds=kinds)
        x, v = data["position"], data["velocity"]
        m = data.get("mass")
        n = len(x)
        lj, eu = self.modules
        sim = md.LJSystem(self.domain, particle_capacity=max(self.particle_capacity, 4 * n),
                          neighbor_capacity=self.neighbor_capacity, send_capacity=max(200000, 2 * n),
                          cutoff=lj["cutoff"], skin=self.neighbor_radius - lj["cutoff"], dt=eu["symbols"]["dt"],
                          half=False, ntypes=1, epsilon=[builtins.float(lj["symbols"]["epsilon"])],
                          sigma6=[builtins.float(lj["symbols"]["sigma6"])], reneighbor_every=self.reneighbor,
                          thermo_every=0, periodic=self.pbc_flags)
        sim.upload(x, v, m)
        sim._call("pairs_hip_setup_cells")
        return sim

    def _dem_system(self):
        from . import dem
        con = self.modules[1]["symbols"]
        grav = self.modules[0]["symbols"]
        dt = self.modules[2]["symbols"]["dt"]
        if con.get("dt") != dt:
            raise UnsupportedModule("the contact model and the integrator must use the same dt")
        grid = [s for s in self.setups if s[0] == "sc_grid"]
        if len(grid) != 1 or self.domain is None:
            raise UnsupportedModule("the DEM path is set up by set_domain and one dem_sc_grid call")
        _, xmax, ymax, zmax, spacing, diameter, dmin, dmax, v0, rho_p, ntypes = grid[0]
        fs = self.feature_props.get("friction_static", (None, None, [0.0]))[2][0]
        fd = self.feature_props.get("friction_dynamic", (None, None, [0.5]))[2][0]
        # kappa, ln e and t_c are the module's symbols; width = the cell spacing
        params = dict(diameter=diameter, gravity=grav["gravity_SI"], density_fluid=grav["densityFluid_SI"],
                      density_particle=grav["densityParticle_SI"], spacing=spacing, initial_velocity=v0, dt=dt,
                      friction=fd, friction_static=fs, restitution=math.exp(con["lnDryResCoeff"]),
                      collision_time=con["collisionTime_SI"])
        sysm = dem.DEMSystem(tuple(d[1] for d in self.domain), periodic=tuple(self.pbc_flags),
                             particle_capacity=min(self.particle_capacity, 4 << 20),
                             contact_capacity=self.neighbor_capacity, send_capacity=200000, params=params,
                             ntypes=ntypes, kappa=con["kappa"], ln_restitution=con["lnDryResCoeff"],
                             cell_width=self.cell_spacing)
        sysm.sc_grid((xmax, ymax, zmax), spacing=spacing, diameter=diameter, initial_velocity=v0, density=rho_p,
                     min_diameter=dmin, max_diameter=dmax)
        for s in self.setups:
            if s[0] == "read":
                sysm.read_particle_data(s[1], s[2], s[3].name)
        sysm.setup()
        return sysm

    def generate(self, run=None, emit_cpp=None, quiet=False):
        """Lower the recognised modules onto the GPU path; run timesteps + 1
        iterations of the reference's schedule (thermo rows printed as the
        generated program prints them, VTK files at the requested frequency);
        with emit_cpp, also write the host C++ driver of the same program.
        Defaults come from run_script (a script calls generate() bare)."""
        run = _SCRIPT["run"] if run is None else run
        emit_cpp = emit_cpp or _SCRIPT["emit_cpp"]
        quiet = quiet or _SCRIPT["quiet"]
        kind, _ = self._plan()
        if emit_cpp:
            with open(emit_cpp, "w") as f:
                f.write(self._cpp(kind))
        if not run:
            return None
        from . import io
        n = self.timesteps
        if kind == "md_euler":
            # the reference's timestep loop (src/pairs/sim/timestep.py:36-61)
            # over the older program's two modules, module by module
            self.system = s = self._md_euler_system()
            for ts in range(n + 1):
                if self.reneighbor > 0 and ((ts + 1) % self.reneighbor == 0 or ts == 0):
                    s.exchange()
                    s.borders()
                    s.build_cell_lists()
                    s.build_neighbor_lists()
                else:
                    s.synchronize()
                s.lennard_jones_older()
                s.euler()
                if self.vtk_file and self.vtk_frequency and ts % self.vtk_frequency == 0:
                    io.vtk_write(s, self.vtk_file + "_local", ts, self.vtk_frequency)
            return s
        if kind == "md":
            self.system = s = self._md_system()
            stops = sorted({0, n + 1} | ({t + 1 for t in range(0, n + 1, self.vtk_frequency)}
                                         if self.vtk_file and self.vtk_frequency else set()))
            for a, b in zip(stops[:-1], stops[1:]):
                rows = s.run(a, b, n)
                for ts in range(a, b):
                    if self.thermo_every and ((ts + 1) % self.thermo_every == 0 or ts == 0):
                        r = rows[len([t for t in range(a, ts) if (t + 1) % self.thermo_every == 0 or t == 0])]
                        self.thermo.append((ts, r[0], r[1]))
                        if not quiet:
                            print(f"{r[0]:g}\t{r[1]:g}", flush=True)
                if self.vtk_file:
                    io.vtk_write(s, self.vtk_file + "_local", b - 1, self.vtk_frequency)
        else:
            self.system = s = self._dem_system()
            stops = sorted({0, n + 1} | ({t + 1 for t in range(0, n + 1, self.vtk_frequency)}
                                         if self.vtk_file and self.vtk_frequency else set()))
            for a, b in zip(stops[:-1], stops[1:]):
                s.run(a, b)
                if self.vtk_file:
                    io.vtk_write(s, self.vtk_file + "_local", b - 1, self.vtk_frequency)
        return s

    def _cpp_md_euler(self):
        """The older API's LJ program (particles from one file, full lists,
        its own pair module, Euler) as a host program on the C ABI: the
        modules called one by one in the reference's timestep order, each
        retried with grown capacities on overflow (the generated resize loop,
        src/pairs/transformations/modules.py:159-205)."""
        n = self.timesteps
        reads = [s for s in self.setups if s[0] == "read"]
        _, path, names, _ = reads[0]
        width = {"mass": 1, "position": 3, "velocity": 3}
        kinds = {"mass": "PAIRS_HIP_PROP_REAL", "position": "PAIRS_HIP_PROP_VECTOR",
                 "velocity": "PAIRS_HIP_PROP_VECTOR"}
        for nm in names:
            if nm not in width:
                raise UnsupportedModule(f"from_file property {nm}: the older program reads mass, position, velocity")
        cols, off = {}, 0
        for nm in names:
            cols[nm] = off
            off += width[nm]
        lj, eu = self.modules
        lo = [d[0] for d in self.domain]
        hi = [d[1] for d in self.domain]
        mcol = f"rows[(size_t) i * {off} + {cols['mass']}]" if "mass" in cols else "1.0"
        return f'''// host program of "{self.ref}" (the older API) for the HIP target (include/pairs_hip.h)
// usage: prog [state.bin [particles.csv]]  (state.bin: nlocal, then positions,
// velocities and forces of the final state in the reference's particle order)
#include <cstdio>
#include <vector>
#include "pairs_hip.h"
#include "pairs_hip_io.h"

// a module call with the resize protocol: grow the capacity it outgrew (x2 the need) and repeat
static int module(pairs_hip_ctx *ctx, int (*fn)(pairs_hip_ctx *)) {{
    for (int attempt = 0; attempt < 8; attempt++) {{
        const int r = fn(ctx);
        if (r <= 0) return r;
        const int need = pairs_hip_last_required(ctx, r);
        int grow[3] = {{0, 0, 0}};
        grow[r == PAIRS_HIP_OVERFLOW_PARTICLES ? 0 : r == PAIRS_HIP_OVERFLOW_NEIGHBORS ? 1 : 2] = 2 * need;
        if (pairs_hip_reserve(ctx, grow[0], grow[1], grow[2])) return -1;
    }}
    return -1;
}}

static int dump(pairs_hip_ctx *ctx, const char *path) {{
    const int n = pairs_hip_count(ctx, PAIRS_HIP_NLOCAL);
    if (n < 0) return 1;
    std::vector<double> x(3 * (size_t) n), v(3 * (size_t) n), f(3 * (size_t) n);
    if (pairs_hip_download(ctx, PAIRS_HIP_POSITION, x.data(), n) != n) return 1;
    if (pairs_hip_download(ctx, PAIRS_HIP_VELOCITY, v.data(), n) != n) return 1;
    if (pairs_hip_download(ctx, PAIRS_HIP_FORCE, f.data(), n) != n) return 1;
    FILE *o = std::fopen(path, "wb");
    if (!o) return 1;
    std::fwrite(&n, sizeof(int), 1, o);
    std::fwrite(x.data(), sizeof(double), x.size(), o);
    std::fwrite(v.data(), sizeof(double), v.size(), o);
    std::fwrite(f.data(), sizeof(double), f.size(), o);
    return std::fclose(o) != 0;
}}

int main(int argc, char **argv) {{
    const char *input = argc > 2 ? argv[2] : {_c_string(path)};
    const int kinds[{len(names)}] = {{{", ".join(kinds[nm] for nm in names)}}};
    const double lo[3] = {{{lo[0]!r}, {lo[1]!r}, {lo[2]!r}}}, hi[3] = {{{hi[0]!r}, {hi[1]!r}, {hi[2]!r}}};
    int need = 0;
    pairs_hip_read_particle_csv(input, {len(names)}, kinds, {names.index("position")}, -1, lo, hi, nullptr, 0, &need);
    std::vector<double> rows((size_t) {off} * (need > 0 ? need : 1));
    const int np = pairs_hip_read_particle_csv(input, {len(names)}, kinds, {names.index("position")}, -1, lo, hi,
                                               rows.data(), need, &need);
    if (np < 0) {{ std::fprintf(stderr, "cannot read %s\\n", input); return 1; }}
    std::vector<double> x(3 * (size_t) np), v(3 * (size_t) np), m(np);
    for (int i = 0; i < np; i++) {{
        for (int d = 0; d < 3; d++) {{
            x[3 * i + d] = rows[(size_t) i * {off} + {cols["position"]} + d];
            v[3 * i + d] = rows[(size_t) i * {off} + {cols["velocity"]} + d];
        }}
        m[i] = {mcol};
    }}
    static double eps[1] = {{{float_(lj["symbols"]["epsilon"])!r}}};
    static double sig6[1] = {{{float_(lj["symbols"]["sigma6"])!r}}};
    pairs_hip_md_config c = {{}};
    c.particle_capacity = {max(self.particle_capacity, 16)};
    c.neighbor_capacity = {self.neighbor_capacity};
    c.send_capacity = 200000;
    c.ncells_capacity = 100000;
    for (int d = 0; d < 3; d++) {{ c.box_lo[d] = lo[d]; c.box_hi[d] = hi[d]; }}
    c.periodic[0] = {int(self.pbc_flags[0])}; c.periodic[1] = {int(self.pbc_flags[1])}; c.periodic[2] = {int(self.pbc_flags[2])};
    c.cell_spacing = {self.neighbor_radius!r};
    c.cutoff = {lj["cutoff"]!r};
    c.dt = {eu["symbols"]["dt"]!r};
    c.half = 0;
    c.ntypes = 1;
    c.epsilon = eps;
    c.sigma6 = sig6;
    c.reneighbor_every = {self.reneighbor};
    c.thermo_every = 0;
    pairs_hip_ctx *ctx;
    if (pairs_hip_create(&ctx, &c)) return 1;
    int bad = pairs_hip_upload(ctx, np, x.data(), v.data(), m.data(), nullptr, nullptr, nullptr) != 0;
    if (!bad) bad = pairs_hip_setup_cells(ctx) != 0;
    for (int ts = 0; !bad && ts <= {n}; ts++) {{
        if ({self.reneighbor} > 0 && ((ts + 1) % {max(self.reneighbor, 1)} == 0 || ts == 0)) {{
            bad = module(ctx, pairs_hip_exchange) || module(ctx, pairs_hip_borders) ||
                  module(ctx, pairs_hip_bin_cells) || module(ctx, pairs_hip_build_neighbors);
        }} else {{
            bad = pairs_hip_sync_ghosts(ctx) != 0;
        }}
        if (!bad) bad = pairs_hip_lj_older(ctx) != 0 || pairs_hip_euler(ctx) != 0;
    }}
    if (!bad && argc > 1) bad = dump(ctx, argv[1]);
    pairs_hip_destroy(ctx);
    return bad;
}}
'''

    def _cpp(self, kind):
        """The host C++ program of this simulation on the C ABI (the HIP
        target's generated main)."""
        if kind == "md_euler":
            return self._cpp_md_euler()
        n = self.timesteps
        if kind == "md":
            (_, nx, ny, nz, rho, temp, nt) = self.setups[0]
            lj, vv = self.modules[1], self.modules[0]
            eps = self.feature_props.get("epsilon", (0, 0, [1.0] * nt * nt))[2]
            sg6 = self.feature_props.get("sigma6", (0, 0, [1.0] * nt * nt))[2]
            a = (4.0 / rho) ** (1.0 / 3.0)
            return f'''// host program of "{self.ref}" for the HIP target (include/pairs_hip.h)
// usage: prog [state.bin]  (state.bin: nlocal, then positions and velocities
// of the final state in the reference's particle order, as doubles)
#include <cstdio>
#include <vector>
#include "pairs_hip.h"

static int dump(pairs_hip_ctx *ctx, const char *path) {{
    const int n = pairs_hip_count(ctx, PAIRS_HIP_NLOCAL);
    if (n < 0) return 1;
    std::vector<double> x(3 * (size_t) n), v(3 * (size_t) n);
    if (pairs_hip_download(ctx, PAIRS_HIP_POSITION, x.data(), n) != n) return 1;
    if (pairs_hip_download(ctx, PAIRS_HIP_VELOCITY, v.data(), n) != n) return 1;
    FILE *f = std::fopen(path, "wb");
    if (!f) return 1;
    std::fwrite(&n, sizeof(int), 1, f);
    std::fwrite(x.data(), sizeof(double), x.size(), f);
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    return std::fclose(f) != 0;
}}

int main(int argc, char **argv) {{
    static double eps[{nt * nt}] = {{{", ".join(repr(float_(x)) for x in eps)}}};
    static double sig6[{nt * nt}] = {{{", ".join(repr(float_(x)) for x in sg6)}}};
    pairs_hip_md_config c = {{}};
    c.particle_capacity = {max(self.particle_capacity, 4 * 4 * nx * ny * nz)};
    c.neighbor_capacity = {self.neighbor_capacity};
    c.send_capacity = {max(200000, 8 * nx * ny * nz)};
    c.ncells_capacity = 100000;
    const double len[3] = {{{nx * a!r}, {ny * a!r}, {nz * a!r}}};
    for (int d = 0; d < 3; d++) {{ c.box_lo[d] = 0.0; c.box_hi[d] = len[d]; c.periodic[d] = {int(self.pbc_flags[0])}; }}
    c.cell_spacing = {self.neighbor_radius!r};
    c.cutoff = {lj["cutoff"]!r};
    c.dt = {vv["symbols"]["dt"]!r};
    c.half = {int(self.half)};
    c.ntypes = {nt};
    c.epsilon = eps;
    c.sigma6 = sig6;
    c.reneighbor_every = {self.reneighbor};
    c.thermo_every = {self.thermo_every};
    pairs_hip_ctx *ctx;
    if (pairs_hip_create(&ctx, &c)) return 1;
    pairs_hip_copper_fcc_lattice(ctx, {nx}, {ny}, {nz}, {rho!r}, {nt});
    pairs_hip_adjust_thermo(ctx, {temp!r});
    pairs_hip_setup_cells(ctx);
    double th[5 * 4096];
    const int rows = pairs_hip_run(ctx, 0, {n + 1}, {n}, th, 4096);
    for (int r = 0; r < rows; r++) std::printf("%g\\t%g\\n", th[5 * r], th[5 * r + 1]);
    int bad = rows < 0;
    if (!bad && argc > 1) bad = dump(ctx, argv[1]);
    pairs_hip_destroy(ctx);
    return bad;
}}
'''
        grav, con, eul = self.modules
        (_, xmax, ymax, zmax, spacing, diameter, dmin, dmax, v0, rho_p, ntypes) = \
            [s for s in self.setups if s[0] == "sc_grid"][0]
        planes = []
        for s in self.setups:
            if s[0] == "read" and s[3].name == "halfspace":
                from . import io
                d = io.read_particle_data(s[1], s[2], self.domain)
                for k in range(len(d["uid"])):
                    planes.append((int(d["uid"][k]), int(d["type"][k]), float_(d["mass"][k]),
                                   [float_(x) for x in d["position"][k]], [float_(x) for x in d["normal"][k]],
                                   int(d["flags"][k])))
        fs = self.feature_props.get("friction_static", (0, 0, [0.0] * ntypes * ntypes))[2]
        fd = self.feature_props.get("friction_dynamic", (0, 0, [0.5] * ntypes * ntypes))[2]
        lines = [f'''// host program of "{self.ref}" for the HIP target (include/pairs_hip_dem.h)
// usage: prog [state.bin]  (state.bin: nlocal, then positions, linear
// velocities and uids of the final state's locals)
#include <cstdio>
#include <vector>
#include "pairs_hip_dem.h"

static int dump(pairs_hip_dem_ctx *d, const char *path) {{
    const int n = pairs_hip_dem_count(d, PAIRS_HIP_DEM_NLOCAL);
    if (n < 0) return 1;
    std::vector<double> x(3 * (size_t) n), v(3 * (size_t) n);
    std::vector<int> uid(n);
    if (pairs_hip_dem_download(d, PAIRS_HIP_DEM_POSITION, x.data(), n) != n) return 1;
    if (pairs_hip_dem_download(d, PAIRS_HIP_DEM_LINEAR_VELOCITY, v.data(), n) != n) return 1;
    if (pairs_hip_dem_download(d, PAIRS_HIP_DEM_UID, uid.data(), n) != n) return 1;
    FILE *f = std::fopen(path, "wb");
    if (!f) return 1;
    std::fwrite(&n, sizeof(int), 1, f);
    std::fwrite(x.data(), sizeof(double), x.size(), f);
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fwrite(uid.data(), sizeof(int), uid.size(), f);
    return std::fclose(f) != 0;
}}

int main(int argc, char **argv) {{
    pairs_hip_dem_config c = {{}};
    c.particle_capacity = {min(self.particle_capacity, 4 << 20)};
    c.contact_capacity = {self.neighbor_capacity};
    c.send_capacity = 200000;
    c.ncells_capacity = 1 << 22;
    c.box_hi[0] = {self.domain[0][1]!r}; c.box_hi[1] = {self.domain[1][1]!r}; c.box_hi[2] = {self.domain[2][1]!r};
    c.periodic[0] = {int(self.pbc_flags[0])}; c.periodic[1] = {int(self.pbc_flags[1])}; c.periodic[2] = {int(self.pbc_flags[2])};
    c.cell_width = {self.cell_spacing!r};
    c.dt = {eul["symbols"]["dt"]!r};
    c.density_particle = {grav["symbols"]["densityParticle_SI"]!r};
    c.density_fluid = {grav["symbols"]["densityFluid_SI"]!r};
    c.gravity = {grav["symbols"]["gravity_SI"]!r};
    c.kappa = {con["symbols"]["kappa"]!r};
    c.ln_restitution = {con["symbols"]["lnDryResCoeff"]!r};
    c.collision_time = {con["symbols"]["collisionTime_SI"]!r};
    c.ntypes = {ntypes};''']
        for k in range(ntypes * ntypes):
            lines.append(f"    c.friction_static[{k}] = {float_(fs[k])!r}; c.friction_dynamic[{k}] = {float_(fd[k])!r};")
        lines.append(f'''    pairs_hip_dem_ctx *d;
    if (pairs_hip_dem_create(&d, &c)) return 1;
    pairs_hip_dem_sc_grid(d, {xmax!r}, {ymax!r}, {zmax!r}, {spacing!r}, {diameter!r}, {dmin!r}, {dmax!r}, {v0!r}, {rho_p!r}, {ntypes});''')
        for (uid, typ, mass, pos, nrm, fl) in planes:
            lines.append(f"    {{ const double p[3] = {{{pos[0]!r}, {pos[1]!r}, {pos[2]!r}}}, n[3] = "
                         f"{{{nrm[0]!r}, {nrm[1]!r}, {nrm[2]!r}}}; "
                         f"pairs_hip_dem_add_halfspace(d, {uid}, {typ}, {mass!r}, p, n, {fl}); }}")
        lines.append(f'''    pairs_hip_dem_setup(d);
    int bad = pairs_hip_dem_run(d, 0, {n + 1}) != 0;
    if (!bad && argc > 1) bad = dump(d, argv[1]);
    pairs_hip_dem_destroy(d);
    return bad;
}}
''')
        return "\n".join(lines)


def float_(x):
    return builtins.float(x)


def _c_string(text):
    """a C string literal of text (backslashes, quotes and control characters escaped)"""
    out = []
    for ch in text:
        if ch in '\\"':
            out.append("\\" + ch)
        elif ord(ch) < 32 or ord(ch) > 126:
            out.append("\\%03o" % ord(ch))
        else:
            out.append(ch)
    return '"' + "".join(out) + '"'


# how generate() behaves inside run_script, and the simulations it made
_SCRIPT = {"run": True, "emit_cpp": None, "quiet": False}
_CREATED = []


def simulation(ref, shapes=None, dims=3, timesteps=100, double_prec=None, use_contact_history=False,
               particle_capacity=800000, neighbor_capacity=100, debug=False):
    """pairs.simulation (src/pairs/__init__.py:9-22).  The older API's call
    (examples/lj_onetype.py:29) names neither shapes nor precision: shapes
    default to point masses and an unstated precision runs in FP64."""
    sim = Simulation(ref, shapes if shapes is not None else [_POINT_MASS], dims, timesteps,
                     True if double_prec is None else double_prec, use_contact_history, particle_capacity,
                     neighbor_capacity, debug)
    _CREATED.append(sim)
    return sim


def run_script(path, argv=(), run=True, emit_cpp=None, quiet=False):
    """Run a DSL script with this module installed as ``pairs``; returns the
    simulations it created (generate() runs them unless run is False)."""
    import runpy
    sys.modules["pairs"] = sys.modules[__name__]
    old, saved = sys.argv, dict(_SCRIPT)
    sys.argv = [path] + list(argv)
    _SCRIPT.update(run=run, emit_cpp=emit_cpp, quiet=quiet)
    del _CREATED[:]
    try:
        runpy.run_path(path, run_name="__main__")
        return list(_CREATED)
    finally:
        sys.argv = old
        _SCRIPT.update(saved)


if __name__ == "__main__":
    if len(sys.argv) < 2:
        print("usage: python -m pairs_amd.dsl <script.py> [args...]", file=sys.stderr)
        sys.exit(2)
    run_script(sys.argv[1], sys.argv[2:])
