"""Owner of one C-ABI plan (one particle set on one mesh) and of the calls
that drive the HIP kernels for it."""
from ctypes import byref, c_void_p

import torch

from . import _lib
from ._lib import call, ptr


def _as_list(x):
    if isinstance(x, torch.Tensor):
        return x.reshape(-1).tolist()
    try:
        return [int(v) for v in x]
    except TypeError:
        return [int(x)]


class Plan:
    def __init__(self, elements, particles):
        _lib.require_gpu(particles._buf, elements.nodes._buf)
        if particles.dtype != elements.nodes.dtype:
            raise ValueError("particles and nodes must share a dtype")
        if particles._ndim != elements.ndim:
            raise ValueError("particle and mesh dimensions differ")
        self.elements = elements
        self.particles = particles
        per_E = particles._has_youngs
        kind, de = particles.material._kernel_coefficients(per_E)
        d = _lib.PlanDesc()
        d.element = elements._element_code
        d.dtype = _lib.dtype_code(particles.dtype)
        d.nx, d.ny = elements._nx, elements._ny
        d.hx, d.hy = elements._hx, elements._hy
        d.nparticles = len(particles)
        d.ld = particles._ld
        d.nld = elements.nodes._ld
        d.material = kind
        d.per_particle_E = 1 if per_E else 0
        for k in range(3):
            d.de[k] = float(de[k])
        self.desc = d
        h = c_void_p()
        with torch.cuda.device(particles.device):
            call("dmpm_plan_create", byref(d), byref(h))
        self.h = h
        self._sig = None
        self.refresh_tables()

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            try:
                _lib.lib().dmpm_plan_destroy(h)
            except Exception:
                pass
            self.h = None

    # -- tables ----------------------------------------------------------------
    def _signature(self):
        el, ps = self.elements, self.particles
        tr = getattr(ps, "_traction_specs", [])
        return (id(el.constraints), len(el.constraints),
                id(el.concentrated_nodal_forces), len(el.concentrated_nodal_forces),
                id(tr), len(tr))

    def refresh_tables(self):
        sig = self._signature()
        if sig == self._sig:
            return
        el, ps = self.elements, self.particles
        tractions = list(getattr(ps, "_traction_specs", []))
        fns, fn_index = [], {}

        def fidx(f):
            if f is None:
                return -1
            if id(f) not in fn_index:
                fn_index[id(f)] = len(fns)
                fns.append(f)
            return fn_index[id(f)]

        nodes, dirs, vals = [], [], []
        for ids, c in el.constraints:
            for i in _as_list(ids):
                nodes.append(i), dirs.append(int(c.dir)), vals.append(float(c.velocity))
        call("dmpm_plan_set_constraints", self.h, len(nodes), _lib.i32_array(nodes),
             _lib.i32_array(dirs), _lib.f64_array(vals))
        f_nodes, f_dirs, f_fn, f_vals = [], [], [], []
        for cnf in el.concentrated_nodal_forces:
            k = fidx(cnf.function)
            for i in _as_list(cnf.node_ids):
                f_nodes.append(i), f_dirs.append(int(cnf.dir)), f_fn.append(k)
                f_vals.append(float(cnf.force))
        t_ids, t_dirs, t_fn, t_vals = [], [], [], []
        for pt in tractions:
            k = fidx(pt.function)
            for i in _as_list(pt.pids):
                t_ids.append(i), t_dirs.append(int(pt.dir)), t_fn.append(k)
                t_vals.append(float(pt.traction))
        kinds, offs, cnts, xs, fs = [], [], [], [], []
        from .functions import Linear
        for f in fns:
            if isinstance(f, Linear):
                kinds.append(_lib.DMPM_FN_LINEAR)
                offs.append(len(xs)), cnts.append(len(f.xvalues))
                xs.extend(float(v) for v in f.xvalues)
                fs.extend(float(v) for v in f.fxvalues)
            else:
                kinds.append(_lib.DMPM_FN_UNIT)
                offs.append(len(xs)), cnts.append(0)
        call("dmpm_plan_set_functions", self.h, len(fns), _lib.i32_array(kinds),
             _lib.i32_array(offs), _lib.i32_array(cnts), _lib.f64_array(xs), _lib.f64_array(fs))
        call("dmpm_plan_set_nodal_forces", self.h, len(f_nodes), _lib.i32_array(f_nodes),
             _lib.i32_array(f_dirs), _lib.i32_array(f_fn), _lib.f64_array(f_vals))
        call("dmpm_plan_set_tractions", self.h, len(t_ids), _lib.i32_array(t_ids),
             _lib.i32_array(t_dirs), _lib.i32_array(t_fn), _lib.f64_array(t_vals))
        self._sig = sig

    # -- argument helpers -----------------------------------------------------
    @property
    def pb(self):
        return ptr(self.particles._buf)

    @property
    def eid(self):
        return ptr(self.particles._eid)

    @property
    def nb(self):
        return ptr(self.elements.nodes._buf)

    @property
    def s(self):
        return _lib.stream_of(self.particles._buf)

    def _gravity(self, gravity):
        g = torch.as_tensor(gravity, dtype=torch.float64).reshape(-1).cpu().tolist()
        if len(g) == 1:
            g = g * 2
        if len(g) < 2:
            g = g + [0.0] * (2 - len(g))
        return _lib.f64_array(g[:2])

    # -- per-method calls ------------------------------------------------------
    def set_particle_element_ids(self):
        call("dmpm_set_particle_element_ids", self.h, self.pb, self.eid, self.s)

    def update_natural_coords(self):
        call("dmpm_update_natural_coords", self.h, self.pb, self.eid, self.s)

    def compute_particle_volume(self, elements):
        ve = float(elements.volume.reshape(-1)[0].item())
        call("dmpm_compute_particle_volume", self.h, self.pb, self.eid, ve, self.s)

    def compute_nodal_mass(self):
        call("dmpm_compute_nodal_mass", self.h, self.pb, self.eid, self.nb, self.s)

    def compute_nodal_momentum(self):
        call("dmpm_compute_nodal_momentum", self.h, self.pb, self.eid, self.nb, self.s)

    def compute_velocity(self):
        call("dmpm_compute_velocity", self.h, self.nb, self.s)

    def apply_boundary_constraints(self):
        self.refresh_tables()
        call("dmpm_apply_boundary_constraints", self.h, self.nb, self.s)

    def compute_external_force(self):
        call("dmpm_compute_external_force", self.h, self.pb, self.eid, self.nb, self.s)

    def compute_body_force(self, gravity):
        call("dmpm_compute_body_force", self.h, self.pb, self.eid, self.nb,
             self._gravity(gravity), self.s)

    def assign_tractions(self, t):
        self.refresh_tables()
        call("dmpm_assign_tractions", self.h, self.pb, float(t), self.s)

    def apply_particle_traction_forces(self):
        call("dmpm_apply_particle_traction_forces", self.h, self.pb, self.eid, self.nb, self.s)

    def apply_concentrated_nodal_forces(self, t):
        self.refresh_tables()
        call("dmpm_apply_concentrated_nodal_forces", self.h, self.nb, float(t), self.s)

    def compute_internal_force(self):
        call("dmpm_compute_internal_force", self.h, self.pb, self.eid, self.nb, self.s)

    def update_nodal_acceleration_velocity(self, dt):
        self.refresh_tables()
        call("dmpm_update_nodal_acceleration_velocity", self.h, self.nb, float(dt), self.s)

    def update_position_velocity(self, elements, dt, velocity_update):
        call("dmpm_update_position_velocity", self.h, self.pb, self.eid, self.nb, float(dt),
             1 if velocity_update else 0, self.s)

    def compute_strain(self, elements, dt):
        call("dmpm_compute_strain", self.h, self.pb, self.eid, self.nb, float(dt), self.s)

    def update_volume(self):
        call("dmpm_update_volume", self.h, self.pb, self.s)

    def compute_stress(self):
        call("dmpm_compute_stress", self.h, self.pb, self.s)

    # -- fused phases ------------------------------------------------------------
    def phase_nodal_kinematics(self):
        self.refresh_tables()
        call("dmpm_phase_nodal_kinematics", self.h, self.pb, self.eid, self.nb, self.s)

    def phase_stress_strain(self, dt):
        call("dmpm_phase_stress_strain", self.h, self.pb, self.eid, self.nb, float(dt), self.s)

    def phase_forces(self, gravity, step, dt):
        self.refresh_tables()
        call("dmpm_phase_forces", self.h, self.pb, self.eid, self.nb, self._gravity(gravity),
             int(step), float(dt), self.s)

    def phase_particle_kinematics(self, dt, velocity_update):
        self.refresh_tables()
        call("dmpm_phase_particle_kinematics", self.h, self.pb, self.eid, self.nb, float(dt),
             1 if velocity_update else 0, self.s)

    def set_option(self, option, value):
        call("dmpm_plan_set_option", self.h, int(option), int(value))

    def set_block(self, cells_per_block):
        call("dmpm_plan_set_block", self.h, int(cells_per_block))

    def run(self, scheme, velocity_update, gravity, dt, step0, nsteps, keep_temporaries=True,
            per_phase=False):
        self.refresh_tables()
        flags = ((1 if keep_temporaries else 0) | (2 if self.particles._has_fext else 0)
                 | (4 if per_phase else 0))
        code = _lib.DMPM_USF if scheme == "usf" else _lib.DMPM_USL
        call("dmpm_run", self.h, self.pb, self.eid, self.nb, code,
             1 if velocity_update else 0, self._gravity(gravity), float(dt), int(step0),
             int(nsteps), flags, self.s)


def plan_for(elements, particles):
    key = id(elements)
    p = particles._plans.get(key)
    if p is None or p.elements is not elements:
        p = Plan(elements, particles)
        particles._plans[key] = p
    return p


def timing_enable(plan, kernel_class, max_launches=0):
    """Records HIP events around launches of one kernel class (see
    ``_lib.KC``; ``"all"`` for every class); ``kernel_class=None`` disables."""
    kc = -1 if kernel_class is None else (-2 if kernel_class == "all"
                                          else _lib.KC[kernel_class])
    call("dmpm_timing_enable", plan.h, kc, int(max_launches))


def timing_read(plan, kernel_class=None):
    """(total milliseconds, launches) of the enabled kernel class, or of
    ``kernel_class`` when given."""
    from ctypes import c_double, c_int64
    ms, n = c_double(), c_int64()
    if kernel_class is None:
        call("dmpm_timing_read", plan.h, byref(ms), byref(n))
    else:
        call("dmpm_timing_read_class", plan.h, _lib.KC[kernel_class], byref(ms), byref(n))
    return ms.value, n.value
