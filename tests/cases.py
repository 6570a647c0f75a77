"""Case builders shared by the tests: one spec dict drives both the oracle
(CPU restatement) and the HIP package, so both see identical inputs.

Spec keys: element ('quad4' | 'linear1d'), nelements, el_len, particles
(list of coordinates), init_velocity (scalar, per-dimension or per-particle),
material, constraints, nodal_forces, tractions, gravity, scheme, dt, nsteps,
velocity_update, optional youngs (per-particle E)."""
import json
import math
import os

import numpy as np

GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")


def load_goldens():
    with open(os.path.join(GOLDEN_DIR, "reference_goldens.json")) as f:
        d = json.load(f)
    d.pop("_source", None)
    for v in d.values():
        v.setdefault("element", "quad4")
    return d


def load_kats():
    with open(os.path.join(GOLDEN_DIR, "reference_kats.json")) as f:
        return json.load(f)


def check_value(stress, chk):
    """The reference's assertion: round(reduce(stress[:, comp]) - value, d) == 0."""
    col = np.asarray(stress, dtype=np.float64).reshape(len(stress), -1)[:, chk["component"]]
    got = float(col.max() if chk["reduce"] == "max" else col.min())
    return got, round(got - chk["value"], chk["decimals"]) == 0.0


def _velocity(spec, n, ndim):
    v = np.asarray(spec.get("init_velocity", 0.0), dtype=np.float64)
    return np.broadcast_to(v.reshape(-1, ndim) if v.size == n * ndim else v, (n, ndim))


# ---------------------------------------------------------------- oracle side
def oracle_function(fspec):
    from oracle import mpm as om
    if fspec is None or fspec["type"] == "Unit":
        return om.Unit(-1)
    return om.Linear(0, fspec["xvalues"], fspec["fxvalues"])


def oracle_material(mspec):
    from oracle import mpm as om
    m = dict(mspec)
    return getattr(om, m.pop("type"))(m)


def build_oracle(spec, dtype=np.float32):
    """Restates the reference's setup (io.py:52-170 for the TOML cases)."""
    from oracle import mpm as om
    cons = [(np.asarray(c["nodes"]), om.Constraint(c["dir"], c["velocity"]))
            for c in spec["constraints"]]
    cnfs = [om.NodalForce(np.asarray(f["nodes"]), oracle_function(f.get("function")),
                          f["dir"], f["force"]) for f in spec["nodal_forces"]]
    el = om.Elements(spec["element"], spec["nelements"], spec["el_len"], cons, cnfs, dtype)
    pts = np.asarray(spec["particles"], dtype=np.float64).reshape(len(spec["particles"]), -1)
    ps = om.Particles(pts, oracle_material(spec["material"]), dtype=dtype)
    ps.velocity[:] = _velocity(spec, len(pts), pts.shape[1]).astype(dtype)
    if spec.get("youngs") is not None:
        ps.youngs_modulus = np.asarray(spec["youngs"], dtype=dtype)
    trs = [om.ParticleTraction(t["pset"], np.asarray(t["pids"]),
                               oracle_function(t.get("function")), t["dir"], t["traction"])
           for t in spec["tractions"]]
    sets = [ps]
    for es in spec.get("extra_sets", []):  # further particle sets on the same mesh
        q = np.asarray(es["particles"], dtype=np.float64).reshape(len(es["particles"]), -1)
        e = om.Particles(q, oracle_material(es["material"]), dtype=dtype)
        e.velocity[:] = _velocity(es, len(q), q.shape[1]).astype(dtype)
        sets.append(e)
    mesh = om.Mesh(el, sets, trs)
    return om.MPMExplicit(mesh, spec["dt"], spec["scheme"], spec["velocity_update"],
                          spec["nsteps"])


build_oracle_golden = build_oracle


# ------------------------------------------------------------------ GPU side
def gpu_function(fspec):
    import diffmpm_amd as d
    if fspec is None or fspec["type"] == "Unit":
        return d.Unit(-1)
    return d.Linear(0, fspec["xvalues"], fspec["fxvalues"])


def gpu_material(mspec):
    import diffmpm_amd as d
    m = dict(mspec)
    return getattr(d, m.pop("type"))(m)


def build_gpu(spec, dtype=None, keep_temporaries=True):
    import torch
    import diffmpm_amd as d
    dtype = dtype or torch.float32
    dev = "cuda"
    cons = [(torch.tensor(c["nodes"]), d.Constraint(c["dir"], c["velocity"]))
            for c in spec["constraints"]]
    cnfs = [d.NodalForce(torch.tensor(f["nodes"]), gpu_function(f.get("function")),
                         f["dir"], f["force"]) for f in spec["nodal_forces"]]
    if spec["element"] == "quad4":
        nx, ny = spec["nelements"]
        el = d.Quadrilateral4Node((nx, ny), nx * ny, spec["el_len"], cons,
                                  concentrated_nodal_forces=cnfs, device=dev, dtype=dtype)
        ndim, Mesh = 2, d.Mesh2D
    else:
        el = d.Linear1D(spec["nelements"], spec["nelements"], spec["el_len"], cons,
                        concentrated_nodal_forces=cnfs, device=dev, dtype=dtype)
        ndim, Mesh = 1, d.Mesh1D
    pts = np.asarray(spec["particles"], dtype=np.float64).reshape(len(spec["particles"]), 1, ndim)
    np_dt = np.float64 if dtype == torch.float64 else np.float32
    ps = d.Particles(torch.from_numpy(pts.astype(np_dt)), gpu_material(spec["material"]),
                     torch.zeros(len(pts), dtype=torch.int32), device=dev, dtype=dtype)
    ps.velocity = torch.from_numpy(
        _velocity(spec, len(pts), ndim).astype(np_dt).reshape(len(pts), 1, ndim))
    if spec.get("youngs") is not None:
        ps.youngs_modulus = torch.as_tensor(np.asarray(spec["youngs"], dtype=np_dt))
    trs = [d.ParticleTraction(t["pset"], torch.tensor(t["pids"]),
                              gpu_function(t.get("function")), t["dir"], t["traction"])
           for t in spec["tractions"]]
    sets = [ps]
    for es in spec.get("extra_sets", []):  # further particle sets on the same mesh
        q = np.asarray(es["particles"], dtype=np.float64).reshape(len(es["particles"]), 1, ndim)
        e = d.Particles(torch.from_numpy(q.astype(np_dt)), gpu_material(es["material"]),
                        torch.zeros(len(q), dtype=torch.int32), device=dev, dtype=dtype)
        e.velocity = torch.from_numpy(
            _velocity(es, len(q), ndim).astype(np_dt).reshape(len(q), 1, ndim))
        sets.append(e)
    mesh = Mesh({"particles": sets, "elements": el, "particle_surface_traction": trs})
    return d.MPMExplicit(mesh, spec["dt"], spec["scheme"], spec["velocity_update"],
                         spec["nsteps"], keep_temporaries=keep_temporaries)


# -------------------------------------------------------- state comparison
PARTICLE_FIELDS = ["loc", "velocity", "momentum", "stress", "strain", "strain_rate", "dstrain",
                   "volume", "density", "mass", "reference_loc", "dvolumetric_strain",
                   "volumetric_strain_centroid", "size", "traction"]
NODE_FIELDS = ["mass", "velocity", "acceleration", "momentum", "f_int", "f_ext"]


def gpu_particle_field(ps, name):
    v = getattr(ps, name).detach().double().cpu().numpy()
    return v.reshape(v.shape[0], -1) if v.ndim > 1 else v


def oracle_particle_field(ps, name):
    v = np.asarray(getattr(ps, name), dtype=np.float64)
    return v.reshape(v.shape[0], -1)


def gpu_node_field(nodes, name):
    v = getattr(nodes, name).detach().double().cpu().numpy()
    return v.reshape(v.shape[0], -1)


def oracle_node_field(nodes, name):
    v = np.asarray(getattr(nodes, name), dtype=np.float64)
    return v.reshape(v.shape[0], -1)


def normwise(a, b):
    """||a - b|| / ||b|| (||a - b|| when b is zero), with both vectors scaled by
    their largest finite magnitude first, so that values near the float range
    (nan_to_num's +-max of a massless node under a nodal force) do not
    overflow the sums of squares."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    fin = np.concatenate([np.abs(a[np.isfinite(a)]), np.abs(b[np.isfinite(b)])])
    s = float(fin.max()) if fin.size else 0.0
    if s == 0.0 or not np.isfinite(s):
        s = 1.0
    den = np.linalg.norm(b / s)
    num = np.linalg.norm((a - b) / s)
    if den == 0:
        return num * s
    return num / den


def load_oracle_state_into_gpu(osim, gsim):
    """Copies the oracle's full particle state (every particle set) and the
    persistent nodal velocity into the GPU objects (for per-step,
    re-synchronised parity)."""
    import torch
    for ops, gps in zip(osim.mesh.particles, gsim.mesh.particles):
        for name in PARTICLE_FIELDS + ["acceleration", "f_ext"]:
            v = np.asarray(getattr(ops, name))
            view = getattr(gps, name)
            setattr(gps, name, torch.from_numpy(v.reshape(view.shape).astype(np.float64)))
        gps._has_fext = bool(np.any(ops.f_ext))
        gps.element_ids = torch.from_numpy(ops.element_ids.astype(np.int32))
    gn, on = gsim.mesh.elements.nodes, osim.mesh.elements.nodes
    for name in NODE_FIELDS:
        v = np.asarray(getattr(on, name))
        setattr(gn, name, torch.from_numpy(v.reshape(getattr(gn, name).shape).astype(np.float64)))


def copy_oracle_state(src, dst):
    """Copies an oracle simulation's particle and nodal state into another
    (of possibly another dtype), for per-step comparisons from one state."""
    sp, dp = src.mesh.particles[0], dst.mesh.particles[0]
    for name in PARTICLE_FIELDS + ["acceleration", "f_ext"]:
        v = getattr(dp, name)
        v[...] = np.asarray(getattr(sp, name)).astype(v.dtype)
    dp.element_ids[...] = sp.element_ids
    sn, dn = src.mesh.elements.nodes, dst.mesh.elements.nodes
    for name in NODE_FIELDS:
        v = getattr(dn, name)
        v[...] = np.asarray(getattr(sn, name)).astype(v.dtype)


# ------------------------------------------------------- synthetic cases
def synthetic_spec(element="quad4", nx=8, ny=6, rows=4, seed=0, scheme="usf",
                   velocity_update=False, jitter=0.2, traction=True, nodal_force=True,
                   gravity=True, E=1e6, nu=0.3, dt=1e-5, nsteps=10, per_particle_E=False,
                   velocity_scale=0.5, drift=None):
    """A small seeded case that exercises every term of the step."""
    from oracle.mpm import gauss_points_2d
    rng = np.random.Generator(np.random.PCG64(seed))
    lin = {"type": "Linear", "xvalues": [0.0, 5 * dt, 1.0], "fxvalues": [0.0, 1.0, 1.0]}
    if element == "quad4":
        h = 1.0 / max(nx, ny)
        pts = gauss_points_2d(nx, ny, h, h, rows=rows, jitter=jitter, seed=seed,
                              dtype=np.float64)
        n = len(pts)
        spec = dict(element="quad4", nelements=[nx, ny], el_len=[h, h], particles=pts.tolist())
        bottom = list(range(nx + 1))
        left = [j * (nx + 1) for j in range(ny + 1)]
        top = [rows * (nx + 1) + i for i in range(nx + 1)]
        spec["constraints"] = [{"nodes": bottom, "dir": 1, "velocity": 0.0},
                               {"nodes": left, "dir": 0, "velocity": 0.0}]
        spec["nodal_forces"] = ([{"nodes": top, "dir": 1, "force": -10.0, "function": lin}]
                                if nodal_force else [])
        spec["gravity"] = [0.0, -9.81] if gravity else [0.0, 0.0]
        vel = velocity_scale * rng.standard_normal((n, 2))
    else:
        h = 1.0 / nx
        g = (1.0 - 1.0 / math.sqrt(3.0)) / 2.0
        base = np.arange(nx)
        pts = np.stack([base + g, base + 1 - g], axis=1).reshape(-1, 1)
        pts = pts + rng.uniform(-jitter * 0.5, jitter * 0.5, pts.shape) * g
        pts = np.clip(pts * h, 0, np.nextafter(np.float32(nx * h), 0))
        n = len(pts)
        spec = dict(element="linear1d", nelements=nx, el_len=h, particles=pts.tolist())
        spec["constraints"] = [{"nodes": [0], "dir": 0, "velocity": 0.0}]
        spec["nodal_forces"] = ([{"nodes": [nx], "dir": 0, "force": 5.0, "function": lin}]
                                if nodal_force else [])
        spec["gravity"] = [-9.81] if gravity else [0.0]
        vel = velocity_scale * rng.standard_normal((n, 1))
    if drift is not None:
        vel = vel + np.asarray(drift, dtype=np.float64)[: vel.shape[1]]
    spec["init_velocity"] = vel.tolist()
    spec["material"] = {"type": "LinearElastic", "density": 1000.0, "youngs_modulus": E,
                        "poisson_ratio": nu}
    tp = sorted(rng.choice(n, size=min(3, n), replace=False).tolist())
    spec["tractions"] = ([{"pset": [0], "pids": tp, "dir": 0, "traction": 50.0,
                           "function": lin}] if traction else [])
    spec.update(scheme=scheme, dt=dt, nsteps=nsteps, velocity_update=velocity_update)
    if per_particle_E:
        spec["youngs"] = (E * np.exp(0.3 * rng.standard_normal(n))).tolist()
    return spec


# ------------------------------------------------------- benchmark configs
def config_c1_spec(nsteps=40):
    """1D bar: Linear1D, 128 elements on [0, 1], 2 Gauss particles per
    element (256), v0 = 0.1 sin(pi x / 2), node 0 fixed, E = 100, rho = 1,
    nu = 0 (the configuration omits nu; LinearElastic needs it), dt = 0.01,
    USF.  c dt / h = 12.8: the rollout is CFL-unstable and turns non-finite."""
    nx = 128
    h = 1.0 / nx
    g = (1.0 - 1.0 / math.sqrt(3.0)) / 2.0
    base = np.arange(nx, dtype=np.float64)
    pts = (np.stack([base + g, base + 1 - g], axis=1).reshape(-1, 1) * h)
    vel = 0.1 * np.sin(np.pi * pts / 2.0)
    return dict(element="linear1d", nelements=nx, el_len=h, particles=pts.tolist(),
                constraints=[{"nodes": [0], "dir": 0, "velocity": 0.0}], nodal_forces=[],
                gravity=[0.0], init_velocity=vel.tolist(),
                material={"type": "LinearElastic", "density": 1.0, "youngs_modulus": 100.0,
                          "poisson_ratio": 0.0},
                tractions=[], scheme="usf", dt=0.01, nsteps=nsteps, velocity_update=False)


def config_c2_spec(nsteps=3):
    """Q4 256 x 256, particles at the Gauss points of the lower 128 rows,
    E = 1e6, nu = 0.3, rho = 1000, bottom nodes fixed in y, left nodes fixed
    in x, -1e3 N per node in y on node row 128 (the top edge of the block)
    ramped by Linear([0, 500 dt, inf], [0, 1, 1]), no gravity, dt = 1e-5,
    USL."""
    from oracle.mpm import gauss_points_2d
    n = 256
    h = 1.0 / n
    dt = 1e-5
    pts = gauss_points_2d(n, n, h, h, rows=128, dtype=np.float64)
    top = [128 * (n + 1) + i for i in range(n + 1)]
    ramp = {"type": "Linear", "xvalues": [0.0, 500 * dt, float("inf")],
            "fxvalues": [0.0, 1.0, 1.0]}
    return dict(element="quad4", nelements=[n, n], el_len=[h, h], particles=pts.tolist(),
                constraints=[{"nodes": list(range(n + 1)), "dir": 1, "velocity": 0.0},
                             {"nodes": [j * (n + 1) for j in range(n + 1)], "dir": 0,
                              "velocity": 0.0}],
                nodal_forces=[{"nodes": top, "dir": 1, "force": -1e3, "function": ramp}],
                gravity=[0.0, 0.0], init_velocity=0.0,
                material={"type": "LinearElastic", "density": 1000.0, "youngs_modulus": 1e6,
                          "poisson_ratio": 0.3},
                tractions=[], scheme="usl", dt=dt, nsteps=nsteps, velocity_update=False)


def gaussian_random_field(n, h, corr, seed):
    from diffmpm_amd.workloads import gaussian_field
    return gaussian_field(n, h, corr, seed)


def bilinear(field, h, pts):
    from diffmpm_amd.workloads import bilinear as b
    return b(field, h, pts)


def config_c4_spec(nsteps=3, seed=0, n=128):
    """Q4 128 x 128, particles at the Gauss points of the lower 64 rows,
    E_p = 1e6 exp(0.3 g_p) with g a unit-variance Gaussian random field of
    correlation length 0.1 (filtered white noise on the node grid, bilinear
    to the particles, seeded), nu = 0.3, rho = 1000, gravity, bottom fixed,
    dt = 1e-5, USF."""
    from oracle.mpm import gauss_points_2d
    h = 1.0 / n
    pts = gauss_points_2d(n, n, h, h, rows=n // 2, dtype=np.float64)
    gfield = gaussian_random_field(n, h, 0.1, seed)
    E = 1e6 * np.exp(0.3 * bilinear(gfield, h, pts))
    bottom = list(range(n + 1))
    return dict(element="quad4", nelements=[n, n], el_len=[h, h], particles=pts.tolist(),
                constraints=[{"nodes": bottom, "dir": 0, "velocity": 0.0},
                             {"nodes": bottom, "dir": 1, "velocity": 0.0}],
                nodal_forces=[], gravity=[0.0, -9.81], init_velocity=0.0,
                material={"type": "LinearElastic", "density": 1000.0, "youngs_modulus": 1e6,
                          "poisson_ratio": 0.3},
                tractions=[], scheme="usf", dt=1e-5, nsteps=nsteps, velocity_update=False,
                youngs=E.tolist())


def gpu_state_into_oracle(gsim, osim):
    """Copies a GPU simulation's particle state (first particle set) and
    nodal state into an oracle simulation of the same mesh, so that both can
    take the next step from identical inputs."""
    gps, ops = gsim.mesh.particles[0], osim.mesh.particles[0]
    for name in PARTICLE_FIELDS + ["acceleration", "f_ext"]:
        cur = getattr(ops, name)
        v = getattr(gps, name).detach().cpu().numpy().reshape(np.shape(cur))
        setattr(ops, name, v.astype(np.asarray(cur).dtype))
    ops.element_ids = gps.element_ids.cpu().numpy().astype(np.int32)
    gn, on = gsim.mesh.elements.nodes, osim.mesh.elements.nodes
    for name in NODE_FIELDS + ["f_damp"]:
        cur = getattr(on, name)
        v = getattr(gn, name).detach().cpu().numpy().reshape(np.shape(cur))
        setattr(on, name, v.astype(np.asarray(cur).dtype))


def fp32_within_reference_precision(got, want32, want64, tol=1e-5, factor=1.0):
    """FP32 parity criterion for quantities whose FP32 rounding the reference
    itself amplifies (gradients of nodal velocities on fine meshes, sums over
    thousands of particles): the HIP result agrees with the FP32 oracle within
    tol, or it is within tol of the FP64 oracle, or no further from the FP64
    oracle than factor times the FP32 oracle (the reference's own arithmetic)
    is.  Returns (ok, err vs FP32 oracle, err vs FP64 oracle, FP32 oracle vs
    FP64)."""
    e32, e64, eo = normwise(got, want32), normwise(got, want64), normwise(want32, want64)
    return (e32 <= tol or e64 <= tol or e64 <= factor * eo), e32, e64, eo
